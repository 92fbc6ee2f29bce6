"""Surrogate-triaged stuck-at fault simulation.

Candidates are (fault, stimulus) pairs: the stuck-at-0/1 faults of every
register crossed with a pool of 256-cycle random stimuli (the GROUP_RANDOM
input mode keyed by a stimulus id).  The plain path simulates every candidate
exactly (rrtl_amd/faultsim.py).

The triaged path

  1. runs the fault-free design once per stimulus (one 64-lane group each, one
     launch) and derives, for every candidate, how the faulted register's
     fault-free value relates to the stuck value;
  2. applies two exact rules: a candidate whose register never leaves the
     stuck value (not excited) equals the golden lane at every cycle, and a
     register with no path to an output (not observable) never changes what
     is compared -- neither can be detected, by any simulation;
  3. scores every candidate with the surrogate MLP (include/rrtl_mlp.h) on
     512 features: a hashed one-hot code of the fault (256: 4 hashes into 4
     blocks of 64); bit statistics of the stimulus (248: the faulted
     register's fault-free waveform averaged over 8 bins of 32 cycles, each
     bin repeated over 16 columns, then for the first 120 primary inputs the
     centred count of ones over the cycles, scaled by 1/4); and 8 excitation
     features;
  4. simulates exactly in stages: each fault's best 2 eligible candidates
     (with the audit, a random 1 % of the eligible candidates scored below
     ``pmin``), then its next 14 for the faults still undetected, then --
     for coverage parity -- every eligible candidate left of the faults still
     undetected.

Every detection the triaged path reports comes from an exact run; the
surrogate only decides the order and, without the last stage, what to give up.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from .faultsim import FAULTS_PER_GROUP, FaultSim, observable_signals, pack
from .program import KIND_INPUT

FAULT_CODE_WIDTH = 256
STIM_STAT_WIDTH = 248
WAVE_BINS = 128                          # columns of the waveform part
WAVE_BINS_DEFAULT = 8                    # distinct bins (each repeated 16 times)
PI_STATS = STIM_STAT_WIDTH - WAVE_BINS   # 120
PI_STAT_SCALE = 0.25
EXCITE_WIDTH = 8
FEATURE_WIDTH = FAULT_CODE_WIDTH + STIM_STAT_WIDTH + EXCITE_WIDTH   # 512
CODE_HASHES = 4
BF16_ONE = 0x3F80
SURROGATE_DIMS = (FEATURE_WIDTH, 1024, 1024, 1024, 1)

K1 = np.uint64(0x9E3779B97F4A7C15)
K2 = np.uint64(0xD1B54A32D192ED03)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def to_bf16_bits(x: np.ndarray) -> np.ndarray:
    """bf16 patterns of float32 values (round to nearest even)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def fault_code_table(n_faults: int, seed: int = 0x5EED) -> np.ndarray:
    """[n_faults][256] bf16 patterns: one 1.0 in each of 4 blocks of 64, at a
    hash of (fault, block)."""
    t = np.zeros((n_faults, FAULT_CODE_WIDTH), dtype=np.uint16)
    f = np.arange(n_faults, dtype=np.uint64)
    block = FAULT_CODE_WIDTH // CODE_HASHES
    for h in range(CODE_HASHES):
        with np.errstate(over="ignore"):
            k = _mix64(np.uint64(seed) ^ (f * K1) ^ np.uint64((h + 1) << 56))
        t[np.arange(n_faults), h * block + (k % np.uint64(block)).astype(np.int64)] = BF16_ONE
    return t


def group_stimulus_ones(seed: int, sig: int, cycles: int, stim_ids: np.ndarray) -> np.ndarray:
    """Number of cycles in [0, cycles) at which the GROUP_RANDOM input ``sig``
    is 1 under each stimulus id: bit 0 of stim_word(seed, sig << 8, c, id),
    the library's stimulus hash (include/rrtl_hip.h, RRTL_IN_GROUP_RANDOM)."""
    c = np.arange(cycles, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64(seed) ^ (np.uint64((sig << 8) + 1) * K1) ^ ((c + np.uint64(1)) * K2))
        w = _mix64(key[:, None] + (np.asarray(stim_ids, dtype=np.uint64)[None, :] + np.uint64(1)) * K1)
    return (w & np.uint64(1)).sum(axis=0).astype(np.int64)


def input_stats(inputs: Sequence[int], stim_ids: np.ndarray, seed: int, cycles: int, width: int = PI_STATS) -> np.ndarray:
    """[len(stim_ids)][width] float32: for input i (the first ``width``), the
    count of ones over the cycles, centred and scaled to unit variance."""
    stim_ids = np.asarray(stim_ids, dtype=np.uint64)
    out = np.zeros((stim_ids.size, width), dtype=np.float32)
    for j, sig in enumerate(list(inputs)[:width]):
        out[:, j] = (group_stimulus_ones(seed, sig, cycles, stim_ids) - cycles / 2) / np.sqrt(cycles / 4)
    return out


def activity_table(golden: np.ndarray, pi_stats: np.ndarray, wave_bins: int = WAVE_BINS,
                   pi_scale: float = PI_STAT_SCALE) -> np.ndarray:
    """[registers * stimuli][248] bf16 patterns, row r * S + s: register r's
    fault-free waveform under stimulus s averaged over ``wave_bins`` equal
    bins (each bin repeated to fill 128 columns), then the stimulus's input
    statistics scaled by ``pi_scale``."""
    S, R, C = golden.shape
    edges = np.linspace(0, C, wave_bins + 1).astype(int)
    wave = np.zeros((S, R, wave_bins), dtype=np.float32)
    for b in range(wave_bins):
        wave[:, :, b] = golden[:, :, edges[b]:max(edges[b + 1], edges[b] + 1)].mean(axis=2)
    out = np.empty((R, S, STIM_STAT_WIDTH), dtype=np.float32)
    out[:, :, :WAVE_BINS] = np.repeat(wave, WAVE_BINS // wave_bins, axis=2).transpose(1, 0, 2)
    out[:, :, WAVE_BINS:] = pi_scale * pi_stats[None, :, :PI_STATS]
    return to_bf16_bits(out.reshape(R * S, STIM_STAT_WIDTH))


def excitation(golden: np.ndarray, reg_index: np.ndarray, stuck: np.ndarray):
    """golden: [stimuli][registers][cycles] fault-free register values after
    every tick.  For fault f (register reg_index[f] stuck at stuck[f]) under
    stimulus s: the number of cycles the fault-free value differs from the
    stuck value ([faults][stimuli] int) and 8 features ([faults][stimuli][8]
    bf16 patterns): excited fraction, excited at all, first excited cycle
    (fraction; 1 if never), toggle rate, log excited count, stuck value,
    fault-free ones fraction, excited at every cycle."""
    S, R, C = golden.shape
    g = golden.astype(bool)
    ones = g.sum(axis=2)                                        # [S][R]
    prev = np.concatenate([np.zeros((S, R, 1), dtype=bool), g[:, :, :-1]], axis=2)
    toggles = (g != prev).sum(axis=2)
    first1 = np.where(g.any(axis=2), g.argmax(axis=2), C)
    first0 = np.where((~g).any(axis=2), (~g).argmax(axis=2), C)
    ri = np.asarray(reg_index, dtype=np.int64)
    v = np.asarray(stuck, dtype=np.int64)[:, None]
    on = ones[:, ri].T                                          # [F][S]
    exc = np.where(v == 0, on, C - on)
    first = np.where(v == 0, first1[:, ri].T, first0[:, ri].T)
    feat = np.empty(exc.shape + (EXCITE_WIDTH,), dtype=np.float32)
    feat[..., 0] = exc / C
    feat[..., 1] = exc > 0
    feat[..., 2] = first / C
    feat[..., 3] = toggles[:, ri].T / C
    feat[..., 4] = np.log2(1.0 + exc) / np.log2(1.0 + C)
    feat[..., 5] = np.broadcast_to(v, exc.shape)
    feat[..., 6] = on / C
    feat[..., 7] = exc == C
    return exc, to_bf16_bits(feat)


@dataclass
class Features:
    reg_index: np.ndarray     # [faults] row of the fault's register in the activity table
    excited: np.ndarray       # [faults][stimuli] excited cycle counts
    extra: np.ndarray         # [faults * stimuli][8] bf16 excitation features
    table: np.ndarray         # [registers * stimuli][248] bf16 activity table
    kernel_ms: float          # device time of the fault-free run


def candidate_features(fs: FaultSim, stim_ids: np.ndarray, wave_bins: int = WAVE_BINS_DEFAULT,
                       pi_scale: float = PI_STAT_SCALE) -> Features:
    """Everything the surrogate reads about the candidates of ``fs.faults``
    under the stimuli ``stim_ids``, from one fault-free launch."""
    faults = fs.faults
    regs, reg_index = np.unique(faults[:, 0], return_inverse=True)
    g, ms = fs.golden_register_bits(stim_ids, regs)
    exc, xf = excitation(g, reg_index, faults[:, 1])
    tab = activity_table(g, input_stats(primary_inputs(fs.program), stim_ids, fs.seed, fs.cycles), wave_bins, pi_scale)
    return Features(reg_index, exc, xf.reshape(-1, EXCITE_WIDTH), tab, ms)


@dataclass
class Labels:
    detected: np.ndarray     # [faults][stimuli] bool, exact
    kernel_ms: float         # device time of the exact runs
    groups: int


def label_all(fs: FaultSim, stim_ids: np.ndarray, stims_per_run: int = 256) -> Labels:
    """Exact detection of every fault under every stimulus (the plain path)."""
    stim_ids = np.asarray(stim_ids, dtype=np.uint64)
    F = len(fs.faults)
    det = np.zeros((F, stim_ids.size), dtype=bool)
    ms, groups = 0.0, 0
    for s0 in range(0, stim_ids.size, stims_per_run):
        ids = stim_ids[s0:s0 + stims_per_run]
        fi = np.tile(np.arange(F), ids.size)
        si = np.repeat(ids, F)
        r = fs.run(fi, si)
        det[:, s0:s0 + ids.size] = r.detected.reshape(ids.size, F).T
        ms += r.kernel_ms
        groups += r.groups
    return Labels(det, ms, groups)


@dataclass
class TriageResult:
    detected: np.ndarray          # [faults] bool: flagged by some exact run
    simulated: np.ndarray         # [faults][stimuli] bool: candidates simulated exactly
    confirmed: np.ndarray         # [faults][stimuli] bool: exact detections
    sim_ms: float                 # device time of the exact launches (audit included)
    groups: int
    launches: int
    stages: List[Dict] = field(default_factory=list)
    audit_candidates: int = 0
    audit_detections: int = 0     # candidates the audit found detected
    audit_new_faults: int = 0     # faults only the audit found


def staged_order(prob: np.ndarray) -> np.ndarray:
    """[faults][stimuli] stimulus indices per fault in decreasing score, by
    octaves of the score, ties broken in one global stimulus order so that
    faults share stimuli (dense 64-lane groups)."""
    F, S = prob.shape
    q = np.where(prob > 0, np.floor(np.log2(np.maximum(prob, 1e-30))), -1e9)
    return np.lexsort((np.broadcast_to(np.arange(S), (F, S)), -q), axis=1)


def triage(fs: FaultSim, prob: Optional[np.ndarray], eligible: np.ndarray, stim_ids: np.ndarray,
           stages: Sequence[int] = (2, 14), pmin: float = 1e-3, tail: bool = True, audit_frac: float = 0.01,
           seed: int = 7) -> TriageResult:
    """Staged exact simulation of the ``eligible`` candidates ([faults][stimuli]
    bool; the rest are provably undetectable).

    Each fault's eligible candidates are taken in the order of ``prob`` (None:
    pool order, the method without a network): stage i simulates the next
    ``stages[i]`` of them, scored at least ``pmin``, for every fault not yet
    detected.  With ``tail`` a last stage simulates every eligible candidate
    left of the faults still undetected, so the coverage equals the plain
    path's; without it faults whose remaining scores are all below ``pmin``
    are given up.  The audit, a random ``audit_frac`` of the eligible
    candidates scored below ``pmin``, runs with the first stage."""
    stim_ids = np.asarray(stim_ids, dtype=np.uint64)
    F, S = eligible.shape
    if prob is None:
        p = np.where(eligible, 1.0, 0.0)
        order = np.broadcast_to(np.arange(S), (F, S))
    else:
        p = np.where(eligible, prob, 0.0)
        order = staged_order(p)
    el_sorted = np.take_along_axis(eligible, order, axis=1)
    rank_sorted = np.cumsum(el_sorted, axis=1) - 1
    rank = np.empty_like(rank_sorted)
    np.put_along_axis(rank, order, rank_sorted, axis=1)
    scored = eligible & (p >= pmin)
    rng = np.random.default_rng(seed)
    below = np.flatnonzero((eligible & ~scored).ravel())
    n_audit = int(round(audit_frac * below.size))
    audit = np.sort(rng.choice(below, n_audit, replace=False)) if n_audit else np.zeros(0, dtype=np.int64)
    detected = np.zeros(F, dtype=bool)
    simulated = np.zeros((F, S), dtype=bool)
    confirmed = np.zeros((F, S), dtype=bool)
    res = TriageResult(detected, simulated, confirmed, 0.0, 0, 0)
    lo = 0
    plan = [(k, False) for k in stages] + ([(S, True)] if tail else [])
    for i, (k, is_tail) in enumerate(plan):
        if is_tail:
            pick = eligible & ~simulated & ~detected[:, None]
        else:
            pick = scored & (rank >= lo) & (rank < lo + k) & ~detected[:, None]
            lo += k
        f, st = np.nonzero(pick)
        n_main = f.size
        if i == 0 and audit.size:
            fa, sa = np.divmod(audit, S)
            f, st = np.concatenate([f, fa]), np.concatenate([st, sa])
        if f.size == 0:
            continue
        r = fs.run(f, stim_ids[st])
        simulated[f, st] = True
        confirmed[f[r.detected], st[r.detected]] = True
        hit = np.zeros(F, dtype=bool)
        hit[f[:n_main][r.detected[:n_main]]] = True
        if i == 0 and audit.size:
            ad = r.detected[n_main:]
            res.audit_candidates = int(ad.size)
            res.audit_detections = int(ad.sum())
            found = np.zeros(F, dtype=bool)
            found[f[n_main:][ad]] = True
            res.audit_new_faults = int((found & ~hit & ~detected).sum())
            hit |= found
        detected |= hit
        res.stages.append({"candidates": int(f.size), "groups": r.groups, "kernel_ms": r.kernel_ms,
                           "detected_faults": int(detected.sum()), "tail": is_tail})
        res.sim_ms += r.kernel_ms
        res.groups += r.groups
        res.launches += 1
    return res


def triage_deferred(fs: FaultSim, score: Callable[[np.ndarray, np.ndarray], np.ndarray], eligible: np.ndarray,
                    stim_ids: np.ndarray, first: int = 2, stages: Sequence[int] = (14,), tail: bool = True) -> TriageResult:
    """The network aimed at what is left after a cheap first stage: each
    fault's first ``first`` eligible candidates in pool order are simulated
    without any scoring (most detectable faults fall there), then only the
    eligible candidates of the faults still undetected are scored
    (``score(faults, stimulus indices)`` -> probabilities) and the later
    stages, and with ``tail`` the parity stage, take them in score order.
    Scoring time is booked by the caller (the ``score`` callback)."""
    stim_ids = np.asarray(stim_ids, dtype=np.uint64)
    F, S = eligible.shape
    rank = np.cumsum(eligible, axis=1) - 1
    pick = eligible & (rank < first)
    f, st = np.nonzero(pick)
    detected = np.zeros(F, dtype=bool)
    simulated = np.zeros((F, S), dtype=bool)
    confirmed = np.zeros((F, S), dtype=bool)
    res = TriageResult(detected, simulated, confirmed, 0.0, 0, 0)
    if f.size:
        r = fs.run(f, stim_ids[st])
        simulated[f, st] = True
        confirmed[f[r.detected], st[r.detected]] = True
        detected[f[r.detected]] = True
        res.stages.append({"candidates": int(f.size), "groups": r.groups, "kernel_ms": r.kernel_ms,
                           "detected_faults": int(detected.sum()), "tail": False, "scored": 0})
        res.sim_ms += r.kernel_ms
        res.groups += r.groups
        res.launches += 1
    left = eligible & ~simulated & ~detected[:, None]
    fl, sl = np.nonzero(left)
    prob = np.zeros((F, S))
    if fl.size:
        prob[fl, sl] = np.maximum(score(fl, sl), 1e-30)
    rest = triage(fs, prob, left, stim_ids, stages, 0.0, tail, 0.0)
    for k in range(len(rest.stages)):
        rest.stages[k]["scored"] = int(fl.size) if k == 0 else 0
    res.stages += rest.stages
    res.detected = detected | rest.detected
    res.simulated = simulated | rest.simulated
    res.confirmed = confirmed | rest.confirmed
    res.sim_ms += rest.sim_ms
    res.groups += rest.groups
    res.launches += rest.launches
    return res


def dropping(fs: FaultSim, stim_ids: np.ndarray, mask: Optional[np.ndarray] = None, max_round: int = 64) -> Dict:
    """Exact simulation with fault dropping and no surrogate: every undetected
    fault is simulated under the next round of stimuli (rounds of 1, 2, 4, ...
    up to ``max_round`` stimuli, in pool order) until the pool is used up.
    ``mask`` [faults][stimuli] limits the candidates (e.g. to excited ones)."""
    stim_ids = np.asarray(stim_ids, dtype=np.uint64)
    F, S = len(fs.faults), stim_ids.size
    detected = np.zeros(F, dtype=bool)
    ms, groups, launches, pos, k = 0.0, 0, 0, 0, 1
    while pos < S:
        todo = np.flatnonzero(~detected)
        if todo.size == 0:
            break
        cols = np.arange(pos, min(S, pos + k))
        fi = np.tile(todo, cols.size)
        si = np.repeat(cols, todo.size)
        if mask is not None:
            keep = mask[fi, si]
            fi, si = fi[keep], si[keep]
        if fi.size:
            r = fs.run(fi, stim_ids[si])
            detected[fi[r.detected]] = True
            ms += r.kernel_ms
            groups += r.groups
            launches += 1
        pos += cols.size
        k = min(2 * k, max_round)
    return {"gpu_ms": ms, "groups": groups, "launches": launches, "detected": detected}


def primary_inputs(program) -> List[int]:
    return [i for i, s in enumerate(program.signals) if s.kind == KIND_INPUT]


def candidate_grid(reg_index: np.ndarray, n_stims: int):
    """(fault index, activity-table row) of every candidate, fault-major: the
    row of fault f under stimulus s is reg_index[f] * n_stims + s."""
    ri = np.asarray(reg_index, dtype=np.uint32)
    fi = np.repeat(np.arange(ri.size, dtype=np.uint32), n_stims)
    si = np.repeat(ri * np.uint32(n_stims), n_stims) + np.tile(np.arange(n_stims, dtype=np.uint32), ri.size)
    return fi, si


def run_experiment(program, n_train: int = 512, n_eval: int = 256, cycles: int = 256, seed: int = 1,
                   epochs: int = 6, batch: int = 8192, lr: float = 1e-3, pmins: Sequence[float] = (1e-3,),
                   stages: Sequence[int] = (2, 14), audit_frac: float = 0.01, predict_rows: int = 65536,
                   weight_decay: float = 0.0,
                   wave_bins: int = WAVE_BINS_DEFAULT, pi_scale: float = PI_STAT_SCALE,
                   log: Callable[[str], None] = print, keep: Optional[Dict] = None) -> Dict:
    """The whole comparison on one design: exact labels for the training pool,
    surrogate training on the GPU, then on a fresh evaluation pool the plain
    path (ground truth), the triaged path with the parity stage and without it
    at each ``pmin``, the same stages in pool order with the exact rules and
    no network, and plain fault dropping with neither."""
    from .faultsim import register_faults
    from .surrogate import Surrogate

    faults = register_faults(program)
    F = len(faults)
    fs = FaultSim(program, faults, cycles=cycles, seed=seed)
    train_ids = np.arange(n_train, dtype=np.uint64)
    eval_ids = np.arange(1_000_000, 1_000_000 + n_eval, dtype=np.uint64)
    out: Dict = {"faults": F, "train_stimuli": n_train, "eval_stimuli": n_eval, "cycles": cycles}
    t0 = time.time()
    # ---- training data: exact labels and features of the training pool
    Lt = label_all(fs, train_ids)
    ft = candidate_features(fs, train_ids, wave_bins, pi_scale)
    codes = fault_code_table(F)
    out["train_label_ms"] = Lt.kernel_ms
    out["train_detect_rate"] = float(Lt.detected.mean())
    log(f"training labels: {Lt.detected.size} candidates, {Lt.kernel_ms:.1f} ms, {time.time() - t0:.1f} s")
    # ---- surrogate training on the GPU
    sur = Surrogate(SURROGATE_DIMS, max_batch=max(batch, predict_rows), seed=seed, lr=lr, weight_decay=weight_decay)
    sur.set_tables(codes, ft.table, EXCITE_WIDTH)
    fi, si = candidate_grid(ft.reg_index, n_train)
    losses = sur.train_pairs(fi, si, Lt.detected.ravel().astype(np.float32), epochs=epochs, batch=batch,
                             seed=seed, extra=ft.extra)
    st = sur.stats()
    out["train"] = {"epochs": epochs, "loss": [float(x) for x in losses], "gpu_ms": st["kernel_ms"],
                    "gemm_ms": st["gemm_ms"], "gemm_tflops": st["gemm_flop"] / max(st["gemm_ms"], 1e-9) / 1e9,
                    "steps": st["steps"]}
    log(f"surrogate trained: loss {losses[0]:.4f} -> {losses[-1]:.4f}, {st['kernel_ms']:.1f} ms")
    del ft
    # ---- plain path on the evaluation pool (ground truth)
    Le = label_all(fs, eval_ids)
    plain_cov = Le.detected.any(axis=1)
    out["plain"] = {"gpu_ms": Le.kernel_ms, "groups": Le.groups, "candidates": int(Le.detected.size),
                    "detected_faults": int(plain_cov.sum()), "coverage": float(plain_cov.mean())}
    log(f"plain: {Le.kernel_ms:.1f} ms, coverage {plain_cov.mean():.4f}")
    # ---- triaged path: features, scores, exact stages
    fe = candidate_features(fs, eval_ids, wave_bins, pi_scale)
    ge_ms = fe.kernel_ms
    # the exact rules: a candidate can only be detected if its register leaves
    # the stuck value (excited) and reaches an output (observable)
    observable = observable_signals(program)[faults[:, 0]]
    eligible = (fe.excited > 0) & observable[:, None]
    assert not (Le.detected & ~eligible).any(), "a candidate outside the exact rules was detected"
    sur.set_tables(codes, fe.table, EXCITE_WIDTH)
    fi, si = candidate_grid(fe.reg_index, n_eval)
    prob = sur.predict_pairs(fi, si, extra=fe.extra).reshape(F, n_eval)
    pst = sur.stats()
    out["features_ms"] = ge_ms
    out["eligible_candidates"] = int(eligible.sum())
    out["observable_faults"] = int(observable.sum())
    out["predict"] = {"gpu_ms": pst["kernel_ms"], "gemm_ms": pst["gemm_ms"],
                      "gemm_tflops": pst["gemm_flop"] / max(pst["gemm_ms"], 1e-9) / 1e9,
                      "brier": float(np.mean((prob - Le.detected) ** 2))}
    if keep is not None:
        keep.update(prob=prob, labels=Le.detected, excited=fe.excited > 0, eligible=eligible,
                    train_labels=Lt.detected)

    def record(name: str, tr: TriageResult, network: bool, score_ms: float = 0.0, **extra) -> Dict:
        gpu = ge_ms + (pst["kernel_ms"] if network else score_ms) + tr.sim_ms
        row = {"gpu_ms": gpu, "sim_ms": tr.sim_ms, "score_ms": pst["kernel_ms"] if network else score_ms,
               "time_vs_plain": gpu / Le.kernel_ms, "launches": tr.launches,
               "groups": tr.groups, "candidates": int(tr.simulated.sum()), "detected_faults": int(tr.detected.sum()),
               "coverage_vs_plain": float(tr.detected.sum() / max(1, plain_cov.sum())),
               "missed_faults": int((plain_cov & ~tr.detected).sum()),
               "unconfirmed_detections": int((tr.confirmed & ~Le.detected).sum()),
               "audit_candidates": tr.audit_candidates, "audit_detections": tr.audit_detections,
               "audit_new_faults": tr.audit_new_faults, "stages": tr.stages, **extra}
        log(f"{name}: {gpu:.1f} ms ({row['time_vs_plain']:.3f} of plain), coverage {row['coverage_vs_plain']:.4f} "
            f"of plain, missed {row['missed_faults']}, unconfirmed {row['unconfirmed_detections']}, "
            f"{row['candidates']} candidates in {tr.launches} launches")
        return row

    out["triage_parity"] = record("triage (network, exact rules, tail)",
                                  triage(fs, prob, eligible, eval_ids, stages, pmins[0], True, audit_frac), True,
                                  pmin=pmins[0])
    out["triage"] = [record(f"triage (network, exact rules, pmin {pm})",
                            triage(fs, prob, eligible, eval_ids, stages, pm, False, audit_frac), True, pmin=pm)
                     for pm in pmins]
    out["rules_dropping"] = record("exact rules + staged dropping, no network",
                                   triage(fs, None, eligible, eval_ids, stages, 0.0, True, 0.0), False)
    # the network scoring only what a pool-order first stage leaves undetected
    scored = {"ms": 0.0, "candidates": 0}

    def score(f, s):
        pr = sur.predict_pairs(f, fe.reg_index[f].astype(np.uint32) * np.uint32(n_eval) + s.astype(np.uint32),
                               extra=fe.extra[f.astype(np.int64) * n_eval + s])
        scored["ms"] += sur.stats()["kernel_ms"]
        scored["candidates"] += int(f.size)
        return pr

    tr = triage_deferred(fs, score, eligible, eval_ids, stages[0], stages[1:], True)
    out["triage_deferred"] = record("pool-order first stage, network on the rest, exact rules, tail", tr, False,
                                    score_ms=scored["ms"], scored_candidates=scored["candidates"])
    d0 = dropping(fs, eval_ids)
    out["dropping"] = {"gpu_ms": d0["gpu_ms"], "groups": d0["groups"], "launches": d0["launches"],
                       "detected_faults": int(d0["detected"].sum()), "time_vs_plain": d0["gpu_ms"] / Le.kernel_ms,
                       "unconfirmed_detections": int((d0["detected"] & ~plain_cov).sum())}
    log(f"dropping, no rules, no network: {d0['gpu_ms']:.1f} ms ({out['dropping']['time_vs_plain']:.3f} of plain), "
        f"detected {out['dropping']['detected_faults']}")
    out["wall_s"] = time.time() - t0
    sur.close()
    return out
