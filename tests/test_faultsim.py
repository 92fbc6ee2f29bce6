"""CPU checks of the fault-simulation pieces: candidate packing, the oracle's
group-shared stimulus against its numpy restatement, and on-trace detection
against the reference harness's per-lane comparison (fault_sim.rs:86-95)."""
from __future__ import annotations

import numpy as np

import oracle
from oracle import stim
from rrtl_amd import designs, lower
from rrtl_amd.faultsim import FAULTS_PER_GROUP, pack, register_faults
from rrtl_amd.program import KIND_INPUT, KIND_OUTPUT, KIND_REG


def small_program():
    return lower.lower_module(designs.synthetic_netlist(gates=600, levels=12, dffs=40, inputs=12, outputs=8, seed=3))


def test_register_faults_order():
    p = small_program()
    f = register_faults(p)
    regs = [i for i, s in enumerate(p.signals) if s.kind == KIND_REG]
    assert f.shape == (2 * len(regs), 2)
    assert list(f[:4].ravel()) == [regs[0], 0, regs[0], 1, regs[1], 0, regs[1], 1]


def test_pack_invariants():
    rng = np.random.default_rng(0)
    stim_id = rng.integers(0, 7, 1000).astype(np.uint64)
    pk = pack(stim_id)
    assert pk.lanes == pk.groups * 64
    assert len(np.unique(pk.lane)) == stim_id.size          # one lane per candidate
    assert np.all(pk.lane % 64 != 0)                        # lane 0 of a group stays golden
    assert np.array_equal(pk.group_key[pk.lane // 64], stim_id)  # group carries the candidate's stimulus
    per_group = np.bincount(pk.lane // 64, minlength=pk.groups)
    assert per_group.max() <= FAULTS_PER_GROUP
    counts = np.bincount(stim_id.astype(np.int64))
    assert pk.groups == int(sum((c + 62) // 63 for c in counts))
    empty = pack(np.zeros(0, dtype=np.uint64))
    assert empty.groups == 0 and empty.lanes == 0


def test_oracle_group_stimulus_matches_restatement():
    """bp_run's GROUP_RANDOM inputs equal a step-by-step replay of the numpy
    restatement (set every input word, tick), trace for trace."""
    p = small_program()
    lanes, cycles, seed = 256, 20, 11
    keys = np.array([5, 9, 5, 1000], dtype=np.uint64)
    ins = [i for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    obs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
    a = oracle.BitParallelRef(p, lanes)
    _, tr = a.run(seed=seed, cycles=cycles, inputs=[(i, oracle.IN_GROUP_RANDOM, 0) for i in ins], observe=obs,
                  observe_post=True, digest=False, trace=True, group_key=keys)
    b = oracle.BitParallelRef(p, lanes)
    for c in range(cycles):
        for i in ins:
            b.set_words(i, stim.group_random_words(seed, i, c, keys))
        b.tick()
        got = np.stack([b.get_words(o) for o in obs])
        assert np.array_equal(got, tr[c]), c
    # groups with the same key saw the same inputs: their golden lanes agree
    assert np.array_equal(tr[:, :, 0] & np.uint64(1), tr[:, :, 2] & np.uint64(1))


def test_detect_from_trace_matches_lane_compare():
    """Bitwise detection over a trace equals the harness's per-lane comparison
    of each lane's outputs with lane 0 after every tick."""
    p = small_program()
    faults = register_faults(p)
    lanes = 64 * 2
    ref = oracle.BitParallelRef(p, lanes)
    for k, (sig, v) in enumerate(faults[: 2 * FAULTS_PER_GROUP]):
        g, j = divmod(k, FAULTS_PER_GROUP)
        ref.set_force(int(sig), g * 64 + 1 + j, int(v))
    ins = [i for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    obs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
    keys = np.array([3, 3], dtype=np.uint64)
    _, tr = ref.run(seed=2, cycles=64, inputs=[(i, oracle.IN_GROUP_RANDOM, 0) for i in ins], observe=obs,
                    observe_post=True, digest=False, trace=True, group_key=keys)
    flags = stim.detect_from_trace(tr)
    # per-lane comparison, one lane at a time
    bits = (tr[:, :, :, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)   # [c, obs, word, bit]
    for w in range(2):
        gold = bits[:, :, w, 0]
        for b in range(64):
            differs = bool(np.any(bits[:, :, w, b] != gold))
            assert flags[w * 64 + b] == differs
    assert not flags[0] and not flags[64]
    assert 0 < flags.sum() < 126


def test_observable_signals_on_hand_built_design():
    """A register feeding only a dead gate, and one feeding an output through
    another register, against the reachability rule."""
    from rrtl_amd import ir
    from rrtl_amd.faultsim import observable_signals
    m = ir.Module("Obs")
    a = m.input("a", 1)
    q_dead = m.reg("q_dead", 1)
    q_mid = m.reg("q_mid", 1)
    q_out = m.reg("q_out", 1)
    dead = m.wire("dead", 1)
    m.assign(dead, q_dead.value() & a.value())          # read by nothing
    m.next(q_dead, a.value())
    m.next(q_mid, a.value() ^ q_mid.value())
    m.next(q_out, q_mid.value())                        # q_mid reaches y only through q_out
    y = m.output("y", 1)
    m.assign(y, q_out.value())
    p = lower.lower_module(m)
    obs = observable_signals(p)
    idx = {s.name: i for i, s in enumerate(p.signals)}
    assert obs[idx["q_out"]] and obs[idx["q_mid"]] and obs[idx["a"]]
    assert not obs[idx["q_dead"]] and not obs[idx["dead"]]


def test_unobservable_register_faults_never_detected_by_oracle():
    """Every fault the rule calls unobservable stays undetected in an oracle
    fault simulation of the small netlist."""
    from rrtl_amd.faultsim import observable_signals
    p = small_program()
    faults = register_faults(p)
    obs = observable_signals(p)[faults[:, 0]]
    assert (~obs).sum() > 0
    lanes = 64 * ((len(faults) + 62) // 63)
    ref = oracle.BitParallelRef(p, lanes)
    for k, (sig, v) in enumerate(faults):
        g, j = divmod(k, FAULTS_PER_GROUP)
        ref.set_force(int(sig), g * 64 + 1 + j, int(v))
    ins = [i for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    outs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
    keys = np.full(lanes // 64, 5, dtype=np.uint64)
    _, tr = ref.run(seed=4, cycles=64, inputs=[(i, oracle.IN_GROUP_RANDOM, 0) for i in ins], observe=outs,
                    observe_post=True, digest=False, trace=True, group_key=keys)
    flags = stim.detect_from_trace(tr)
    lane = np.array([(k // FAULTS_PER_GROUP) * 64 + 1 + k % FAULTS_PER_GROUP for k in range(len(faults))])
    det = flags[lane]
    assert not (det & ~obs).any()
    assert det.any()


class _TableFaultSim:
    """Stands in for FaultSim with a fixed exact label matrix (CPU only)."""

    def __init__(self, labels, stim_ids):
        self.labels = labels
        self.col = {int(s): i for i, s in enumerate(stim_ids)}
        self.faults = np.zeros((labels.shape[0], 2), dtype=np.uint32)
        self.runs = []

    def run(self, fi, stim):
        from rrtl_amd.faultsim import RunResult, pack
        cols = np.array([self.col[int(s)] for s in stim])
        self.runs.append((np.asarray(fi).copy(), cols))
        return RunResult(self.labels[fi, cols], 1.0, pack(stim).groups, 0)


def test_staged_triage_logic():
    """Stages, parity tail, give-up without the tail, and the pool order of the
    method without a network, on a synthetic label matrix."""
    from rrtl_amd import triage
    rng = np.random.default_rng(0)
    F, S = 50, 40
    rate = rng.choice([0.0, 0.02, 0.3, 0.9], F)
    labels = rng.random((F, S)) < rate[:, None]
    eligible = labels | (rng.random((F, S)) < 0.7)
    ids = np.arange(100, 100 + S, dtype=np.uint64)
    plain = labels.any(axis=1)
    prob = np.where(labels, 0.6, rate[:, None] * 0.5 + 1e-4)     # an informative score
    for p_, tail in ((prob, True), (None, True), (prob, False)):
        fs = _TableFaultSim(labels, ids)
        r = triage.triage(fs, p_, eligible, ids, stages=(2, 5), pmin=1e-3, tail=tail, audit_frac=0.0)
        assert not (r.confirmed & ~labels).any()
        assert not (r.simulated & ~eligible).any()
        if tail:
            assert np.array_equal(r.detected, plain)
        else:
            assert not (r.detected & ~plain).any()
        # a fault is never simulated again after a stage that detected it
        seen_det = np.zeros(F, dtype=bool)
        for fi, cols in fs.runs:
            assert not seen_det[fi].any()
            seen_det[fi[labels[fi, cols]]] = True
    # without a network the first stage takes each fault's first eligible stimuli in pool order
    fs = _TableFaultSim(labels, ids)
    triage.triage(fs, None, eligible, ids, stages=(2,), pmin=0.0, tail=False, audit_frac=0.0)
    fi, cols = fs.runs[0]
    for f in range(F):
        assert list(np.sort(cols[fi == f])) == list(np.flatnonzero(eligible[f])[:2])
