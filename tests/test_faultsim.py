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
