"""Fault simulation on the GPU (group-shared stimulus, on-chip detection)
against the bit-parallel oracle's trace, lane for lane."""
import numpy as np
import pytest

import oracle
from oracle import stim
from rrtl_amd import designs, lower
from rrtl_amd.backend import IN_GROUP_RANDOM, HipError, HipSimulator
from rrtl_amd.faultsim import FaultSim, pack, register_faults
from rrtl_amd.program import KIND_INPUT, KIND_OUTPUT, KIND_REG, KIND_WIRE
from conftest import flush_print

pytestmark = pytest.mark.gpu


def oracle_detect(p, faults, fault_idx, stim_id, cycles, seed, threads=0):
    """Per-candidate detection from the oracle's trace of the same packing."""
    pk = pack(stim_id)
    ref = oracle.BitParallelFast(p, pk.lanes, threads=threads) if threads else oracle.BitParallelRef(p, pk.lanes)
    for k, f in enumerate(fault_idx):
        ref.set_force(int(faults[f, 0]), int(pk.lane[k]), int(faults[f, 1]))
    ins = [i for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    obs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
    out = ref.run(seed=seed, cycles=cycles, inputs=[(i, IN_GROUP_RANDOM, 0) for i in ins], observe=obs,
                  observe_post=True, digest=False, trace=True, group_key=pk.group_key)
    return stim.detect_from_trace(out[1])[pk.lane]


@pytest.mark.parametrize("tile_words", [2, 4, 8])
def test_fault_detect_matches_oracle_small(tile_words):
    flush_print(f"case: fault detect small netlist tile_words={tile_words}")
    p = lower.lower_module(designs.synthetic_netlist(gates=2000, levels=20, dffs=150, inputs=24, outputs=16, seed=7))
    faults = register_faults(p)
    rng = np.random.default_rng(tile_words)
    n = 900
    fault_idx = rng.integers(0, len(faults), n)
    stim_id = rng.integers(0, 9, n).astype(np.uint64)
    fs = FaultSim(p, faults, cycles=96, seed=5, tile_words=tile_words)
    got = fs.run(fault_idx, stim_id)
    want = oracle_detect(p, faults, fault_idx, stim_id, 96, 5)
    assert np.array_equal(got.detected, want)
    assert 0 < want.sum() < n
    assert got.kernel_ms > 0


def test_fault_detect_matches_oracle_20k_netlist():
    """The 20k-gate netlist: every register fault under two stimuli (8064
    faulty lanes), 256 cycles."""
    flush_print("case: fault detect 20k netlist, 4000 faults x 2 stimuli")
    p = lower.lower_module(designs.synthetic_netlist())
    faults = register_faults(p)
    fault_idx = np.tile(np.arange(len(faults)), 2)
    stim_id = np.repeat(np.array([17, 4242], dtype=np.uint64), len(faults))
    fs = FaultSim(p, faults, cycles=256, seed=1)
    got = fs.run(fault_idx, stim_id)
    want = oracle_detect(p, faults, fault_idx, stim_id, 256, 1, threads=16)
    assert np.array_equal(got.detected, want)
    assert 0.3 < want.mean() < 0.95


def test_stage_and_commit_clamps_match_oracle():
    """Forces on wires (clamped after the stage that writes them) and on
    registers (clamped at the commit) in one run, traced against the oracle."""
    flush_print("case: wire and register forces, traced")
    p = lower.lower_module(designs.synthetic_netlist(gates=1500, levels=16, dffs=100, inputs=16, outputs=12, seed=9))
    lanes, cycles, seed = 640, 40, 3
    wires = [i for i, s in enumerate(p.signals) if s.kind == KIND_WIRE][::37]
    regs = [i for i, s in enumerate(p.signals) if s.kind == KIND_REG][::7]
    rng = np.random.default_rng(1)
    forces = [(s, int(rng.integers(0, lanes)), int(rng.integers(0, 2))) for s in wires + regs for _ in range(3)]
    ins = [(i, IN_GROUP_RANDOM, 0) for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    obs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
    keys = rng.integers(0, 1 << 40, (lanes + 63) // 64).astype(np.uint64)
    sim = HipSimulator(p, lanes, engine="gate", tile_words=2)
    f = np.array(forces, dtype=np.int64)
    sim.set_forces(f[:, 0], f[:, 1], f[:, 2])
    sim.run(cycles, seed=seed, inputs=ins, observe=obs, observe_post=True, trace=True, group_key=keys)
    ref = oracle.BitParallelRef(p, lanes)
    for s, l, v in forces:
        ref.set_force(s, l, v)
    _, tr = ref.run(seed=seed, cycles=cycles, inputs=ins, observe=obs, observe_post=True, digest=False, trace=True,
                    group_key=keys)
    assert np.array_equal(stim.u32_rows_to_u64(sim.trace().reshape(cycles, len(obs), -1)), tr)
    ref_state = np.stack([ref.get_words(s) for s in regs])
    dev_state = np.stack([stim.u32_rows_to_u64(sim.get_signal_bits(s)) for s in regs])
    assert np.array_equal(dev_state, ref_state)


def test_detect_needs_two_word_tiles():
    p = lower.lower_module(designs.synthetic_netlist(gates=300, levels=6, dffs=20, inputs=8, outputs=4, seed=2))
    sim = HipSimulator(p, 128, engine="gate", tile_words=1)
    obs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
    with pytest.raises(HipError) as e:
        sim.run(4, observe=obs, detect=True)
    assert e.value.code == "E_HIP_UNSUPPORTED"
