"""Generated code split between several waves per 64 lanes
(rrtl_hip_options.lane_waves): every wave of a workgroup runs its part of each
evaluation on the same lanes, values cross between waves through LDS and
barriers order the parts.  Results must be those of one wave per lane set:
random designs against the packed oracle, picorv32 against the host
batch-SIMD engine signal by signal and cycle by cycle, halting and lane
compaction against the unsplit kernel and the countdown's closed form."""
import random

import numpy as np
import pytest

import oracle
from conftest import flush_print
from designgen import random_design
from rrtl_amd import divergent, lower, riscv, workloads
from rrtl_amd import picorv32 as pv
from rrtl_amd.backend import HipError, HipSimulator
from rrtl_amd.program import KIND_INPUT

pytestmark = pytest.mark.gpu


def make_split(p, lanes, waves, **kw):
    """The engine with `waves` waves per 64 lanes, or None when the design's
    code cannot be split (interpreted records, a reload after a store)."""
    try:
        sim = HipSimulator(p, lanes, engine="word", codegen="on", lane_waves=waves, **kw)
    except HipError as e:
        if "lane_waves" in str(e):
            return None
        raise
    assert sim.launch_info()["lane_waves"] == waves
    return sim


@pytest.mark.parametrize("waves", [2, 4])
@pytest.mark.parametrize("seed", list(range(12)))
def test_random_designs_split_match_packed_oracle(seed, waves):
    flush_print(f"case: split random design seed={seed} waves={waves}")
    # one-limb values only: wider ones keep interpreted records, which are not split
    p = lower.lower_module(random_design(seed, async_resets=False, widths=(1, 4, 8, 13, 16, 31, 32),
                                         one_limb_memory=True))
    lanes = 160
    sim = make_split(p, lanes, waves)
    if sim is None:
        pytest.skip("the design's generated code cannot be split")
    ref = oracle.PackedRef(p, lanes)
    rng = random.Random(seed)
    inputs = [i for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    for c in range(16):
        for i in inputs:
            s = p.signals[i]
            if s.name == "rst":
                v = np.full((1, lanes), 1 if c < 2 else rng.getrandbits(1), np.uint32)
            elif s.name == "rstn":
                v = np.full((1, lanes), 0 if c < 2 else 1, np.uint32)
            else:
                v = np.array([[rng.getrandbits(32) for _ in range(lanes)] for _ in range(s.limbs)], np.uint32)
            sim.set_signal_limbs(i, v, settle=False)
            ref.set_signal_limbs(i, v, eval=False)
        if c % 3 == 0:
            sim.eval_combinational(); ref.eval_combinational()
        if c % 4 == 1:
            sim.tick_from_evaluated(); ref.tick_from_evaluated()
        else:
            sim.tick(); ref.tick()
        for i in range(len(p.signals)):
            assert np.array_equal(sim.get_signal_limbs(i), ref.get_signal_limbs(i)), (seed, c, p.signals[i].name)
        for k in range(len(p.memories)):
            assert np.array_equal(sim.get_memory_limbs(k), ref.get_memory_limbs(k)), (seed, c, "memory")


@pytest.mark.parametrize("waves,wave_lanes", [(2, 0), (4, 0), (4, 32), (4, 16), (2, 32)])
def test_picorv32_split_every_signal_every_cycle(waves, wave_lanes):
    """300 lanes (a partial last workgroup), a different 512-instruction
    program per lane, 3000 cycles: per-lane digests of every signal at every
    cycle and the final memories equal the host engine's; with 64 lanes per
    workgroup, or (wave_lanes) 32 or 16, the waves' other threads idle."""
    flush_print(f"case: split picorv32 digest, 300 lanes x 3000 cycles, waves={waves} wave_lanes={wave_lanes}")
    p = pv.program()
    lanes, cycles = 300, 3000
    rng = random.Random(17)
    progs = [pv.random_program(rng, n=512) for _ in range(lanes)]
    ins = pv.inputs(p)
    obs = [i for i, s in enumerate(p.signals) if s.kind != KIND_INPUT]
    sim = make_split(p, lanes, waves, wave_lanes=wave_lanes)
    assert sim is not None
    assert sim.launch_info()["wave_lanes"] == (wave_lanes or 64)
    sim.set_memory_limbs("mem", pv.images(progs))
    sim.run(cycles, seed=3, inputs=ins, observe=obs)
    cpu = oracle.SimdCpu(p, lanes, threads=8)
    cpu.set_memory_limbs(p.memory_index("mem"), pv.images(progs))
    dg, _ = cpu.run(0, cycles, ins, obs)
    assert np.array_equal(sim.digest(), dg)
    for m in ("mem", "cpu.cpuregs"):
        assert np.array_equal(sim.get_memory_limbs(m), cpu.get_memory_limbs(p.memory_index(m))), m


@pytest.mark.parametrize("compact", [False, True])
def test_picorv32_split_halting_equals_unsplit(compact):
    """Uneven halting with and without lane compaction: halt cycles, trace of
    the program counter, registers and memory equal the one-wave kernel's."""
    flush_print(f"case: split picorv32 halting, 700 lanes, compact={compact}")
    p = pv.program()
    lanes, cap = 700, 3000
    rng = random.Random(23)
    progs = [pv.random_program(rng, n=64, backward=0.2) for _ in range(lanes)]
    obs = [p.signal_index(n) for n in ("cpu.reg_pc", "cpu.mem_addr", "trap")]
    out = {}
    for waves in (1, 4):
        sim = HipSimulator(p, lanes, engine="word", codegen="on", lane_waves=waves, compact_lanes=compact,
                           ticks_per_launch=128 if compact else 0)
        assert sim.launch_info()["lane_waves"] == waves
        sim.set_memory_limbs("mem", pv.images(progs))
        sim.run(cap, inputs=pv.inputs(p), observe=obs, halt="trap", trace=not compact)
        out[waves] = (sim.halt_cycles(), sim.get_memory_limbs("cpu.cpuregs"), sim.get_memory_limbs("mem"),
                      sim.digest(), sim.snapshot()[0], None if compact else sim.trace())
        sim.close()
    for x, y in zip(out[1], out[4]):
        if x is not None:
            assert np.array_equal(x, y)
    assert 0 < int((out[4][0] < cap).sum()) < lanes


def test_countdown_forced_split_with_compaction_matches_closed_form():
    """A design too small to gain from the split, split anyway: halt cycles
    and accumulators equal the closed form with lane compaction on."""
    flush_print("case: split countdown, 5000 lanes, compaction")
    p = divergent.program()
    lanes, din, cap = 5000, 0x9E3779B1, 1000
    seeds = divergent.run_lengths(lanes, short=40, long=600, seed=3)
    seeds[:5] = [0, 1, 2, 63, 999]
    sim = make_split(p, lanes, 2, compact_lanes=True, ticks_per_launch=64)
    assert sim is not None
    sim.set_signal("seed", seeds)
    sim.set_signal("din", [din] * lanes)
    sim.run(cap, inputs=divergent.run_inputs(p), observe=["acc", "left"], halt="done")
    assert np.array_equal(sim.halt_cycles(), np.maximum(seeds.astype(np.int64), 1) + 1)
    assert np.array_equal(sim.get_signal_limbs("acc")[0], divergent.closed_form(seeds, din))


def test_rv32i_core_split_equals_unsplit():
    flush_print("case: split rv32i core, 1000 lanes")
    job = workloads.rv32i_job(1000, 1500, 7, n_insns=128, distinct=64)
    res = {}
    for waves in (1, 2, 4):
        sim = HipSimulator(job.program, job.lanes, engine="word", codegen="on", lane_waves=waves)
        job.setup(sim)
        sim.run(**dict(job.run, observe=[job.program.signal_index("pc")]))
        res[waves] = (sim.digest(), sim.halt_cycles(), sim.get_memory_limbs("regs"), sim.get_memory_limbs("mem"))
        sim.close()
    for waves in (2, 4):
        for x, y in zip(res[1], res[waves]):
            assert np.array_equal(x, y), waves


def test_auto_split_at_16384_lanes():
    """The engine splits picorv32 by itself when the batch gives fewer than
    one 64-lane wave per SIMD; the split run equals the one-wave run."""
    flush_print("case: automatic split picorv32, 16384 lanes x 1500 cycles")
    job = workloads.picorv32_job(16384, 1500, 1, n_insns=512, distinct=64)
    res = {}
    for waves in (0, 1):
        sim = job.make(codegen="on", lane_waves=waves)
        info = sim.launch_info()
        if waves == 0:
            assert info["lane_waves"] == 4, info
        sim.run(**dict(job.run, observe=[job.program.signal_index("cpu.reg_pc")]))
        res[waves] = (sim.digest(), sim.halt_cycles(), sim.get_memory_limbs("cpu.cpuregs"))
        sim.close()
    for x, y in zip(res[0], res[1]):
        assert np.array_equal(x, y)
