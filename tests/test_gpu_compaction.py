"""Lane compaction of halting runs (rrtl_hip_options.compact_lanes): lanes
that halted are dropped between launches and the live ones packed into full
waves.  Compacted and uncompacted runs give the same halt cycles, final state
and per-lane observation digests, and the divergent countdown matches its
closed form."""
import random

import numpy as np
import pytest

from conftest import flush_print
from rrtl_amd import divergent
from rrtl_amd import picorv32 as pv
from rrtl_amd.backend import HipSimulator

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("codegen", ["on", "off"])
def test_divergent_countdown_compacted_equals_uncompacted(codegen):
    flush_print(f"case: compaction countdown, 5000 lanes, codegen={codegen}")
    p = divergent.program()
    lanes, din, cap = 5000, 0x9E3779B1, 1000
    seeds = divergent.run_lengths(lanes, short=40, long=600, seed=3)
    seeds[:7] = [0, 1, 2, 63, 64, 65, 999]
    out = {}
    for compact in (False, True):
        sim = HipSimulator(p, lanes, engine="word", codegen=codegen, compact_lanes=compact, ticks_per_launch=64)
        sim.set_signal("seed", seeds)
        sim.set_signal("din", [din] * lanes)
        sim.run(cap, inputs=divergent.run_inputs(p), observe=["acc", "left"], halt="done")
        out[compact] = (sim.halt_cycles(), sim.get_signal_limbs("acc"), sim.get_signal_limbs("left"),
                        sim.digest(), sim.stats())
        sim.close()
    a, b = out[False], out[True]
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    assert a[4]["lane_cycles"] == b[4]["lane_cycles"] == int(a[0].sum())
    assert b[4]["launches"] == (cap + 63) // 64
    want_halt = np.maximum(seeds.astype(np.int64), 1) + 1
    assert np.array_equal(a[0], want_halt)
    assert np.array_equal(a[1][0], divergent.closed_form(seeds, din))
    assert not a[2].any()


@pytest.mark.parametrize("codegen", ["on", "off"])
def test_picorv32_compacted_equals_uncompacted(codegen):
    flush_print(f"case: compaction picorv32, 700 lanes, codegen={codegen}")
    p = pv.program()
    lanes, cap = 700, 3000
    rng = random.Random(21)
    progs = [pv.random_program(rng, n=64, backward=0.2) for _ in range(lanes)]
    obs = [p.signal_index(n) for n in ("cpu.reg_pc", "cpu.mem_addr", "trap")]
    out = {}
    for compact in (False, True):
        sim = HipSimulator(p, lanes, engine="word", codegen=codegen, compact_lanes=compact, ticks_per_launch=100)
        sim.set_memory_limbs("mem", pv.images(progs))
        sim.run(cap, inputs=pv.inputs(p), observe=obs, halt="trap")
        out[compact] = (sim.halt_cycles(), sim.get_memory_limbs("cpu.cpuregs"), sim.get_memory_limbs("mem"),
                        sim.digest(), sim.snapshot()[0])
        sim.close()
    for x, y in zip(out[False], out[True]):
        assert np.array_equal(x, y)
    halt = out[True][0]
    assert 0 < int((halt < cap).sum()) < lanes
