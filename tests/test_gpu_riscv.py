"""RV32I core on the word engine: per-lane random programs halting on ecall
(uneven halt cycles), against the instruction-set simulator, and the
sum-program closed form."""
import random

import numpy as np
import pytest

from rrtl_amd import lower, riscv
from rrtl_amd.backend import IN_CONST, HipSimulator
from rrtl_amd.riscv import Iss, load_image
from conftest import flush_print

pytestmark = pytest.mark.gpu


def core_inputs(p):
    return [(p.signal_index("clk"), IN_CONST, 1), (p.signal_index("rst"), IN_CONST, 0)]


def load_programs(sim, progs):
    img = np.array([load_image(pr) for pr in progs], dtype=np.uint32)
    sim.set_memory_limbs("mem", img.reshape(len(progs), -1, 1))


def test_sum_program_closed_form():
    p = lower.lower_module(riscv.rv32i_core())
    limits = [1, 2, 6, 10, 33]
    sim = HipSimulator(p, len(limits), engine="word")
    load_programs(sim, [riscv.sum_program(n) for n in limits])
    sim.run(2000, inputs=core_inputs(p), halt="trap")
    a0 = sim.get_signal("a0")
    mem = sim.get_memory_limbs("mem")
    for l, n in enumerate(limits):
        assert a0[l] == sum(range(1, n))
        assert mem[l, 0x100 >> 2, 0] == sum(range(1, n))
    assert all(t == 1 for t in sim.get_signal("trap"))


@pytest.mark.parametrize("codegen,wave_lanes", [("on", 0), ("off", 0), ("on", 4), ("on", 8), ("on", 32), ("off", 16)])
@pytest.mark.parametrize("lanes,n,cap,seed", [(300, 64, 1500, 1), (1024, 512, 4000, 2)])
def test_random_programs_match_iss(lanes, n, cap, seed, codegen, wave_lanes):
    """Uneven halting: each lane stops after the cycle whose edge raised trap
    (its ecall, so the cycle count equals the instructions the ISS retired,
    the ecall included); lanes that loop run to the cap.  Halt cycles,
    registers and memory equal the ISS's."""
    flush_print(f"case: rv32i random programs lanes={lanes} n={n} cap={cap} codegen={codegen} wave_lanes={wave_lanes}")
    p = lower.lower_module(riscv.rv32i_core())
    rng = random.Random(seed)
    progs = [riscv.random_program(rng, n=n) for _ in range(lanes)]
    sim = HipSimulator(p, lanes, engine="word", codegen=codegen, wave_lanes=wave_lanes)
    load_programs(sim, progs)
    sim.run(cap, inputs=core_inputs(p), halt="trap")
    halt = sim.halt_cycles()
    regs = sim.get_memory_limbs("regs")
    mem = sim.get_memory_limbs("mem")
    halted = 0
    for l, pr in enumerate(progs):
        s = Iss(load_image(pr))
        s.run(cap)
        if s.trap and s.cycles < cap:
            assert halt[l] == s.cycles, l
            halted += 1
        else:
            assert halt[l] == cap, l
        assert [int(x) for x in regs[l, :, 0]] == s.regs, l
        assert [int(x) for x in mem[l, :, 0]] == s.mem, l
    assert 0 < halted < lanes  # uneven: some lanes halt, some loop to the cap
    assert sim.stats()["lane_cycles"] == int(halt.sum())
