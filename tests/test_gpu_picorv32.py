"""picorv32 (the reference's core and SoC, imported from its Verilog by
rrtl_amd.verilog) on the word engine, generated kernel and record
interpreter: the reference's sum known answer on every lane, every signal at
every cycle against the host batch-SIMD engine (per-lane digests), halt
cycles against the packed oracle, and architectural state against the RV32I
instruction-set simulator, up to the full 16,384-lane, 512-instruction
configuration."""
import os
import random

import numpy as np
import pytest

import oracle
from conftest import flush_print
from rrtl_amd import picorv32 as pv
from rrtl_amd.backend import HipSimulator
from rrtl_amd.program import KIND_INPUT

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def observed(p):
    return [i for i, s in enumerate(p.signals) if s.kind != KIND_INPUT]


def host_run(p, progs, cycles, ins, obs):
    cpu = oracle.SimdCpu(p, len(progs), threads=THREADS)
    cpu.set_memory_limbs(p.memory_index("mem"), pv.images(progs))
    dg, _ = cpu.run(0, cycles, ins, obs)
    return cpu, dg


@pytest.mark.parametrize("codegen", ["on", "off"])
def test_sum_known_answer_every_lane(codegen):
    """picorv32_gpu.rs:24-36, :85-91: 8 cycles of reset, 3000 cycles, then
    result == 15, done == 1 and trap == 0 on every lane; the whole run equals
    the host engine signal by signal, cycle by cycle."""
    flush_print(f"case: picorv32 sum known answer, 2048 lanes, codegen={codegen}")
    p = pv.program()
    lanes, cycles = 2048, 8 + 3000
    progs = [pv.sum_program()] * lanes
    ins = pv.inputs(p, reset_cycles=8)
    obs = observed(p)
    sim = HipSimulator(p, lanes, engine="word", codegen=codegen)
    sim.set_memory_limbs("mem", pv.images(progs))
    sim.run(cycles, inputs=ins, observe=obs)
    assert set(sim.get_signal("result")) == {15}
    assert set(sim.get_signal("done")) == {1}
    assert set(sim.get_signal("trap")) == {0}
    cpu, dg = host_run(p, progs, cycles, ins, obs)
    assert np.array_equal(sim.digest(), dg)
    for m in ("mem", "cpu.cpuregs"):
        assert np.array_equal(sim.get_memory_limbs(m), cpu.get_memory_limbs(p.memory_index(m))), m


@pytest.mark.parametrize("codegen", ["on", "off"])
def test_random_programs_every_signal_every_cycle(codegen):
    """1024 lanes, a different 512-instruction program on each, 4000 cycles
    without halting: per-lane digests of every signal at every cycle and the
    final memories equal the host engine's."""
    flush_print(f"case: picorv32 random programs digest, 1024 lanes x 4000 cycles, codegen={codegen}")
    p = pv.program()
    lanes, cycles = 1024, 4000
    rng = random.Random(11)
    progs = [pv.random_program(rng, n=512) for _ in range(lanes)]
    ins = pv.inputs(p)
    obs = observed(p)
    sim = HipSimulator(p, lanes, engine="word", codegen=codegen)
    sim.set_memory_limbs("mem", pv.images(progs))
    sim.run(cycles, seed=3, inputs=ins, observe=obs)
    cpu, dg = host_run(p, progs, cycles, ins, obs)
    assert np.array_equal(sim.digest(), dg)
    for m in ("mem", "cpu.cpuregs"):
        assert np.array_equal(sim.get_memory_limbs(m), cpu.get_memory_limbs(p.memory_index(m))), m


def oracle_halts(p, progs, cap):
    """The device loop on the packed oracle: per cycle drive inputs, settle,
    edge, settle; a lane halts after the cycle whose edge raised trap (the
    halt signal is a combinational output, read after the post-edge settle)."""
    lanes = len(progs)
    ref = oracle.PackedRef(p, lanes)
    ref.set_memory_limbs(p.memory_index("mem"), pv.images(progs))
    clk, rst, trap = p.signal_index("clk"), p.signal_index("resetn"), p.signal_index("trap")
    ref.set_signal(clk, 1, eval=False)
    halt = np.full(lanes, cap, dtype=np.int64)
    for c in range(cap):
        ref.set_signal(rst, 0 if c < pv.RESET_CYCLES else 1, eval=False)
        ref.eval_combinational()
        ref.tick_from_evaluated()
        ref.eval_combinational()
        t = ref.get_signal_limbs(trap)[0]
        newly = (t == 1) & (halt == cap)
        halt[newly] = c + 1
        if (halt < cap).all():
            break
    return halt


@pytest.mark.parametrize("codegen", ["on", "off"])
def test_halting_programs_against_oracle_and_iss(codegen):
    """Uneven halting: halt cycles equal the packed oracle's; halted lanes'
    registers and memory equal the ISS after the program's ecall; lanes that
    loop run to the cap; simulated lane-cycles = sum of halt cycles."""
    flush_print(f"case: picorv32 halting programs, 128 lanes, cap 3000, codegen={codegen}")
    p = pv.program()
    lanes, cap = 128, 3000
    rng = random.Random(5)
    progs = [pv.random_program(rng, n=64, backward=0.2) for _ in range(lanes)]
    sim = HipSimulator(p, lanes, engine="word", codegen=codegen)
    sim.set_memory_limbs("mem", pv.images(progs))
    sim.run(cap, inputs=pv.inputs(p), halt="trap")
    halt = sim.halt_cycles()
    want = oracle_halts(p, progs, cap)
    assert np.array_equal(halt, want)
    assert sim.stats()["lane_cycles"] == int(halt.sum())
    regs = sim.get_memory_limbs("cpu.cpuregs")[:, :, 0]
    mem = sim.get_memory_limbs("mem")[:, :, 0]
    halted = 0
    for l, pr in enumerate(progs):
        if halt[l] >= cap:
            continue
        s = pv.iss_run(pr, 100_000)
        assert s.trap, l
        assert [int(x) for x in regs[l]] == s.regs, l
        assert [int(x) for x in mem[l]] == s.mem, l
        halted += 1
    assert 0 < halted < lanes


def test_full_configuration_against_iss():
    """The full configuration (16,384 lanes, 512-instruction programs, 256
    distinct programs dealt round-robin, cap 20,000 cycles for test time):
    every lane's registers and memory equal the ISS run for that lane's
    program (ecall reached within the cap) or equal the lanes sharing its
    program (loops), and halted lanes report trap."""
    from rrtl_amd import workloads
    flush_print("case: picorv32 full configuration vs ISS, 16384 lanes, cap 20000")
    job = workloads.picorv32_job(16384, 20_000, seed=1, backward=0.2)
    sim = job.make()
    sim.run(**job.run)
    halt = sim.halt_cycles()
    regs = sim.get_memory_limbs("cpu.cpuregs")[:, :, 0]
    mem = sim.get_memory_limbs("mem")[:, :, 0]
    trap = np.array(sim.get_signal("trap"))
    rng = random.Random(1)
    progs = [pv.random_program(rng, n=512, backward=0.2) for _ in range(256)]
    halted_programs = 0
    for k, pr in enumerate(progs):
        lanes = np.arange(k, 16384, 256)
        # lanes with one program behave identically
        assert (halt[lanes] == halt[k]).all() and (regs[lanes] == regs[k]).all() and (mem[lanes] == mem[k]).all()
        if halt[k] < 20_000:
            s = pv.iss_run(pr, 1_000_000)
            assert s.trap and trap[k] == 1, k
            assert [int(x) for x in regs[k]] == s.regs, k
            assert [int(x) for x in mem[k]] == s.mem, k
            halted_programs += 1
    assert halted_programs > 0
    assert sim.stats()["lane_cycles"] == int(halt.sum())
