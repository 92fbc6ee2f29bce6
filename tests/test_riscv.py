"""RV32I core workload: encoder known answers, the ISS against closed forms,
and the core (lowered, on the packed oracle) against the ISS cycle by cycle."""
import random

import numpy as np
import pytest

import oracle
from rrtl_amd import lower, riscv
from rrtl_amd.riscv import Iss, asm, load_image


def test_encodings_known_answers():
    # standard RV32I encodings
    assert asm("addi", 1, 0, 5) == 0x00500093
    assert asm("lui", 5, 0x12345) == 0x123452B7
    assert asm("jal", 1, 8) == 0x008000EF
    assert asm("sw", 2, 3, 4) == 0x0021A223
    assert asm("beq", 1, 2, 8) == 0x00208463
    assert asm("ecall") == 0x00000073
    assert asm("sub", 3, 1, 2) == 0x402081B3
    assert asm("srai", 4, 5, 3) == 0x4032D213
    assert asm("lbu", 6, 7, -1) == 0xFFF3C303


def test_iss_sum_closed_form():
    for limit in (1, 2, 6, 50):
        s = Iss(load_image(riscv.sum_program(limit)))
        s.run(10_000)
        assert s.trap
        assert s.regs[10] == sum(range(1, limit))
        assert s.mem[0x100 >> 2] == sum(range(1, limit))


def test_iss_instruction_semantics():
    prog = [asm("addi", 1, 0, -1), asm("srli", 2, 1, 28), asm("srai", 3, 1, 28), asm("slti", 4, 1, 0),
            asm("sltiu", 5, 1, 0), asm("lui", 6, 0x80000), asm("slt", 7, 6, 0), asm("sltu", 8, 6, 0),
            asm("sb", 1, 0, 0x201), asm("lb", 9, 0, 0x201), asm("lbu", 10, 0, 0x201), asm("lhu", 11, 0, 0x200),
            asm("auipc", 12, 1), asm("ecall")]
    s = Iss(load_image(prog))
    s.run(100)
    r = s.regs
    assert r[1] == 0xFFFFFFFF and r[2] == 0xF and r[3] == 0xFFFFFFFF
    assert r[4] == 1 and r[5] == 0 and r[6] == 0x80000000 and r[7] == 1 and r[8] == 0
    assert r[9] == 0xFFFFFFFF and r[10] == 0xFF and r[11] == 0xFF00 and r[12] == 48 + 0x1000
    assert s.cycles == len(prog)


def _core_vs_iss(lanes, cycles, seed):
    p = lower.lower_module(riscv.rv32i_core())
    rng = random.Random(seed)
    progs = [riscv.random_program(rng, n=48) for _ in range(lanes)]
    ref = oracle.PackedRef(p, lanes)
    mem_idx = [m.name for m in p.memories].index("mem")
    img = np.array([load_image(pr) for pr in progs], dtype=np.uint64)
    # lane-major memory limbs: [lanes][words][limbs]
    ref.set_memory_limbs(mem_idx, img.astype(np.uint32).reshape(lanes, -1, 1))
    ref.set_signal(p.signal_index("clk"), [1] * lanes)
    ref.set_signal(p.signal_index("rst"), [0] * lanes)
    iss = [Iss(load_image(pr)) for pr in progs]
    pc_i, trap_i, a0_i = p.signal_index("pc"), p.signal_index("trap"), p.signal_index("a0")
    for c in range(cycles):
        ref.eval_combinational()
        pcs, traps, a0s = ref.get_signal(pc_i), ref.get_signal(trap_i), ref.get_signal(a0_i)
        for l in range(lanes):
            assert pcs[l] == iss[l].pc and traps[l] == int(iss[l].trap) and a0s[l] == iss[l].regs[10], (l, c)
        ref.tick_from_evaluated()
        for s in iss:
            s.step()
    regs = ref.get_memory_limbs([m.name for m in p.memories].index("regs"))
    mem = ref.get_memory_limbs(mem_idx)
    return regs, mem, iss


def test_core_matches_iss_on_packed_oracle():
    regs, mem, iss = _core_vs_iss(lanes=6, cycles=160, seed=3)
    for l, s in enumerate(iss):
        assert [int(x) for x in regs[l, :, 0]] == s.regs, l
        assert [int(x) for x in mem[l, :, 0]] == s.mem, l


def test_host_simd_engine_matches_iss():
    """The host batch-SIMD restatement (the RV32I CPU baseline) on random
    programs, lane-sharded over threads: registers and memory equal the ISS's
    after the same number of cycles (the core holds after trap)."""
    from rrtl_amd.backend import IN_CONST
    p = lower.lower_module(riscv.rv32i_core())
    lanes, cycles = 37, 150
    rng = random.Random(9)
    progs = [riscv.random_program(rng, n=64) for _ in range(lanes)]
    cpu = oracle.SimdCpu(p, lanes, threads=3, shard_lanes=8)
    img = np.array([load_image(pr) for pr in progs], dtype=np.uint32)
    cpu.set_memory_limbs(p.memory_index("mem"), img.reshape(lanes, -1, 1))
    ins = [(p.signal_index("clk"), IN_CONST, 1), (p.signal_index("rst"), IN_CONST, 0)]
    cpu.run(1, cycles, ins, [])
    regs = cpu.get_memory_limbs(p.memory_index("regs"))
    mem = cpu.get_memory_limbs(p.memory_index("mem"))
    for l, pr in enumerate(progs):
        s = Iss(load_image(pr))
        s.run(cycles)
        assert [int(x) for x in regs[l, :, 0]] == s.regs, l
        assert [int(x) for x in mem[l, :, 0]] == s.mem, l
