"""An RV32I core as a word-level design, its assembler and its random programs.

The per-lane CPU workload of the reference is picorv32 running per-lane
programs, ``ecall`` halting through ``trap`` (bench/sv/picorv32.v,
bench/sv/picorv32_soc.v with a 1024-word memory; crates/rrtl-sim-ir/examples/
picorv32_batch_bench.rs, crates/rrtl-gpu-sim/examples/picorv32_gpu.rs).  The
reference reaches picorv32 through its SystemVerilog frontend, which this
package does not have; ``rv32i_core`` is a single-cycle RV32I core written
directly in the design IR with the same memory map (one unified 4 KiB word
memory per lane, code from address 0) and the same halting rule (``ecall`` or
``ebreak`` raises ``trap`` and the core stops).  Its architectural results are
pinned by ``Iss``, an instruction-set simulator written from the RV32I base
ISA specification, one instruction per cycle like the core.

Memory addresses are taken modulo 4 KiB (``addr[11:2]`` selects the word),
in the core and in ``Iss`` alike.
"""
from __future__ import annotations

import random
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

from . import ir
from .ir import concat, lit, mux

XLEN = 32
MEM_WORDS = 1024

# opcodes (RV32I base)
OP_LUI, OP_AUIPC, OP_JAL, OP_JALR = 0x37, 0x17, 0x6F, 0x67
OP_BRANCH, OP_LOAD, OP_STORE, OP_IMM, OP_REG, OP_FENCE, OP_SYSTEM = 0x63, 0x03, 0x23, 0x13, 0x33, 0x0F, 0x73


# ---------------------------------------------------------------- the core
def _shift(v: ir.Expr, sh: ir.Expr, kind: str) -> ir.Expr:
    """Barrel shifter over 5 stages; kind in {'sll', 'srl', 'sra'}."""
    for k in range(5):
        n = 1 << k
        if kind == "sll":
            moved = concat(v.slice(0, XLEN - n), lit(0, n))
        else:
            fill = v.bit(XLEN - 1).sext(n) if kind == "sra" else lit(0, n)
            moved = concat(fill, v.slice(n, XLEN - n))
        v = mux(sh.bit(k), moved, v)
    return v


def rv32i_core(mem_words: int = MEM_WORDS) -> ir.Module:
    """Single-cycle RV32I core with a unified per-lane word memory.

    Ports: ``clk``, ``rst`` (synchronous: pc <- 0, trap <- 0); outputs ``trap``
    (1 after ``ecall``/``ebreak``; the core then holds), ``pc`` and ``a0``
    (x10).  Memories: ``mem`` (code and data, ``mem_words`` x 32) and ``regs``
    (x0..x31; x0 reads as 0 and is never written)."""
    assert mem_words & (mem_words - 1) == 0
    aw = (mem_words - 1).bit_length()
    m = ir.Module("Rv32i")
    clk = m.input("clk", 1)
    rst = m.input("rst", 1)
    trap_o = m.output("trap", 1)
    pc_o = m.output("pc", XLEN)
    a0_o = m.output("a0", XLEN)
    mem = m.mem("mem", XLEN, mem_words)
    regs = m.mem("regs", XLEN, 32)
    pc = m.reg("pc_r", XLEN)
    trap = m.reg("trap_r", 1)
    for r in (pc, trap):
        m.clock(r, clk)
        m.reset(r, rst, 0)

    def w(name, width, expr):
        s = m.wire(name, width)
        m.assign(s, expr)
        return s.value()

    insn = w("insn", XLEN, mem.read(pc.value().slice(2, aw)))
    opcode = insn.slice(0, 7)
    rd = insn.slice(7, 5)
    f3 = insn.slice(12, 3)
    rs1 = insn.slice(15, 5)
    rs2 = insn.slice(20, 5)
    f7b5 = insn.bit(30)
    zero32 = lit(0, XLEN)
    x1 = w("rs1_val", XLEN, mux(rs1.eq(lit(0, 5)), zero32, regs.read(rs1)))
    x2 = w("rs2_val", XLEN, mux(rs2.eq(lit(0, 5)), zero32, regs.read(rs2)))
    imm_i = insn.slice(20, 12).sext(XLEN)
    imm_s = concat(insn.slice(25, 7), insn.slice(7, 5)).sext(XLEN)
    imm_b = concat(insn.bit(31), insn.bit(7), insn.slice(25, 6), insn.slice(8, 4), lit(0, 1)).sext(XLEN)
    imm_u = concat(insn.slice(12, 20), lit(0, 12))
    imm_j = concat(insn.bit(31), insn.slice(12, 8), insn.bit(20), insn.slice(21, 10), lit(0, 1)).sext(XLEN)

    def is_op(code):
        return opcode.eq(lit(code, 7))

    is_lui, is_auipc, is_jal, is_jalr = is_op(OP_LUI), is_op(OP_AUIPC), is_op(OP_JAL), is_op(OP_JALR)
    is_branch, is_load, is_store = is_op(OP_BRANCH), is_op(OP_LOAD), is_op(OP_STORE)
    is_imm, is_reg, is_system = is_op(OP_IMM), is_op(OP_REG), is_op(OP_SYSTEM)

    # ALU: operand b is the immediate for OP-IMM, rs2 for OP
    b = w("alu_b", XLEN, mux(is_imm, imm_i, x2))
    sh = b.slice(0, 5)
    sub = is_reg & f7b5
    add_sub = mux(sub, x1 - b, x1 + b)
    slt = (x1.cast(True).lt(b.cast(True))).zext(XLEN)
    sltu = x1.lt(b).zext(XLEN)
    srl_sra = mux(f7b5, _shift(x1, sh, "sra"), _shift(x1, sh, "srl"))

    def f3is(v):
        return f3.eq(lit(v, 3))

    alu = w("alu", XLEN,
            mux(f3is(0), add_sub,
                mux(f3is(1), _shift(x1, sh, "sll"),
                    mux(f3is(2), slt,
                        mux(f3is(3), sltu,
                            mux(f3is(4), x1 ^ b,
                                mux(f3is(5), srl_sra,
                                    mux(f3is(6), x1 | b, x1 & b))))))))
    # branches
    eq = x1.eq(x2)
    lt_s = x1.cast(True).lt(x2.cast(True))
    lt_u = x1.lt(x2)
    take = w("take", 1,
             mux(f3is(0), eq, mux(f3is(1), ~eq, mux(f3is(4), lt_s, mux(f3is(5), ~lt_s, mux(f3is(6), lt_u, ~lt_u))))))
    # memory: the word at addr[aw+1:2]; sub-word access by addr[1:0]
    addr = w("addr", XLEN, x1 + mux(is_store, imm_s, imm_i))
    waddr = addr.slice(2, aw)
    word = w("mem_word", XLEN, mem.read(waddr))
    boff = addr.slice(0, 2)
    byte = w("load_byte", 8,
             mux(boff.eq(lit(0, 2)), word.slice(0, 8),
                 mux(boff.eq(lit(1, 2)), word.slice(8, 8), mux(boff.eq(lit(2, 2)), word.slice(16, 8), word.slice(24, 8)))))
    half = w("load_half", 16, mux(addr.bit(1), word.slice(16, 16), word.slice(0, 16)))
    load_val = w("load_val", XLEN,
                 mux(f3is(0), byte.sext(XLEN),
                     mux(f3is(1), half.sext(XLEN),
                         mux(f3is(4), byte.zext(XLEN), mux(f3is(5), half.zext(XLEN), word)))))
    sb = x2.slice(0, 8)
    sb_word = mux(boff.eq(lit(0, 2)), concat(word.slice(8, 24), sb),
                  mux(boff.eq(lit(1, 2)), concat(word.slice(16, 16), sb, word.slice(0, 8)),
                      mux(boff.eq(lit(2, 2)), concat(word.slice(24, 8), sb, word.slice(0, 16)),
                          concat(sb, word.slice(0, 24)))))
    sh_word = mux(addr.bit(1), concat(x2.slice(0, 16), word.slice(0, 16)), concat(word.slice(16, 16), x2.slice(0, 16)))
    store_val = w("store_val", XLEN, mux(f3is(0), sb_word, mux(f3is(1), sh_word, x2)))
    # next pc and write-back
    run = w("run", 1, ~trap.value())
    pc4 = pc + lit(4, XLEN)
    jalr_t = concat((x1 + imm_i).slice(1, XLEN - 1), lit(0, 1))
    next_pc = w("next_pc", XLEN,
                mux(is_jal, pc + imm_j, mux(is_jalr, jalr_t, mux(is_branch & take, pc + imm_b, pc4))))
    wb = w("wb", XLEN,
           mux(is_lui, imm_u, mux(is_auipc, pc + imm_u, mux(is_jal | is_jalr, pc4, mux(is_load, load_val, alu)))))
    writes_rd = is_lui | is_auipc | is_jal | is_jalr | is_load | is_imm | is_reg
    halt_now = w("halt_now", 1, is_system)   # ecall / ebreak
    m.next(pc, mux(run & ~halt_now, next_pc, pc.value()))
    m.next(trap, trap.value() | halt_now)
    m.mem_write(regs, run & writes_rd & rd.ne(lit(0, 5)), rd, wb)
    m.mem_write(mem, run & is_store, waddr, store_val)
    m.assign(trap_o, trap)
    m.assign(pc_o, pc)
    m.assign(a0_o, regs.read(lit(10, 5)))
    return m


# ---------------------------------------------------------------- assembler
def _r(f7, rs2, rs1, f3, rd, op):
    return (f7 << 25) | (rs2 << 20) | (rs1 << 15) | (f3 << 12) | (rd << 7) | op


def _i(imm, rs1, f3, rd, op):
    return ((imm & 0xFFF) << 20) | (rs1 << 15) | (f3 << 12) | (rd << 7) | op


def _s(imm, rs2, rs1, f3, op):
    imm &= 0xFFF
    return ((imm >> 5) << 25) | (rs2 << 20) | (rs1 << 15) | (f3 << 12) | ((imm & 0x1F) << 7) | op


def _b(imm, rs2, rs1, f3):
    imm &= 0x1FFF
    return (((imm >> 12) & 1) << 31) | (((imm >> 5) & 0x3F) << 25) | (rs2 << 20) | (rs1 << 15) | (f3 << 12) | \
        (((imm >> 1) & 0xF) << 8) | (((imm >> 11) & 1) << 7) | OP_BRANCH


def _u(imm20, rd, op):
    return ((imm20 & 0xFFFFF) << 12) | (rd << 7) | op


def _j(imm, rd):
    imm &= 0x1FFFFF
    return (((imm >> 20) & 1) << 31) | (((imm >> 1) & 0x3FF) << 21) | (((imm >> 11) & 1) << 20) | \
        (((imm >> 12) & 0xFF) << 12) | (rd << 7) | OP_JAL


ALU_R = {"add": (0, 0), "sub": (0x20, 0), "sll": (0, 1), "slt": (0, 2), "sltu": (0, 3), "xor": (0, 4),
         "srl": (0, 5), "sra": (0x20, 5), "or": (0, 6), "and": (0, 7)}
ALU_I = {"addi": 0, "slti": 2, "sltiu": 3, "xori": 4, "ori": 6, "andi": 7}
SHIFT_I = {"slli": (0, 1), "srli": (0, 5), "srai": (0x20, 5)}
LOADS = {"lb": 0, "lh": 1, "lw": 2, "lbu": 4, "lhu": 5}
STORES = {"sb": 0, "sh": 1, "sw": 2}
BRANCHES = {"beq": 0, "bne": 1, "blt": 4, "bge": 5, "bltu": 6, "bgeu": 7}


def asm(op: str, *a) -> int:
    """Encode one RV32I instruction: asm('addi', rd, rs1, imm), asm('lw', rd,
    rs1, imm), asm('sw', rs2, rs1, imm), asm('beq', rs1, rs2, offset),
    asm('jal', rd, offset), asm('jalr', rd, rs1, imm), asm('lui', rd, imm20),
    asm('ecall')."""
    if op in ALU_R:
        f7, f3 = ALU_R[op]
        rd, rs1, rs2 = a
        return _r(f7, rs2, rs1, f3, rd, OP_REG)
    if op in ALU_I:
        rd, rs1, imm = a
        return _i(imm, rs1, ALU_I[op], rd, OP_IMM)
    if op in SHIFT_I:
        f7, f3 = SHIFT_I[op]
        rd, rs1, sh = a
        return _r(f7, sh & 31, rs1, f3, rd, OP_IMM)
    if op in LOADS:
        rd, rs1, imm = a
        return _i(imm, rs1, LOADS[op], rd, OP_LOAD)
    if op in STORES:
        rs2, rs1, imm = a
        return _s(imm, rs2, rs1, STORES[op], OP_STORE)
    if op in BRANCHES:
        rs1, rs2, off = a
        return _b(off, rs2, rs1, BRANCHES[op])
    if op == "jal":
        rd, off = a
        return _j(off, rd)
    if op == "jalr":
        rd, rs1, imm = a
        return _i(imm, rs1, 0, rd, OP_JALR)
    if op == "lui":
        rd, imm20 = a
        return _u(imm20, rd, OP_LUI)
    if op == "auipc":
        rd, imm20 = a
        return _u(imm20, rd, OP_AUIPC)
    if op == "ecall":
        return OP_SYSTEM
    if op == "ebreak":
        return (1 << 20) | OP_SYSTEM
    raise ValueError(f"unknown instruction {op}")


# ---------------------------------------------------------------- programs
DATA_BASE = 0x800        # data region: words 512..1023
BASE_REG = 31            # holds DATA_BASE in random programs, never overwritten


def random_program(rng: random.Random, n: int = 512, alu: float = 0.60, mem: float = 0.25,
                   backward: Optional[float] = None, data_base: int = DATA_BASE,
                   data_span: int = 0x7FC) -> List[int]:
    """``n`` random RV32I instructions (ALU ``alu``, load/store ``mem``, branch
    the rest; in-range targets), preceded by setting the data base register
    and followed by ``ecall``.  Branch targets stay inside the program, so a
    program may loop until the cycle cap: lanes halt unevenly.  ``backward``
    (None: targets uniform over the program) is the probability that a branch
    goes backward; loads and stores address [data_base, data_base + data_span)."""
    lo = ((data_base & 0xFFF) ^ 0x800) - 0x800     # addi sign-extends its immediate
    prog = [asm("lui", BASE_REG, ((data_base - lo) >> 12) & 0xFFFFF), asm("addi", BASE_REG, BASE_REG, lo)]
    start = len(prog)
    total = start + n + 1
    regs = list(range(1, 31))
    for k in range(n):
        u = rng.random()
        pos = start + k
        if u < alu:
            rd, rs1, rs2 = rng.choice(regs), rng.choice(regs + [0]), rng.choice(regs + [0])
            kind = rng.random()
            if kind < 0.45:
                prog.append(asm(rng.choice(list(ALU_R)), rd, rs1, rs2))
            elif kind < 0.85:
                prog.append(asm(rng.choice(list(ALU_I)), rd, rs1, rng.randrange(-2048, 2048)))
            elif kind < 0.95:
                prog.append(asm(rng.choice(list(SHIFT_I)), rd, rs1, rng.randrange(32)))
            else:
                prog.append(asm(rng.choice(["lui", "auipc"]), rd, rng.randrange(1 << 20)))
        elif u < alu + mem:
            off = rng.randrange(0, data_span)
            if rng.random() < 0.5:
                op = rng.choice(list(LOADS))
                al = {"lb": 1, "lbu": 1, "lh": 2, "lhu": 2, "lw": 4}[op]
                prog.append(asm(op, rng.choice(regs), BASE_REG, off - off % al))
            else:
                op = rng.choice(list(STORES))
                al = {"sb": 1, "sh": 2, "sw": 4}[op]
                prog.append(asm(op, rng.choice(regs + [0]), BASE_REG, off - off % al))
        else:
            if backward is None:
                target = rng.randrange(start, total)
            elif rng.random() < backward and pos > start:
                target = rng.randrange(start, pos)
            else:
                target = rng.randrange(pos + 1, total)
            prog.append(asm(rng.choice(list(BRANCHES)), rng.choice(regs + [0]), rng.choice(regs + [0]),
                            (target - pos) * 4))
    prog.append(asm("ecall"))
    return prog


def sum_program(limit: int) -> List[int]:
    """sum 1..limit-1 into a0 and store it at 0x100, then ecall (the
    architectural closed form of the reference's picorv32 example)."""
    return [
        asm("addi", 10, 0, 0),          # a0 = 0
        asm("addi", 11, 0, 1),          # i = 1
        asm("addi", 12, 0, limit),      # n
        asm("bge", 11, 12, 16),         # loop: if i >= n goto done
        asm("add", 10, 10, 11),
        asm("addi", 11, 11, 1),
        asm("jal", 0, -12),
        asm("sw", 10, 0, 0x100),        # done
        asm("ecall"),
    ]


# ---------------------------------------------------------------- ISS
def _sx(v: int, bits: int) -> int:
    v &= (1 << bits) - 1
    return v - (1 << bits) if v >> (bits - 1) else v


@dataclass
class Iss:
    """RV32I instruction-set simulator, one instruction per cycle, with the
    core's memory map (word memory, addresses modulo its size)."""
    mem: List[int]
    regs: List[int] = field(default_factory=lambda: [0] * 32)
    pc: int = 0
    trap: bool = False
    cycles: int = 0

    def step(self) -> None:
        if self.trap:
            return
        M = len(self.mem)
        insn = self.mem[(self.pc >> 2) % M]
        op, rd, f3 = insn & 0x7F, (insn >> 7) & 31, (insn >> 12) & 7
        rs1, rs2, f7 = (insn >> 15) & 31, (insn >> 20) & 31, insn >> 25
        x1, x2 = self.regs[rs1], self.regs[rs2]
        imm_i = _sx(insn >> 20, 12)
        imm_s = _sx(((insn >> 25) << 5) | ((insn >> 7) & 31), 12)
        imm_b = _sx((((insn >> 31) & 1) << 12) | (((insn >> 7) & 1) << 11) | (((insn >> 25) & 0x3F) << 5) |
                    (((insn >> 8) & 0xF) << 1), 13)
        imm_u = insn & 0xFFFFF000
        imm_j = _sx((((insn >> 31) & 1) << 20) | (((insn >> 12) & 0xFF) << 12) | (((insn >> 20) & 1) << 11) |
                    (((insn >> 21) & 0x3FF) << 1), 21)
        pc4 = (self.pc + 4) & 0xFFFFFFFF
        npc, wb = pc4, None
        if op == OP_LUI:
            wb = imm_u
        elif op == OP_AUIPC:
            wb = (self.pc + imm_u) & 0xFFFFFFFF
        elif op == OP_JAL:
            wb, npc = pc4, (self.pc + imm_j) & 0xFFFFFFFF
        elif op == OP_JALR:
            wb, npc = pc4, (x1 + imm_i) & 0xFFFFFFFE
        elif op == OP_BRANCH:
            s1, s2 = _sx(x1, 32), _sx(x2, 32)
            take = {0: x1 == x2, 1: x1 != x2, 4: s1 < s2, 5: s1 >= s2, 6: x1 < x2, 7: x1 >= x2}.get(f3, x1 >= x2)
            if take:
                npc = (self.pc + imm_b) & 0xFFFFFFFF
        elif op == OP_LOAD:
            a = (x1 + imm_i) & 0xFFFFFFFF
            word = self.mem[(a >> 2) % M]
            byte = (word >> (8 * (a & 3))) & 0xFF
            half = (word >> (16 * ((a >> 1) & 1))) & 0xFFFF
            wb = {0: _sx(byte, 8) & 0xFFFFFFFF, 1: _sx(half, 16) & 0xFFFFFFFF, 4: byte, 5: half}.get(f3, word)
        elif op == OP_STORE:
            a = (x1 + imm_s) & 0xFFFFFFFF
            i = (a >> 2) % M
            word = self.mem[i]
            if f3 == 0:
                sh = 8 * (a & 3)
                word = (word & ~(0xFF << sh) & 0xFFFFFFFF) | ((x2 & 0xFF) << sh)
            elif f3 == 1:
                sh = 16 * ((a >> 1) & 1)
                word = (word & ~(0xFFFF << sh) & 0xFFFFFFFF) | ((x2 & 0xFFFF) << sh)
            else:
                word = x2
            self.mem[i] = word
        elif op in (OP_IMM, OP_REG):
            b = imm_i & 0xFFFFFFFF if op == OP_IMM else x2
            sh = b & 31
            if f3 == 0:
                wb = (x1 - b) if (op == OP_REG and (f7 >> 5) & 1) else (x1 + b)
            elif f3 == 1:
                wb = x1 << sh
            elif f3 == 2:
                wb = int(_sx(x1, 32) < _sx(b, 32))
            elif f3 == 3:
                wb = int(x1 < b)
            elif f3 == 4:
                wb = x1 ^ b
            elif f3 == 5:
                wb = (_sx(x1, 32) >> sh) if (f7 >> 5) & 1 else (x1 >> sh)
            elif f3 == 6:
                wb = x1 | b
            else:
                wb = x1 & b
            wb &= 0xFFFFFFFF
        elif op == OP_SYSTEM:
            self.trap = True
            self.cycles += 1
            return
        if wb is not None and rd != 0:
            self.regs[rd] = wb & 0xFFFFFFFF
        self.pc = npc
        self.cycles += 1

    def run(self, max_cycles: int) -> None:
        while not self.trap and self.cycles < max_cycles:
            self.step()


def load_image(prog: Sequence[int], words: int = MEM_WORDS) -> List[int]:
    img = [0] * words
    for i, v in enumerate(prog):
        img[i] = v & 0xFFFFFFFF
    return img
