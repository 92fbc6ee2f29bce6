"""Workload designs named by the batch-simulation configurations.

* ``counter`` -- the README 8-bit counter (clk/rst/en, uint8 count, out=count),
  the design of crates/rrtl-core/src/lib.rs:6818-6834.
* ``ready_valid_fifo`` -- a 16-entry, 32-bit ready/valid FIFO backed by a memory.
* ``crc32_rtl`` -- the word-level byte-parallel CRC-32 accumulator of
  bench/sv/crc32.sv (reset to 0xFFFFFFFF, 8 serial LFSR steps per byte).
* ``synthetic_netlist`` -- the levelized random gate netlist (see its docstring).
"""
from __future__ import annotations

from typing import Dict, List, Tuple

from . import ir
from .ir import lit, mux, uint


def counter(width: int = 8) -> ir.Module:
    m = ir.Module("Counter")
    clk = m.input("clk", 1)
    rst = m.input("rst", 1)
    en = m.input("en", 1)
    out = m.output("out", width)
    count = m.reg("count", width)
    m.clock(count, clk)
    m.reset(count, rst, 0)
    m.next(count, mux(en.value(), count + lit(1, width), count.value()))
    m.assign(out, count)
    return m


def ready_valid_fifo(depth: int = 16, width: int = 32) -> ir.Module:
    """Synchronous ready/valid FIFO: ``in_valid``/``in_ready``/``in_data`` push,
    ``out_valid``/``out_ready``/``out_data`` pop; storage is a ``depth``-entry memory."""
    assert depth & (depth - 1) == 0, "depth must be a power of two"
    aw = max(1, (depth - 1).bit_length())
    m = ir.Module("Fifo")
    clk = m.input("clk", 1)
    rst = m.input("rst", 1)
    in_valid = m.input("in_valid", 1)
    in_data = m.input("in_data", width)
    out_ready = m.input("out_ready", 1)
    in_ready = m.output("in_ready", 1)
    out_valid = m.output("out_valid", 1)
    out_data = m.output("out_data", width)
    count_o = m.output("count", aw + 1)
    mem = m.mem("storage", width, depth)
    wptr = m.reg("wptr", aw)
    rptr = m.reg("rptr", aw)
    cnt = m.reg("cnt", aw + 1)
    push = m.wire("push", 1)
    pop = m.wire("pop", 1)
    for r in (wptr, rptr, cnt):
        m.clock(r, clk)
        m.reset(r, rst, 0)
    full = cnt.value().eq(lit(depth, aw + 1))
    empty = cnt.value().eq(lit(0, aw + 1))
    m.assign(in_ready, ~full)
    m.assign(out_valid, ~empty)
    m.assign(push, in_valid.value() & ~full)
    m.assign(pop, out_ready.value() & ~empty)
    m.assign(out_data, mem.read(rptr))
    m.assign(count_o, cnt)
    one_a = lit(1, aw)
    m.next(wptr, mux(push.value(), wptr + one_a, wptr.value()))
    m.next(rptr, mux(pop.value(), rptr + one_a, rptr.value()))
    one_c = lit(1, aw + 1)
    inc = cnt + one_c
    dec = cnt - one_c
    m.next(cnt, mux(push.value() & ~pop.value(), inc,
                    mux(pop.value() & ~push.value(), dec, cnt.value())))
    m.mem_write(mem, push, wptr, in_data)
    return m


CRC32_POLY = 0x04C11DB7


def crc32_rtl() -> ir.Module:
    """bench/sv/crc32.sv as word-level RTL: c = crc; 8 x { c = (c<<1) ^ ({32{c[31]^din[i]}} & POLY) }."""
    m = ir.Module("crc32")
    clk = m.input("clk", 1)
    rst = m.input("rst", 1)
    din = m.input("din", 8)
    crc = m.reg("crc", 32)
    m.clock(crc, clk)
    m.reset(crc, rst, 0xFFFFFFFF)
    c = crc.value()
    for i in range(8):
        fb = c.bit(31) ^ din.value().bit(i)
        shifted = ir.concat(c.slice(0, 31), lit(0, 1))
        sel = mux(fb, lit(CRC32_POLY, 32), lit(0, 32))
        c = shifted ^ sel
    m.next(crc, c)
    out = m.output("crc_out", 32)
    m.assign(out, crc)
    return m


def crc32_step(crc: int, byte: int) -> int:
    """Closed form of one clock of bench/sv/crc32.sv (non-reflected MSB-first CRC-32)."""
    c = crc
    for i in range(8):
        fb = ((c >> 31) ^ (byte >> i)) & 1
        c = ((c << 1) & 0xFFFFFFFF) ^ (CRC32_POLY if fb else 0)
    return c


# ----------------------------------------------------------- gate netlists
class SplitMix64:
    """Deterministic generator (Steele et al., OOPSLA 2014) so netlists do not
    depend on numpy's stream stability."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def below(self, n: int) -> int:
        return (self.next() >> 11) * n >> 53


def synthetic_netlist(gates: int = 20000, levels: int = 64, dffs: int = 2000, inputs: int = 256,
                      outputs: int = 128, seed: int = 1,
                      mix=(("and", 40), ("or", 30), ("xor", 20), ("nand", 10))) -> ir.Module:
    """Levelized random gate netlist (the 20k-gate workload at its defaults).

    * ``inputs`` primary inputs ``pi{i}`` and ``dffs`` flops ``q{i}`` (no reset)
      form level 0;
    * ``gates`` two-input gates ``g{k}`` spread evenly over ``levels`` levels,
      op drawn from ``mix`` (percent weights: AND/OR/XOR/NAND 40/30/20/10);
      fanin A is drawn uniformly from the previous level (so every level is
      exactly one deeper), fanin B uniformly from everything earlier (all lower
      levels, flops and inputs);
    * each flop's D is a uniformly drawn gate; each output ``po{j}`` reads a
      uniformly drawn gate of the last eight levels.
    All draws come from ``SplitMix64(seed)``.
    """
    rng = SplitMix64(seed)
    m = ir.Module("SynthNetlist")
    pis = [m.input(f"pi{i}", 1) for i in range(inputs)]
    qs = [m.reg(f"q{i}", 1) for i in range(dffs)]
    level0 = pis + qs
    per = [gates // levels + (1 if l < gates % levels else 0) for l in range(levels)]
    table = []
    for name, w in mix:
        table += [name] * w
    earlier: List[ir.Signal] = list(level0)
    prev: List[ir.Signal] = list(level0)
    all_gates: List[ir.Signal] = []
    last_levels: List[ir.Signal] = []
    k = 0
    for l in range(levels):
        cur = []
        for _ in range(per[l]):
            op = table[rng.below(len(table))]
            a = prev[rng.below(len(prev))].value()
            b = earlier[rng.below(len(earlier))].value()
            g = m.wire(f"g{k}", 1)
            k += 1
            if op == "and":
                e = a & b
            elif op == "or":
                e = a | b
            elif op == "xor":
                e = a ^ b
            else:
                e = ~(a & b)
            m.assign(g, e)
            cur.append(g)
        earlier.extend(cur)
        all_gates.extend(cur)
        if l >= levels - 8:
            last_levels.extend(cur)
        prev = cur
    for q in qs:
        m.next(q, all_gates[rng.below(len(all_gates))])
    for j in range(outputs):
        o = m.output(f"po{j}", 1)
        m.assign(o, last_levels[rng.below(len(last_levels))])
    return m


def gate_chain_design(depth: int = 16, cones: int = 64) -> ir.Module:
    """NAND/NOR/XNOR cones of bench ``bitparallel_bench.rs:10-33``: one flop per cone
    fed back through ``depth`` gates mixing inputs ``a``/``b``."""
    m = ir.Module("Gates")
    clk = m.input("clk", 1)
    a = m.input("a", 1)
    b = m.input("b", 1)
    for c in range(cones):
        acc = m.reg(f"g{c}", 1)
        m.clock(acc, clk)
        x = acc.value()
        for d in range(depth):
            r = (c + d) % 3
            if r == 0:
                x = ~(x & a.value())
            elif r == 1:
                x = ~(x | b.value())
            else:
                x = ~(x ^ a.value())
        m.next(acc, x)
        o = m.output(f"o{c}", 1)
        m.assign(o, acc)
    return m
