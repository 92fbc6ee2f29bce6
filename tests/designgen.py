"""Random typed RTL designs for differential tests (engines vs the oracles).

Every design uses the whole expression set of the reference's ``Expr``
(crates/rrtl-ir/src/lib.rs:230-280) at widths that exercise one-limb, limb-
boundary and multi-limb paths, registers with sync and async, active-high and
active-low resets, and optionally a memory with reads and writes.
"""
from __future__ import annotations

import random
from typing import List

from rrtl_amd import ir
from rrtl_amd.ir import BitType, concat, lit, mux, uint

WIDTHS = [1, 3, 8, 16, 31, 32, 33, 63, 64, 65, 96, 127, 128]


def random_design(seed: int, n_inputs: int = 4, n_regs: int = 4, n_wires: int = 8, depth: int = 3,
                  widths=WIDTHS, with_memory: bool = True, async_resets: bool = True,
                  one_limb_memory: bool = False) -> ir.Module:
    """``one_limb_memory``: the memory's words and read addresses fit one limb
    (by default a 40-bit memory and 33-bit addresses, with the saturating
    usize decode, are possible)."""
    rng = random.Random(seed)
    addr_widths = [2, 4, 5] if one_limb_memory else [2, 4, 5, 33]
    m = ir.Module(f"Rand{seed}")
    rst = m.input("rst", 1)
    rstn = m.input("rstn", 1)
    ins = [m.input(f"in{i}", BitType(rng.choice(widths), rng.random() < 0.3)) for i in range(n_inputs)]
    regs = [m.reg(f"r{i}", BitType(rng.choice(widths), rng.random() < 0.3)) for i in range(n_regs)]
    mem = None
    if with_memory:
        mem = m.mem("ram", rng.choice([8, 32] if one_limb_memory else [8, 32, 40]), rng.choice([4, 16]))
    pool: List[ir.Expr] = [s.value() for s in ins + regs]

    def leaf(ty: BitType) -> ir.Expr:
        cands = [e for e in pool if e.ty == ty]
        if cands and rng.random() < 0.8:
            return rng.choice(cands)
        if rng.random() < 0.6:
            return fit(rng.choice(pool), ty)
        return lit(rng.getrandbits(ty.width), ty.width, ty.signed)

    def fit(e: ir.Expr, ty: BitType) -> ir.Expr:
        if e.ty.width > ty.width:
            e = e.trunc(ty.width) if rng.random() < 0.5 else e.slice(rng.randrange(e.ty.width - ty.width + 1), ty.width)
        elif e.ty.width < ty.width:
            e = e.sext(ty.width) if rng.random() < 0.5 else e.zext(ty.width)
        if e.ty.signed != ty.signed:
            e = e.cast(ty.signed)
        return e

    def gen(ty: BitType, d: int) -> ir.Expr:
        if d == 0:
            return leaf(ty)
        k = rng.randrange(13)
        if k == 0:
            return ~gen(ty, d - 1)
        if k in (1, 2, 3):
            a, b = gen(ty, d - 1), gen(ty, d - 1)
            return [a & b, a | b, a ^ b][k - 1]
        if k in (4, 5, 6):
            a, b = gen(ty, d - 1), gen(ty, d - 1)
            return [a + b, a - b, a * b][k - 4]
        if k == 7:
            ot = BitType(rng.choice(widths), rng.random() < 0.5)
            a, b = gen(ot, d - 1), gen(ot, d - 1)
            c = [a.eq(b), a.ne(b), a.lt(b)][rng.randrange(3)]
            return fit(c, ty)
        if k == 8:
            c = fit(gen(uint(rng.choice([1, 8])), d - 1), uint(1)) if rng.random() < 0.5 else gen(uint(1), d - 1)
            if c.ty != uint(1):
                c = c.bit(0)
            return mux(c, gen(ty, d - 1), gen(ty, d - 1))
        if k == 9 and ty.width >= 2:
            w1 = rng.randrange(1, ty.width)
            return fit(concat(gen(uint(w1), d - 1), gen(uint(ty.width - w1), d - 1)), ty)
        if k == 10:
            src = gen(BitType(rng.choice(widths), False), d - 1)
            return fit(src, ty)
        if k == 11 and mem is not None:
            a = gen(uint(rng.choice(addr_widths)), d - 1)
            return fit(mem.read(a), ty)
        return leaf(ty)

    wires = []
    for i in range(n_wires):
        ty = BitType(rng.choice(widths), rng.random() < 0.3)
        w = m.wire(f"w{i}", ty)
        m.assign(w, gen(ty, depth))
        wires.append(w)
        pool.append(w.value())
    for i, r in enumerate(regs):
        m.next(r, gen(r.ty, depth))
        kind = rng.choice(["none", "sync", "async"] if async_resets else ["none", "sync"])
        if kind != "none":
            low = rng.random() < 0.5
            m.reset(r, rstn if low else rst, rng.getrandbits(r.width), kind, "low" if low else "high")
    if mem is not None:
        en = gen(uint(1), 1)
        m.mem_write(mem, en, gen(uint(rng.choice([2, 4, 5])), 1), gen(mem.ty, 2))
    for i in range(2):
        ty = BitType(rng.choice(widths), False)
        o = m.output(f"out{i}", ty)
        m.assign(o, gen(ty, depth))
    return m


def random_gate_design(seed: int, n_inputs: int = 24, n_regs: int = 40, n_wires: int = 300,
                       observe_regs: bool = True, async_resets: bool = False) -> ir.Module:
    """Random 1-bit netlist for the gate engine's tile compiler: registers with
    sync (or async) resets of both polarities and without; register-to-register
    shift chains (a next value that is another register is staged before the
    commit); next values drawn from early and late wires (registers read only
    early lend their slot; a late next value may land in its own register's
    slot); outputs that are registers (observed register slots keep the
    observation, commit and inputs in separate barrier phases) or wires."""
    rng = random.Random(seed)
    m = ir.Module(f"RandGates{seed}")
    clk = m.input("clk", 1)
    rst = m.input("rst", 1)
    rstn = m.input("rstn", 1)
    ins = [m.input(f"d{i}", 1) for i in range(n_inputs)]
    regs = [m.reg(f"q{i}", 1) for i in range(n_regs)]
    for q in regs:
        m.clock(q, clk)
    pool: List[ir.Expr] = [s.value() for s in ins + regs]
    wires = []
    for w in range(n_wires):
        # mostly recent values (depth), sometimes anything earlier
        def pick():
            if rng.random() < 0.6 and len(pool) > 8:
                return pool[-1 - rng.randrange(min(len(pool), 24))]
            return rng.choice(pool)
        a, b = pick(), pick()
        k = rng.randrange(8)
        e = [a & b, a | b, a ^ b, ~(a & b), ~(a | b), ~(a ^ b), ~a, mux(pick(), a, b)][k]
        s = m.wire(f"w{w}", 1)
        m.assign(s, e)
        wires.append(s)
        pool.append(s.value())
    for i, q in enumerate(regs):
        r = rng.random()
        if r < 0.15 and i > 0:
            m.next(q, regs[rng.randrange(n_regs)].value())   # shift chain (possibly itself)
        elif r < 0.55:
            m.next(q, wires[rng.randrange(n_wires // 4)].value())   # early next value
        else:
            m.next(q, wires[n_wires - 1 - rng.randrange(n_wires // 4)].value())   # late next value
        kind = "async" if async_resets and rng.random() < 0.3 else "sync"
        r = rng.random()
        if r < 0.3:
            m.reset(q, rst, rng.randrange(2), kind, "high")
        elif r < 0.5:
            m.reset(q, rstn, rng.randrange(2), kind, "low")
    for j in range(12):
        o = m.output(f"o{j}", 1)
        src = regs[rng.randrange(n_regs)] if observe_regs and j % 3 == 0 else wires[n_wires - 1 - rng.randrange(40)]
        m.assign(o, src.value())
    return m


def fusion_design(width: int = 8) -> ir.Module:
    """Every fusable pattern of the word engine's record encoder at one width
    (NOT over AND/OR/XOR, over OR-of-AND and AND-of-OR with one or two inner
    cells; OR(AND), AND(OR), XOR(XOR), ADD(ADD), ADD(MUL) with the inner value
    on either side, SUB(MUL) with the product as minuend), each on
    inputs of its own so that the lowering's shared subexpressions do not give
    an inner value a second reader, plus cases that must stay unfused: an inner
    value with two readers, an inner value of another width and a stored inner
    value."""
    m = ir.Module(f"Fuse{width}")
    clk = m.input("clk", 1)
    k = [0]

    def trio():
        k[0] += 1
        return tuple(m.input(f"{n}{k[0]}", width).value() for n in "abc")

    exprs = {}
    A, B, _ = trio(); exprs["nand"] = ~(A & B)
    A, B, _ = trio(); exprs["nor"] = ~(A | B)
    A, B, _ = trio(); exprs["xnor"] = ~(A ^ B)
    A, B, C = trio(); exprs["aoi"] = ~((A & B) | C)
    A, B, C = trio(); exprs["oai"] = ~((A | B) & C)
    A, B, C = trio(); exprs["andor"] = (A & B) | C
    A, B, C = trio(); exprs["orand"] = C & (A | B)
    A, B, C = trio(); exprs["xor3"] = C ^ (A ^ B)
    A, B, C = trio(); exprs["add3"] = (A + B) + C
    A, B, C = trio(); exprs["muladd"] = C + (A * B)
    A, B, C = trio(); exprs["mulsub"] = (A * B) - C
    A, B, C = trio(); exprs["submul"] = C - (A * B)   # subtrahend product: stays unfused
    A, B, C = trio(); D = m.input(f"d{k[0]}", width).value(); exprs["aoi22"] = ~((A & B) | (C & D))
    A, B, C = trio(); D = m.input(f"d{k[0]}", width).value(); exprs["oai22"] = ~((A | B) & (C | D))
    A, B, C = trio()
    shared = A & B
    exprs["shared1"] = shared | C
    exprs["shared2"] = shared ^ C
    A, B, C = trio()
    exprs["widthchange"] = (A.zext(width + 1) & B.zext(width + 1)).trunc(width) | C
    A, B, C = trio()
    stored = m.wire("stored", width)
    m.assign(stored, A + B)
    exprs["storedsum"] = stored.value() + C
    A, B, _ = trio()
    r = m.reg("acc", width)
    m.clock(r, clk)
    m.next(r, (r.value() * A) + B)  # a register's next value through MULADD
    exprs["acc"] = r.value()
    for name, e in exprs.items():
        o = m.output(f"o_{name}", width)
        m.assign(o, e)
    return m
