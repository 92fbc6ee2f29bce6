"""The specializer (rrtl_amd/specialize.py, restating crates/rrtl-sim-ir/src/
specialize.rs): every pass leaves the program observationally equivalent --
the packed oracle gives the same per-lane digest of every signal at every
cycle and the same final state -- while folding, copying and removing what the
reference's rules say it may."""
import random

import numpy as np
import pytest

import oracle
from designgen import random_design
from harness import packed_run
from rrtl_amd import ir, lower, specialize
from rrtl_amd.backend import IN_CONST, IN_RANDOM
from rrtl_amd.ir import lit, mux
from rrtl_amd.program import KIND_INPUT, OP_LIT, OP_MUL


def same_behaviour(p, q, lanes=37, cycles=25, seed=4, inputs=None):
    ins = inputs or [(i, IN_RANDOM, 0) for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    obs = [i for i, s in enumerate(p.signals) if s.kind != KIND_INPUT]
    ra, _, da = packed_run(p, lanes, seed, cycles, ins, obs)
    rb, _, db = packed_run(q, lanes, seed, cycles, ins, obs)
    assert np.array_equal(da, db)
    va, ma = ra.snapshot()
    vb, mb = rb.snapshot()
    assert np.array_equal(va, vb) and np.array_equal(ma, mb)


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
def test_random_designs_keep_their_behaviour(seed):
    p = lower.lower_module(random_design(seed))
    q, st = specialize.specialize_program(p)
    assert st.instrs_after <= st.instrs_before
    same_behaviour(p, q)


def constant_heavy():
    """Every identity of try_fold (specialize.rs:846) on live values."""
    m = ir.Module("Consts")
    clk = m.input("clk", 1)
    a = m.input("a", 16).value()
    b = m.input("b", 16).value()
    r = m.reg("r", 16)
    m.clock(r, clk)
    ones, zero, one = lit(0xFFFF, 16), lit(0, 16), lit(1, 16)
    outs = {
        "and1": a & ones, "and0": a & zero, "or0": a | zero, "or1": a | ones, "xor0": zero ^ a,
        "add0": a + zero, "sub0": a - zero, "subself": (a + b) - (a + b), "mul0": a * zero, "mul1": one * a,
        "eqself": (a ^ b).eq(a ^ b).zext(16), "neself": (a ^ b).ne(a ^ b).zext(16),
        "ltself": (a + b).lt(a + b).zext(16), "muxc": mux(lit(1, 1), a, b), "muxeq": mux(a.bit(0), b, b),
        "folded": (lit(3, 16) * lit(5, 16) + lit(2, 16)) ^ a, "notc": ~lit(0x00F0, 16) & a,
        "slicec": lit(0xABCD, 16).slice(4, 8).zext(16) + a, "sextc": lit(0x80, 8).sext(16) | b,
    }
    for k, e in outs.items():
        o = m.output(f"o_{k}", 16)
        m.assign(o, e)
    m.next(r, (r.value() * one) + (a & zero))
    return m


def test_identities_fold_and_copy():
    p = lower.lower_module(constant_heavy())
    q, st = specialize.specialize_program(p)
    assert st.folded >= 8 and st.copies >= 8 and st.dead > 0
    assert st.instrs_after < st.instrs_before
    same_behaviour(p, q)


def test_freeze_signals_folds_a_constant_mode():
    """A mode input held at a constant: freezing it folds the logic it
    selects; the frozen program equals the original while the mode holds."""
    m = ir.Module("Mode")
    clk = m.input("clk", 1)
    mode = m.input("mode", 1)
    a = m.input("a", 32).value()
    r = m.reg("r", 32)
    m.clock(r, clk)
    m.next(r, mux(mode.value(), r.value() * a + lit(7, 32), r.value() ^ a))
    o = m.output("o", 32)
    m.assign(o, mux(mode.value(), r.value() + a, r.value() - a))
    p = lower.lower_module(m)
    q, st = specialize.freeze_signals_program(p, {p.signal_index("mode"): 0})
    assert st.dead > 0 and st.instrs_after < st.instrs_before
    ins = [(p.signal_index("mode"), IN_CONST, 0), (p.signal_index("a"), IN_RANDOM, 0),
           (p.signal_index("clk"), IN_CONST, 1)]
    same_behaviour(p, q, inputs=ins)


def test_strength_reduction_of_wide_constant_multiplies():
    """x * C at 64 bits with C of few non-adjacent-form digits becomes
    shifts and adds; the result is the same product mod 2^64."""
    m = ir.Module("Wide")
    clk = m.input("clk", 1)
    x = m.input("x", 64).value()
    consts = [0x10001, 0xFFFF_FFFF_FFFF_FFFF, 7, 0x8000_0000_0000_0001, 0x1234_5678_9ABC_DEF1]
    for k, c in enumerate(consts):
        o = m.output(f"p{k}", 64)
        m.assign(o, x * lit(c, 64))
    p = lower.lower_module(m)
    q, st = specialize.strength_reduce_program(p)
    assert st.reduced == 4  # the last constant has too many digits
    muls = sum(1 for b in q.blocks for pk in b.packets for i in pk.instrs if i.op == OP_MUL)
    assert muls == 1
    same_behaviour(p, q, lanes=20, cycles=6)
    for c in consts:
        digits = specialize.naf(c, 64)
        assert sum(s << k for k, s in digits) % (1 << 64) == c


def test_naf_is_minimal_and_non_adjacent():
    rng = random.Random(1)
    for _ in range(200):
        c = rng.getrandbits(40)
        d = specialize.naf(c, 64)
        assert sum(s << k for k, s in d) == c
        ks = [k for k, _ in d]
        assert all(b - a >= 2 for a, b in zip(ks, ks[1:]))
