"""Pins the CPU oracles before they are trusted as checkers.

* Every machine op of the packed oracle against Python big-int closed forms
  (exhaustive at small widths, random at limb-boundary widths) -- the limb
  helpers of crates/rrtl-sim-ir/src/lib.rs:12942-13196.
* Scalar oracle vs packed oracle on random multi-limb designs with memories
  (two independent restatements: fixpoint scalar vs levelized packed).
* Closed forms: counter, crc32 (bench/sv/crc32.sv).
* Known answers from the reference: crc32 0x53535102 (1 reset + 200 x 0xAB,
  restated closed form) and the fault campaign 64/64 detected with latency
  min 1 / max 6 / mean 2.1 (docs/rrtl-optimizations.md:1412-1416).
* The fast host engine (timed CPU baseline) against the faithful engine.
* numpy stimulus/digest helpers against the C ones.
"""
import itertools
import os
import random

import numpy as np
import pytest

import oracle
from oracle import stim
from oracle.scalar_sim import ScalarSimulator
from rrtl_amd import designs, ir, lower, netlist
from rrtl_amd.ir import BitType, concat, lit, mux, uint
from rrtl_amd.program import KIND_INPUT, KIND_OUTPUT, KIND_REG
from designgen import random_design

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _signed(v, w):
    v &= (1 << w) - 1
    return v - (1 << w) if v >> (w - 1) & 1 else v


def closed_form(op, a, b, wa, wb, wo, sa=False):
    m = (1 << wo) - 1
    if op == "and": return a & b & m
    if op == "or": return (a | b) & m
    if op == "xor": return (a ^ b) & m
    if op == "add": return (a + b) & m
    if op == "sub": return (a - b) & m
    if op == "mul": return (a * b) & m
    if op == "eq": return int(a == b)
    if op == "ne": return int(a != b)
    if op == "lt":
        return int(_signed(a, wa) < _signed(b, wa)) if sa else int(a < b)
    raise ValueError(op)


def _binop_design(op, w, signed):
    m = ir.Module(f"B_{op}_{w}")
    a = m.input("a", BitType(w, signed))
    b = m.input("b", BitType(w, signed))
    wo = 1 if op in ("eq", "ne", "lt") else w
    o = m.output("o", wo)
    e = {"and": lambda: a & b, "or": lambda: a | b, "xor": lambda: a ^ b, "add": lambda: a + b,
         "sub": lambda: a - b, "mul": lambda: a * b, "eq": lambda: a.value().eq(b),
         "ne": lambda: a.value().ne(b), "lt": lambda: a.value().lt(b)}[op]()
    m.assign(o, e.cast(False) if e.ty.signed else e)
    return m, wo


def _run_pairs(m, pairs):
    p = lower.lower_module(m)
    lanes = len(pairs)
    ref = oracle.PackedRef(p, lanes)
    ref.set_signal(p.signal_index("a"), [x for x, _ in pairs], eval=False)
    ref.set_signal(p.signal_index("b"), [y for _, y in pairs])
    return ref.get_signal(p.signal_index("o"))


@pytest.mark.parametrize("op", ["and", "or", "xor", "add", "sub", "mul", "eq", "ne", "lt"])
@pytest.mark.parametrize("signed", [False, True])
def test_binary_ops_exhaustive_small_widths(op, signed):
    for w in (1, 3, 5):
        m, wo = _binop_design(op, w, signed)
        pairs = list(itertools.product(range(1 << w), repeat=2))
        got = _run_pairs(m, pairs)
        exp = [closed_form(op, x, y, w, w, wo, signed) for x, y in pairs]
        assert got == exp, (op, w, signed)


@pytest.mark.parametrize("op", ["and", "or", "xor", "add", "sub", "mul", "eq", "ne", "lt"])
def test_binary_ops_random_wide(op):
    rng = random.Random(sum(map(ord, op)))
    for w in (31, 32, 33, 63, 64, 65, 96, 127, 128, 200):
        for signed in (False, True):
            m, wo = _binop_design(op, w, signed)
            pairs = [(rng.getrandbits(w), rng.getrandbits(w)) for _ in range(64)]
            pairs += [(0, 0), ((1 << w) - 1, 1), (1 << (w - 1), (1 << (w - 1)) - 1), ((1 << w) - 1, (1 << w) - 1)]
            got = _run_pairs(m, pairs)
            exp = [closed_form(op, x, y, w, w, wo, signed) for x, y in pairs]
            assert got == exp, (op, w, signed)


def test_unary_and_width_ops():
    rng = random.Random(3)
    for w in (1, 7, 32, 33, 64, 100):
        m = ir.Module("U")
        a = m.input("a", w)
        outs = {}
        def out(name, e):
            o = m.output(name, e.ty.width)
            m.assign(o, e.cast(False) if e.ty.signed else e)
            outs[name] = o
        out("not", ~a)
        out("sext", a.value().sext(w + 37))
        out("zext", a.value().zext(w + 37))
        if w > 1:
            out("trunc", a.value().trunc(w - 1))
            out("slice", a.value().slice(w // 3, w - w // 3))
        out("cat", concat(a.value(), lit(5, 3), a.value()))
        p = lower.lower_module(m)
        vals = [rng.getrandbits(w) for _ in range(40)] + [0, (1 << w) - 1, 1 << (w - 1)]
        ref = oracle.PackedRef(p, len(vals))
        ref.set_signal(p.signal_index("a"), vals)
        msk = (1 << w) - 1
        exp = {
            "not": [~v & msk for v in vals],
            "sext": [_signed(v, w) & ((1 << (w + 37)) - 1) for v in vals],
            "zext": vals,
            "cat": [(v << (w + 3)) | (5 << w) | v for v in vals],
        }
        if w > 1:
            exp["trunc"] = [v & ((1 << (w - 1)) - 1) for v in vals]
            exp["slice"] = [(v >> (w // 3)) & ((1 << (w - w // 3)) - 1) for v in vals]
        for k, e in exp.items():
            assert ref.get_signal(p.signal_index(k)) == e, (w, k)


def test_memory_read_write_and_out_of_range():
    m = ir.Module("Mem")
    addr = m.input("addr", 6)
    waddr = m.input("waddr", 6)
    we = m.input("we", 1)
    wd = m.input("wd", 40)
    mem = m.mem("ram", 40, 20)
    o = m.output("o", 40)
    m.assign(o, mem.read(addr))
    m.mem_write(mem, we, waddr, wd)
    p = lower.lower_module(m)
    ref = oracle.PackedRef(p, 4)
    model = [[0] * 20 for _ in range(4)]
    rng = random.Random(8)
    for c in range(60):
        wa = [rng.randrange(64) for _ in range(4)]
        wv = [rng.getrandbits(40) for _ in range(4)]
        en = [rng.getrandbits(1) for _ in range(4)]
        ra = [rng.randrange(64) for _ in range(4)]
        ref.set_signal(p.signal_index("waddr"), wa, eval=False)
        ref.set_signal(p.signal_index("wd"), wv, eval=False)
        ref.set_signal(p.signal_index("we"), en, eval=False)
        ref.set_signal(p.signal_index("addr"), ra)
        assert ref.get_signal(p.signal_index("o")) == [model[l][ra[l]] if ra[l] < 20 else 0 for l in range(4)]
        ref.tick()
        for l in range(4):
            if en[l] and wa[l] < 20:
                model[l][wa[l]] = wv[l]


@pytest.mark.parametrize("seed", range(40))
def test_scalar_and_packed_oracles_agree(seed):
    """Two independent restatements agree on designs without async resets (with async
    resets the reference's own PackedSimulator evaluates Comb before AsyncResetComb,
    lib.rs:9278-9281, while the scalar Simulator iterates to a fixpoint)."""
    m = random_design(seed, async_resets=False)
    p = lower.lower_module(m)
    lanes = 3
    ref = oracle.PackedRef(p, lanes)
    scs = [ScalarSimulator(random_design(seed, async_resets=False)) for _ in range(lanes)]
    rng = random.Random(seed + 100)
    for cyc in range(10):
        for s in m.signals:
            if s.kind != "input":
                continue
            if s.name == "rst":
                vals = [1 if cyc < 1 else rng.getrandbits(1) for _ in range(lanes)]
            elif s.name == "rstn":
                vals = [0 if cyc < 1 else 1 for _ in range(lanes)]
            else:
                vals = [rng.getrandbits(s.width) for _ in range(lanes)]
            ref.set_signal(p.signal_index(s.name), vals, eval=False)
            for l, sc in enumerate(scs):
                sc.set_raw(sc.m.signal(s.name), vals[l])
        ref.eval_combinational()
        for sc in scs:
            sc.eval_combinational()
        for s in m.signals:
            assert ref.get_signal(p.signal_index(s.name)) == [sc.get(sc.m.signal(s.name)) for sc in scs], (cyc, s.name)
        ref.tick()
        for sc in scs:
            sc.tick()


def test_counter_closed_form_on_packed_oracle():
    p = lower.lower_module(designs.counter())
    lanes = 256
    ref = oracle.PackedRef(p, lanes)
    rng = np.random.default_rng(0)
    count = np.zeros(lanes, np.int64)
    for c in range(300):
        en = rng.integers(0, 2, lanes)
        ref.set_signal(p.signal_index("rst"), [1 if c == 0 else 0] * lanes, eval=False)
        ref.set_signal(p.signal_index("en"), en.tolist())
        ref.tick()
        count = np.zeros_like(count) if c == 0 else (count + en) % 256
        assert ref.get_signal(p.signal_index("out")) == count.tolist()


def crc32_program():
    with open(os.path.join(GOLDEN, "crc32_gates.json")) as f:
        return lower.lower_module(netlist.import_yosys_netlist(f.read(), "crc32"))


def test_crc32_rtl_matches_closed_form_and_gate_netlist():
    """The word-level crc32.sv restatement, the Yosys gate netlist fixture and the
    closed form agree; 0x53535102 after 1 reset + 200 x 0xAB."""
    pr = lower.lower_module(designs.crc32_rtl())
    ref = oracle.PackedRef(pr, 1)
    ref.set_signal(pr.signal_index("rst"), [1], eval=False)
    ref.set_signal(pr.signal_index("din"), [0xAB])
    ref.tick()
    ref.set_signal(pr.signal_index("rst"), [0])
    ref.tick_many(200)
    assert ref.get_signal(pr.signal_index("crc_out")) == [0x53535102]
    pg = crc32_program()
    bp = oracle.BitParallelRef(pg, 1)
    bp.set_signal_all(pg.signal_index("clk"), 1)
    bp.set_signal_all(pg.signal_index("rst"), 1)
    for i in range(8):
        bp.set_signal_all(pg.signal_index(f"din_{i}"), (0xAB >> i) & 1)
    bp.tick()
    bp.set_signal_all(pg.signal_index("rst"), 0)
    bp.tick_many(200)
    assert sum(bp.get_bit(pg.signal_index(f"crc_{b}"), 0) << b for b in range(32)) == 0x53535102


def test_fault_sim_known_answer_on_oracle():
    pg = crc32_program()
    regs = [i for i, s in enumerate(pg.signals) if s.kind == KIND_REG]
    faults = [(r, v) for r in regs for v in (0, 1)]
    lanes = len(faults) + 1
    bp = oracle.BitParallelRef(pg, lanes)
    for f, (sig, v) in enumerate(faults):
        bp.set_force(sig, f + 1, v)
    crc = [pg.signal_index(f"crc_{b}") for b in range(32)]
    din = [pg.signal_index(f"din_{i}") for i in range(8)]
    bp.set_signal_all(pg.signal_index("clk"), 1)
    det = [None] * len(faults)
    for cyc in range(96):
        dv = ((cyc * 2654435761) >> 13) & 0xFF
        bp.set_signal_all(pg.signal_index("rst"), cyc < 1)
        for i in range(8):
            bp.set_signal_all(din[i], (dv >> i) & 1)
        bp.tick()
        words = np.stack([bp.get_words(c) for c in crc])
        bits = ((words[:, :, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)).reshape(32, -1)[:, :lanes]
        for f in range(len(faults)):
            if det[f] is None and not np.array_equal(bits[:, f + 1], bits[:, 0]):
                det[f] = cyc
    lat = [d + 1 for d in det if d is not None]
    assert (len(lat), min(lat), max(lat), round(sum(lat) / len(lat), 1)) == (64, 1, 6, 2.1)


def test_fast_engine_matches_faithful_engine():
    p = lower.lower_module(designs.synthetic_netlist(gates=2000, levels=16, dffs=200, inputs=40, outputs=24, seed=3))
    lanes = 700
    ins = [(i, oracle.IN_RANDOM, 0) for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    obs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
    for post in (False, True):
        a = oracle.BitParallelRef(p, lanes)
        b = oracle.BitParallelFast(p, lanes, threads=4)
        for lane in (3, 300, 699):
            a.set_force(obs[0] - 5, lane, 1)
            b.set_force(obs[0] - 5, lane, 1)
        da, ta = a.run(seed=5, cycles=30, inputs=ins, observe=obs, observe_post=post, digest=True, trace=True)
        db, tb, _ = b.run(seed=5, cycles=30, inputs=ins, observe=obs, observe_post=post, digest=True, trace=True)
        assert np.array_equal(ta, tb)
        assert np.array_equal(da, db)
        for r in [i for i, s in enumerate(p.signals) if s.kind == KIND_REG][::11]:
            assert np.array_equal(a.get_words(r), b.get_words(r))


@pytest.mark.parametrize("W", [1, 4, 8])
def test_aot_engine_matches_faithful_engine(W):
    """The AOT host engine (straight-line C per design, oracle/bp_aot.py) at
    64-, 256- and 512-lane groups against the faithful bit-parallel engine:
    output signatures over two consecutive runs and every register; on a
    synthetic netlist and on a random 1-bit design with sync resets of both
    polarities (lanes not a multiple of the group)."""
    from designgen import random_gate_design
    from oracle import bp_aot
    cases = [lower.lower_module(designs.synthetic_netlist(gates=1500, levels=14, dffs=150, inputs=40, outputs=24,
                                                          seed=4)),
             lower.lower_module(random_gate_design(11, observe_regs=True, async_resets=False))]
    for p in cases:
        lanes = 900
        clk = [i for i, s in enumerate(p.signals) if s.name == "clk"]
        ins = [(i, oracle.IN_CONST, 1) for i in clk] + \
              [(i, oracle.IN_RANDOM, 0) for i, s in enumerate(p.signals) if s.kind == KIND_INPUT and i not in clk]
        obs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
        ref = oracle.BitParallelRef(p, lanes)
        aot = bp_aot.BitParallelAot(p, lanes, ins, obs, group_words=W, threads=3)
        assert 0 < aot.gates
        for run in range(2):
            dr, _ = ref.run(seed=7 + run, cycles=25, inputs=ins, observe=obs, digest=True)
            da, _ = aot.run(7 + run, 25)
            assert np.array_equal(da, dr), (p.top, run)
        for r in [i for i, s in enumerate(p.signals) if s.kind == KIND_REG]:
            assert np.array_equal(aot.get_words(r), ref.get_words(r)), p.signals[r].name


def test_numpy_stimulus_and_digest_match_c():
    """oracle/stim.py against oracle/stim.h through a one-input design."""
    m = ir.Module("Pass")
    i = m.input("i", 1)
    o = m.output("o", 1)
    m.assign(o, i)
    p = lower.lower_module(m)
    lanes = 64 * 5
    bp = oracle.BitParallelRef(p, lanes)
    dg, tr = bp.run(seed=77, cycles=9, inputs=[(0, oracle.IN_RANDOM, 0)], observe=[1], digest=True, trace=True,
                    word_base=3)
    S = np.zeros((64, 5), np.uint64)
    for c in range(9):
        w = stim.random_input_words(77, 0, c, 5, base=3)
        assert np.array_equal(tr[c, 0], w)
        S = stim.sig_digest_cycle(S, c, w.reshape(1, 5))
    assert np.array_equal(S, dg)
    # per-lane expansion used by the word-level harness
    lanes_bits = stim.random_input_lanes(77, 0, 1, 4, 64 * 8)[0]
    w = stim.random_input_words(77, 0, 4, 8)
    exp = ((w[np.arange(512) // 64] >> (np.arange(512) % 64).astype(np.uint64)) & np.uint64(1)).astype(np.uint32)
    assert np.array_equal(lanes_bits, exp)


def test_synthetic_netlist_shape():
    m = designs.synthetic_netlist()
    kinds = {}
    for s in m.signals:
        kinds[s.kind] = kinds.get(s.kind, 0) + 1
    assert kinds == {"input": 256, "reg": 2000, "wire": 20000, "output": 128}
    ops = {}
    for w in m.signals:
        if w.kind == "wire":
            e = m.assigns[w]
            k = "nand" if e.op == "not" else e.op
            ops[k] = ops.get(k, 0) + 1
    frac = {k: v / 20000 for k, v in ops.items()}
    assert abs(frac["and"] - 0.4) < 0.02 and abs(frac["or"] - 0.3) < 0.02
    assert abs(frac["xor"] - 0.2) < 0.02 and abs(frac["nand"] - 0.1) < 0.02
    p = lower.lower_module(m)
    assert p.is_gate_level()
    assert len(p.comb.packets) >= 64


@pytest.mark.parametrize("seed", range(8))
def test_simd_engine_matches_packed_oracle(seed):
    """The host batch-SIMD baseline (lane shards, one-limb vector paths) against the packed oracle."""
    from harness import packed_run
    m = random_design(seed + 200, async_resets=seed % 2 == 0)
    p = lower.lower_module(m)
    lanes, cycles = 150, 20
    ins = []
    for i, s in enumerate(p.signals):
        if s.kind != KIND_INPUT:
            continue
        if s.name == "rst":
            ins.append((i, oracle.IN_PULSE, oracle.pulse_arg(2, 1, 0)))
        elif s.name == "rstn":
            ins.append((i, oracle.IN_CONST, 1))
        else:
            ins.append((i, oracle.IN_RANDOM, 0))
    obs = [i for i, s in enumerate(p.signals) if s.kind != KIND_INPUT]
    for post in (False, True):
        simd = oracle.SimdCpu(p, lanes, threads=3, shard_lanes=64)
        dg, _ = simd.run(seed, cycles, ins, obs, observe_post=post)
        ref, _, od = packed_run(p, lanes, seed, cycles, ins, obs, observe_post=post)
        assert np.array_equal(dg, od)
        for i in range(len(p.signals)):
            assert np.array_equal(simd.get_signal_limbs(i), ref.get_signal_limbs(i)), p.signals[i].name


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_gate_design_bitparallel_matches_packed(seed):
    """The two CPU oracles of the gate engine and the word engine agree, cycle
    by cycle, on random 1-bit designs with sync resets of both polarities,
    register shift chains and observed registers: the bit-parallel restatement
    runs its device-style loop (inputs, settle, PRE observation, edge) and the
    packed restatement is driven through the same steps one API call at a time."""
    from designgen import random_gate_design
    p = lower.lower_module(random_gate_design(seed, observe_regs=True, async_resets=False))
    lanes, cycles, rseed = 256, 24, 5 + seed  # whole 64-lane words: no padding lanes
    clk = p.signal_index("clk")
    driven = [i for i, s in enumerate(p.signals) if s.kind == KIND_INPUT and i != clk]
    ins = [(clk, oracle.IN_CONST, 1)] + [(i, oracle.IN_RANDOM, 0) for i in driven]
    obs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
    bp = oracle.BitParallelRef(p, lanes)
    _, tr = bp.run(seed=rseed, cycles=cycles, inputs=ins, observe=obs, digest=False, trace=True)
    pk = oracle.PackedRef(p, lanes)
    pk.set_signal(clk, [1] * lanes)
    for c in range(cycles):
        for i in driven:
            pk.set_signal_limbs(i, stim.random_input_lanes(rseed, i, 1, c, lanes), eval=False)
        pk.eval_combinational()
        for j, o in enumerate(obs):
            bits = pk.get_signal_limbs(o)[0].astype(np.uint64)
            words = np.zeros(tr.shape[-1], np.uint64)
            for lane in range(lanes):
                words[lane // 64] |= bits[lane] << np.uint64(lane % 64)
            assert np.array_equal(words, tr[c, j]), (c, p.signals[o].name)
        pk.tick_from_evaluated()
