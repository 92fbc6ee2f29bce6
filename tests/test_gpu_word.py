"""Word-level engine on the GPU against the packed oracle, the scalar oracle and
closed forms (differential method of the reference's tests, e.g.
crates/rrtl-gpu-sim/src/interp.rs:2081 interp_runner_matches_simd_cpu)."""
import random

import numpy as np
import pytest

import oracle
from oracle import stim
from oracle.scalar_sim import ScalarSimulator
from rrtl_amd import designs, lower
from rrtl_amd.backend import IN_CONST, IN_PULSE, IN_RANDOM, HipSimulator
from rrtl_amd.program import KIND_INPUT, KIND_OUTPUT
from conftest import flush_print
from designgen import random_design
from harness import input_limbs, packed_run

pytestmark = pytest.mark.gpu


def counter_inputs(p):
    return [(p.signal_index("clk"), IN_CONST, 1), (p.signal_index("rst"), IN_PULSE, oracle.pulse_arg(1, 1, 0)),
            (p.signal_index("en"), IN_RANDOM, 0)]


def test_counter_4096x1000_matches_oracle_and_closed_form():
    """README counter, 4096 lanes x 1000 cycles: rst=1 at cycle 0, en ~ Bernoulli(0.5)."""
    flush_print("case: counter 4096x1000 word engine")
    p = lower.lower_module(designs.counter())
    lanes, cycles, seed = 4096, 1000, 11
    sim = HipSimulator(p, lanes, engine="word")
    out = p.signal_index("out")
    sim.run(cycles, seed=seed, inputs=counter_inputs(p), observe=[out], trace=True)
    tr = sim.trace().reshape(cycles, 1, lanes)
    # closed form: out(c) = count before edge c; count' = rst ? 0 : count + en (mod 256)
    count = np.zeros(lanes, dtype=np.int64)
    en_idx = p.signal_index("en")
    for c in range(cycles):
        assert np.array_equal(tr[c, 0], count.astype(np.uint32)), f"cycle {c}"
        en = stim.random_input_lanes(seed, en_idx, 1, c, lanes)[0].astype(np.int64)
        count = np.zeros_like(count) if c == 0 else (count + en) % 256
    # packed oracle replays the same stimulus
    ref, otr, od = packed_run(p, lanes, seed, cycles, counter_inputs(p), [out])
    assert np.array_equal(tr, otr)
    assert np.array_equal(sim.digest(), od)
    assert sim.get_signal("count") == ref.get_signal(p.signal_index("count"))


def test_counter_scalar_oracle_lanes():
    """A few lanes cycle by cycle against the scalar Simulator restatement."""
    p = lower.lower_module(designs.counter())
    lanes, cycles, seed = 64, 200, 3
    sim = HipSimulator(p, lanes, engine="word")
    sim.run(cycles, seed=seed, inputs=counter_inputs(p), observe=["out"], trace=True)
    tr = sim.trace().reshape(cycles, lanes)
    en_idx = p.signal_index("en")
    for lane in (0, 17, 63):
        m = designs.counter()
        sc = ScalarSimulator(m)
        sc.set(m.signal("clk"), 1)
        for c in range(cycles):
            sc.set(m.signal("rst"), 1 if c == 0 else 0)
            en = int(stim.random_input_lanes(seed, en_idx, 1, c, lanes)[0, lane])
            sc.set(m.signal("en"), en)
            assert tr[c, lane] == sc.get(m.signal("out")), (lane, c)
            sc.tick()


@pytest.mark.parametrize("codegen", ["on", "off", "off-menu"])
@pytest.mark.parametrize("seed", range(12))
def test_random_designs_match_packed_oracle(seed, codegen):
    """Generated kernel ("on"), record interpreter with macro records ("off")
    and with the fusion menu instead ("off-menu") alike."""
    flush_print(f"case: random design seed={seed} codegen={codegen}")
    m = random_design(seed)
    p = lower.lower_module(m)
    lanes = 96
    sim = HipSimulator(p, lanes, engine="word", codegen=codegen.split("-")[0],
                       interp_macros="off" if codegen == "off-menu" else "on")
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


@pytest.mark.parametrize("n", [0, 1, 2, 3, 17, 64, 257])
def test_tick_many_equals_repeated_tick(n):
    """Comb fusion invariant (crates/rrtl-sim-ir/src/lib.rs:14634)."""
    m = random_design(101, async_resets=False)
    p = lower.lower_module(m)
    lanes = 64
    a = HipSimulator(p, lanes, engine="word")
    b = HipSimulator(p, lanes, engine="word")
    ref = oracle.PackedRef(p, lanes)
    rng = np.random.default_rng(5)
    for i, s in enumerate(p.signals):
        if s.kind == KIND_INPUT:
            v = rng.integers(0, 2**32, size=(s.limbs, lanes), dtype=np.uint32)
            if s.name in ("rst", "rstn"):
                v[:] = 0 if s.name == "rst" else 1
            for x in (a, b):
                x.set_signal_limbs(i, v)
            ref.set_signal_limbs(i, v)
    a.tick_many(n)
    for _ in range(n):
        b.tick()
    ref.tick_many(n)
    va, ma = a.snapshot()
    vb, mb = b.snapshot()
    vr, mr = ref.snapshot()
    assert np.array_equal(va, vb) and np.array_equal(ma, mb)
    assert np.array_equal(va, vr) and np.array_equal(ma, mr)


@pytest.mark.parametrize("codegen", ["on", "off"])
def test_random_stimulus_run_matches_oracle_all_widths(codegen):
    """Device-side RANDOM/CONST/PULSE stimulus and PRE/POST observation."""
    for seed, post in ((7, False), (8, True)):
        flush_print(f"case: random stimulus run seed={seed} codegen={codegen}")
        m = random_design(seed, async_resets=False)
        p = lower.lower_module(m)
        lanes, cycles = 130, 24
        ins = []
        for i, s in enumerate(p.signals):
            if s.kind != KIND_INPUT:
                continue
            if s.name == "rst":
                ins.append((i, IN_PULSE, oracle.pulse_arg(2, 1, 0)))
            elif s.name == "rstn":
                ins.append((i, IN_CONST, 1))
            else:
                ins.append((i, IN_RANDOM, 0))
        obs = [i for i, s in enumerate(p.signals) if s.kind != KIND_INPUT]
        sim = HipSimulator(p, lanes, engine="word", codegen=codegen)
        sim.run(cycles, seed=seed, inputs=ins, observe=obs, observe_post=post, trace=True)
        ref, otr, od = packed_run(p, lanes, seed, cycles, ins, obs, observe_post=post)
        assert np.array_equal(sim.trace().reshape(otr.shape), otr)
        assert np.array_equal(sim.digest(), od)
        vs, ms = sim.snapshot()
        vr, mr = ref.snapshot()
        assert np.array_equal(vs, vr) and np.array_equal(ms, mr)


@pytest.mark.parametrize("block_threads", [0, 64])
def test_storage_modes_generated_equals_interpreter(block_threads):
    """A design whose signals and scratch exceed a block's LDS share at 256
    lanes per block runs with both in HBM; at 64 lanes per block its scratch
    fits in LDS.  The generated kernel and the interpreter give the same trace,
    digest and final state in either storage mode, and the packed oracle
    agrees."""
    flush_print(f"case: storage modes block_threads={block_threads}")
    m = designs.synthetic_netlist(gates=400, levels=16, dffs=40, inputs=64, outputs=32, seed=4)
    p = lower.lower_module(m)
    lanes, cycles, seed = 200, 12, 9
    ins = [(i, IN_RANDOM, 0) for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    obs = [i for i, s in enumerate(p.signals) if s.kind != KIND_INPUT][:40]
    out = {}
    for cg in ("on", "off"):
        sim = HipSimulator(p, lanes, engine="word", codegen=cg, block_threads=block_threads)
        sim.run(cycles, seed=seed, inputs=ins, observe=obs, trace=True)
        out[cg] = (sim.trace(), sim.digest(), sim.snapshot()[0])
        sim.close()
    for a, b in zip(out["on"], out["off"]):
        assert np.array_equal(a, b)
    ref, otr, od = packed_run(p, lanes, seed, cycles, ins, obs)
    assert np.array_equal(out["on"][0].reshape(otr.shape), otr)
    assert np.array_equal(out["on"][1], od)


@pytest.mark.parametrize("codegen", ["on", "off"])
def test_fifo_ready_valid(codegen):
    flush_print(f"case: fifo codegen={codegen}")
    p = lower.lower_module(designs.ready_valid_fifo())
    lanes, cycles, seed = 512, 300, 21
    ins = [(p.signal_index("clk"), IN_CONST, 1), (p.signal_index("rst"), IN_PULSE, oracle.pulse_arg(1, 1, 0)),
           (p.signal_index("in_valid"), IN_RANDOM, 0), (p.signal_index("out_ready"), IN_RANDOM, 0),
           (p.signal_index("in_data"), IN_RANDOM, 0)]
    obs = [p.signal_index(n) for n in ("in_ready", "out_valid", "out_data", "count")]
    sim = HipSimulator(p, lanes, engine="word", codegen=codegen)
    sim.run(cycles, seed=seed, inputs=ins, observe=obs, trace=True)
    ref, otr, od = packed_run(p, lanes, seed, cycles, ins, obs)
    tr = sim.trace().reshape(otr.shape)
    assert np.array_equal(tr, otr)
    assert np.array_equal(sim.digest(), od)
    # closed form: occupancy follows accepted pushes and pops
    cnt = np.zeros(lanes, np.int64)
    for c in range(cycles):
        assert np.array_equal(tr[c, 3], cnt.astype(np.uint32)), c
        iv = stim.random_input_lanes(seed, p.signal_index("in_valid"), 1, c, lanes)[0].astype(bool)
        orr = stim.random_input_lanes(seed, p.signal_index("out_ready"), 1, c, lanes)[0].astype(bool)
        push = iv & (cnt < 16)
        pop = orr & (cnt > 0)
        cnt = np.zeros_like(cnt) if c == 0 else cnt + push - pop


def test_snapshot_restore_continue():
    m = random_design(55, async_resets=False)
    p = lower.lower_module(m)
    lanes = 40
    a = HipSimulator(p, lanes, engine="word")
    rng = np.random.default_rng(1)
    for i, s in enumerate(p.signals):
        if s.kind == KIND_INPUT and s.name not in ("rst", "rstn"):
            a.set_signal_limbs(i, rng.integers(0, 2**32, size=(s.limbs, lanes), dtype=np.uint32))
    a.set_signal("rstn", 1)
    a.tick_many(5)
    v, mm = a.snapshot()
    b = HipSimulator(p, lanes, engine="word")
    b.restore(v, mm)
    a.tick_many(9)
    b.tick_many(9)
    assert all(np.array_equal(x, y) for x, y in zip(a.snapshot(), b.snapshot()))


def test_halting_lanes_stop_counting():
    """A per-lane down-counter that halts at zero (divergent_batch.rs halting rule)."""
    from rrtl_amd import ir
    from rrtl_amd.ir import lit, mux
    mod = ir.Module("Down")
    start = mod.input("start", 16)
    load = mod.input("load", 1)
    cnt = mod.reg("cnt", 16)
    done = mod.wire("done", 1)
    mod.assign(done, cnt.value().eq(lit(0, 16)) & ~load.value())
    mod.next(cnt, mux(load.value(), start.value(), mux(done.value(), cnt.value(), cnt - lit(1, 16))))
    p = lower.lower_module(mod)
    lanes = 300
    sim = HipSimulator(p, lanes, engine="word")
    starts = np.arange(lanes, dtype=np.uint32) % 97 + 1
    sim.set_signal_limbs("start", starts.reshape(1, lanes))
    sim.set_signal("load", 1)
    sim.tick()
    sim.set_signal("load", 0)
    sim.run(200, inputs=[], observe=[], halt="done")
    h = sim.halt_cycles()
    assert np.array_equal(h, starts.astype(np.uint64))
    assert sim.stats()["lane_cycles"] == int(starts.sum())
    assert all(v == 0 for v in sim.get_signal("cnt"))


def test_errors_are_reported():
    p = lower.lower_module(designs.counter())
    from rrtl_amd.backend import HipError
    with pytest.raises(HipError) as e:
        HipSimulator(p, 0, engine="word")
    assert e.value.code == "E_SIM_IR_LANES"
    sim = HipSimulator(p, 8, engine="word")
    with pytest.raises(HipError) as e:
        sim.get_signal_limbs(999)
    assert e.value.code == "E_SIM_IR_SIGNAL"
    # the library's own check, below the binding
    buf = np.zeros(64, np.uint32)
    assert sim.L.rrtl_hip_get_signal_limbs(sim.h, 999, buf.ctypes.data) == 1
    assert sim.L.rrtl_hip_last_error(sim.h).decode().startswith("E_SIM_IR_SIGNAL")
    with pytest.raises(HipError) as e:
        sim.set_force(0, 0, 1)
    assert e.value.code == "E_HIP_UNSUPPORTED"


@pytest.mark.parametrize("width", [8, 32])
@pytest.mark.parametrize("codegen", ["on", "off", "off-macros"])
def test_fused_records_match_packed_oracle(width, codegen):
    """Every fused record form (and the forms that must stay unfused) against
    the packed oracle, in the record interpreter (with the fusion menu, or
    with macro records in its place) and the generated kernel."""
    from designgen import fusion_design
    flush_print(f"case: fused records width={width} codegen={codegen}")
    p = lower.lower_module(fusion_design(width))
    lanes = 128
    sim = HipSimulator(p, lanes, engine="word", codegen=codegen.split("-")[0],
                       interp_macros="on" if codegen == "off-macros" else "off")
    ref = oracle.PackedRef(p, lanes)
    rng = random.Random(width)
    ins = [i for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    outs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
    for c in range(6):
        for i in ins:
            v = np.array([[1 if p.signals[i].name == "clk" else rng.getrandbits(width) for _ in range(lanes)]], np.uint32)
            sim.set_signal_limbs(i, v, settle=False)
            ref.set_signal_limbs(i, v, eval=False)
        sim.eval_combinational(); ref.eval_combinational()
        for o in outs:
            assert np.array_equal(sim.get_signal_limbs(o), ref.get_signal_limbs(o)), (c, p.signals[o].name)
        sim.tick_from_evaluated(); ref.tick_from_evaluated()
