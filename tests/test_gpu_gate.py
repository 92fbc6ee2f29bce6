"""Bit-sliced gate engine on the GPU against the bit-parallel oracle, the crc32
known answers and the fault-simulation known answers
(crates/rrtl-sim-ir/examples/fault_sim.rs; docs/rrtl-optimizations.md:1412-1416)."""
import os

import numpy as np
import pytest

import oracle
from oracle import stim
from rrtl_amd import designs, lower, netlist
from rrtl_amd.backend import IN_CONST, IN_PULSE, IN_RANDOM, HipSimulator
from rrtl_amd.program import KIND_INPUT, KIND_OUTPUT, KIND_REG
from conftest import flush_print

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def valid_lanes(a: np.ndarray, lanes: int) -> np.ndarray:
    """Zero the bits of lanes past ``lanes`` in u64 lane words (last axis): the
    oracle pads to 64-lane words, the device to 32-lane words."""
    a = np.array(a, dtype=np.uint64, copy=True)
    words = a.shape[-1]
    for w in range(words):
        lo = 64 * w
        if lo >= lanes:
            a[..., w] = 0
        elif lanes - lo < 64:
            a[..., w] &= np.uint64((1 << (lanes - lo)) - 1)
    return a


def crc32_program():
    with open(os.path.join(GOLDEN, "crc32_gates.json")) as f:
        return lower.lower_module(netlist.import_yosys_netlist(f.read(), "crc32"))


def read_crc(sim, p, lane=0):
    return sum(int(sim.get_signal(p.signal_index(f"crc_{b}"))[lane]) << b for b in range(32))


def test_crc32_known_answer():
    """1 reset cycle then 200 cycles of din=0xAB give 0x53535102 (closed form of crc32.sv)."""
    flush_print("case: crc32 KAT gate engine")
    p = crc32_program()
    sim = HipSimulator(p, 64, engine="gate")
    assert sim.kind == "gate"
    sim.set_signal("clk", 1)
    sim.set_signal("rst", 1)
    for i in range(8):
        sim.set_signal(f"din_{i}", (0xAB >> i) & 1)
    sim.tick()
    sim.set_signal("rst", 0)
    sim.tick_many(200)
    v = 0xFFFFFFFF
    for _ in range(200):
        v = designs.crc32_step(v, 0xAB)
    assert v == 0x53535102
    assert read_crc(sim, p) == v
    assert len(set(sim.get_signal("crc_5"))) == 1  # every lane identical


def test_fault_sim_known_answer():
    """64 stuck-at faults on the 32 flops, lane 0 golden: 64/64 detected, latency 1/6/2.1."""
    flush_print("case: fault-sim KAT gate engine")
    p = crc32_program()
    regs = [i for i, s in enumerate(p.signals) if s.kind == KIND_REG]
    faults = [(r, v) for r in regs for v in (0, 1)]
    lanes = len(faults) + 1
    sim = HipSimulator(p, lanes, engine="gate")
    for f, (sig, v) in enumerate(faults):
        sim.set_force(sig, f + 1, v)
    crc = [p.signal_index(f"crc_{b}") for b in range(32)]
    din = [p.signal_index(f"din_{i}") for i in range(8)]
    sim.set_signal("clk", 1)
    detected = [None] * len(faults)
    gold = 0xFFFFFFFF
    for cyc in range(96):
        dv = ((cyc * 2654435761) >> 13) & 0xFF
        sim.set_signal("rst", 1 if cyc < 1 else 0, settle=False)
        for i in range(8):
            sim.set_signal(din[i], (dv >> i) & 1, settle=False)
        sim.tick()
        words = np.stack([sim.get_signal_bits(c) for c in crc])  # [32, words32]
        bits = ((words[:, :, None] >> np.arange(32, dtype=np.uint32)) & 1).reshape(32, -1)[:, :lanes]
        g = bits[:, 0]
        gold = 0xFFFFFFFF if cyc < 1 else designs.crc32_step(gold, dv)
        assert sum(int(g[b]) << b for b in range(32)) == gold
        for f in range(len(faults)):
            if detected[f] is None and not np.array_equal(bits[:, f + 1], g):
                detected[f] = cyc
    lat = [d + 1 for d in detected if d is not None]
    assert len(lat) == 64
    assert (min(lat), max(lat), round(sum(lat) / len(lat), 1)) == (1, 6, 2.1)


def test_fault_sim_device_run_post_observation():
    """The same campaign as one device-resident run with POST observation and a trace."""
    p = crc32_program()
    regs = [i for i, s in enumerate(p.signals) if s.kind == KIND_REG]
    faults = [(r, v) for r in regs for v in (0, 1)]
    lanes = len(faults) + 1
    crc = [p.signal_index(f"crc_{b}") for b in range(32)]
    ins = [(p.signal_index("clk"), IN_CONST, 1), (p.signal_index("rst"), IN_PULSE, oracle.pulse_arg(1, 1, 0))]
    ins += [(p.signal_index(f"din_{i}"), IN_RANDOM, 0) for i in range(8)]
    sim = HipSimulator(p, lanes, engine="gate")
    ref = oracle.BitParallelRef(p, lanes)
    for f, (sig, v) in enumerate(faults):
        sim.set_force(sig, f + 1, v)
        ref.set_force(sig, f + 1, v)
    sim.run(96, seed=9, inputs=ins, observe=crc, observe_post=True, trace=True)
    dg, tr = ref.run(seed=9, cycles=96, inputs=ins, observe=crc, observe_post=True, digest=True, trace=True)
    got = stim.u32_rows_to_u64(sim.trace().reshape(96, 32, -1))
    assert np.array_equal(valid_lanes(got, lanes), valid_lanes(tr, lanes))
    assert np.array_equal(valid_lanes(stim.u32_rows_to_u64(sim.digest()), lanes), valid_lanes(dg, lanes))


def small_netlist_program(**kw):
    m = designs.synthetic_netlist(**kw)
    return lower.lower_module(m)


def netlist_inputs(p):
    return [(i, IN_RANDOM, 0) for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]


def netlist_outputs(p):
    return [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]


@pytest.mark.parametrize("tile_words,stage_gates", [(1, 2), (2, 2), (4, 2), (8, 2), (1, 3), (2, 3), (4, 3), (8, 3), (1, 4), (2, 4), (4, 4), (8, 4)])
def test_synthetic_netlist_trace_matches_oracle(tile_words, stage_gates):
    flush_print(f"case: small synthetic netlist tile_words={tile_words} stage_gates={stage_gates}")
    p = small_netlist_program(gates=3000, levels=24, dffs=300, inputs=48, outputs=32, seed=4)
    lanes, cycles, seed = 1000, 40, 17
    ins, obs = netlist_inputs(p), netlist_outputs(p)
    sim = HipSimulator(p, lanes, engine="gate", tile_words=tile_words, stage_gates=stage_gates)
    sim.run(cycles, seed=seed, inputs=ins, observe=obs, trace=True)
    ref = oracle.BitParallelRef(p, lanes)
    dg, tr = ref.run(seed=seed, cycles=cycles, inputs=ins, observe=obs, digest=True, trace=True)
    got = stim.u32_rows_to_u64(sim.trace().reshape(cycles, len(obs), -1))
    assert np.array_equal(got, tr)
    assert np.array_equal(stim.u32_rows_to_u64(sim.digest()), dg)
    regs = [i for i, s in enumerate(p.signals) if s.kind == KIND_REG]
    for r in regs[::7] + [i for i, s in enumerate(p.signals) if s.name.startswith("g")][::97]:
        assert np.array_equal(stim.u32_rows_to_u64(sim.get_signal_bits(r))[: ref.words], ref.get_words(r)), p.signals[r].name


def test_gate_engine_api_matches_oracle():
    """eval/tick/tick_many/tick_from_evaluated through the exact HBM kernels and the tiled kernel."""
    p = lower.lower_module(designs.gate_chain_design(depth=16, cones=64))
    lanes = 1024
    sim = HipSimulator(p, lanes, engine="gate")
    ref = oracle.BitParallelRef(p, lanes)
    a = np.array([(l & 1) for l in range(lanes)], np.uint32).reshape(1, lanes)
    b = np.array([((l >> 1) & 1) for l in range(lanes)], np.uint32).reshape(1, lanes)
    sim.set_signal_limbs("a", a); sim.set_signal_limbs("b", b); sim.set_signal("clk", 1)
    pa = np.zeros(ref.words, np.uint64)
    pb = np.zeros(ref.words, np.uint64)
    for l in range(lanes):
        pa[l // 64] |= np.uint64(int(a[0, l]) << (l % 64))
        pb[l // 64] |= np.uint64(int(b[0, l]) << (l % 64))
    ref.set_words(p.signal_index("a"), pa); ref.set_words(p.signal_index("b"), pb)
    ref.set_signal_all(p.signal_index("clk"), 1)
    outs = [p.signal_index(f"o{c}") for c in range(64)]

    def same():
        for o in outs:
            assert np.array_equal(stim.u32_rows_to_u64(sim.get_signal_bits(o))[: ref.words], ref.get_words(o))
    for _ in range(3):
        sim.tick(); ref.tick()
    same()
    sim.tick_many(37); ref.tick_many(37)
    same()
    sim.eval_combinational(); ref.settle()
    sim.tick_from_evaluated(); ref.tick()  # reference tick = settle, edge, settle
    sim.eval_combinational()
    same()


def test_full_size_netlist_sampled_words():
    """The 20k-gate workload at full shape (65,536 lanes) for 300 cycles; four lane
    words checked against the oracle, every lane against a second GPU run with a
    different tile shape and one pass per tile (no work queue)."""
    flush_print("case: 20k-gate netlist 65536 lanes x 300 cycles")
    p = small_netlist_program()
    lanes, cycles, seed = 65536, 300, 1
    ins, obs = netlist_inputs(p), netlist_outputs(p)
    sim = HipSimulator(p, lanes, engine="gate")
    sim.run(cycles, seed=seed, inputs=ins, observe=obs)
    d1 = stim.u32_rows_to_u64(sim.digest())
    sim2 = HipSimulator(p, lanes, engine="gate", tile_words=1, tile_chunks=1)
    sim2.run(cycles, seed=seed, inputs=ins, observe=obs)
    assert np.array_equal(d1, stim.u32_rows_to_u64(sim2.digest()))
    for w64 in (0, 313, 777, 1023):
        ref = oracle.BitParallelRef(p, 64)
        dg, _ = ref.run(seed=seed, cycles=cycles, inputs=ins, observe=obs, word_base=w64, digest=True)
        assert np.array_equal(d1[:, w64], dg[:, 0]), w64


def test_full_size_netlist_every_lane_against_host_aot_engine():
    """The 20k-gate workload at full shape: 65,536 lanes x 2,000 cycles, every
    lane's output signature and every register against the host AOT engine
    (straight-line C per design, pinned to the faithful bit-parallel oracle
    in the oracle tests), in two consecutive runs."""
    from oracle import bp_aot
    flush_print("case: 20k-gate netlist 65536 lanes x 2000 cycles, every lane vs host AOT engine")
    p = small_netlist_program()
    lanes, cycles = 65536, 2000
    ins, obs = netlist_inputs(p), netlist_outputs(p)
    regs = [i for i, s in enumerate(p.signals) if s.kind == KIND_REG]
    sim = HipSimulator(p, lanes, engine="gate")
    aot = bp_aot.BitParallelAot(p, lanes, ins, obs, group_words=8, threads=max(1, min(16, os.cpu_count() or 1)))
    for seed in (7, 8):
        sim.run(cycles, seed=seed, inputs=ins, observe=obs)
        da, _ = aot.run(seed, cycles)
        assert np.array_equal(stim.u32_rows_to_u64(sim.digest()), da), seed
    for r in regs:
        assert np.array_equal(stim.u32_rows_to_u64(sim.get_signal_bits(r)), aot.get_words(r)), p.signals[r].name


def test_work_queue_error_in_an_earlier_launch_fails_the_run(monkeypatch):
    """The work queue's error flag is cleared once per run, not per launch: an
    error raised in the first of three launches (injected from the host before
    that launch) still fails the run; without the injection the run succeeds."""
    from rrtl_amd.backend import HipError
    flush_print("case: work queue error flag over three launches")
    p = small_netlist_program(gates=3000, levels=24, dffs=300, inputs=48, outputs=32, seed=5)
    ins, obs = netlist_inputs(p), netlist_outputs(p)
    monkeypatch.setenv("RRTL_HIP_INJECT_QUEUE_ERROR", "0")
    sim = HipSimulator(p, 1000, engine="gate", tile_chunks=3, ticks_per_launch=16)
    with pytest.raises(HipError, match="E_HIP_INTERNAL"):
        sim.run(40, seed=1, inputs=ins, observe=obs)
    sim.close()
    monkeypatch.delenv("RRTL_HIP_INJECT_QUEUE_ERROR")
    ok = HipSimulator(p, 1000, engine="gate", tile_chunks=3, ticks_per_launch=16)
    ok.run(40, seed=1, inputs=ins, observe=obs)
    assert ok.stats()["launches"] == 3


@pytest.mark.parametrize("chunks,ticks", [(5, 0), (3, 16), (40, 0)])
def test_work_queue_chunks_match_single_pass(chunks, ticks):
    """The work-queue schedule (each tile's cycles cut into chunks that continue
    from HBM) gives the same trace, signature and state as one pass per tile and
    as the oracle, with a stuck-at force clamped on the way."""
    flush_print(f"case: work queue chunks={chunks} ticks_per_launch={ticks}")
    p = small_netlist_program(gates=3000, levels=24, dffs=300, inputs=48, outputs=32, seed=5)
    lanes, cycles, seed = 1000, 40, 23
    ins, obs = netlist_inputs(p), netlist_outputs(p)
    forced = [i for i, s in enumerate(p.signals) if s.kind == KIND_REG][3]
    runs = []
    for ch in (1, chunks):
        sim = HipSimulator(p, lanes, engine="gate", tile_chunks=ch, ticks_per_launch=ticks)
        sim.set_force(forced, 77, 1)
        sim.run(cycles, seed=seed, inputs=ins, observe=obs, trace=True)
        runs.append(sim)
    a, b = runs
    assert np.array_equal(a.trace(), b.trace())
    assert np.array_equal(a.digest(), b.digest())
    va, _ = a.snapshot()
    vb, _ = b.snapshot()
    assert np.array_equal(va, vb)
    ref = oracle.BitParallelRef(p, lanes)
    ref.set_force(forced, 77, 1)
    dg, tr = ref.run(seed=seed, cycles=cycles, inputs=ins, observe=obs, digest=True, trace=True)
    assert np.array_equal(stim.u32_rows_to_u64(b.trace().reshape(cycles, len(obs), -1)), tr)
    assert np.array_equal(stim.u32_rows_to_u64(b.digest()), dg)


@pytest.mark.parametrize("seed,observe_regs,async_resets,post,tile_words,stage_gates", [
    (1, True, False, False, 2, 2), (2, False, False, False, 2, 2), (3, True, True, False, 2, 3),
    (4, False, True, False, 4, 2), (5, False, False, True, 2, 2), (6, True, False, False, 8, 4),
    (7, False, False, False, 1, 3)])
def test_random_gate_design_trace_matches_oracle(seed, observe_regs, async_resets, post, tile_words, stage_gates):
    """Random 1-bit designs that exercise the tile compiler's register handling
    (resets, staged shift chains, lent register slots, next values written into
    their own register, observed registers): every observed value of every
    cycle, the signature and the final registers against the bit-parallel
    oracle, over two runs in a row."""
    from designgen import random_gate_design
    flush_print(f"case: random gate design seed={seed} observe_regs={observe_regs} async={async_resets} "
                f"post={post} tile_words={tile_words} stage_gates={stage_gates}")
    p = lower.lower_module(random_gate_design(seed, observe_regs=observe_regs, async_resets=async_resets))
    lanes, cycles, rseed = 1000, 40, 11 + seed
    clk = p.signal_index("clk")
    ins = [(clk, IN_CONST, 1)] + [(i, IN_RANDOM, 0) for i, s in enumerate(p.signals)
                                  if s.kind == KIND_INPUT and i != clk]
    obs = netlist_outputs(p)
    sim = HipSimulator(p, lanes, engine="gate", tile_words=tile_words, stage_gates=stage_gates)
    ref = oracle.BitParallelRef(p, lanes)
    # the engine settles at construction (the SimBackend constructor contract,
    # crates/rrtl-sim-ir/src/lib.rs:9058); BitParallelSimulator::new does not
    # (bitparallel.rs:48-90): with async resets that the all-zero initial
    # inputs assert, the two start from different register values
    ref.settle()
    for run in range(2):
        sim.run(cycles, seed=rseed + run, inputs=ins, observe=obs, observe_post=post, trace=True)
        dg, tr = ref.run(seed=rseed + run, cycles=cycles, inputs=ins, observe=obs, observe_post=post, digest=True,
                         trace=True)
        got = stim.u32_rows_to_u64(sim.trace().reshape(cycles, len(obs), -1))
        assert np.array_equal(got, tr), f"run {run}"
        assert np.array_equal(stim.u32_rows_to_u64(sim.digest()), dg), f"run {run}"
        for r in [i for i, s in enumerate(p.signals) if s.kind == KIND_REG]:
            assert np.array_equal(stim.u32_rows_to_u64(sim.get_signal_bits(r))[: ref.words], ref.get_words(r)), \
                (run, p.signals[r].name)


def test_register_held_tables_overflow_to_memory():
    """More inputs than threads (300 > 256), more observed signals than wave 0
    holds in registers (300 > 256) and more reset-free commits than the
    threads hold (2,600 > 2,048): the overflow goes through the tables in
    memory, with the same results as the oracle."""
    flush_print("case: register-held table overflow (300 inputs, 300 outputs, 2600 flops)")
    p = small_netlist_program(gates=6000, levels=16, dffs=2600, inputs=300, outputs=300, seed=9)
    lanes, cycles, seed = 640, 20, 23
    ins, obs = netlist_inputs(p), netlist_outputs(p)
    sim = HipSimulator(p, lanes, engine="gate", tile_words=2)
    sim.run(cycles, seed=seed, inputs=ins, observe=obs, trace=True)
    ref = oracle.BitParallelRef(p, lanes)
    dg, tr = ref.run(seed=seed, cycles=cycles, inputs=ins, observe=obs, digest=True, trace=True)
    got = stim.u32_rows_to_u64(sim.trace().reshape(cycles, len(obs), -1))
    assert np.array_equal(got, tr)
    assert np.array_equal(stim.u32_rows_to_u64(sim.digest()), dg)
    for r in [i for i, s in enumerate(p.signals) if s.kind == KIND_REG][::13]:
        assert np.array_equal(stim.u32_rows_to_u64(sim.get_signal_bits(r))[: ref.words], ref.get_words(r)), p.signals[r].name
