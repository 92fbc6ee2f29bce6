"""Benchmark: lane-cycles/s of the bit-sliced gate engine on one MI355X.

Workload (the 20k-gate synthetic netlist): 20,000 two-input gates (AND/OR/XOR/
NAND 40/30/20/10 %), 64 levels, 2,000 flops, 256 primary inputs, 128 outputs,
65,536 lanes (bit-sliced, 32 lanes per u32 word), 10,000 cycles of uniform
random input bits generated on chip, every output folded into a per-lane
signature every cycle.  One step = one device-resident run of all lanes over
all cycles.  Inputs are resident in HBM before the timed region.

Printed: one JSON line with the metric, the roofline of the dominant kernel
(algorithmic bytes per lane-cycle = (inputs + 2 x flop bits + observed outputs)
/ 8 = (256 + 4000 + 128) / 8 = 548 B, divided by the kernel's average duration
measured with HIP events on the engine's stream) and the CPU baseline (the
oracle's group-major bit-parallel host engine, all allowed host threads, on a
bounded sample of the same workload), plus a bit-exactness check of sampled
lanes of the timed run against that host engine.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

B_ALG = (256 + 2 * 2000 + 128) / 8.0   # bytes per lane-cycle
HBM_PEAK_GBS = 8000.0


def build_program():
    from rrtl_amd import designs, lower
    return lower.lower_module(designs.synthetic_netlist())


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lanes", type=int, default=65536)
    ap.add_argument("--cycles", type=int, default=0,
                    help="cycles per run (gates: 10,000; rv32i: the cycle cap, 100,000)")
    ap.add_argument("--tile-words", type=int, default=0)
    ap.add_argument("--stage-gates", type=int, default=0, help="gate positions per thread per stage (0: automatic)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device with torch (scale reference for HBM counters)")
    ap.add_argument("--workload", choices=("gates", "counter", "mixed", "rv32i", "rv32i-core", "faultsim", "divergent"),
                    default="gates",
                    help="gates: the headline line; counter: the README counter, 4,096 lanes x 1,000 cycles, "
                         "out observed every cycle; mixed: the four-design batch, persistent launch vs one "
                         "launch per design; rv32i: picorv32 (imported from the reference's Verilog) running "
                         "per-lane random RV32I programs on the word engine; rv32i-core: the same programs on this "
                         "package's single-cycle RV32I core; faultsim: surrogate-triaged stuck-at fault "
                         "simulation on the 20k-gate netlist; divergent: lanes halting at data-dependent cycles, "
                         "with and without lane compaction")
    ap.add_argument("--epochs", type=int, default=6, help="faultsim: surrogate training epochs")
    ap.add_argument("--no-autotune", action="store_true", help="mixed: default launch shapes instead of the autotuner's")
    ap.add_argument("--skip-persistent", action="store_true", help="mixed: leave out the persistent work-queue kernel")
    ap.add_argument("--wave-lanes", type=int, default=0,
                    help="rv32i: active lanes per wave (0: automatic)")
    ap.add_argument("--lane-waves", type=int, default=0,
                    help="rv32i: waves per 64 lanes for the generated kernel (0: automatic, 1: off, 2, 4)")
    ap.add_argument("--codegen", choices=("auto", "on", "off"), default="auto",
                    help="rv32i: generated per-design kernel (auto/on) or the record interpreter (off)")
    args = ap.parse_args()
    if not args.cycles:
        args.cycles = {"rv32i": 100000, "rv32i-core": 100000, "counter": 1000}.get(args.workload, 10000)
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.workload == "mixed":
        return bench_mixed(args)
    if args.workload == "faultsim":
        return bench_faultsim(args)
    if args.workload in ("rv32i", "rv32i-core"):
        return bench_rv32i(args)
    if args.workload == "divergent":
        return bench_divergent(args)
    if args.workload == "counter":
        return bench_counter(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        return 2

    from rrtl_amd.backend import IN_RANDOM, HipSimulator
    from rrtl_amd.program import KIND_INPUT, KIND_OUTPUT

    p = build_program()
    ins = [(i, IN_RANDOM, 0) for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    obs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
    # the host AOT engines compile (gcc, about a minute) while the GPU part runs
    aot_pending = None if args.no_cpu_baseline else start_aot_builds(p, ins, obs)
    sim = HipSimulator(p, args.lanes, engine="gate", tile_words=args.tile_words, stage_gates=args.stage_gates)

    def step():
        sim.run(args.cycles, seed=args.seed, inputs=ins, observe=obs)

    for _ in range(args.warmup):
        step()
    sim.synchronize()
    t0 = time.perf_counter()
    kernel_ms = []
    for _ in range(args.steps):
        step()
        kernel_ms.append(sim.stats()["kernel_ms"])
    sim.synchronize()
    wall = time.perf_counter() - t0
    lane_cycles = args.lanes * args.cycles
    value = lane_cycles * args.steps / wall
    info = sim.tile_info()
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    achieved = lane_cycles * B_ALG / avg_kernel_s / 1e9

    # bit-exactness: sampled lane words replay the same history (warmup + timed
    # runs, each continuing the last) on the host engine; the last run's output
    # signatures and the final flop state must match
    check = None
    if not args.no_check:
        import oracle
        from oracle import stim
        from rrtl_amd.program import KIND_REG
        dg = stim.u32_rows_to_u64(sim.digest())
        regs = [i for i, s in enumerate(p.signals) if s.kind == KIND_REG]
        words64 = args.lanes // 64
        sample = sorted({0, words64 // 2, words64 - 8})
        def replay(base):
            ref = oracle.BitParallelFast(p, 512, threads=1)
            for _ in range(args.warmup + args.steps):
                d, _, _ = ref.run(seed=args.seed, cycles=args.cycles, inputs=ins, observe=obs, word_base=base,
                                  digest=True)
            return d, {r: ref.get_words(r) for r in regs[::97]}
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(len(sample)) as ex:
            replays = list(ex.map(replay, sample))
        ok = True
        for base, (d, words) in zip(sample, replays):
            ok &= bool(np.array_equal(d, dg[:, base:base + 8]))
            for r, w in words.items():
                got = stim.u32_rows_to_u64(sim.get_signal_bits(r))[base:base + 8]
                ok &= bool(np.array_equal(got, w))
        check = {"bit_exact": ok, "sample": f"{len(sample)} x 512 lanes x {args.warmup + args.steps} x {args.cycles} "
                                            "cycles: output signatures of the last run and sampled flops"}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = gates_cpu_baseline(p, ins, obs, args, aot_pending, int(info["gates"]))

    traffic = measured_traffic(args.lanes, args.cycles, info["tile_words"])
    line = {
        "metric": "lane-cycles/sec at 1 GPU (bit-exact) + % HBM roofline vs host batch-SIMD",
        "value": value, "unit": "lane-cycles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32 (bit-sliced, 32 lanes/word)", "data": "synthetic (on-device counter-hash inputs)",
        "config": {"workload": "20k-gate levelized netlist: 20000 gates AND/OR/XOR/NAND 40/30/20/10, 64 levels, "
                               "2000 DFFs, 256 PI, 128 PO; 65536 lanes x 10000 cycles",
                   "lanes": args.lanes, "cycles": args.cycles, "engine": "gate", "tile_words": info["tile_words"],
                   "live_slots": info["live_slots"], "levels": info["levels"], "gates_per_cycle": info["gates"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes"] if traffic else None, "traffic_detail": traffic,
                     "b_alg_bytes_per_lane_cycle": B_ALG, "kernel_ms_avg": avg_kernel_s * 1e3,
                     "note": "B_alg is the notional traffic of streaming every live bit through HBM; the state "
                             "stays in LDS, so the measured traffic is far lower and the kernel is bound on "
                             "chip (on_chip_roofs)"},
        "on_chip_roofs": on_chip_roofs(lane_cycles / avg_kernel_s, info["tile_words"]),
        "cpu_baseline": cpu,
        "check": check,
    }
    print(json.dumps(line), flush=True)
    return 0


AOT_GROUP_WORDS = (1, 8)   # the reference's 64-lane groups, and 512-lane groups (AVX-512 vectors)


def host_threads() -> int:
    return max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), os.cpu_count() or 1))


def start_aot_builds(p, ins, obs):
    """Compiles the host AOT engines (oracle/bp_aot.py) in the background."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import bp_aot
    ex = ThreadPoolExecutor(len(AOT_GROUP_WORDS))
    futs = {W: ex.submit(bp_aot.BitParallelAot, p, 64 * W, ins, obs, W, 1) for W in AOT_GROUP_WORDS}
    ex.shutdown(wait=False)
    return futs


def gates_cpu_baseline(p, ins, obs, args, aot_pending, gpu_gates: int) -> dict:
    """Host engines on bounded samples of the same workload (about
    --cpu-seconds each): the interpreted group-major engine and the AOT engine
    (straight-line C per design, the reference's BitParallelAot) at 64- and
    512-lane groups, all host threads, plus the fastest at one thread.  Each
    sample's signature is compared with a fresh GPU run of the same lanes and
    cycles.  The baseline value is the fastest all-thread engine."""
    import oracle
    from oracle import bp_aot
    from rrtl_amd.backend import HipSimulator
    threads = host_threads()

    def gpu_digest(lanes, cyc):
        g = HipSimulator(p, lanes, engine="gate", tile_words=args.tile_words, stage_gates=args.stage_gates)
        g.run(cyc, seed=args.seed, inputs=ins, observe=obs)
        d = stim_u64(g.digest())
        g.close()
        return d

    def timed(make, run, lanes):
        probe = 10
        _, secs = run(make(), probe)
        cyc = max(probe, min(args.cycles, int(probe * args.cpu_seconds / max(secs, 1e-6))))
        d, secs = run(make(), cyc)
        return {"value": lanes * cyc / secs, "sample": f"{lanes} lanes x {cyc} cycles ({secs:.1f} s)",
                "bit_exact_vs_gpu": bool(np.array_equal(d, gpu_digest(lanes, cyc)))}

    engines = {}

    def run_interp(e, cyc):
        d, _, secs = e.run(seed=args.seed, cycles=cyc, inputs=ins, observe=obs, digest=True)
        return d, secs
    interp = timed(lambda: oracle.BitParallelFast(p, args.lanes, threads=threads), run_interp, args.lanes)
    interp.update(threads=threads, gates_per_cycle=int(oracle.BitParallelFast(p, 64, threads=1).gate_count),
                  engine="interpreted gate list (group-major, 512-lane groups)")
    engines["interpreted"] = interp

    def run_aot(e, cyc):
        return e.run(args.seed, cyc)
    for W in AOT_GROUP_WORDS:
        gates = aot_pending[W].result().gates  # waits for the compile
        r = timed(lambda W=W: bp_aot.BitParallelAot(p, args.lanes, ins, obs, W, threads), run_aot, args.lanes)
        r.update(threads=threads, gates_per_cycle=gates,
                 engine=f"AOT straight-line C (gcc -O3 -march=native), {64 * W}-lane groups")
        engines[f"aot_{64 * W}"] = r
    best = max(engines, key=lambda k: engines[k]["value"])
    one = 4096
    if best.startswith("aot"):
        W = int(best[4:]) // 64
        r1 = timed(lambda: bp_aot.BitParallelAot(p, one, ins, obs, W, 1), run_aot, one)
    else:
        r1 = timed(lambda: oracle.BitParallelFast(p, one, threads=1), run_interp, one)
    r1.update(threads=1, engine=engines[best]["engine"])
    engines[best + "_1thread"] = r1
    b = engines[best]
    return {"value": b["value"], "unit": "lane-cycles/s", "cores": threads, "kind": "port", "engine": best,
            "sample": b["sample"], "bit_exact_vs_gpu": all(e["bit_exact_vs_gpu"] for e in engines.values()),
            "gpu_gates_per_cycle": gpu_gates, "engines": engines}


LDS_PEAK_BYTES_PER_CLK_CU = 128   # LDS bandwidth per CU per clock (CDNA3 / CDNA4 documentation)
VALU_WAVE_INSTS_PER_CLK_CU = 1    # four SIMDs, one wave64 VALU instruction each per four clocks
DEVICE_CUS, DEVICE_CLK_GHZ = 256, 2.4


def on_chip_roofs(lane_cycles_per_s: float, tile_words: int):
    """The units that bound the tile kernel: LDS bytes and VALU wave-instructions
    per lane-cycle from the committed SQ counters of the same kernel on this
    workload (profiles/gate_tile_kernel_20k_netlist_sq.json), times the
    measured rate, against the device's LDS bandwidth and VALU issue rate at
    2.4 GHz.  None when no such profile exists."""
    path = os.path.join(ROOT, "profiles", "gate_tile_kernel_20k_netlist_sq.json")
    if not os.path.exists(path) or tile_words != 2:
        return None
    with open(path) as f:
        m = json.load(f)
    c, lc = m["counters"], m["lanes"] * m["cycles"]
    lds_b = c["SQ_INSTS_LDS"] * m["lds_bytes_per_wave_instruction"] / lc
    valu = c["SQ_INSTS_VALU"] / lc
    lds_peak = LDS_PEAK_BYTES_PER_CLK_CU * DEVICE_CUS * DEVICE_CLK_GHZ * 1e9
    valu_peak = VALU_WAVE_INSTS_PER_CLK_CU * DEVICE_CUS * DEVICE_CLK_GHZ * 1e9
    return {
        "lds": {"achieved": lds_b * lane_cycles_per_s / 1e12, "peak": lds_peak / 1e12, "unit": "TB/s",
                "frac": lds_b * lane_cycles_per_s / lds_peak, "bytes_per_lane_cycle": lds_b,
                "bank_conflict_cycles_per_lane_cycle": c["SQ_LDS_BANK_CONFLICT"] / lc},
        "valu": {"achieved": valu * lane_cycles_per_s / 1e12, "peak": valu_peak / 1e12,
                 "unit": "T wave-instructions/s", "frac": valu * lane_cycles_per_s / valu_peak,
                 "wave_instructions_per_lane_cycle": valu},
        "profiled": m["workload"],
    }


def measured_traffic(lanes: int, cycles: int, tile_words: int):
    """HBM bytes of one launch from the committed rocprofv3 FETCH_SIZE and
    WRITE_SIZE of the timed kernel on this workload, each scaled by what the
    same counter reads for a device copy of known size, per lane-cycle, times
    this launch's lane-cycles.  None when no such profile exists."""
    path = os.path.join(ROOT, "profiles", "gate_tile_kernel_20k_netlist_hbm.json")
    if not os.path.exists(path) or tile_words != 2:
        return None
    with open(path) as f:
        m = json.load(f)
    cal = m["copy_calibration"]
    fetch_scale = cal["bytes"] / 1024.0 / cal["fetch_size_kib"]
    write_scale = cal["bytes"] / 1024.0 / cal["write_size_kib"]
    per_lane_cycle = 1024.0 * (m["fetch_size_kib"] * fetch_scale + m["write_size_kib"] * write_scale) / (
        m["lanes"] * m["cycles"])
    return {"bytes": per_lane_cycle * lanes * cycles, "bytes_per_lane_cycle": per_lane_cycle,
            "fetch_scale": fetch_scale, "write_scale": write_scale, "profiled": m["workload"]}


def bench_mixed(args) -> int:
    """Four designs (counter, 20k-gate netlist, FIFO, RV32I core; lanes uniform
    in [1k, 64k], cycles log-uniform in [1e3, 1e5], seed --seed) in one graph
    launch -- the persistent work-queue kernel, or every design's own kernels
    overlapped on their own streams -- against one launch per design on fresh
    engines with the same specs.  Device time from HIP events."""
    from rrtl_amd import autotune, workloads
    from rrtl_amd.backend import Batch
    jobs = workloads.mixed_workload(args.seed)
    # launch shape per design from the autotuner (tile width / workgroup size,
    # cycles per launch), used by every mode alike
    picks = [{} for _ in jobs]
    if not args.no_autotune:
        for k, j in enumerate(jobs):
            run_kw = {key: v for key, v in j.run.items() if key != "cycles"}
            rec = autotune.autotune(j.program, j.lanes, j.run["cycles"], j.name, engine=j.engine,
                                    run_kwargs=run_kw, setup=j.setup)
            picks[k] = {key: v for key, v in rec.options().items() if key != "engine"}
    res = {}
    modes = ("per_design", "concurrent") if args.skip_persistent else ("per_design", "persistent", "concurrent")
    for mode in modes:
        # fresh engines from the same state; each mode's run plans are compiled
        # before its timed launches (inside the batch call, outside the events)
        sims = [j.make(**pk) for j, pk in zip(jobs, picks)]
        batch = Batch([(s, j.run) for s, j in zip(sims, jobs)])
        t0 = time.perf_counter()
        st = batch.run(mode)
        res[mode] = {"kernel_ms": st["kernel_ms"], "wall_ms": (time.perf_counter() - t0) * 1e3,
                     "launches": st["launches"], "lane_cycles": st["lane_cycles"]}
        if mode == "per_design":
            res[mode]["design_ms"] = {j.name: s.stats()["kernel_ms"] for s, j in zip(sims, jobs)}
        if mode == "persistent":
            res[mode]["kernel"] = batch.describe()
        batch.close()
        for s in sims:
            s.close()
    # the batched form is the faster of the two single-graph-launch modes
    best = min((m for m in ("persistent", "concurrent") if m in res), key=lambda m: res[m]["kernel_ms"])
    lc = res[best]["lane_cycles"]
    line = {"metric": "mixed batch lane-cycles/s: one graph launch vs one launch per design",
            "value": lc / (res[best]["kernel_ms"] / 1e3), "unit": "lane-cycles/s", "n_gpus": 1,
            "higher_is_better": True, "data": "synthetic",
            "speedup_vs_per_design": res["per_design"]["kernel_ms"] / res[best]["kernel_ms"],
            "config": {"workload": "mixed: " + ", ".join(f"{j.name} {j.lanes} lanes x {j.run['cycles']} cycles"
                                                          for j in jobs), "seed": args.seed, "batched_mode": best,
                       "launch_shapes": {j.name: pk for j, pk in zip(jobs, picks)}},
            "modes": res}
    print(json.dumps(line), flush=True)
    return 0


def bench_counter(args) -> int:
    """The README counter (clk / rst / en, 8-bit count, out = count): 4,096
    lanes x 1,000 cycles, rst at cycle 0, en a random bit per lane and cycle,
    out folded into each lane's digest every cycle.  Word engine; device time
    from HIP events over K runs after W untimed ones; a fresh engine's run from
    the initial state is checked against the host batch-SIMD engine's digests.
    Launch- and latency-bound: 1,000 dependent cycles over 64 waves."""
    from rrtl_amd import workloads
    lanes = 4096
    job = workloads.counter_job(lanes, args.cycles, args.seed)
    sim = job.make()
    for _ in range(args.warmup):
        sim.run(**job.run)
    ms = 0.0
    for _ in range(args.steps):
        sim.run(**job.run)
        ms += sim.stats()["kernel_ms"]
    sim.close()
    chk = job.make()
    chk.run(**job.run)
    gpu_digest = chk.digest()
    chk.close()
    lane_cycles = lanes * args.cycles * args.steps
    value = lane_cycles / (ms / 1e3)
    b_alg = (1 + 1 + 2 * 8 + 8) / 8   # rst, en, count read and written, out
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        threads = host_threads()
        eng = oracle.SimdCpu(job.program, lanes, threads=threads)
        dg, secs = eng.run(args.seed, args.cycles, job.run["inputs"], job.run["observe"])
        cpu = {"value": lanes * args.cycles / secs, "unit": "lane-cycles/s", "cores": threads, "kind": "port",
               "sample": f"{lanes} lanes x {args.cycles} cycles, out observed every cycle ({secs * 1e3:.1f} ms)",
               "bit_exact_vs_gpu": bool(np.array_equal(dg, gpu_digest))}
    line = {"metric": "counter: lane-cycles/s, every lane's out folded into its digest every cycle",
            "value": value, "unit": "lane-cycles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms / args.steps, "higher_is_better": True, "data": "synthetic (on-device counter-hash en)",
            "dtype": "u32",
            "config": {"workload": f"README counter, {lanes} lanes x {args.cycles} cycles", "lanes": lanes,
                       "cycles": args.cycles, "engine": "word"},
            "roofline": {"bound": "hbm", "achieved": value * b_alg / 1e9, "peak": 8000.0, "unit": "GB/s",
                         "frac": value * b_alg / 8e12, "traffic": None,
                         "note": "3.25 B per lane-cycle of algorithmic bytes; the run is launch- and "
                                 "latency-bound (1,000 dependent cycles on 64 waves), not HBM-bound"},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    return 0


def bench_rv32i(args) -> int:
    """Per-lane random RV32I programs (512 instructions: ALU 60 %, load/store
    25 %, branch 15 %, in-range targets), 16,384 lanes, halting on ecall, cycle
    cap --cycles, on picorv32 (the reference's core and SoC imported from its
    Verilog) or, with --workload rv32i-core, on the single-cycle core."""
    from rrtl_amd import workloads
    if args.workload == "rv32i":
        job = workloads.picorv32_job(16384, args.cycles, args.seed)
        design = "picorv32_soc (bench/sv/picorv32.v + picorv32_soc.v, imported)"
    else:
        job = workloads.rv32i_job(16384, args.cycles, args.seed)
        design = "single-cycle rv32i core"
    sim = job.make(codegen=args.codegen, lane_waves=args.lane_waves, wave_lanes=args.wave_lanes)
    sim.run(**dict(job.run, cycles=1))   # compile the run plan
    sim = job.make(codegen=args.codegen, lane_waves=args.lane_waves, wave_lanes=args.wave_lanes)
    launch = sim.launch_info()
    t0 = time.perf_counter()
    sim.run(**job.run)
    sim.synchronize()
    wall = time.perf_counter() - t0
    st = sim.stats()
    halt = sim.halt_cycles()
    value = st["lane_cycles"] / (st["kernel_ms"] / 1e3)
    # issue-bound path: the lowered program's operations per simulated cycle
    # (every machine instruction and effect of the four blocks) against the
    # int32 VALU peak (256 CUs x 64 lanes x 2.4 GHz)
    from rrtl_amd.program import OP_LIT, OP_SIGNAL
    ops = sum(sum(1 for i in pk.instrs if i.op not in (OP_LIT, OP_SIGNAL)) + len(pk.effects)
              for b in job.program.blocks for pk in b.packets)
    valu_peak = 256 * 64 * 2.4e9 / 1e12
    achieved = value * ops / 1e12
    cpu = None
    if not args.no_cpu_baseline:
        cpu = rv32i_cpu_baseline(job, args)
    uneven = rv32i_uneven_halting(args) if args.workload == "rv32i" else None
    line = {"metric": "RV32I per-lane programs: simulated lane-cycles/s (halted lanes stop)",
            "value": value, "unit": "lane-cycles/s", "n_gpus": 1,
            "higher_is_better": True, "data": "synthetic (random RV32I programs)", "dtype": "u32",
            "wall_ms": wall * 1e3, "kernel_ms": st["kernel_ms"],
            "config": {"workload": f"{design}, 16384 lanes, 512-instruction random programs, cap {args.cycles}",
                       "codegen": args.codegen, "launch": launch,
                       "halt_cycles_min": int(halt.min()), "halt_cycles_max": int(halt.max()),
                       "lanes_at_cap": int((halt >= args.cycles).sum())},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": valu_peak, "unit": "Tops/s",
                         "frac": achieved / valu_peak, "traffic": None, "ops_per_lane_cycle": ops},
            "cpu_baseline": cpu, "uneven_halting": uneven}
    print(json.dumps(line), flush=True)
    return 0


def rv32i_uneven_halting(args) -> dict:
    """picorv32 with programs whose branches mostly jump forward (5 % of
    in-range targets backward), so that most lanes reach ecall at uneven
    cycles: the same 16,384-lane run under the cycle cap without and with lane
    compaction between launches (automatic launch lengths: 1,024 cycles, then
    each twice the last); halt cycles must agree."""
    from rrtl_amd import workloads
    out = {"workload": f"picorv32, 16384 lanes, 512-instruction programs with 5 % backward branches, cap {args.cycles}"}
    halts = {}
    for compact in (False, True):
        job = workloads.picorv32_job(16384, args.cycles, args.seed, backward=0.05)
        kw = dict(codegen=args.codegen, lane_waves=args.lane_waves, wave_lanes=args.wave_lanes)
        if compact:
            kw.update(compact_lanes=True)
        sim = job.make(**kw)
        sim.run(**dict(job.run, cycles=1))   # compile the run plan
        sim = job.make(**kw)
        sim.run(**job.run)
        st = sim.stats()
        h = sim.halt_cycles()
        halts[compact] = h
        key = "compacted" if compact else "uncompacted"
        out[key] = {"kernel_ms": st["kernel_ms"], "lane_cycles_per_s": st["lane_cycles"] / (st["kernel_ms"] / 1e3),
                    "launches": st["launches"]}
        sim.close()
    h = halts[False]
    out.update(halt_cycles_equal=bool(np.array_equal(halts[False], halts[True])),
               halt_cycles_min=int(h.min()), halt_cycles_median=int(np.median(h)), halt_cycles_max=int(h.max()),
               lanes_at_cap=int((h >= args.cycles).sum()),
               compaction_speedup=out["uncompacted"]["kernel_ms"] / out["compacted"]["kernel_ms"])
    return out


def bench_divergent(args) -> int:
    """The divergent countdown (rrtl_amd/divergent.py, the shape of the
    reference's divergent-batch example): 2,097,152 lanes, each halting when its
    counter, loaded with its run length, reaches 0; 1/8 of the lanes (scattered)
    run 20,000 cycles, the rest 1,000.  Simulated lane-cycles (halted lanes stop
    counting) per second of device time, with lane compaction between
    512-cycle launches and without it; both runs give the same results."""
    from rrtl_amd import divergent
    from rrtl_amd.backend import HipSimulator
    p = divergent.program()
    lanes, short, long, din = 1 << 21, 1000, 20000, 0x9E3779B1
    seeds = divergent.run_lengths(lanes, short, long, args.seed)
    ins = divergent.run_inputs(p)
    res, finals = {}, {}
    for compact in (False, True):
        def make():
            sim = HipSimulator(p, lanes, engine="word", compact_lanes=compact, ticks_per_launch=512)
            sim.set_signal_limbs("seed", seeds.reshape(1, -1))
            sim.set_signal_limbs("din", np.full((1, lanes), din, dtype=np.uint32))
            return sim
        warm = make()
        warm.run(1, inputs=ins, halt="done")  # compiles the run plan
        warm.close()
        sim = make()
        sim.run(long + 2, inputs=ins, halt="done")
        st = sim.stats()
        res["compacted" if compact else "uncompacted"] = {
            "kernel_ms": st["kernel_ms"], "launches": st["launches"], "lane_cycles": st["lane_cycles"],
            "lane_cycles_per_s": st["lane_cycles"] / (st["kernel_ms"] / 1e3)}
        finals[compact] = (sim.halt_cycles(), sim.get_signal_limbs("acc"))
        sim.close()
    same = all(np.array_equal(a, b) for a, b in zip(finals[False], finals[True]))
    sample = np.arange(0, lanes, 997)
    closed = bool(np.array_equal(finals[True][1][0][sample], divergent.closed_form(seeds[sample], din)))
    c, u = res["compacted"], res["uncompacted"]
    line = {"metric": "divergent countdown: simulated lane-cycles/s (halted lanes stop), lane compaction",
            "value": c["lane_cycles_per_s"], "unit": "lane-cycles/s", "n_gpus": 1, "higher_is_better": True,
            "data": "synthetic", "dtype": "u32",
            "speedup_vs_uncompacted": u["kernel_ms"] / c["kernel_ms"],
            "config": {"workload": f"countdown, {lanes} lanes, 1/8 scattered {long} cycles, rest {short} cycles, "
                                   f"512-cycle launches", "seed": args.seed},
            "modes": res, "check": {"compacted_equals_uncompacted": same, "closed_form_sample": closed}}
    print(json.dumps(line), flush=True)
    return 0


def rv32i_cpu_baseline(job, args) -> dict:
    """The host batch-SIMD engine on all 16,384 lanes for a bounded number of
    cycles (no halting: the reference's engine steps every lane), and the GPU
    on the same lanes and cycles for a bit-exactness check of registers,
    memory, pc and trap."""
    import oracle
    p = job.program
    threads = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), os.cpu_count() or 1))
    ins = job.run["inputs"]

    class Images:   # captures the job's memory images
        lanes = job.lanes

        def set_memory_limbs(self, name, v):
            self.mem = v
    img = Images()
    job.setup(img)

    def cpu_run(cycles):
        eng = oracle.SimdCpu(p, job.lanes, threads=threads)
        eng.set_memory_limbs(p.memory_index("mem"), img.mem)
        _, secs = eng.run(job.run["seed"], cycles, ins, [])
        return eng, secs
    probe = 10
    _, secs = cpu_run(probe)
    cyc = max(probe, min(args.cycles, int(probe * args.cpu_seconds / max(secs, 1e-6))))
    eng, secs = cpu_run(cyc)
    gpu = job.make(codegen=args.codegen)
    gpu.run(cyc, seed=job.run["seed"], inputs=ins)
    same = all(np.array_equal(gpu.get_memory_limbs(m), eng.get_memory_limbs(p.memory_index(m))) for m in job.check_mems)
    for name in job.check_sigs:
        same &= bool(np.array_equal(gpu.get_signal_limbs(name), eng.get_signal_limbs(p.signal_index(name))))
    gpu.close()
    return {"value": job.lanes * cyc / secs, "unit": "lane-cycles/s", "cores": threads, "kind": "port",
            "sample": f"{job.lanes} lanes x {cyc} cycles, every lane stepped ({secs:.1f} s)", "bit_exact_vs_gpu": same,
            "checked": list(job.check_mems + job.check_sigs)}


def bench_faultsim(args) -> int:
    """Stuck-at faults on all 2,000 flops of the 20k-gate netlist (4,000
    faults) x 256 random 256-cycle stimuli = 1,024,000 candidates.  The
    surrogate (512-1024-1024-1024-1, bf16 matrix-core kernels, fp32
    accumulation) is trained on the GPU on 2,048,000 exactly simulated
    candidates of 512 other stimuli; on the evaluation pool the plain path
    (every candidate exact) is the reference and each method's GPU time is the
    device time of its launches (features, scoring, exact stages)."""
    from rrtl_amd import triage
    p = build_program()
    out = triage.run_experiment(p, epochs=args.epochs, pmins=(1e-3,), seed=args.seed,
                                log=lambda m: print(m, file=sys.stderr, flush=True))
    par = out["triage_parity"]
    pred = out["predict"]
    line = {"metric": "surrogate-triaged fault simulation: GPU time to the plain path's coverage (fraction of plain)",
            "value": par["time_vs_plain"], "unit": "fraction of plain-path GPU time", "n_gpus": 1,
            "higher_is_better": False, "data": "synthetic (20k-gate netlist, random stimuli)", "dtype": "u32 / bf16",
            "plain_ms": out["plain"]["gpu_ms"], "triage_ms": par["gpu_ms"],
            "coverage_vs_plain": par["coverage_vs_plain"], "unconfirmed_detections": par["unconfirmed_detections"],
            "roofline": {"bound": "mfma", "kernel": "gemm_nt_kernel (surrogate scoring)",
                         "achieved": pred["gemm_tflops"], "peak": 2500.0, "unit": "TFLOP/s",
                         "frac": pred["gemm_tflops"] / 2500.0, "traffic": None},
            "config": {"workload": "faultsim: 4000 flop stuck-at faults x 256 stimuli x 256 cycles on the 20k-gate "
                                   "netlist; surrogate trained on 512 other stimuli", "epochs": args.epochs},
            "methods": {k: out[k] for k in ("plain", "triage_parity", "triage_deferred", "rules_dropping", "dropping")},
            "triage_no_tail": out["triage"], "train": out["train"], "predict": pred,
            "features_ms": out["features_ms"], "eligible_candidates": out["eligible_candidates"]}
    print(json.dumps(line, default=float), flush=True)
    return 0


def calibrate_copy(nbytes: int) -> int:
    """A plain device-to-device copy of known size (the runtime's copy kernel):
    rocprofv3's FETCH_SIZE / WRITE_SIZE for it give the counters' scale against
    algorithmic bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"calibrate_copy_bytes": nbytes, "copies": 3}), flush=True)
    return 0


def stim_u64(a):
    from oracle import stim
    return stim.u32_rows_to_u64(a)


if __name__ == "__main__":
    sys.exit(main())
