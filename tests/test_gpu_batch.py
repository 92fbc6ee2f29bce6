"""Multi-design batches: one persistent graph-captured launch, or every
design's own kernels overlapped on their own streams (captured into one graph
or issued directly), must give every design exactly what its own run gives
(digests, halt cycles, final state)."""
import numpy as np
import pytest

import oracle
from rrtl_amd import designs, lower
from rrtl_amd.backend import IN_CONST, IN_PULSE, IN_RANDOM, Batch, HipSimulator
from rrtl_amd.program import KIND_INPUT, KIND_OUTPUT
from conftest import flush_print
from designgen import random_design

pytestmark = pytest.mark.gpu


def mixed_jobs():
    jobs = []
    p = lower.lower_module(designs.counter())
    jobs.append((p, 2000, "word", dict(cycles=3000, seed=1, observe=[p.signal_index("out")],
                                       inputs=[(p.signal_index("clk"), IN_CONST, 1),
                                               (p.signal_index("rst"), IN_PULSE, oracle.pulse_arg(1, 1, 0)),
                                               (p.signal_index("en"), IN_RANDOM, 0)])))
    g = lower.lower_module(designs.synthetic_netlist(gates=4000, levels=32, dffs=400, inputs=64, outputs=32, seed=7))
    jobs.append((g, 4096, "gate", dict(cycles=500, seed=2,
                                       inputs=[(i, IN_RANDOM, 0) for i, s in enumerate(g.signals) if s.kind == KIND_INPUT],
                                       observe=[i for i, s in enumerate(g.signals) if s.kind == KIND_OUTPUT])))
    f = lower.lower_module(designs.ready_valid_fifo())
    jobs.append((f, 1000, "word", dict(cycles=700, seed=3,
                                       inputs=[(f.signal_index("clk"), IN_CONST, 1),
                                               (f.signal_index("rst"), IN_PULSE, oracle.pulse_arg(1, 1, 0)),
                                               (f.signal_index("in_valid"), IN_RANDOM, 0),
                                               (f.signal_index("out_ready"), IN_RANDOM, 0),
                                               (f.signal_index("in_data"), IN_RANDOM, 0)],
                                       observe=[f.signal_index(n) for n in ("out_valid", "out_data", "count")])))
    from rrtl_amd import workloads
    cpu = workloads.rv32i_job(700, 900, 5, n_insns=96, distinct=64)
    jobs.append((cpu.program, cpu.lanes, "word", dict(cpu.run, observe=[cpu.program.signal_index("pc")]), cpu.setup))
    r = lower.lower_module(random_design(77, async_resets=False))
    jobs.append((r, 300, "word", dict(cycles=150, seed=4,
                                      inputs=[(i, IN_RANDOM, 0) for i, s in enumerate(r.signals) if s.kind == KIND_INPUT],
                                      observe=[i for i, s in enumerate(r.signals) if s.kind != KIND_INPUT])))
    return jobs


def make(job, **opts):
    sim = HipSimulator(job[0], job[1], engine=job[2], **opts)
    if len(job) > 4 and job[4]:
        job[4](sim)
    return sim


@pytest.mark.parametrize("mode", ["persistent", "concurrent", "streams"])
def test_batch_matches_individual_runs(mode):
    flush_print(f"case: mixed batch {mode} vs individual")
    jobs = mixed_jobs()
    solo = [make(j) for j in jobs]
    both = [make(j) for j in jobs]
    batch = Batch([(s, j[3]) for s, j in zip(both, jobs)])
    for rep in range(2):  # the second launch replays the captured graph
        for s, j in zip(solo, jobs):
            s.run(**j[3])
        st = batch.run(mode)
        if mode == "streams":
            assert st["launches"] >= len(jobs)
        else:
            assert st["launches"] == 1
        assert st["lane_cycles"] == sum(s.stats()["lane_cycles"] for s in solo)
        for a, b, j in zip(solo, both, jobs):
            assert np.array_equal(a.digest(), b.digest()), (rep, a.kind)
            if j[3].get("halt"):
                assert np.array_equal(a.halt_cycles(), b.halt_cycles()), rep
            va, ma = a.snapshot()
            vb, mb = b.snapshot()
            assert np.array_equal(va, vb) and np.array_equal(ma, mb), (rep, a.kind)


def test_batch_per_design_mode_matches():
    jobs = mixed_jobs()[:2]
    sims = [make(j) for j in jobs]
    ref = [make(j) for j in jobs]
    batch = Batch([(s, j[3]) for s, j in zip(sims, jobs)])
    st = batch.run("per_design")
    assert st["launches"] == 2
    for s, j in zip(ref, jobs):
        s.run(**j[3])
    for a, b in zip(sims, ref):
        assert np.array_equal(a.digest(), b.digest())


def c4_oracle_jobs():
    """The mixed configuration's four designs at sizes the CPU oracles cover:
    the counter, the 20k-gate netlist, the FIFO and picorv32 (halting)."""
    from rrtl_amd import picorv32 as pv, workloads
    import random
    jobs = []
    c = workloads.counter_job(1500, 800, seed=11)
    jobs.append((c.program, c.lanes, "word", c.run, None))
    g = workloads.netlist_job(1024, 120, seed=12)
    jobs.append((g.program, g.lanes, "gate", g.run, None))
    f = workloads.fifo_job(900, 600, seed=13)
    jobs.append((f.program, f.lanes, "word", f.run, None))
    p = pv.program()
    rng = random.Random(14)
    progs = [pv.random_program(rng, n=64, backward=0.2) for _ in range(96)]
    jobs.append((p, 96, "word", dict(cycles=2500, seed=15, inputs=pv.inputs(p), observe=[], halt="trap"),
                 lambda sim: sim.set_memory_limbs("mem", pv.images(progs)), progs))
    return jobs


@pytest.mark.parametrize("mode,compact", [("persistent", False), ("concurrent", False), ("concurrent", True)])
def test_c4_designs_in_one_batch_match_the_cpu_oracles(mode, compact):
    """Every design of the batch against the CPU oracles: the word designs'
    per-lane digests and final state against the packed oracle replaying the
    same stimulus, the netlist's signatures against the host bit-parallel
    engine, picorv32's halt cycles against the packed oracle and its
    registers and memory against the RV32I ISS.  With ``compact`` picorv32
    packs its live lanes between 200-cycle launches inside the captured graph."""
    from harness import packed_run
    from oracle import stim
    from test_gpu_picorv32 import oracle_halts
    from rrtl_amd import picorv32 as pv
    flush_print(f"case: C4 designs batch {mode} compact={compact} vs CPU oracles")
    jobs = c4_oracle_jobs()
    sims = [make(j[:5]) for j in jobs[:3]]
    sims.append(make(jobs[3][:5], **(dict(compact_lanes=True, ticks_per_launch=200) if compact else {})))
    batch = Batch([(s, j[3]) for s, j in zip(sims, jobs)])
    batch.run(mode)
    # counter and FIFO: packed oracle replay of the same run
    for k in (0, 2):
        p, lanes, _, run, _ = jobs[k]
        ref, _, dg = packed_run(p, lanes, run["seed"], run["cycles"], run["inputs"], run["observe"])
        assert np.array_equal(sims[k].digest(), dg), p.top
        v, m = sims[k].snapshot()
        vr, mr = ref.snapshot()
        assert np.array_equal(v, vr) and np.array_equal(m, mr), p.top
    # netlist: host bit-parallel engine signatures and flops
    p, lanes, _, run, _ = jobs[1]
    ref = oracle.BitParallelFast(p, lanes, threads=4)
    dg, _, _ = ref.run(seed=run["seed"], cycles=run["cycles"], inputs=run["inputs"], observe=run["observe"], digest=True)
    assert np.array_equal(stim.u32_rows_to_u64(sims[1].digest()), dg)
    from rrtl_amd.program import KIND_REG
    for r in [i for i, s in enumerate(p.signals) if s.kind == KIND_REG][::37]:
        assert np.array_equal(stim.u32_rows_to_u64(sims[1].get_signal_bits(r)), ref.get_words(r))
    # picorv32: halt cycles (packed oracle) and architectural state (ISS)
    p, lanes, _, run, _, progs = jobs[3]
    halt = sims[3].halt_cycles()
    assert np.array_equal(halt, oracle_halts(p, progs, run["cycles"]))
    regs = sims[3].get_memory_limbs("cpu.cpuregs")[:, :, 0]
    mem = sims[3].get_memory_limbs("mem")[:, :, 0]
    for l, pr in enumerate(progs):
        if halt[l] < run["cycles"]:
            s = pv.iss_run(pr, 100_000)
            assert [int(x) for x in regs[l]] == s.regs and [int(x) for x in mem[l]] == s.mem, l
    batch.close()
