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


def make(job):
    sim = HipSimulator(job[0], job[1], engine=job[2])
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
