"""The autotuner on the device: every candidate launch shape runs, the pick is
stored and looked up like the reference's GPU autotune recommendations
(crates/rrtl-gpu-sim/src/lib.rs:55-190), and an engine built from the pick
reproduces the oracle's results."""
import numpy as np
import pytest

import oracle
from oracle import stim
from rrtl_amd import designs, divergent, lower
from rrtl_amd.autotune import (autotune, find_recommendation, load_recommendations, options_from_recommendations,
                               save_recommendations)
from rrtl_amd.backend import IN_RANDOM, HipSimulator
from rrtl_amd.program import KIND_INPUT, KIND_OUTPUT
from conftest import flush_print

pytestmark = pytest.mark.gpu


def test_gate_autotune_pick_matches_bit_parallel_oracle(tmp_path):
    flush_print("case: autotune gate engine, synthetic netlist 3000 gates, 1000 lanes")
    p = lower.lower_module(designs.synthetic_netlist(gates=3000, levels=24, dffs=300, inputs=48, outputs=32, seed=4))
    lanes, cycles, seed = 1000, 300, 23
    ins = [(i, IN_RANDOM, 0) for i, s in enumerate(p.signals) if s.kind == KIND_INPUT]
    obs = [i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]
    kw = {"seed": seed, "inputs": ins, "observe": obs}
    rec = autotune(p, lanes, cycles, "synthetic-3000", run_kwargs=kw, repeats=1)
    assert rec.engine == "gate"
    ran = [c for c in rec.candidates if "error" not in c]
    assert len(rec.candidates) == 12 and len(ran) == 12, rec.candidates
    assert rec.autotune_metric_ns == min(c["ns_per_lane_cycle"] for c in ran)
    assert rec.tile_words in (1, 2, 4, 8) and rec.stage_gates in (2, 3, 4)

    path = str(tmp_path / "tune.json")
    save_recommendations(path, [rec])
    items = load_recommendations(path)
    assert find_recommendation(items, "synthetic-3000", lanes, cycles) == rec
    opts = options_from_recommendations(items, "synthetic-3000", 2 * lanes, cycles, mode="nearest")
    assert opts == rec.options()

    sim = HipSimulator(p, lanes, **opts)
    sim.run(cycles, **kw)
    dg, _ = oracle.BitParallelRef(p, lanes).run(seed=seed, cycles=cycles, inputs=ins, observe=obs, digest=True)
    assert np.array_equal(stim.u32_rows_to_u64(sim.digest()), dg)
    sim.close()


def test_word_autotune_halting_run_pick_matches_closed_form():
    flush_print("case: autotune word engine, halting countdown, 20000 lanes")
    p = divergent.program()
    lanes, din, cap = 20000, 0x9E3779B1, 2600
    seeds = divergent.run_lengths(lanes, short=50, long=1200, seed=5)

    def setup(sim):
        sim.set_signal("seed", seeds)
        sim.set_signal("din", [din] * lanes)

    kw = {"inputs": divergent.run_inputs(p), "observe": ["acc"], "halt": "done"}
    rec = autotune(p, lanes, cap, "countdown", run_kwargs=kw, setup=setup, repeats=1)
    assert rec.engine == "word"
    compacting = [c for c in rec.candidates if c.get("compact_lanes")]
    assert len(compacting) == 6 and all("error" not in c for c in compacting), rec.candidates
    assert all("error" not in c for c in rec.candidates), rec.candidates
    assert any(c.get("lane_waves", 1) > 1 for c in rec.candidates)

    sim = HipSimulator(p, lanes, **rec.options())
    setup(sim)
    sim.run(cap, **kw)
    assert np.array_equal(sim.halt_cycles(), seeds.astype(np.int64) + 1)
    assert np.array_equal(sim.get_signal_limbs("acc")[0], divergent.closed_form(seeds, din))
    flush_print(f"pick: {rec.options()} {rec.autotune_metric_ns:.4f} ns/lane-cycle")
    sim.close()
