"""The HIP arm of the runtime (rrtl_amd/runtime.py) on the GPU: a batch split
into lane shards, each on its own HIP engine, runs with checkpoints at a
cadence and equals one unsharded HIP engine and the packed oracle; a
checkpoint restored into a fresh sharded runtime continues identically."""
import numpy as np
import pytest

import oracle
from conftest import flush_print
from designgen import random_design
from rrtl_amd import lower, picorv32 as pv, runtime

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("engine,codegen", [("word", "on"), ("word", "off")])
def test_sharded_hip_runtime_with_checkpoints(engine, codegen):
    flush_print(f"case: sharded runtime engine={engine} codegen={codegen}")
    p = lower.lower_module(random_design(9))
    lanes = 300
    opts = runtime.HipOptions(engine=engine, codegen=codegen)
    rt = runtime.LaneShardedRuntime(p, runtime.plan_shards(lanes, [128, 100, 72], opts))
    one = runtime.LaneShardedRuntime(p, runtime.plan_shards(lanes, [lanes], opts))
    ref = oracle.PackedRef(p, lanes)
    rng = np.random.default_rng(1)
    for i, s in enumerate(p.signals):
        v = rng.integers(0, 2**32, size=(s.limbs, lanes), dtype=np.uint32)
        rt.set_signal_limbs(i, v)
        one.set_signal_limbs(i, v)
        ref.set_signal_limbs(i, v)
    cps = {}
    rep = rt.tick_many_with_checkpoints(40, runtime.RuntimeCheckpointCadence(16, include_initial=True),
                                        lambda r, n, cp: cps.__setitem__(n, cp.to_json()))
    assert rep.checkpoints_emitted == 4 and sorted(cps) == [0, 16, 32, 40]
    one.tick_many(40)
    ref.tick_many(40)
    for i in range(len(p.signals)):
        got = rt.get_signal_limbs(i)
        assert np.array_equal(got, one.get_signal_limbs(i)), p.signals[i].name
        assert np.array_equal(got, ref.get_signal_limbs(i)), p.signals[i].name
    fresh = runtime.LaneShardedRuntime(p, runtime.plan_shards(lanes, [128, 100, 72], opts))
    fresh.restore_checkpoint(runtime.RuntimeCheckpoint.from_json(cps[16]))
    fresh.tick_many(24)
    for i in range(len(p.signals)):
        assert np.array_equal(fresh.get_signal_limbs(i), ref.get_signal_limbs(i)), p.signals[i].name
    for r in (rt, one, fresh):
        r.close()


def test_picorv32_shards_reach_the_known_answer():
    """picorv32 sum program over four shards with a checkpoint restore in the
    middle of the run: result 15 on every lane."""
    flush_print("case: sharded runtime picorv32")
    p = pv.program()
    lanes = 1024
    opts = runtime.HipOptions(engine="word")
    shards = runtime.plan_shards(lanes, [256, 256, 384, 128], opts)
    rt = runtime.LaneShardedRuntime(p, shards)
    rt.set_memory_limbs(p.memory_index("mem"), pv.images([pv.sum_program()] * lanes))
    clk, rst = p.signal_index("clk"), p.signal_index("resetn")
    rt.set_signal_limbs(clk, np.ones((1, lanes), np.uint32))
    rt.set_signal_limbs(rst, np.zeros((1, lanes), np.uint32))
    rt.tick_many(8)
    rt.set_signal_limbs(rst, np.ones((1, lanes), np.uint32))
    cps = {}
    rt.tick_many_with_checkpoints(3000, runtime.RuntimeCheckpointCadence(1000), lambda r, n, cp: cps.__setitem__(n, cp))
    assert set(rt.get_signal_limbs(p.signal_index("result"))[0]) == {15}
    again = runtime.LaneShardedRuntime(p, shards)
    again.restore_checkpoint(cps[1000])
    again.tick_many(2000)
    assert set(again.get_signal_limbs(p.signal_index("result"))[0]) == {15}
    assert set(again.get_signal_limbs(p.signal_index("done"))[0]) == {1}
    rt.close()
    again.close()
