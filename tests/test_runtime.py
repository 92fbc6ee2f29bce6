"""The lane-sharded runtime harness (rrtl_amd/runtime.py) on the CPU: the
same harness the HIP arm uses, with the packed oracle as every shard's engine.
Sharded runs equal one unsharded engine; the checkpoint cadence follows the
reference's rule (crates/rrtl-runtime/src/lib.rs:7450-7521); checkpoints
round-trip through JSON and restore into a fresh runtime that then continues
exactly like the uninterrupted run; bad checkpoints are rejected."""
import numpy as np
import pytest

import oracle
from designgen import random_design
from rrtl_amd import lower, runtime
from rrtl_amd.program import KIND_INPUT


class OracleShard:
    """The runtime's engine methods over the packed oracle (test side only)."""

    def __init__(self, p, lanes):
        self.ref = oracle.PackedRef(p, lanes)

    def snapshot_storage(self): return self.ref.snapshot()
    def restore_storage(self, v, m): self.ref.restore(v, m)
    def set_signal_limbs(self, sig, limbs): self.ref.set_signal_limbs(sig, limbs)
    def get_signal_limbs(self, sig): return self.ref.get_signal_limbs(sig)
    def set_memory_limbs(self, mem, vals): self.ref.set_memory_limbs(mem, vals)
    def get_memory_limbs(self, mem): return self.ref.get_memory_limbs(mem)
    def eval_combinational(self): self.ref.eval_combinational()
    def tick_many(self, n): self.ref.tick_many(n)


def make(p, sizes):
    shards = runtime.plan_shards(sum(sizes), sizes, runtime.HipOptions())
    return runtime.LaneShardedRuntime(p, shards, engine_factory=lambda prog, s: OracleShard(prog, s.lanes))


def seed_state(rt, p, lanes, seed):
    rng = np.random.default_rng(seed)
    for i, s in enumerate(p.signals):
        rt.set_signal_limbs(i, rng.integers(0, 2**32, size=(s.limbs, lanes), dtype=np.uint32))


@pytest.mark.parametrize("steps,every,initial,final,want", [
    (10, 4, False, True, [("cadence", 4), ("cadence", 8), ("final", 10)]),
    (8, 4, False, True, [("cadence", 4), ("cadence", 8)]),
    (8, 4, True, False, [("initial", 0), ("cadence", 4), ("cadence", 8)]),
    (3, 5, False, True, [("final", 3)]),
    (0, 5, False, True, [("final", 0)]),
    (0, 5, True, True, [("initial", 0)]),
])
def test_checkpoint_cadence_follows_the_reference(steps, every, initial, final, want):
    p = lower.lower_module(random_design(2))
    rt = make(p, [5, 3])
    seen = []
    rep = rt.tick_many_with_checkpoints(steps, runtime.RuntimeCheckpointCadence(every, initial, final),
                                        lambda r, n, cp: seen.append((r, n)))
    assert seen == want
    assert (rep.requested_steps, rep.completed_steps, rep.checkpoints_emitted) == (steps, steps, len(want))


def test_sharded_equals_unsharded_and_checkpoints_resume():
    p = lower.lower_module(random_design(6))
    lanes = 19
    a, b = make(p, [7, 5, 7]), make(p, [lanes])
    seed_state(a, p, lanes, 3)
    seed_state(b, p, lanes, 3)
    cps = {}
    a.tick_many_with_checkpoints(23, runtime.RuntimeCheckpointCadence(8), lambda r, n, cp: cps.__setitem__(n, cp))
    b.tick_many(23)
    for i in range(len(p.signals)):
        assert np.array_equal(a.get_signal_limbs(i), b.get_signal_limbs(i)), p.signals[i].name
    for m in range(len(p.memories)):
        assert np.array_equal(a.get_memory_limbs(m), b.get_memory_limbs(m))
    # resume from the step-8 checkpoint (through JSON) in a fresh runtime
    cp = runtime.RuntimeCheckpoint.from_json(cps[8].to_json())
    c = make(p, [7, 5, 7])
    c.restore_checkpoint(cp)
    c.tick_many(15)
    for i in range(len(p.signals)):
        assert np.array_equal(c.get_signal_limbs(i), b.get_signal_limbs(i)), p.signals[i].name


def test_bad_checkpoints_are_rejected():
    p = lower.lower_module(random_design(6))
    rt = make(p, [4, 4])
    cp = rt.checkpoint()
    text = cp.to_json().replace('"format_version": 1', '"format_version": 99')
    with pytest.raises(runtime.RuntimeError_) as e:
        runtime.RuntimeCheckpoint.from_json(text)
    assert e.value.code == "E_RUNTIME_CHECKPOINT_VERSION"
    with pytest.raises(runtime.RuntimeError_) as e:
        make(p, [5, 3]).restore_checkpoint(cp)
    assert e.value.code == "E_RUNTIME_SNAPSHOT_SHAPE"
    other = lower.lower_module(random_design(7))
    with pytest.raises(runtime.RuntimeError_):
        make(other, [4, 4]).restore_checkpoint(cp)
    with pytest.raises(runtime.RuntimeError_) as e:
        runtime.plan_shards(8, [4, 3], runtime.HipOptions())
    assert e.value.code == "E_RUNTIME_SHARDS"
    with pytest.raises(runtime.RuntimeError_) as e:
        rt.tick_many_with_checkpoints(4, runtime.RuntimeCheckpointCadence(0), lambda *a: None)
    assert e.value.code == "E_RUNTIME_CHECKPOINT_CADENCE"
