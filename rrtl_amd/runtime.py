"""Runtime integration: the HIP arm of the reference's lane runtime.

The reference's runtime (crates/rrtl-runtime/src/lib.rs) drives batch engines
through one dispatch surface, ``RuntimeEngine`` (:7785-8100: health,
snapshot/restore of ``PackedSimulatorStorage``, u128 and limb signal access,
memory limbs, eval/tick/tick_many), chooses them by ``RuntimeBackend``
(:29-46), splits a batch into lane shards (``RuntimeShardInfo`` :168,
``DistributedRuntime`` :6785-6880) and checkpoints at a step cadence
(``RuntimeCheckpointCadence`` / ``tick_many_with_checkpoints`` :1565-1655,
:7134-7237; ``RuntimeCheckpoint`` :579 with a format version).

Here ``HipRuntimeEngine`` is the ``RuntimeEngine::Hip`` arm over the C ABI of
include/rrtl_hip.h, and ``LaneShardedRuntime`` splits a batch into contiguous
lane ranges, one engine per shard (on one GPU here; a shard per GPU is the
multi-GPU layout), with snapshots, checkpoints (JSON, versioned, validated
before a restore) and a step cadence.  Shard engines come from a factory, so
any object with the engine methods can serve as a shard (the tests run the
same harness over the CPU oracle).
"""
from __future__ import annotations

import base64
import json
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from .backend import HipSimulator
from .program import Program

CHECKPOINT_FORMAT_VERSION = 1


class RuntimeError_(ValueError):
    """A runtime error with an ``E_RUNTIME_*`` code."""

    def __init__(self, code: str, message: str):
        super().__init__(f"{code}: {message}")
        self.code = code


@dataclass(frozen=True)
class HipOptions:
    """The ``RuntimeBackend::Hip`` arm's options (rrtl_hip_options)."""
    engine: str = "auto"
    device: int = 0
    codegen: str = "auto"
    tile_words: int = 0
    ticks_per_launch: int = 0
    wave_lanes: int = 0
    compact_lanes: bool = False

    def to_dict(self) -> dict:
        return dict(self.__dict__)


@dataclass(frozen=True)
class RuntimeShardInfo:
    worker_id: str
    node: str
    backend: HipOptions
    start_lane: int
    lanes: int


class HipRuntimeEngine:
    """``RuntimeEngine`` over one HIP engine (lanes of one shard)."""

    def __init__(self, program: Program, lanes: int, options: HipOptions = HipOptions()):
        self.program = program
        self.options = options
        self.sim = HipSimulator(program, lanes, engine=options.engine, device=options.device,
                                codegen=options.codegen, tile_words=options.tile_words,
                                ticks_per_launch=options.ticks_per_launch, wave_lanes=options.wave_lanes,
                                compact_lanes=options.compact_lanes)
        self.lanes = lanes

    def health(self) -> dict:
        return {"backend": "hip", "kind": self.sim.kind, "lanes": self.lanes}

    def snapshot_storage(self) -> Tuple[np.ndarray, np.ndarray]:
        return self.sim.snapshot()

    def restore_storage(self, values: np.ndarray, memories: np.ndarray) -> None:
        self.sim.restore(values, memories)

    def set_signal(self, sig: int, lane_values: Sequence[int]) -> None:
        if self.program.signals[sig].width > 128:
            raise RuntimeError_("E_RUNTIME_WIDE_SIGNAL", "signals wider than 128 bits need the limbs API")
        self.sim.set_signal(sig, lane_values)

    def set_signal_limbs(self, sig: int, limbs: np.ndarray) -> None:
        self.sim.set_signal_limbs(sig, limbs)

    def get_signal(self, sig: int) -> List[int]:
        if self.program.signals[sig].width > 128:
            raise RuntimeError_("E_RUNTIME_WIDE_SIGNAL", "signals wider than 128 bits need the limbs API")
        return self.sim.get_signal(sig)

    def get_signal_limbs(self, sig: int) -> np.ndarray:
        return self.sim.get_signal_limbs(sig)

    def set_memory_limbs(self, mem: int, values: np.ndarray) -> None:
        self.sim.set_memory_limbs(mem, values)

    def get_memory_limbs(self, mem: int) -> np.ndarray:
        return self.sim.get_memory_limbs(mem)

    def eval_combinational(self) -> None:
        self.sim.eval_combinational()

    def tick(self) -> None:
        self.sim.tick()

    def tick_many(self, steps: int) -> None:
        self.sim.tick_many(steps)

    def close(self) -> None:
        self.sim.close()


@dataclass
class RuntimeShardSnapshot:
    shard: RuntimeShardInfo
    values: np.ndarray     # word-major (offset + limb) * lanes + lane, the shard's lanes
    memories: np.ndarray   # lane-major per lane


@dataclass
class RuntimeSnapshot:
    total_lanes: int
    program_top: str
    signal_words_per_lane: int
    memory_words_per_lane: int
    shards: List[RuntimeShardSnapshot]


@dataclass
class RuntimeCheckpoint:
    format_version: int
    module_name: str
    topology: List[RuntimeShardInfo]
    snapshot: RuntimeSnapshot

    def to_json(self) -> str:
        def arr(a: np.ndarray) -> str:
            return base64.b64encode(np.ascontiguousarray(a, dtype=np.uint32).tobytes()).decode()
        s = self.snapshot
        return json.dumps({
            "format_version": self.format_version, "module_name": self.module_name,
            "topology": [{"worker_id": t.worker_id, "node": t.node, "backend": t.backend.to_dict(),
                          "start_lane": t.start_lane, "lanes": t.lanes} for t in self.topology],
            "snapshot": {"total_lanes": s.total_lanes, "program_top": s.program_top,
                         "signal_words_per_lane": s.signal_words_per_lane,
                         "memory_words_per_lane": s.memory_words_per_lane,
                         "shards": [{"worker_id": sh.shard.worker_id, "values": arr(sh.values),
                                     "memories": arr(sh.memories)} for sh in s.shards]}})

    @staticmethod
    def from_json(text: str) -> "RuntimeCheckpoint":
        d = json.loads(text)
        if d.get("format_version") != CHECKPOINT_FORMAT_VERSION:
            raise RuntimeError_("E_RUNTIME_CHECKPOINT_VERSION",
                                f"unsupported runtime checkpoint format version {d.get('format_version')}, "
                                f"expected {CHECKPOINT_FORMAT_VERSION}")
        topo = [RuntimeShardInfo(t["worker_id"], t["node"], HipOptions(**t["backend"]), t["start_lane"], t["lanes"])
                for t in d["topology"]]
        s = d["snapshot"]
        by_id = {t.worker_id: t for t in topo}

        def arr(x: str) -> np.ndarray:
            return np.frombuffer(base64.b64decode(x), dtype=np.uint32).copy()
        shards = [RuntimeShardSnapshot(by_id[sh["worker_id"]], arr(sh["values"]), arr(sh["memories"]))
                  for sh in s["shards"]]
        return RuntimeCheckpoint(d["format_version"], d["module_name"], topo,
                                 RuntimeSnapshot(s["total_lanes"], s["program_top"], s["signal_words_per_lane"],
                                                 s["memory_words_per_lane"], shards))


@dataclass(frozen=True)
class RuntimeCheckpointCadence:
    every_steps: int
    include_initial: bool = False
    include_final: bool = True


INITIAL, CADENCE, FINAL = "initial", "cadence", "final"


@dataclass
class RuntimeCheckpointRunReport:
    requested_steps: int
    completed_steps: int
    checkpoints_emitted: int


def plan_shards(total_lanes: int, shard_lanes: Sequence[int], backend: HipOptions, node: str = "local"
                ) -> List[RuntimeShardInfo]:
    """Contiguous lane ranges of the given sizes (they must cover the batch)."""
    if sum(shard_lanes) != total_lanes or any(n <= 0 for n in shard_lanes):
        raise RuntimeError_("E_RUNTIME_SHARDS", "shard lane counts must be positive and add up to the total")
    out, start = [], 0
    for k, n in enumerate(shard_lanes):
        out.append(RuntimeShardInfo(f"w{k}", node, backend, start, n))
        start += n
    return out


class LaneShardedRuntime:
    """A batch of ``total_lanes`` split into lane shards, each on its own engine
    (``DistributedRuntime`` with in-process workers)."""

    def __init__(self, program: Program, shards: Sequence[RuntimeShardInfo],
                 engine_factory: Optional[Callable[[Program, RuntimeShardInfo], object]] = None):
        self.program = program
        self.shards = list(shards)
        self.total_lanes = sum(s.lanes for s in self.shards)
        starts = [s.start_lane for s in self.shards]
        if starts != sorted(starts) or any(a.start_lane + a.lanes != b.start_lane for a, b in zip(shards, shards[1:])):
            raise RuntimeError_("E_RUNTIME_SHARDS", "shards must be contiguous lane ranges in order")
        make = engine_factory or (lambda p, s: HipRuntimeEngine(p, s.lanes, s.backend))
        self.engines = [make(program, s) for s in self.shards]
        self.completed_steps = 0

    def _split(self, lane_values: np.ndarray, axis: int = -1):
        for s, e in zip(self.shards, self.engines):
            yield e, np.take(lane_values, np.arange(s.start_lane, s.start_lane + s.lanes), axis=axis)

    # ---- RuntimeEngine surface over the whole batch
    def set_signal_limbs(self, sig: int, limbs: np.ndarray) -> None:
        for e, part in self._split(np.asarray(limbs, dtype=np.uint32).reshape(-1, self.total_lanes)):
            e.set_signal_limbs(sig, np.ascontiguousarray(part))

    def get_signal_limbs(self, sig: int) -> np.ndarray:
        return np.concatenate([e.get_signal_limbs(sig) for e in self.engines], axis=1)

    def set_memory_limbs(self, mem: int, values: np.ndarray) -> None:
        for e, part in self._split(np.asarray(values, dtype=np.uint32), axis=0):
            e.set_memory_limbs(mem, np.ascontiguousarray(part))

    def get_memory_limbs(self, mem: int) -> np.ndarray:
        return np.concatenate([e.get_memory_limbs(mem) for e in self.engines], axis=0)

    def eval_combinational(self) -> None:
        for e in self.engines:
            e.eval_combinational()

    def tick_many(self, steps: int) -> None:
        for e in self.engines:
            e.tick_many(steps)
        self.completed_steps += steps

    # ---- snapshots and checkpoints
    def snapshot(self) -> RuntimeSnapshot:
        shards = []
        for s, e in zip(self.shards, self.engines):
            v, m = e.snapshot_storage()
            shards.append(RuntimeShardSnapshot(s, np.asarray(v, dtype=np.uint32).copy(),
                                               np.asarray(m, dtype=np.uint32).copy()))
        p = self.program
        return RuntimeSnapshot(self.total_lanes, p.top, p.total_signal_words, p.total_memory_words_per_lane, shards)

    def validate_snapshot(self, snap: RuntimeSnapshot) -> None:
        p = self.program
        if snap.program_top != p.top or snap.signal_words_per_lane != p.total_signal_words or \
                snap.memory_words_per_lane != p.total_memory_words_per_lane:
            raise RuntimeError_("E_RUNTIME_SNAPSHOT_PROGRAM", "snapshot was taken from another program")
        if snap.total_lanes != self.total_lanes or len(snap.shards) != len(self.shards):
            raise RuntimeError_("E_RUNTIME_SNAPSHOT_SHAPE", "snapshot lanes or shard count differ from the runtime's")
        for s, sh in zip(self.shards, snap.shards):
            if (sh.shard.start_lane, sh.shard.lanes) != (s.start_lane, s.lanes):
                raise RuntimeError_("E_RUNTIME_SNAPSHOT_SHAPE", f"shard {s.worker_id} covers other lanes")
            if sh.values.size != p.total_signal_words * s.lanes or \
                    sh.memories.size != p.total_memory_words_per_lane * s.lanes:
                raise RuntimeError_("E_RUNTIME_SNAPSHOT_SHAPE", f"shard {s.worker_id} storage has the wrong size")

    def restore_snapshot(self, snap: RuntimeSnapshot) -> None:
        self.validate_snapshot(snap)
        for e, sh in zip(self.engines, snap.shards):
            e.restore_storage(sh.values, sh.memories)

    def checkpoint(self) -> RuntimeCheckpoint:
        return RuntimeCheckpoint(CHECKPOINT_FORMAT_VERSION, self.program.top, list(self.shards), self.snapshot())

    def restore_checkpoint(self, cp: RuntimeCheckpoint) -> None:
        if cp.format_version != CHECKPOINT_FORMAT_VERSION:
            raise RuntimeError_("E_RUNTIME_CHECKPOINT_VERSION", "unsupported checkpoint format version")
        if cp.module_name != self.program.top:
            raise RuntimeError_("E_RUNTIME_CHECKPOINT_MODULE", "checkpoint of another module")
        self.restore_snapshot(cp.snapshot)

    def tick_many_with_checkpoints(self, steps: int, cadence: RuntimeCheckpointCadence,
                                   on_checkpoint: Callable[[str, int, RuntimeCheckpoint], None]
                                   ) -> RuntimeCheckpointRunReport:
        """Runs ``steps`` steps up to each multiple of ``cadence.every_steps``,
        handing a checkpoint to ``on_checkpoint(reason, completed_steps,
        checkpoint)`` at the start (include_initial), at every multiple and at
        the end (include_final, unless a checkpoint was just emitted there), as
        tick_many_with_checkpoint_sink (crates/rrtl-runtime/src/lib.rs:7450-7521)."""
        if cadence.every_steps <= 0:
            raise RuntimeError_("E_RUNTIME_CHECKPOINT_CADENCE", "every_steps must be positive")
        emitted, done, last_emit = 0, 0, None
        if cadence.include_initial:
            on_checkpoint(INITIAL, 0, self.checkpoint())
            emitted += 1
            last_emit = 0
        while done < steps:
            n = min(steps - done, cadence.every_steps - done % cadence.every_steps)
            self.tick_many(n)
            done += n
            if done % cadence.every_steps == 0:
                on_checkpoint(CADENCE, done, self.checkpoint())
                emitted += 1
                last_emit = done
        if cadence.include_final and last_emit != done:
            on_checkpoint(FINAL, done, self.checkpoint())
            emitted += 1
        return RuntimeCheckpointRunReport(steps, done, emitted)

    def close(self) -> None:
        for e in self.engines:
            if hasattr(e, "close"):
                e.close()
