"""Stuck-at fault simulation on the bit-sliced gate engine.

The reference's harness (crates/rrtl-sim-ir/examples/fault_sim.rs:19-130)
puts one stuck-at-0/1 fault on a flop per lane with
``BitParallelSimulator::set_force`` (re-applied after every commit), keeps
lane 0 fault-free, drives the same stimulus into every lane, and calls a fault
detected at the first tick after which the lane's outputs differ from lane 0's.

Here a *candidate* is a (fault, stimulus) pair.  A run packs candidates into
64-lane groups that share one stimulus: lane 0 of every group is the golden
lane and lanes 1..63 carry one fault each (``set_forces``).  The stimulus of a
group is the engine's GROUP_RANDOM input mode keyed by the candidate's stimulus
id, so every group that carries stimulus ``s`` sees the same 256-cycle input
sequence.  Detection is accumulated on chip (``detect`` runs): a lane is
flagged once any observed output differs from its group's lane 0 after a
tick.  Every detection reported by this module comes from such a run.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from .backend import IN_GROUP_RANDOM, HipSimulator
from .program import EF_CAPTURE, EF_STORE, KIND_INPUT, KIND_OUTPUT, KIND_REG, OP_LIT, OP_SIGNAL, Program

FAULTS_PER_GROUP = 63  # lane 0 of each 64-lane group is the golden lane


def register_faults(program: Program) -> np.ndarray:
    """Stuck-at-0 and stuck-at-1 on every register, in signal order: rows of
    (signal, stuck value) -- the fault list of fault_sim.rs:41-46."""
    regs = [i for i, s in enumerate(program.signals) if s.kind == KIND_REG]
    return np.array([(r, v) for r in regs for v in (0, 1)], dtype=np.uint32).reshape(-1, 2)


def observable_signals(program: Program, observe: Optional[Sequence[int]] = None) -> np.ndarray:
    """[signals] bool: the signal can influence an observed signal (default:
    every output port) through stores and register captures, over any number
    of cycles.  A register outside this set cannot be detected by any stimulus:
    a fault on it never reaches what the comparison reads."""
    n = len(program.signals)
    srcs: List[set] = [set() for _ in range(n)]   # dst signal -> signals its value reads
    for blk in program.blocks:
        val: Dict[int, frozenset] = {}   # value ids are numbered per block
        for pk in blk.packets:
            for ins in pk.instrs:
                if ins.op == OP_SIGNAL:
                    val[ins.dst] = frozenset((ins.args[0],))
                elif ins.op == OP_LIT:
                    val[ins.dst] = frozenset()
                else:
                    acc = set()
                    for a in ins.args:
                        acc |= val.get(a, frozenset())
                    val[ins.dst] = frozenset(acc)
            for e in pk.effects:
                if e.kind in (EF_STORE, EF_CAPTURE):
                    srcs[e.dst] |= val.get(e.value, frozenset())
                    if e.kind == EF_CAPTURE and e.reset is not None:
                        srcs[e.dst].add(e.reset.signal)
    seen = np.zeros(n, dtype=bool)
    stack = list(observe) if observe is not None else [i for i, s in enumerate(program.signals) if s.kind == KIND_OUTPUT]
    for s in stack:
        seen[s] = True
    while stack:
        d = stack.pop()
        for s in srcs[d]:
            if not seen[s]:
                seen[s] = True
                stack.append(s)
    return seen


@dataclass
class Packing:
    lanes: int
    group_key: np.ndarray     # [groups] u64 stimulus id of each group
    lane: np.ndarray          # [n] lane of each candidate
    groups: int


def pack(stim_id: np.ndarray) -> Packing:
    """Groups candidates by stimulus id, 63 to a 64-lane group (lane 0 golden).
    Candidate order within a stimulus is kept."""
    stim_id = np.asarray(stim_id, dtype=np.uint64)
    n = stim_id.size
    order = np.argsort(stim_id, kind="stable")
    s_sorted = stim_id[order]
    # rank of each candidate within its stimulus
    first = np.ones(n, dtype=bool)
    first[1:] = s_sorted[1:] != s_sorted[:-1]
    start_idx = np.flatnonzero(first)
    run_id = np.cumsum(first) - 1
    rank = np.arange(n) - start_idx[run_id]
    # groups per stimulus and the first group of each
    counts = np.diff(np.append(start_idx, n))
    gper = (counts + FAULTS_PER_GROUP - 1) // FAULTS_PER_GROUP
    gfirst = np.concatenate([[0], np.cumsum(gper)[:-1]]).astype(np.int64)
    groups = int(gper.sum()) if n else 0
    g = gfirst[run_id] + rank // FAULTS_PER_GROUP
    lane_sorted = g * 64 + 1 + rank % FAULTS_PER_GROUP
    lane = np.empty(n, dtype=np.int64)
    lane[order] = lane_sorted
    group_key = np.repeat(s_sorted[start_idx], gper) if n else np.zeros(0, dtype=np.uint64)
    return Packing(lanes=groups * 64, group_key=group_key.astype(np.uint64), lane=lane, groups=groups)


@dataclass
class RunResult:
    detected: np.ndarray      # [n] bool per candidate
    kernel_ms: float          # device time of the run
    groups: int
    lane_cycles: int


class FaultSim:
    """Exact fault simulation of (fault, stimulus) candidates on one design.

    ``faults``: rows (signal, stuck value); inputs are driven GROUP_RANDOM from
    ``seed``; ``observe`` defaults to every output port, compared after each
    tick (fault_sim.rs:76-95)."""

    def __init__(self, program: Program, faults: np.ndarray, cycles: int = 256, seed: int = 1,
                 observe: Optional[Sequence[int]] = None, tile_words: int = 2, **engine_opts):
        self.program = program
        self.faults = np.asarray(faults, dtype=np.uint32).reshape(-1, 2)
        self.cycles = cycles
        self.seed = seed
        self.inputs = [i for i, s in enumerate(program.signals) if s.kind == KIND_INPUT]
        self.observe = list(observe) if observe is not None else \
            [i for i, s in enumerate(program.signals) if s.kind == KIND_OUTPUT]
        self.opts = dict(tile_words=tile_words, **engine_opts)

    def run(self, fault_idx: np.ndarray, stim_id: np.ndarray) -> RunResult:
        fault_idx = np.asarray(fault_idx, dtype=np.int64)
        stim_id = np.asarray(stim_id, dtype=np.uint64)
        if fault_idx.shape != stim_id.shape:
            raise ValueError("fault_idx and stim_id differ in shape")
        if fault_idx.size == 0:
            return RunResult(np.zeros(0, dtype=bool), 0.0, 0, 0)
        if fault_idx.min() < 0 or fault_idx.max() >= len(self.faults):
            raise ValueError("fault index out of range")
        pk = pack(stim_id)
        sim = HipSimulator(self.program, pk.lanes, engine="gate", **self.opts)
        try:
            f = self.faults[fault_idx]
            sim.set_forces(f[:, 0], pk.lane, f[:, 1])
            sim.run(self.cycles, seed=self.seed, inputs=[(i, IN_GROUP_RANDOM, 0) for i in self.inputs],
                    observe=self.observe, observe_post=True, detect=True, group_key=pk.group_key)
            flags = sim.detect_flags()
            st = sim.stats()
        finally:
            sim.close()
        return RunResult(flags[pk.lane], st["kernel_ms"], pk.groups, st["lane_cycles"])

    def golden_register_bits(self, stim_ids: np.ndarray, regs: Sequence[int]):
        """Fault-free value of each register after every tick under each
        stimulus, from one launch with one 64-lane group per stimulus and the
        registers traced: ([stimuli][registers][cycles] uint8, kernel ms)."""
        ids = np.asarray(stim_ids, dtype=np.uint64)
        regs = [int(r) for r in regs]
        sim = HipSimulator(self.program, 64 * ids.size, engine="gate", **self.opts)
        try:
            sim.run(self.cycles, seed=self.seed, inputs=[(i, IN_GROUP_RANDOM, 0) for i in self.inputs],
                    observe=regs, observe_post=True, trace=True, group_key=ids)
            tr = sim.trace().reshape(self.cycles, len(regs), -1)
            st = sim.stats()
        finally:
            sim.close()
        bits = (tr[:, :, 0::2] & np.uint32(1)).astype(np.uint8)     # lane 0 of each group
        return np.ascontiguousarray(bits.transpose(2, 1, 0)), st["kernel_ms"]
