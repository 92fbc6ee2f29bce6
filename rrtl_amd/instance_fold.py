"""Instance folding: a design that is N independent instances of one child
module runs as the child alone at N x L lanes, instance i on the lane block
[i L, (i + 1) L).  The program shrinks N-fold (generated code stays small,
every lane runs the same straight-line code) and a batch of few stimulus
lanes fills the GPU with instances instead of leaving it idle.

Mirrors the reference's structural analysis ``analyze_instance_fold``
(crates/rrtl-sim-ir/src/instance_fold.rs:41-103) and its GPU runner
``InstanceFoldPlan`` / ``InstanceFoldSimulator``
(crates/rrtl-gpu-sim/src/instance_fold.rs:23-140), on the Verilog source this
package imports: the top must be a pure structural wrapper (port and wire
declarations and instances only), every instance of the same defined module
with the same parameters, each connection a plain signal, no signal driven by
two instance outputs, no instance output feeding an instance input, no inout
ports.  Shared inputs (a common clock) are allowed and drive every block.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import lower, verilog
from .program import KIND_INPUT, KIND_OUTPUT, Program


class InstanceFoldError(ValueError):
    pass


@dataclass
class FoldConnection:
    top_signal: str
    instance: int
    port: str
    direction: str          # "input" or "output"


@dataclass
class InstanceFold:
    child_module: str
    num_instances: int
    instance_names: List[str]
    connections: List[FoldConnection]
    child_params: Dict[str, int] = field(default_factory=dict)


def _port_directions(m: verilog.ModuleAst) -> Dict[str, str]:
    return {it[5]: it[1] for it in m.items if it[0] == "port"}


def _const_params(params: Dict[str, tuple], pos: List[tuple], child: verilog.ModuleAst) -> Optional[Dict[str, int]]:
    """Parameter overrides of an instance as integers, or None when one is not
    a plain number."""
    out: Dict[str, int] = {}
    names = [p[0] for p in child.params if not p[1]]
    for k, e in list(zip(names, pos)) + list(params.items()):
        if e[0] != "num":
            return None
        out[k] = int(e[1].value)
    return out


def analyze_instance_fold(text: str, top: str, defines: Optional[Dict[str, str]] = None) -> Optional[InstanceFold]:
    """The fold plan when ``top`` is a flat wrapper of N >= 2 independent
    instances of one module defined in ``text``; None otherwise."""
    mods = {m.name: m for m in verilog.parse(text, defines)}
    m = mods.get(top)
    if m is None:
        return None
    insts = [it for it in m.items if it[0] == "inst"]
    if len(insts) < 2:
        return None
    for it in m.items:
        if it[0] in ("port", "inst"):
            continue
        if it[0] == "decl" and it[1] in ("wire", "logic", "tri") and it[6] is None and not it[5]:
            continue
        return None                              # logic of its own (or anything unrecognised)
    child_name = insts[0][1]
    child = mods.get(child_name)
    if child is None or any(it[1] != child_name for it in insts):
        return None
    params = _const_params(insts[0][2], insts[0][3], child)
    if params is None or any(_const_params(it[2], it[3], child) != params for it in insts):
        return None
    dirs = _port_directions(child)
    top_ports = set(m.ports)
    conns: List[FoldConnection] = []
    ins, outs = set(), set()
    for k, it in enumerate(insts):
        _, _, _, _, name, named, pos, _ = it
        bound = dict(zip(child.ports, pos))
        bound.update(named)
        for port, e in bound.items():
            if port not in dirs:
                return None
            if e is None:
                continue
            if e[0] != "id":
                return None
            d = dirs[port]
            if d == "inout":
                return None
            sig = e[1]
            conns.append(FoldConnection(sig, k, port, d))
            if d == "input":
                ins.add(sig)
            elif sig in outs:
                return None                      # two instances drive one signal
            else:
                outs.add(sig)
    if ins & outs:
        return None                              # data flows between instances
    if not (ins | outs) <= top_ports:
        return None                              # an internal wire joins instances
    return InstanceFold(child_name, len(insts), [it[4] for it in insts], conns, params)


class InstanceFoldPlan:
    """The fold plus the child lowered on its own (its port signals serve every
    lane block)."""

    def __init__(self, text: str, top: str, defines: Optional[Dict[str, str]] = None):
        fold = analyze_instance_fold(text, top, defines)
        if fold is None:
            raise InstanceFoldError(f"E_INSTANCE_FOLD: `{top}` is not a flat wrapper of independent identical instances")
        self.fold = fold
        imp = verilog.import_verilog(text, fold.child_module, params=fold.child_params or None, defines=defines)
        self.child: Program = lower.lower_module(imp.module)
        self._by_top: Dict[str, List[FoldConnection]] = {}
        for c in fold.connections:
            self._by_top.setdefault(c.top_signal, []).append(c)

    @property
    def num_instances(self) -> int:
        return self.fold.num_instances

    def port_signal(self, port: str) -> int:
        """The child program's signal index of an input or output port."""
        for i, s in enumerate(self.child.signals):
            if s.kind in (KIND_INPUT, KIND_OUTPUT) and s.name.rsplit(".", 1)[-1] == port:
                return i
        raise InstanceFoldError(f"E_INSTANCE_FOLD: no such port `{port}`")

    def connections(self, top_signal: str) -> List[FoldConnection]:
        c = self._by_top.get(top_signal)
        if not c:
            raise InstanceFoldError(f"E_INSTANCE_FOLD: `{top_signal}` is not connected to an instance")
        return c


class InstanceFoldSimulator:
    """The folded child on one engine at ``num_instances x base_lanes`` lanes;
    top-level signals are read and written per lane block, so the object
    stands in for the flattened design at ``base_lanes`` lanes."""

    def __init__(self, text: str, top: str, base_lanes: int, defines: Optional[Dict[str, str]] = None,
                 engine=None, **engine_opts):
        self.plan = InstanceFoldPlan(text, top, defines)
        self.base_lanes = base_lanes
        if engine is None:
            from .backend import HipSimulator
            engine = HipSimulator
        self.sim = engine(self.plan.child, self.total_lanes, **engine_opts)

    @property
    def num_instances(self) -> int:
        return self.plan.num_instances

    @property
    def total_lanes(self) -> int:
        return self.plan.num_instances * self.base_lanes

    def instance_lane(self, inst: int, local: int) -> int:
        return inst * self.base_lanes + local

    def set_port_limbs(self, port: str, limbs: np.ndarray) -> None:
        """A child port across all lanes ([limbs][total_lanes] u32)."""
        self.sim.set_signal_limbs(self.plan.port_signal(port), np.asarray(limbs, dtype=np.uint32), settle=False)

    def get_port_limbs(self, port: str) -> np.ndarray:
        return self.sim.get_signal_limbs(self.plan.port_signal(port))

    def set_top_limbs(self, top_signal: str, limbs: np.ndarray) -> None:
        """A top-level input ([limbs][base_lanes]): its lane values go to the
        lane block of every instance it is connected to."""
        limbs = np.asarray(limbs, dtype=np.uint32).reshape(-1, self.base_lanes)
        for c in self.plan.connections(top_signal):
            if c.direction != "input":
                raise InstanceFoldError(f"E_INSTANCE_FOLD: `{top_signal}` is an instance output")
            sig = self.plan.port_signal(c.port)
            cur = self.sim.get_signal_limbs(sig)
            lo = c.instance * self.base_lanes
            cur[:, lo:lo + self.base_lanes] = limbs
            self.sim.set_signal_limbs(sig, cur, settle=False)

    def get_top_limbs(self, top_signal: str) -> np.ndarray:
        """A top-level signal driven by an instance output ([limbs][base_lanes])."""
        for c in self.plan.connections(top_signal):
            if c.direction == "output":
                lo = c.instance * self.base_lanes
                return self.sim.get_signal_limbs(self.plan.port_signal(c.port))[:, lo:lo + self.base_lanes]
        raise InstanceFoldError(f"E_INSTANCE_FOLD: `{top_signal}` is not driven by an instance")

    def eval_combinational(self) -> None:
        self.sim.eval_combinational()

    def tick(self) -> None:
        self.sim.tick()

    def tick_many(self, steps: int) -> None:
        self.sim.tick_many(steps)

    def close(self) -> None:
        if hasattr(self.sim, "close"):
            self.sim.close()
