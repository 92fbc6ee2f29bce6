"""Lowering of a flat design to the packed machine program.

Restates the reference's two lowering passes:

* ``lower_to_packed_program`` (crates/rrtl-sim-ir/src/lib.rs:592-660) with
  ``Lowering::allocate_contexts`` (:12338-12410: signals get word-major u32-limb
  offsets in declaration order, memories get per-lane offsets) and the
  schedulers ``schedule_assignments`` (:12734-12790, comb assignments levelized by
  signal dependency, a dependency cycle counts as level 0),
  ``schedule_async_resets`` (:12791), ``schedule_tick_next`` (:12816) and
  ``schedule_memory_writes`` (:12832);
* ``lower_to_machine_program`` / ``lower_packets_to_block`` /
  ``MachinePacketLowerer`` / ``schedule_machine_instrs`` (:661-676, :2997-3178):
  every op-packet's expression trees become SSA instructions with structural
  memoisation, value ids count up across the block, instructions are levelled
  into machine packets and the op-packet's effects ride on its last machine
  packet.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

from . import ir
from .program import (
    Block, Effect, Instr, Packet, PMemory, PReset, Program, PSignal, EF_CAPTURE,
    EF_MEMWRITE, EF_STORE, KIND_OF, OP_CONCAT, OP_LIT, OP_LT, OP_MEMREAD, OP_OF,
    OP_SIGNAL, OP_SLICE, POL_HIGH, POL_LOW, RESET_ASYNC, RESET_SYNC, limbs, to_limbs,
)

# A packed expression node is (op, width, signed, attr, *children) with its
# children given as node ids: structurally equal trees get one id
# (hash-consing), which is the structural equality ``PackedExpr`` derives
# (``#[derive(Hash, Eq)]``) without re-walking shared sub-trees on every hash.
PExpr = int


class _Nodes:
    def __init__(self):
        self.ids: Dict[tuple, int] = {}
        self.nodes: List[tuple] = []

    def intern(self, node: tuple) -> int:
        got = self.ids.get(node)
        if got is None:
            got = len(self.nodes)
            self.ids[node] = got
            self.nodes.append(node)
        return got


# packed ops: ("assign", dst, expr) | ("capture", dst, expr, PReset|None)
#             | ("memwrite", mem, en, addr, data)


def _lower_expr(e: ir.Expr, sig_index: Dict[ir.Signal, int],
                mem_index: Dict[ir.Memory, int], cache: Dict[int, PExpr], nodes: _Nodes) -> PExpr:
    """``Lowering::lower_raw_expr`` (lib.rs:12460-12560).  Each node keeps its own
    type; for validated designs this equals the type the reference threads down
    (rhs of a binary op takes the lhs type, a mux condition is uint(1))."""
    got = cache.get(id(e))
    if got is not None:
        return got
    stack = [(e, False)]
    while stack:
        x, ready = stack.pop()
        if id(x) in cache:
            continue
        if not ready:
            stack.append((x, True))
            stack.extend((a, False) for a in x.args if id(a) not in cache)
            continue
        w, sg = x.ty.width, x.ty.signed
        kids = tuple(cache[id(a)] for a in x.args)
        if x.op == "lit":
            node = ("lit", w, sg, tuple(to_limbs(x.attr, w)))
        elif x.op == "sig":
            node = ("sig", w, sg, sig_index[x.attr])
        elif x.op == "memread":
            node = ("memread", w, sg, mem_index[x.attr]) + kids
        elif x.op == "slice":
            node = ("slice", w, sg, x.attr) + kids
        elif x.op == "lt":
            node = ("lt", w, sg, int(x.args[0].ty.signed)) + kids
        else:
            node = (x.op, w, sg, None) + kids
        cache[id(x)] = nodes.intern(node)
    return cache[id(e)]


def _signal_deps(pe: PExpr, nodes: _Nodes, out: set) -> None:
    stack, seen = [pe], set()
    while stack:
        k = stack.pop()
        if k in seen:
            continue
        seen.add(k)
        x = nodes.nodes[k]
        if x[0] == "sig":
            out.add(x[3])
        stack.extend(x[4:])


def schedule_assignments(assigns: List[Tuple[int, PExpr]], nodes: _Nodes) -> List[List[tuple]]:
    """``schedule_assignments`` + ``assignment_level`` (lib.rs:12734-12790)."""
    dst_to = {dst: i for i, (dst, _) in enumerate(assigns)}
    deps = []
    for dst, pe in assigns:
        s: set = set()
        _signal_deps(pe, nodes, s)
        deps.append([dst_to[x] for x in s if x in dst_to])
    memo: Dict[int, int] = {}

    def level(i: int) -> int:
        # iterative DFS with the reference's cycle guard (a back edge yields 0)
        if i in memo:
            return memo[i]
        visiting = set()
        stack = [(i, 0)]
        visiting.add(i)
        best: Dict[int, int] = {i: 0}
        while stack:
            node, k = stack[-1]
            ds = deps[node]
            if k < len(ds):
                stack[-1] = (node, k + 1)
                d = ds[k]
                if d == node:
                    continue
                if d in memo:
                    best[node] = max(best[node], memo[d] + 1)
                elif d in visiting:
                    # cycle: assignment_level returns 0 for the revisited node,
                    # so this edge contributes 0 + 1
                    best[node] = max(best[node], 1)
                else:
                    visiting.add(d)
                    best[d] = 0
                    stack.append((d, 0))
            else:
                stack.pop()
                visiting.discard(node)
                memo[node] = best[node]
                if stack:
                    parent = stack[-1][0]
                    best[parent] = max(best[parent], best[node] + 1)
        return memo[i]

    levels: List[List[tuple]] = []
    for i, (dst, pe) in enumerate(assigns):
        lv = level(i)
        while len(levels) <= lv:
            levels.append([])
        levels[lv].append(("assign", dst, pe))
    return [ops for ops in levels if ops]


class _PacketLowerer:
    """``MachinePacketLowerer`` (lib.rs:3060-3138): post-order SSA with memo."""

    def __init__(self, next_value: int, nodes: _Nodes):
        self.next_value = next_value
        self.nodes = nodes
        self.memo: Dict[PExpr, int] = {}
        self.instrs: List[Instr] = []

    def lower(self, root: PExpr) -> int:
        got = self.memo.get(root)
        if got is not None:
            return got
        stack = [(root, False)]
        while stack:
            pe, ready = stack.pop()
            if pe in self.memo:
                continue
            x = self.nodes.nodes[pe]
            kids = x[4:]
            if not ready:
                stack.append((pe, True))
                # children pushed in reverse so they are lowered left to right
                stack.extend((k, False) for k in reversed(kids) if k not in self.memo)
                continue
            op, w, sg, attr = x[0], x[1], x[2], x[3]
            if op == "lit":
                args = tuple(attr)
            elif op == "sig":
                args = (attr,)
            elif op == "memread":
                args = (attr, self.memo[kids[0]])
            elif op == "slice":
                args = (self.memo[kids[0]], attr)
            elif op == "lt":
                args = (self.memo[kids[0]], self.memo[kids[1]], attr)
            else:
                args = tuple(self.memo[k] for k in kids)
            v = self.next_value
            self.next_value += 1
            self.instrs.append(Instr(OP_OF[op], v, w, bool(sg), args))
            self.memo[pe] = v
        return self.memo[root]


def _instr_deps(i: Instr) -> Tuple[int, ...]:
    if i.op in (OP_LIT, OP_SIGNAL):
        return ()
    if i.op == OP_SLICE:
        return (i.args[0],)
    if i.op == OP_LT:
        return i.args[:2]
    if i.op == OP_MEMREAD:
        return (i.args[1],)
    return i.args


def schedule_machine_instrs(instrs: List[Instr]) -> List[Packet]:
    """``schedule_machine_instrs`` (lib.rs:3140-3160): level = 1 + max dep level."""
    where = {ins.dst: k for k, ins in enumerate(instrs)}
    lvl: List[int] = []
    for ins in instrs:  # post-order: deps precede users
        l = 0
        for d in _instr_deps(ins):
            k = where.get(d)
            if k is not None:
                l = max(l, lvl[k] + 1)
        lvl.append(l)
    packets: List[Packet] = []
    for ins, l in zip(instrs, lvl):
        while len(packets) <= l:
            packets.append(Packet())
        packets[l].instrs.append(ins)
    return [p for p in packets if p.instrs or p.effects]


def lower_packets_to_block(op_packets: List[List[tuple]], nodes: _Nodes) -> Block:
    """``lower_packets_to_block`` (lib.rs:2997-3048)."""
    block = Block()
    next_value = 0
    for ops in op_packets:
        low = _PacketLowerer(next_value, nodes)
        effects: List[Effect] = []
        for op in ops:
            if op[0] == "assign":
                effects.append(Effect(EF_STORE, op[1], low.lower(op[2])))
            elif op[0] == "capture":
                effects.append(Effect(EF_CAPTURE, op[1], low.lower(op[2]), reset=op[3]))
            else:
                _, mem, en, addr, data = op
                e = low.lower(en)
                a = low.lower(addr)
                d = low.lower(data)
                effects.append(Effect(EF_MEMWRITE, mem, enable=e, addr=a, data=d))
        next_value = low.next_value
        pkts = schedule_machine_instrs(low.instrs)
        if not pkts:
            pkts = [Packet()]
        pkts[-1].effects.extend(effects)
        block.packets.extend(pkts)
    block.n_values = next_value
    return block


def lower_module(m: ir.Module) -> Program:
    """``lower_to_packed_program`` followed by ``lower_to_machine_program``."""
    m.validate()
    sig_index: Dict[ir.Signal, int] = {}
    signals: List[PSignal] = []
    offset = 0
    for s in m.signals:
        lb = limbs(s.width)
        sig_index[s] = len(signals)
        signals.append(PSignal(s.name, s.width, s.ty.signed, offset, lb, KIND_OF[s.kind]))
        offset += lb
    mem_index: Dict[ir.Memory, int] = {}
    memories: List[PMemory] = []
    moff = 0
    for mem in m.memories:
        lb = limbs(mem.ty.width)
        mem_index[mem] = len(memories)
        memories.append(PMemory(mem.name, mem.ty.width, mem.ty.signed, mem.depth,
                                mem.addr_width, moff, lb))
        moff += mem.depth * lb

    cache: Dict[int, PExpr] = {}
    nodes = _Nodes()
    assigns = [(sig_index[d], _lower_expr(m.assigns[d], sig_index, mem_index, cache, nodes))
               for d in m.assign_order]

    regs = []
    for r in m.registers:
        nxt = _lower_expr(m.nexts[r], sig_index, mem_index, cache, nodes)
        rs = m.resets.get(r)
        preset = None
        if rs is not None:
            preset = PReset(sig_index[rs.signal], to_limbs(rs.value, r.width),
                            RESET_ASYNC if rs.kind == "async" else RESET_SYNC,
                            POL_LOW if rs.polarity == "low" else POL_HIGH)
        regs.append((sig_index[r], r, nxt, preset))

    comb = schedule_assignments(assigns, nodes)
    async_ops = [("capture", si, nodes.intern(("lit", r.width, r.ty.signed, tuple(pr.value))), pr)
                 for si, r, _, pr in regs if pr is not None and pr.kind == RESET_ASYNC]
    async_reset_comb = [async_ops] if async_ops else []
    tick_ops = [("capture", si, nxt, pr) for si, _, nxt, pr in regs]
    tick_next = [tick_ops] if tick_ops else []
    writes = [("memwrite", mem_index[w.mem],
               _lower_expr(w.enable, sig_index, mem_index, cache, nodes),
               _lower_expr(w.addr, sig_index, mem_index, cache, nodes),
               _lower_expr(w.data, sig_index, mem_index, cache, nodes)) for w in m.mem_writes]
    tick_commit = [writes] if writes else []

    blocks = tuple(lower_packets_to_block(p, nodes) for p in
                   (async_reset_comb, comb, tick_next, tick_commit))
    return Program(m.name, signals, memories, blocks, offset, moff)
