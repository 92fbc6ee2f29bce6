"""Design specialization of a packed machine program.

Restates the reference's specializer (crates/rrtl-sim-ir/src/specialize.rs):

* ``specialize_program`` (:49, ``specialize_block`` :694, ``try_fold`` :846):
  per block, a forward pass of constant folding, the algebraic identities
  (x & 0, x & ~0, x | 0, x | ~0, x ^ 0, x + 0, x - 0, x - x, x * 0, x * 1,
  x == x, x != x, x < x, mux on a constant or between equal values) and copy
  propagation; then dead-code elimination from the block's effects and
  compaction of the value ids;
* ``freeze_signals_program`` (:87): reads of signals known to hold a value
  become literals, then the static pass folds what they controlled (the
  result equals the input program only while those signals hold those values);
* ``strength_reduce_program`` (:363-530): a multiply of a value wider than 32
  bits by a constant with at most four non-adjacent-form digits becomes
  shifts (a concatenation with zeros, truncated) and adds/subtracts.

Every pass returns a program observationally equivalent to its input for
every signal and memory (tests/test_specialize.py checks that on the packed
oracle).  The passes work on this package's ``Program`` and are its own code;
the rules are the reference's.
"""
from __future__ import annotations

from dataclasses import dataclass, replace
from typing import Dict, List, Optional, Tuple

from .program import (Block, Effect, Instr, OP_ADD, OP_AND, OP_CAST, OP_CONCAT, OP_EQ, OP_LIT, OP_LT, OP_MEMREAD,
                      OP_MUL, OP_MUX, OP_NE, OP_NOT, OP_OR, OP_SEXT, OP_SIGNAL, OP_SLICE, OP_SUB, OP_TRUNC,
                      OP_XOR, OP_ZEXT, EF_MEMWRITE, Packet, Program, limbs, to_limbs)


@dataclass
class SpecializeStats:
    instrs_before: int = 0
    instrs_after: int = 0
    folded: int = 0    # instructions evaluated to a literal
    copies: int = 0    # instructions reduced to a copy of an operand
    dead: int = 0      # instructions removed as unused


def _mask(w: int) -> int:
    return (1 << w) - 1


def _val(limb_list) -> int:
    return sum(int(x) << (32 * k) for k, x in enumerate(limb_list))


def _sval(v: int, w: int) -> int:
    v &= _mask(w)
    return v - (1 << w) if w and v >> (w - 1) & 1 else v


def operands(ins: Instr) -> Tuple[int, ...]:
    """Value ids an instruction reads."""
    if ins.op in (OP_LIT, OP_SIGNAL):
        return ()
    if ins.op == OP_SLICE:
        return (ins.args[0],)
    if ins.op == OP_LT:
        return tuple(ins.args[:2])
    if ins.op == OP_MEMREAD:
        return (ins.args[1],)
    return tuple(ins.args)


def _with_operands(ins: Instr, f) -> Instr:
    a = ins.args
    if ins.op in (OP_LIT, OP_SIGNAL):
        return ins
    if ins.op == OP_SLICE:
        return replace(ins, args=(f(a[0]), a[1]))
    if ins.op == OP_LT:
        return replace(ins, args=(f(a[0]), f(a[1]), a[2]))
    if ins.op == OP_MEMREAD:
        return replace(ins, args=(a[0], f(a[1])))
    return replace(ins, args=tuple(f(x) for x in a))


def _effect_operands(e: Effect) -> Tuple[int, ...]:
    return (e.enable, e.addr, e.data) if e.kind == EF_MEMWRITE else (e.value,)


def _with_effect_operands(e: Effect, f) -> Effect:
    if e.kind == EF_MEMWRITE:
        return replace(e, enable=f(e.enable), addr=f(e.addr), data=f(e.data))
    return replace(e, value=f(e.value))


def _try_fold(ins: Instr, const: Dict[int, int], width: Dict[int, int], canon) -> Optional[Tuple[str, int]]:
    """("const", value) or ("copy", value id) for an instruction, as try_fold (:846)."""
    w = ins.width
    ones = _mask(w)
    c = lambda v: const.get(v)  # noqa: E731
    op, a = ins.op, ins.args
    if op in (OP_LIT, OP_SIGNAL, OP_MEMREAD, OP_CONCAT):
        return None
    if op == OP_NOT:
        return ("const", ~c(a[0]) & ones) if c(a[0]) is not None else None
    if op in (OP_AND, OP_OR, OP_XOR, OP_ADD, OP_SUB, OP_MUL):
        x, y = c(a[0]), c(a[1])
        if x is not None and y is not None:
            r = {OP_AND: x & y, OP_OR: x | y, OP_XOR: x ^ y, OP_ADD: x + y, OP_SUB: x - y, OP_MUL: x * y}[op]
            return ("const", r & ones)
        if op == OP_AND:
            if x == ones:
                return ("copy", a[1])
            if y == ones:
                return ("copy", a[0])
            if x == 0 or y == 0:
                return ("const", 0)
        elif op == OP_OR:
            if x == 0:
                return ("copy", a[1])
            if y == 0:
                return ("copy", a[0])
            if x == ones or y == ones:
                return ("const", ones)
        elif op in (OP_XOR, OP_ADD):
            if x == 0:
                return ("copy", a[1])
            if y == 0:
                return ("copy", a[0])
        elif op == OP_SUB:
            if y == 0:
                return ("copy", a[0])
            if canon(a[0]) == canon(a[1]):
                return ("const", 0)
        elif op == OP_MUL:
            if x == 0 or y == 0:
                return ("const", 0)
            if x == 1:
                return ("copy", a[1])
            if y == 1:
                return ("copy", a[0])
        return None
    if op in (OP_EQ, OP_NE):
        x, y = c(a[0]), c(a[1])
        if x is not None and y is not None:
            return ("const", int((x == y) == (op == OP_EQ)))
        if canon(a[0]) == canon(a[1]):
            return ("const", int(op == OP_EQ))
        return None
    if op == OP_LT:
        x, y = c(a[0]), c(a[1])
        if x is not None and y is not None:
            if a[2]:
                return ("const", int(_sval(x, width[a[0]]) < _sval(y, width[a[1]])))
            return ("const", int(x < y))
        if canon(a[0]) == canon(a[1]):
            return ("const", 0)
        return None
    if op == OP_MUX:
        x = c(a[0])
        if x is not None:
            return ("copy", a[1] if x & 1 else a[2])
        if canon(a[1]) == canon(a[2]):
            return ("copy", a[1])
        return None
    if op == OP_SLICE:
        x = c(a[0])
        return ("const", (x >> a[1]) & ones) if x is not None else None
    if op in (OP_ZEXT, OP_TRUNC, OP_CAST):
        x = c(a[0])
        return ("const", x & ones) if x is not None else None
    if op == OP_SEXT:
        x = c(a[0])
        return ("const", _sval(x, width[a[0]]) & ones) if x is not None else None
    return None


def _specialize_block(bk: Block, st: SpecializeStats) -> Block:
    st.instrs_before += sum(len(p.instrs) for p in bk.packets)
    copy_of: Dict[int, int] = {}
    const: Dict[int, int] = {}
    width: Dict[int, int] = {}

    def canon(v: int) -> int:
        while v in copy_of:
            v = copy_of[v]
        return v

    packets: List[Packet] = []
    for pk in bk.packets:
        instrs = []
        for ins in pk.instrs:
            ins = _with_operands(ins, canon)
            width[ins.dst] = ins.width
            if ins.op == OP_LIT:
                const[ins.dst] = _val(ins.args) & _mask(ins.width)
                instrs.append(ins)
                continue
            r = _try_fold(ins, const, width, canon)
            if r is not None and r[0] == "const":
                ins = Instr(OP_LIT, ins.dst, ins.width, ins.signed, tuple(to_limbs(r[1], ins.width)))
                const[ins.dst] = r[1]
                st.folded += 1
            elif r is not None:
                copy_of[ins.dst] = canon(r[1])
                if canon(r[1]) in const:
                    const[ins.dst] = const[canon(r[1])]
                st.copies += 1
            instrs.append(ins)
        packets.append(Packet(instrs, [_with_effect_operands(e, canon) for e in pk.effects]))
    # dead-code elimination from the effects backwards
    live = {v for pk in packets for e in pk.effects for v in _effect_operands(e)}
    for pk in reversed(packets):
        for ins in reversed(pk.instrs):
            if ins.dst in live:
                live.update(operands(ins))
    kept = []
    for pk in packets:
        ins = [i for i in pk.instrs if i.dst in live]
        st.dead += len(pk.instrs) - len(ins)
        if ins or pk.effects:
            kept.append(Packet(ins, pk.effects))
    # compact value ids in definition order
    remap: Dict[int, int] = {}
    for pk in kept:
        for i in pk.instrs:
            remap[i.dst] = len(remap)
    f = remap.__getitem__
    out = Block([Packet([replace(_with_operands(i, f), dst=f(i.dst)) for i in pk.instrs],
                        [_with_effect_operands(e, f) for e in pk.effects]) for pk in kept], len(remap))
    st.instrs_after += sum(len(p.instrs) for p in out.packets)
    return out


def specialize_program(p: Program) -> Tuple[Program, SpecializeStats]:
    """Constant folding, identities, copy propagation, DCE and value-id
    compaction over every block (``specialize_program``, specialize.rs:49)."""
    st = SpecializeStats()
    blocks = tuple(_specialize_block(b, st) for b in p.blocks)
    return replace(p, blocks=blocks), st


def freeze_signals_program(p: Program, frozen: Dict[int, int]) -> Tuple[Program, SpecializeStats]:
    """Reads of the signals in ``frozen`` (index -> value) become literals, then
    ``specialize_program`` (``freeze_signals_program``, specialize.rs:87)."""
    blocks = []
    for b in p.blocks:
        pks = []
        for pk in b.packets:
            ins = []
            for i in pk.instrs:
                if i.op == OP_SIGNAL and i.args[0] in frozen:
                    i = Instr(OP_LIT, i.dst, i.width, i.signed, tuple(to_limbs(frozen[i.args[0]] & _mask(i.width),
                                                                                  i.width)))
                ins.append(i)
            pks.append(Packet(ins, list(pk.effects)))
        blocks.append(Block(pks, b.n_values))
    return specialize_program(replace(p, blocks=tuple(blocks)))


STRENGTH_MAX_TERMS = 4


def naf(c: int, width: int) -> List[Tuple[int, int]]:
    """Non-adjacent form of c within width bits: (shift, sign) pairs with
    c = sum(sign * 2^shift) mod 2^width, shifts below width."""
    out, k = [], 0
    while c and k <= width + 1:
        if c & 1:
            d = -1 if c & 3 == 3 else 1
            c = c + 1 if d == -1 else c - 1
            if k < width:
                out.append((k, d))
        c >>= 1
        k += 1
    return out


@dataclass
class StrengthStats:
    reduced: int = 0
    zeroed: int = 0


def strength_reduce_program(p: Program) -> Tuple[Program, StrengthStats]:
    """Multiplies wider than 32 bits by a constant of at most four NAF digits
    become shift/add chains (``strength_reduce_program``, specialize.rs:392);
    a multiply by a constant 0 becomes the literal 0."""
    st = StrengthStats()
    blocks = []
    for b in p.blocks:
        const: Dict[int, int] = {}
        nxt = b.n_values
        pks = []
        for pk in b.packets:
            ins: List[Instr] = []
            for i in pk.instrs:
                if i.op == OP_LIT:
                    const[i.dst] = _val(i.args) & _mask(i.width)
                if i.op == OP_MUL:
                    x, y = const.get(i.args[0]), const.get(i.args[1])
                    if (x is None) != (y is None):
                        var, c = (i.args[1], x) if x is not None else (i.args[0], y)
                        w = i.width
                        c &= _mask(w)
                        terms = naf(c, w)
                        if c == 0:
                            ins.append(Instr(OP_LIT, i.dst, w, i.signed, tuple(to_limbs(0, w))))
                            st.zeroed += 1
                            continue
                        if w > 32 and terms and len(terms) <= STRENGTH_MAX_TERMS:
                            acc = None
                            for k, (sh, sign) in enumerate(terms):
                                term = var
                                if sh:
                                    z = nxt
                                    cat = nxt + 1
                                    tr = nxt + 2
                                    nxt += 3
                                    ins.append(Instr(OP_LIT, z, sh, False, tuple(to_limbs(0, sh))))
                                    ins.append(Instr(OP_CONCAT, cat, w + sh, False, (var, z)))
                                    ins.append(Instr(OP_TRUNC, tr, w, i.signed, (cat,)))
                                    term = tr
                                last = k == len(terms) - 1
                                if acc is None:
                                    if sign < 0:
                                        z0 = nxt
                                        nxt += 1
                                        ins.append(Instr(OP_LIT, z0, w, False, tuple(to_limbs(0, w))))
                                        dst = i.dst if last else nxt
                                        if not last:
                                            nxt += 1
                                        ins.append(Instr(OP_SUB, dst, w, i.signed, (z0, term)))
                                        acc = dst
                                    elif last:
                                        ins.append(Instr(OP_TRUNC if sh else OP_ZEXT, i.dst, w, i.signed, (term,)))
                                        acc = i.dst
                                    else:
                                        acc = term
                                else:
                                    dst = i.dst if last else nxt
                                    if not last:
                                        nxt += 1
                                    ins.append(Instr(OP_ADD if sign > 0 else OP_SUB, dst, w, i.signed, (acc, term)))
                                    acc = dst
                            st.reduced += 1
                            continue
                ins.append(i)
            pks.append(Packet(ins, list(pk.effects)))
        blocks.append(Block(pks, nxt))
    out = replace(p, blocks=tuple(blocks))
    # new value ids were appended after the block's; renumber in definition order
    return specialize_program(out)[0] if st.reduced else out, st
