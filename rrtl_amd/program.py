"""Packed machine program: in-memory form and the versioned on-disk blob.

The in-memory classes restate the reference's lowered program
(``PackedMachineProgram`` and friends, crates/rrtl-sim-ir/src/lib.rs:67-323):
signals with word-major u32-limb layouts, memories, and four streams
(``async_reset_comb``, ``comb``, ``tick_next``, ``tick_commit``) of packets, each
packet a list of SSA instructions followed by its effects.

The reference's ``PackedProgram`` derives only ``Clone, Debug, PartialEq, Eq``
(lib.rs:66) and has no serialised form, so the blob below is this library's
own format.  It is the drop-in boundary: a Rust caller serialises its
``PackedMachineProgram`` into exactly these words and hands them to
``rrtl_hip_create``.  Layout (little-endian u32 words)::

    header   : 'RRTL' 'PRG1' version n_signals n_memories
               total_signal_words total_memory_words_per_lane n_names_bytes
    signals  : n_signals  x [width signed offset limbs kind]
    memories : n_memories x [width signed depth addr_width offset limbs]
    4 blocks : [n_values n_packets] then per packet
               [n_instrs n_effects] instrs... effects...
    instr    : [op dst width signed nargs args...]
    effect   : STORE   [0 dst value]
               CAPTURE [1 dst value has_reset rsig rkind rpol nlimbs limbs...]
               MEMWRITE[2 memory enable addr data]
    names    : n_names_bytes of NUL-separated signal names then memory names,
               padded to a word boundary.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

MAGIC = (0x4C545252, 0x31475250)  # b"RRTL", b"PRG1"
VERSION = 1

# PackedInstrKind (lib.rs:266-303), numbered in declaration order
OP_LIT, OP_SIGNAL, OP_NOT, OP_AND, OP_OR, OP_XOR, OP_ADD, OP_SUB, OP_MUL, \
    OP_EQ, OP_NE, OP_LT, OP_MUX, OP_SLICE, OP_ZEXT, OP_SEXT, OP_TRUNC, OP_CAST, \
    OP_CONCAT, OP_MEMREAD = range(20)
OP_NAMES = ["lit", "sig", "not", "and", "or", "xor", "add", "sub", "mul", "eq", "ne",
            "lt", "mux", "slice", "zext", "sext", "trunc", "cast", "concat", "memread"]
OP_OF = {n: i for i, n in enumerate(OP_NAMES)}

# PackedEffect (lib.rs:305-323)
EF_STORE, EF_CAPTURE, EF_MEMWRITE = 0, 1, 2

# PackedSignalKind (lib.rs:142-148)
KIND_INPUT, KIND_OUTPUT, KIND_WIRE, KIND_REG = 0, 1, 2, 3
KIND_OF = {"input": KIND_INPUT, "output": KIND_OUTPUT, "wire": KIND_WIRE, "reg": KIND_REG}

RESET_SYNC, RESET_ASYNC = 0, 1
POL_HIGH, POL_LOW = 0, 1

BLOCK_NAMES = ("async_reset_comb", "comb", "tick_next", "tick_commit")


def limbs(width: int) -> int:
    """``limbs`` (lib.rs:12942): u32 limbs of a value, at least one."""
    return max(1, (width + 31) // 32)


def to_limbs(value: int, width: int) -> List[int]:
    n = limbs(width)
    value &= (1 << width) - 1
    return [(value >> (32 * i)) & 0xFFFFFFFF for i in range(n)]


def from_limbs(ls) -> int:
    v = 0
    for i, l in enumerate(ls):
        v |= int(l) << (32 * i)
    return v


@dataclass
class PSignal:
    name: str
    width: int
    signed: bool
    offset: int
    limbs: int
    kind: int


@dataclass
class PMemory:
    name: str
    width: int
    signed: bool
    depth: int
    addr_width: int
    offset: int
    limbs: int


@dataclass
class PReset:
    signal: int
    value: List[int]  # limbs, fitted to the register width
    kind: int
    polarity: int


@dataclass
class Instr:
    op: int
    dst: int
    width: int
    signed: bool
    args: Tuple[int, ...]


@dataclass
class Effect:
    kind: int
    dst: int = 0            # STORE/CAPTURE: signal; MEMWRITE: memory
    value: int = 0          # STORE/CAPTURE value id
    reset: Optional[PReset] = None
    enable: int = 0
    addr: int = 0
    data: int = 0


@dataclass
class Packet:
    instrs: List[Instr] = field(default_factory=list)
    effects: List[Effect] = field(default_factory=list)


@dataclass
class Block:
    packets: List[Packet] = field(default_factory=list)
    n_values: int = 0


@dataclass
class Program:
    top: str
    signals: List[PSignal]
    memories: List[PMemory]
    blocks: Tuple[Block, Block, Block, Block]
    total_signal_words: int
    total_memory_words_per_lane: int

    # lookup helpers --------------------------------------------------------
    def signal_index(self, name: str) -> int:
        for i, s in enumerate(self.signals):
            if s.name == name:
                return i
        raise KeyError(f"E_SIM_IR_SIGNAL: no signal `{name}`")

    def memory_index(self, name: str) -> int:
        for i, m in enumerate(self.memories):
            if m.name == name:
                return i
        raise KeyError(f"E_SIM_IR_MEMORY: no memory `{name}`")

    @property
    def async_reset_comb(self) -> Block: return self.blocks[0]
    @property
    def comb(self) -> Block: return self.blocks[1]
    @property
    def tick_next(self) -> Block: return self.blocks[2]
    @property
    def tick_commit(self) -> Block: return self.blocks[3]

    def instr_count(self) -> int:
        return sum(len(p.instrs) for b in self.blocks for p in b.packets)

    def is_gate_level(self) -> bool:
        """True when every signal is 1 bit, there are no memories and every op
        is bitwise (the ``BitParallelSimulator::new`` scope, bitparallel.rs:47-85)."""
        if self.memories or any(s.width != 1 for s in self.signals):
            return False
        ok = {OP_AND, OP_OR, OP_XOR, OP_NOT, OP_MUX, OP_SIGNAL, OP_LIT}
        for b in self.blocks:
            for p in b.packets:
                if any(i.op not in ok for i in p.instrs):
                    return False
                if any(e.kind == EF_MEMWRITE for e in p.effects):
                    return False
        return True

    # serialisation -----------------------------------------------------------
    def to_blob(self) -> bytes:
        w: List[int] = []
        names = b"\0".join([s.name.encode() for s in self.signals] +
                           [m.name.encode() for m in self.memories]) + b"\0"
        names += b"\0" * ((-len(names)) % 4)
        w += [MAGIC[0], MAGIC[1], VERSION, len(self.signals), len(self.memories),
              self.total_signal_words, self.total_memory_words_per_lane, len(names)]
        for s in self.signals:
            w += [s.width, int(s.signed), s.offset, s.limbs, s.kind]
        for m in self.memories:
            w += [m.width, int(m.signed), m.depth, m.addr_width, m.offset, m.limbs]
        for b in self.blocks:
            w += [b.n_values, len(b.packets)]
            for p in b.packets:
                w += [len(p.instrs), len(p.effects)]
                for i in p.instrs:
                    w += [i.op, i.dst, i.width, int(i.signed), len(i.args), *i.args]
                for e in p.effects:
                    if e.kind == EF_STORE:
                        w += [EF_STORE, e.dst, e.value]
                    elif e.kind == EF_CAPTURE:
                        r = e.reset
                        if r is None:
                            w += [EF_CAPTURE, e.dst, e.value, 0, 0, 0, 0, 0]
                        else:
                            w += [EF_CAPTURE, e.dst, e.value, 1, r.signal, r.kind, r.polarity,
                                  len(r.value), *r.value]
                    else:
                        w += [EF_MEMWRITE, e.dst, e.enable, e.addr, e.data]
        return struct.pack(f"<{len(w)}I", *[x & 0xFFFFFFFF for x in w]) + names

    @staticmethod
    def from_blob(blob: bytes, top: str = "top") -> "Program":
        n_words = len(blob) // 4
        words = struct.unpack(f"<{n_words}I", blob[: n_words * 4])
        pos = 0

        def take(n=1):
            nonlocal pos
            if pos + n > len(words):
                raise ValueError("E_PROGRAM_TRUNCATED: program blob ends early")
            v = words[pos:pos + n]
            pos += n
            return v

        hdr = take(8)
        if (hdr[0], hdr[1]) != MAGIC or hdr[2] != VERSION:
            raise ValueError("E_PROGRAM_MAGIC: not an RRTL packed program blob")
        nsig, nmem, tsw, tmw, nbytes = hdr[3], hdr[4], hdr[5], hdr[6], hdr[7]
        sigs, mems = [], []
        for _ in range(nsig):
            wd, sg, off, lb, kd = take(5)
            sigs.append(PSignal("", wd, bool(sg), off, lb, kd))
        for _ in range(nmem):
            wd, sg, dp, aw, off, lb = take(6)
            mems.append(PMemory("", wd, bool(sg), dp, aw, off, lb))
        blocks = []
        for _ in range(4):
            nval, npk = take(2)
            blk = Block(n_values=nval)
            for _ in range(npk):
                ni, ne = take(2)
                pk = Packet()
                for _ in range(ni):
                    op, dst, wd, sg, na = take(5)
                    pk.instrs.append(Instr(op, dst, wd, bool(sg), tuple(take(na))))
                for _ in range(ne):
                    (k,) = take(1)
                    if k == EF_STORE:
                        d, v = take(2)
                        pk.effects.append(Effect(EF_STORE, d, v))
                    elif k == EF_CAPTURE:
                        d, v, has, rs, rk, rp, nl = take(7)
                        lv = list(take(nl))
                        pk.effects.append(Effect(EF_CAPTURE, d, v,
                                                 PReset(rs, lv, rk, rp) if has else None))
                    elif k == EF_MEMWRITE:
                        m, en, ad, da = take(4)
                        pk.effects.append(Effect(EF_MEMWRITE, m, enable=en, addr=ad, data=da))
                    else:
                        raise ValueError(f"E_PROGRAM_EFFECT: unknown effect kind {k}")
                blk.packets.append(pk)
            blocks.append(blk)
        raw = blob[pos * 4: pos * 4 + nbytes]
        names = raw.split(b"\0")
        for i, s in enumerate(sigs):
            s.name = names[i].decode() if i < len(names) else f"s{i}"
        for j, m in enumerate(mems):
            k = nsig + j
            m.name = names[k].decode() if k < len(names) else f"m{j}"
        return Program(top, sigs, mems, tuple(blocks), tsw, tmw)
