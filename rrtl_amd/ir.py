"""Typed RTL design IR: the subset of RRTL's builder that the batch engines consume.

Restates the expression and module model of the reference:
  * ``BitType`` / ``Expr`` / ``Reset`` -- crates/rrtl-ir/src/lib.rs:15-38, :230-280, :453-472
  * ``ModuleBuilder`` (input/output/wire/reg/mem, assign, next, reset, mem write)
    -- crates/rrtl-core/src/lib.rs:1079-2600
  * type rules of ``expr_type`` -- crates/rrtl-core/src/lib.rs:5424-5560
    (bitwise ops need equal types, arithmetic/compare need equal numeric types,
    mux condition is uint(1), slices stay in range).

Designs are flat (one module); instance flattening is not needed by the
workloads this package runs.  Every expression node carries its ``BitType`` and
is hashable structurally so the machine lowering can memoise common
sub-expressions exactly like ``MachinePacketLowerer``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple, Union


class DesignError(ValueError):
    """A design rule violation, carrying an ``E_*`` diagnostic code."""

    def __init__(self, code: str, message: str):
        super().__init__(f"{code}: {message}")
        self.code = code


@dataclass(frozen=True)
class BitType:
    width: int
    signed: bool = False

    def __post_init__(self):
        if self.width <= 0:
            raise DesignError("E_WIDTH_ZERO", "bit width must be positive")


def uint(width: int) -> BitType:
    return BitType(width, False)


def sint(width: int) -> BitType:
    return BitType(width, True)


def mask(width: int) -> int:
    return (1 << width) - 1


def encode_value(value: int, ty: BitType) -> int:
    """Two's-complement encode ``value`` into ``ty.width`` bits (``encode_limbs``)."""
    return value & mask(ty.width)


# Signal kinds (PackedSignalKind, crates/rrtl-sim-ir/src/lib.rs:142-148)
INPUT, OUTPUT, WIRE, REG = "input", "output", "wire", "reg"


@dataclass(eq=False)
class Signal:
    module: "Module"
    index: int  # declaration order among non-memory signals
    name: str
    ty: BitType
    kind: str

    @property
    def width(self) -> int:
        return self.ty.width

    def value(self) -> "Expr":
        return Expr("sig", self.ty, (), self)

    # allow signals to be used directly in expressions
    def __and__(self, o): return self.value() & o
    def __or__(self, o): return self.value() | o
    def __xor__(self, o): return self.value() ^ o
    def __add__(self, o): return self.value() + o
    def __sub__(self, o): return self.value() - o
    def __mul__(self, o): return self.value() * o
    def __invert__(self): return ~self.value()
    def __hash__(self): return id(self)
    def __repr__(self): return f"Signal({self.name}:{self.kind}:{self.ty.width})"


@dataclass(eq=False)
class Memory:
    module: "Module"
    index: int
    name: str
    ty: BitType  # data type
    depth: int
    addr_width: int

    def read(self, addr: "ExprLike") -> "Expr":
        a = as_expr(addr)
        return Expr("memread", self.ty, (a,), self)

    def __hash__(self): return id(self)
    def __repr__(self): return f"Memory({self.name}:{self.depth}x{self.ty.width})"


@dataclass(frozen=True)
class Reset:
    """Register reset (``Reset`` in crates/rrtl-ir/src/lib.rs:453-472)."""
    signal: Signal
    value: int
    kind: str = "sync"          # "sync" | "async"
    polarity: str = "high"      # "high" | "low"


ExprLike = Union["Expr", Signal]


def as_expr(x: ExprLike) -> "Expr":
    if isinstance(x, Expr):
        return x
    if isinstance(x, Signal):
        return x.value()
    raise DesignError("E_EXPR", f"cannot use {x!r} as an expression (wrap ints with lit())")


class Expr:
    """A typed expression node.  ``op`` is one of the ``Expr`` variants of
    crates/rrtl-ir/src/lib.rs:230-280 (lower-case), ``attr`` holds the variant's
    non-expression payload (literal value, signal, lsb, memory)."""

    __slots__ = ("op", "ty", "args", "attr", "_hash")

    def __init__(self, op: str, ty: BitType, args: Tuple["Expr", ...], attr=None):
        self.op = op
        self.ty = ty
        self.args = tuple(args)
        self.attr = attr
        self._hash = hash((op, ty, self.args, attr if not isinstance(attr, (Signal, Memory)) else id(attr)))

    def __hash__(self):
        return self._hash

    def __eq__(self, other):
        if self is other:
            return True
        if not isinstance(other, Expr) or self._hash != other._hash:
            return False
        return (self.op == other.op and self.ty == other.ty and self.args == other.args
                and (self.attr is other.attr or self.attr == other.attr))

    def __repr__(self):
        if self.op == "lit":
            return f"lit({self.attr},{self.ty.width})"
        if self.op == "sig":
            return self.attr.name
        return f"{self.op}({', '.join(map(repr, self.args))})"

    @property
    def width(self) -> int:
        return self.ty.width

    # bitwise ----------------------------------------------------------------
    def _same(self, other: ExprLike, op: str, numeric: bool) -> "Expr":
        o = as_expr(other)
        if o.ty.width != self.ty.width or (numeric and o.ty.signed != self.ty.signed):
            raise DesignError("E_TYPE_MISMATCH", f"{op}: {self.ty} vs {o.ty}")
        if not numeric and o.ty != self.ty:
            raise DesignError("E_TYPE_MISMATCH", f"{op}: {self.ty} vs {o.ty}")
        return Expr(op, self.ty, (self, o))

    def __and__(self, o): return self._same(o, "and", False)
    def __or__(self, o): return self._same(o, "or", False)
    def __xor__(self, o): return self._same(o, "xor", False)
    def __add__(self, o): return self._same(o, "add", True)
    def __sub__(self, o): return self._same(o, "sub", True)
    def __mul__(self, o): return self._same(o, "mul", True)
    def __invert__(self): return Expr("not", self.ty, (self,))

    def _cmp(self, other: ExprLike, op: str) -> "Expr":
        o = as_expr(other)
        if o.ty != self.ty:
            raise DesignError("E_TYPE_MISMATCH", f"{op}: {self.ty} vs {o.ty}")
        return Expr(op, uint(1), (self, o))

    def eq(self, o): return self._cmp(o, "eq")
    def ne(self, o): return self._cmp(o, "ne")
    def lt(self, o): return self._cmp(o, "lt")

    def slice(self, lsb: int, width: int) -> "Expr":
        if width <= 0 or lsb < 0 or lsb + width > self.ty.width:
            raise DesignError("E_SLICE_RANGE", f"slice [{lsb}+:{width}] of {self.ty.width} bits")
        return Expr("slice", uint(width), (self,), lsb)

    def bit(self, i: int) -> "Expr":
        return self.slice(i, 1)

    def zext(self, width: int) -> "Expr":
        if width < self.ty.width:
            raise DesignError("E_ZEXT_WIDTH", "zext must not narrow")
        return Expr("zext", BitType(width, self.ty.signed), (self,), width)

    def sext(self, width: int) -> "Expr":
        if width < self.ty.width:
            raise DesignError("E_SEXT_WIDTH", "sext must not narrow")
        return Expr("sext", BitType(width, self.ty.signed), (self,), width)

    def trunc(self, width: int) -> "Expr":
        if width > self.ty.width or width <= 0:
            raise DesignError("E_TRUNC_WIDTH", "trunc must narrow")
        return Expr("trunc", BitType(width, self.ty.signed), (self,), width)

    def cast(self, signed: bool) -> "Expr":
        return Expr("cast", BitType(self.ty.width, signed), (self,), signed)


def lit(value: int, width: int, signed: bool = False) -> Expr:
    ty = BitType(width, signed)
    lo = -(1 << (width - 1)) if signed else 0
    hi = mask(width) if not signed else mask(width - 1)
    if not (lo <= value <= hi) and not (0 <= value <= mask(width)):
        raise DesignError("E_LIT_VALUE_WIDTH", f"literal {value} does not fit in {width} bits")
    return Expr("lit", ty, (), encode_value(value, ty))


def mux(cond: ExprLike, then_expr: ExprLike, else_expr: ExprLike) -> Expr:
    c, t, e = as_expr(cond), as_expr(then_expr), as_expr(else_expr)
    if c.ty != uint(1):
        raise DesignError("E_MUX_COND_WIDTH", f"mux condition must be uint(1), found {c.ty}")
    if t.ty != e.ty:
        raise DesignError("E_TYPE_MISMATCH", f"mux arms {t.ty} vs {e.ty}")
    return Expr("mux", t.ty, (c, t, e))


def concat(*parts: ExprLike) -> Expr:
    """MSB-first concatenation (``Expr::Concat``)."""
    ps = tuple(as_expr(p) for p in parts)
    if not ps:
        raise DesignError("E_CONCAT_EMPTY", "concat needs at least one part")
    return Expr("concat", uint(sum(p.ty.width for p in ps)), ps)


@dataclass
class MemWrite:
    mem: Memory
    enable: Expr
    addr: Expr
    data: Expr


class Module:
    """A flat RTL module built like ``ModuleBuilder`` (rrtl-core/src/lib.rs:1079)."""

    def __init__(self, name: str):
        self.name = name
        self.signals: List[Signal] = []
        self.memories: List[Memory] = []
        self.assigns: Dict[Signal, Expr] = {}
        self.assign_order: List[Signal] = []
        self.nexts: Dict[Signal, Expr] = {}
        self.resets: Dict[Signal, Reset] = {}
        self.clocks: Dict[Signal, Signal] = {}
        self.registers: List[Signal] = []
        self.mem_writes: List[MemWrite] = []
        self._names: Dict[str, object] = {}

    # declarations ----------------------------------------------------------
    def _declare(self, name: str, ty: Union[BitType, int], kind: str) -> Signal:
        if name in self._names:
            raise DesignError("E_DUPLICATE_SIGNAL", f"signal `{name}` declared twice")
        if isinstance(ty, int):
            ty = uint(ty)
        s = Signal(self, len(self.signals), name, ty, kind)
        self.signals.append(s)
        self._names[name] = s
        if kind == REG:
            self.registers.append(s)
        return s

    def input(self, name, ty): return self._declare(name, ty, INPUT)
    def output(self, name, ty): return self._declare(name, ty, OUTPUT)
    def wire(self, name, ty): return self._declare(name, ty, WIRE)
    def reg(self, name, ty): return self._declare(name, ty, REG)

    def mem(self, name: str, ty: Union[BitType, int], depth: int, addr_width: Optional[int] = None) -> Memory:
        if name in self._names:
            raise DesignError("E_DUPLICATE_SIGNAL", f"signal `{name}` declared twice")
        if isinstance(ty, int):
            ty = uint(ty)
        if depth <= 0:
            raise DesignError("E_MEM_DEPTH", "memory depth must be positive")
        aw = addr_width if addr_width is not None else max(1, (depth - 1).bit_length())
        m = Memory(self, len(self.memories), name, ty, depth, aw)
        self.memories.append(m)
        self._names[name] = m
        return m

    def signal(self, name: str):
        return self._names[name]

    # behaviour -------------------------------------------------------------
    def assign(self, dst: Signal, expr: ExprLike) -> None:
        e = as_expr(expr)
        if dst.kind not in (OUTPUT, WIRE):
            raise DesignError("E_ASSIGN_TARGET", f"cannot assign {dst.kind} `{dst.name}`")
        if dst in self.assigns:
            raise DesignError("E_MULTIPLE_DRIVERS", f"`{dst.name}` assigned twice")
        if e.ty.width != dst.ty.width:
            raise DesignError("E_TYPE_MISMATCH", f"assign `{dst.name}`: {dst.ty} <- {e.ty}")
        self.assigns[dst] = e
        self.assign_order.append(dst)

    def next(self, reg: Signal, expr: ExprLike) -> None:
        e = as_expr(expr)
        if reg.kind != REG:
            raise DesignError("E_NEXT_TARGET", f"`{reg.name}` is not a register")
        if e.ty.width != reg.ty.width:
            raise DesignError("E_TYPE_MISMATCH", f"next `{reg.name}`: {reg.ty} <- {e.ty}")
        self.nexts[reg] = e

    def clock(self, reg: Signal, clk: Signal) -> None:
        self.clocks[reg] = clk

    def reset(self, reg: Signal, signal: Signal, value: int, kind: str = "sync",
              polarity: str = "high") -> None:
        if reg.kind != REG:
            raise DesignError("E_RESET_TARGET", f"`{reg.name}` is not a register")
        if kind not in ("sync", "async") or polarity not in ("high", "low"):
            raise DesignError("E_RESET_KIND", f"bad reset kind/polarity {kind}/{polarity}")
        self.resets[reg] = Reset(signal, value, kind, polarity)

    def mem_write(self, mem: Memory, enable: ExprLike, addr: ExprLike, data: ExprLike) -> None:
        en, a, d = as_expr(enable), as_expr(addr), as_expr(data)
        if en.ty.width != 1:
            raise DesignError("E_MEM_ENABLE", "memory write enable must be 1 bit")
        if d.ty.width != mem.ty.width:
            raise DesignError("E_TYPE_MISMATCH", f"memory `{mem.name}` data {mem.ty} <- {d.ty}")
        self.mem_writes.append(MemWrite(mem, en, a, d))

    # validation (validate_design, rrtl-core/src/lib.rs:3417) -----------------
    def validate(self) -> None:
        for s in self.signals:
            if s.kind in (OUTPUT, WIRE) and s not in self.assigns:
                raise DesignError("E_UNDRIVEN", f"{s.kind} `{s.name}` has no driver")
            if s.kind == REG and s not in self.nexts:
                raise DesignError("E_REG_NEXT", f"register `{s.name}` has no next value")
        for e in list(self.assigns.values()) + list(self.nexts.values()):
            _check_owner(self, e)


def _check_owner(m: Module, e: Expr) -> None:
    stack = [e]
    seen = set()
    while stack:
        x = stack.pop()
        if id(x) in seen:
            continue
        seen.add(id(x))
        if x.op in ("sig", "memread") and x.attr.module is not m:
            raise DesignError("E_FOREIGN_SIGNAL", f"`{x.attr.name}` belongs to another module")
        stack.extend(x.args)


def expr_signals(e: Expr) -> List[Signal]:
    """Signals read by ``e`` (used for levelization and cone analysis)."""
    out, stack, seen = [], [e], set()
    while stack:
        x = stack.pop()
        if id(x) in seen:
            continue
        seen.add(id(x))
        if x.op == "sig":
            out.append(x.attr)
        stack.extend(x.args)
    return out
