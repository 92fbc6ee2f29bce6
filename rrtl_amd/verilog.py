"""Verilog-2005 subset importer: source text -> a flat ``ir.Module``.

The reference reaches hand-written RTL through its SystemVerilog frontend
(``import_sv``, crates/rrtl-sv-frontend/src/lib.rs:25; preprocessor
preprocess.rs, parser parser.rs, lowering lower.rs).  This is this package's
own importer for the subset the reference's benchmark designs use
(bench/sv/picorv32.v with bench/sv/picorv32_soc.v, crc32.sv, workload.sv):

* preprocessor: ``define`` (with arguments), ``ifdef/ifndef/elsif/else/endif``,
  ``undef``, ``timescale``/``default_nettype`` (ignored);
* modules with ``#(parameter ...)`` headers, ANSI and non-ANSI ports,
  ``parameter``/``localparam``, ``wire``/``reg``/``logic``/``integer``
  declarations (with initialisers on nets), unpacked-array memories;
* ``assign``, ``always @*``/``always_comb``, ``always @(posedge clk)``/
  ``always_ff``, ``initial`` (ignored), empty tasks, ``generate if``/``for``,
  named-port module instances with parameter overrides (flattened);
* statements: ``begin/end``, ``if/else``, ``case/casez/casex``, blocking and
  non-blocking assignments to whole variables, bit and part selects,
  concatenations and memory words, ``for`` loops over integer variables
  (unrolled), system tasks (ignored);
* expressions with the Verilog-2005 sizing rules (IEEE 1364-2005 5.4 and
  5.5): context-determined widths, signedness from all operands,
  ``$signed``/``$unsigned``, shifts with self-determined amounts.

Semantics chosen where Verilog leaves freedom (both CPU oracles and the HIP
engines run the same lowered program, so they do not affect parity):
``x``/``z`` bits are 0; a variable of an ``always @*`` block that a path
leaves unassigned is 0 on that path (``full_case`` don't-care, no latch);
a clocked block's blocking-assigned variable is a register whose next value is
its last assignment; out-of-range reads give 0.

Elaboration folds constants (parameters fold away disabled features), and the
returned module keeps only logic that reaches a top-level output, a memory or
a name in ``keep``: combinational variables are inlined into their readers, so
the lowered program computes every register's next value in one packet with
common sub-expressions shared (the reference's ``specialize_program``,
crates/rrtl-sim-ir/src/specialize.rs:49, removes dead and constant logic the
same way).
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from . import ir
from .ir import BitType, DesignError, Expr, uint


class VerilogError(DesignError):
    def __init__(self, message: str, line: Optional[int] = None):
        where = f" (line {line})" if line is not None else ""
        super().__init__("E_SV_IMPORT", message + where)


# =============================================================== preprocessor
def _strip_comments(text: str) -> str:
    out, i, n = [], 0, len(text)
    while i < n:
        c = text[i]
        if c == '"':
            j = i + 1
            while j < n and text[j] != '"':
                j += 2 if text[j] == "\\" else 1
            out.append(text[i:j + 1])
            i = j + 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            if j < 0:
                raise VerilogError("unterminated block comment")
            out.append("\n" * text.count("\n", i, j))  # keep line numbers
            i = j + 2
        else:
            out.append(c)
            i += 1
    return "".join(out)


_DIRECTIVES = ("define", "undef", "ifdef", "ifndef", "elsif", "else", "endif", "timescale",
               "default_nettype", "resetall", "celldefine", "endcelldefine", "include")


def _split_args(s: str, i: int) -> Tuple[List[str], int]:
    """Macro call arguments from s[i] == '(' up to the matching ')'."""
    depth, args, cur, j = 0, [], [], i
    while j < len(s):
        c = s[j]
        if c in "([{":
            depth += 1
            if depth > 1:
                cur.append(c)
        elif c in ")]}":
            depth -= 1
            if depth == 0:
                args.append("".join(cur).strip())
                return args, j + 1
            cur.append(c)
        elif c == "," and depth == 1:
            args.append("".join(cur).strip())
            cur = []
        elif c == '"':
            k = s.index('"', j + 1)
            cur.append(s[j:k + 1])
            j = k
        else:
            cur.append(c)
        j += 1
    raise VerilogError("unterminated macro arguments")


def _expand(line: str, macros: Dict[str, Tuple[Optional[List[str]], str]], depth: int = 0) -> str:
    if "`" not in line:
        return line
    if depth > 64:
        raise VerilogError("macro expansion too deep")
    out, i = [], 0
    while i < len(line):
        c = line[i]
        if c != "`":
            out.append(c)
            i += 1
            continue
        m = re.match(r"`([A-Za-z_]\w*)", line[i:])
        if not m:
            raise VerilogError(f"stray backtick in {line.strip()!r}")
        name = m.group(1)
        if name not in macros:
            raise VerilogError(f"undefined macro `{name}")
        params, body = macros[name]
        i += m.end()
        if params is not None:
            k = i
            while k < len(line) and line[k] in " \t":
                k += 1
            if k >= len(line) or line[k] != "(":
                raise VerilogError(f"macro `{name} needs arguments")
            args, i = _split_args(line, k)
            if len(args) != len(params):
                raise VerilogError(f"macro `{name} takes {len(params)} arguments, got {len(args)}")
            for p, a in zip(params, args):
                body = re.sub(r"\b%s\b" % re.escape(p), lambda _m, a=a: a, body)
        out.append(_expand(body, macros, depth + 1))
    return "".join(out)


def preprocess(text: str, defines: Optional[Dict[str, str]] = None) -> str:
    """Runs the directives; returns the source with one output line per input line."""
    text = _strip_comments(text)
    macros: Dict[str, Tuple[Optional[List[str]], str]] = {k: (None, v) for k, v in (defines or {}).items()}
    lines = text.split("\n")
    out: List[str] = []
    # stack entries: [active, taken, parent_active]
    stack: List[List[bool]] = []
    active = True
    i = 0
    while i < len(lines):
        line = lines[i]
        i += 1
        s = line.strip()
        m = re.match(r"`(\w+)", s)
        if m and m.group(1) in _DIRECTIVES:
            d = m.group(1)
            rest = s[m.end():].strip()
            if d in ("ifdef", "ifndef"):
                name = rest.split()[0]
                cond = (name in macros) == (d == "ifdef")
                stack.append([active and cond, cond, active])
                active = active and cond
            elif d == "elsif":
                if not stack:
                    raise VerilogError("`elsif without `ifdef", i)
                top = stack[-1]
                cond = (not top[1]) and rest.split()[0] in macros
                top[0] = top[2] and cond
                top[1] = top[1] or cond
                active = top[0]
            elif d == "else":
                if not stack:
                    raise VerilogError("`else without `ifdef", i)
                top = stack[-1]
                top[0] = top[2] and not top[1]
                top[1] = True
                active = top[0]
            elif d == "endif":
                if not stack:
                    raise VerilogError("`endif without `ifdef", i)
                active = stack.pop()[2]
            elif active and d == "define":
                body = rest
                while body.endswith("\\") and i < len(lines):
                    body = body[:-1] + " " + lines[i].strip()
                    out.append("")
                    i += 1
                mm = re.match(r"([A-Za-z_]\w*)(\(([^)]*)\))?\s*(.*)$", body)
                if not mm:
                    raise VerilogError(f"bad `define {body!r}", i)
                params = [p.strip() for p in mm.group(3).split(",")] if mm.group(2) else None
                macros[mm.group(1)] = (params, mm.group(4))
            elif active and d == "undef":
                macros.pop(rest.split()[0], None)
            elif active and d == "include":
                raise VerilogError("`include is not supported: concatenate the sources instead", i)
            out.append("")
            continue
        out.append(_expand(line, macros) if active else "")
    if stack:
        raise VerilogError("unterminated `ifdef")
    return "\n".join(out)


# ===================================================================== lexer
@dataclass
class Tok:
    kind: str   # id sys num str op eof
    val: object
    line: int


_OPS = sorted(["<<<", ">>>", "===", "!==", "==", "!=", "<=", ">=", "&&", "||", "<<", ">>", "~&", "~|", "~^", "^~",
               "**", "+:", "-:", "->", "::", "@*", "(", ")", "[", "]", "{", "}", ";", ",", ":", "?", "=", "<", ">",
               "+", "-", "*", "/", "%", "&", "|", "^", "~", "!", "@", "#", ".", "'"], key=len, reverse=True)
_NUM_BASED = re.compile(r"(\d[\d_]*)?\s*'\s*([sS]?)([bBoOdDhH])\s*([0-9a-fA-FxXzZ?_]+)")
_NUM_DEC = re.compile(r"\d[\d_]*")
_IDENT = re.compile(r"[A-Za-z_][A-Za-z0-9_$]*")
_SYS = re.compile(r"\$[A-Za-z_][A-Za-z0-9_$]*")


@dataclass(frozen=True)
class Num:
    width: Optional[int]    # None: unsized
    signed: bool
    value: int              # x/z/? bits are 0
    xz: int = 0             # mask of x/z/? bits (casez/casex labels)
    based: bool = False


def _parse_based(size: Optional[str], s: str, base: str, digits: str, line: int) -> Num:
    digits = digits.replace("_", "")
    bits = {"b": 1, "o": 3, "h": 4}.get(base.lower())
    signed = bool(s)
    if bits is None:  # decimal
        if any(c in "xXzZ?" for c in digits):
            value, xz, nd = 0, -1, 32
        else:
            value, xz = int(digits), 0
            nd = max(1, value.bit_length())
    else:
        value = xz = 0
        for c in digits:
            value <<= bits
            xz <<= bits
            if c in "xXzZ?":
                xz |= (1 << bits) - 1
            else:
                value |= int(c, 16)
        nd = bits * len(digits)
    if size is not None:
        width = int(size.replace("_", ""))
        if width <= 0:
            raise VerilogError("zero-width literal", line)
        value &= (1 << width) - 1
        xz &= (1 << width) - 1
    else:
        width = None
        if nd > 32:
            width = nd  # an unsized literal wider than an integer keeps its digits
    if xz == -1:
        xz = (1 << (width or 32)) - 1
    return Num(width, signed, value, xz, True)


def lex(text: str) -> List[Tok]:
    toks: List[Tok] = []
    i, n, line = 0, len(text), 1
    while i < n:
        c = text[i]
        if c == "\n":
            line += 1
            i += 1
            continue
        if c in " \t\r\f\v":
            i += 1
            continue
        if text.startswith("(*", i):
            # attribute (* ... *); "(*)" is the sensitivity list of always @(*)
            j = i + 2
            while j < n and text[j] in " \t":
                j += 1
            if j < n and text[j] != ")":
                k = text.find("*)", i + 2)
                if k < 0:
                    raise VerilogError("unterminated attribute", line)
                line += text.count("\n", i, k)
                i = k + 2
                continue
        if c == '"':
            j = i + 1
            buf = []
            while j < n and text[j] != '"':
                if text[j] == "\\" and j + 1 < n:
                    buf.append({"n": "\n", "t": "\t", "\\": "\\", '"': '"'}.get(text[j + 1], text[j + 1]))
                    j += 2
                else:
                    buf.append(text[j])
                    j += 1
            toks.append(Tok("str", "".join(buf), line))
            i = j + 1
            continue
        m = _NUM_BASED.match(text, i)
        if m and (c.isdigit() or c == "'"):
            toks.append(Tok("num", _parse_based(m.group(1), m.group(2), m.group(3), m.group(4), line), line))
            line += text.count("\n", i, m.end())
            i = m.end()
            continue
        if c.isdigit():
            m = _NUM_DEC.match(text, i)
            v = int(m.group(0).replace("_", ""))
            toks.append(Tok("num", Num(None, True, v), line))
            i = m.end()
            continue
        if c == "\\":  # escaped identifier
            j = i + 1
            while j < n and not text[j].isspace():
                j += 1
            toks.append(Tok("id", text[i + 1:j], line))
            i = j
            continue
        m = _IDENT.match(text, i)
        if m:
            toks.append(Tok("id", m.group(0), line))
            i = m.end()
            continue
        m = _SYS.match(text, i)
        if m:
            toks.append(Tok("sys", m.group(0), line))
            i = m.end()
            continue
        for op in _OPS:
            if text.startswith(op, i):
                toks.append(Tok("op", op, line))
                i += len(op)
                break
        else:
            raise VerilogError(f"unexpected character {c!r}", line)
    toks.append(Tok("eof", None, line))
    return toks


# ====================================================================== AST
# Expressions are tuples: ("num", Num) ("str", text) ("id", name)
# ("sel", base, kind, a, b) with kind "bit" (a), "part" (a=msb, b=lsb),
# "up"/"down" (a=base, b=width); base is an expression (an id or a sel)
# ("un", op, x) ("bin", op, x, y) ("tern", c, x, y) ("cat", [x...])
# ("rep", count, [x...]) ("call", name, [args])
# Statements: ("blk", [s...]) ("if", c, s, s|None) ("case", kind, sel, [(labels|None, s)])
# ("asg", blocking, lv, x) ("for", init, cond, step, body) ("nop",) ("task", name)
# Lvalues are expressions of ids, sels and cats.


@dataclass
class ModuleAst:
    name: str
    line: int
    params: List[tuple] = field(default_factory=list)   # (name, local, range, signed, kind, expr)
    ports: List[str] = field(default_factory=list)
    items: List[tuple] = field(default_factory=list)


class Parser:
    def __init__(self, toks: List[Tok]):
        self.t = toks
        self.i = 0

    # --- token helpers
    def peek(self, k: int = 0) -> Tok:
        return self.t[min(self.i + k, len(self.t) - 1)]

    def at(self, val, k: int = 0) -> bool:
        tk = self.peek(k)
        return tk.kind in ("op", "id") and tk.val == val

    def next(self) -> Tok:
        tk = self.t[self.i]
        self.i += 1
        return tk

    def expect(self, val) -> Tok:
        tk = self.next()
        if tk.val != val or tk.kind not in ("op", "id"):
            raise VerilogError(f"expected {val!r}, found {tk.val!r}", tk.line)
        return tk

    def accept(self, val) -> bool:
        if self.at(val):
            self.i += 1
            return True
        return False

    def ident(self) -> str:
        tk = self.next()
        if tk.kind != "id":
            raise VerilogError(f"expected an identifier, found {tk.val!r}", tk.line)
        return tk.val

    # --- source
    def source(self) -> List[ModuleAst]:
        mods = []
        while self.peek().kind != "eof":
            if self.at("module") or self.at("macromodule"):
                mods.append(self.module())
            else:
                tk = self.peek()
                raise VerilogError(f"expected `module`, found {tk.val!r}", tk.line)
        return mods

    def module(self) -> ModuleAst:
        line = self.next().line
        m = ModuleAst(self.ident(), line)
        if self.accept("#"):
            self.expect("(")
            kind = "parameter"
            while not self.at(")"):
                if self.at("parameter") or self.at("localparam"):
                    kind = self.next().val
                m.params.extend(self.param_decl(kind == "localparam", header=True))
                if not self.accept(","):
                    break
            self.expect(")")
        if self.accept("("):
            if not self.at(")"):
                self.port_list(m)
            self.expect(")")
        self.expect(";")
        while not self.at("endmodule"):
            self.item(m.items, m)
        self.expect("endmodule")
        return m

    def port_list(self, m: ModuleAst) -> None:
        direction = kind = None
        signed, rng = False, None
        while True:
            if self.peek().val in ("input", "output", "inout") and self.peek().kind == "id":
                direction = self.next().val
                kind, signed, rng = "wire", False, None
                if self.peek().val in ("wire", "reg", "logic") and self.peek().kind == "id":
                    kind = self.next().val
                if self.accept("signed"):
                    signed = True
                if self.at("["):
                    rng = self.range_()
            name = self.ident()
            m.ports.append(name)
            if direction is not None:
                m.items.append(("port", direction, kind, signed, rng, name, self.peek().line))
            if not self.accept(","):
                return

    def range_(self) -> Tuple[tuple, tuple]:
        self.expect("[")
        a = self.expr()
        self.expect(":")
        b = self.expr()
        self.expect("]")
        return (a, b)

    def param_decl(self, local: bool, header: bool = False) -> List[tuple]:
        kind, signed, rng = None, False, None
        if self.peek().val in ("integer", "logic", "reg", "bit") and self.peek().kind == "id":
            kind = self.next().val
        if self.accept("signed"):
            signed = True
        if self.at("["):
            rng = self.range_()
        out = []
        while True:
            name = self.ident()
            self.expect("=")
            out.append((name, local, rng, signed, kind, self.expr()))
            if header:
                return out
            if not self.accept(","):
                return out
            if self.peek().val in ("parameter", "localparam"):
                return out

    def item(self, items: List[tuple], m: ModuleAst) -> None:
        tk = self.peek()
        v = tk.val
        if tk.kind == "op" and v == ";":
            self.next()
            return
        if tk.kind != "id":
            raise VerilogError(f"unexpected {v!r} in module body", tk.line)
        if v in ("parameter", "localparam"):
            self.next()
            for p in self.param_decl(v == "localparam"):
                m.params.append(p)
            self.expect(";")
        elif v in ("input", "output", "inout"):
            self.next()
            kind, signed, rng = "wire", False, None
            if self.peek().val in ("wire", "reg", "logic") and self.peek().kind == "id":
                kind = self.next().val
            if self.accept("signed"):
                signed = True
            if self.at("["):
                rng = self.range_()
            while True:
                items.append(("port", v, kind, signed, rng, self.ident(), tk.line))
                if not self.accept(","):
                    break
            self.expect(";")
        elif v in ("wire", "reg", "logic", "integer", "genvar", "bit", "tri"):
            self.next()
            signed, rng = v == "integer", None
            if self.accept("signed"):
                signed = True
            if self.at("["):
                rng = self.range_()
            while True:
                name = self.ident()
                dims = []
                while self.at("["):
                    dims.append(self.range_())
                init = self.expr() if self.accept("=") else None
                items.append(("decl", v, signed, rng, name, dims, init, tk.line))
                if not self.accept(","):
                    break
            self.expect(";")
        elif v == "assign":
            self.next()
            while True:
                lv = self.lvalue()
                self.expect("=")
                items.append(("assign", lv, self.expr(), tk.line))
                if not self.accept(","):
                    break
            self.expect(";")
        elif v in ("always", "always_comb", "always_ff", "always_latch"):
            self.next()
            if v == "always_latch":
                raise VerilogError("always_latch is not supported", tk.line)
            sens = ("comb", None)
            if v != "always_comb":
                sens = self.sensitivity(tk.line)
            items.append(("always", sens, self.stmt(), tk.line))
        elif v == "initial":
            self.next()
            items.append(("initial", self.stmt(), tk.line))
        elif v == "task":
            self.next()
            name = self.ident()
            self.expect(";")
            body = []
            while not self.at("endtask"):
                body.append(self.stmt())
            self.expect("endtask")
            items.append(("task", name, body, tk.line))
        elif v == "function":
            raise VerilogError("functions are not supported", tk.line)
        elif v == "generate":
            self.next()
            while not self.at("endgenerate"):
                self.item(items, m)
            self.expect("endgenerate")
        elif v == "if":
            items.append(self.gen_if())
        elif v == "for":
            items.append(self.gen_for())
        elif v == "begin":
            self.next()
            if self.accept(":"):
                self.ident()
            while not self.at("end"):
                self.item(items, m)
            self.expect("end")
        else:
            items.append(self.instance())

    def gen_block(self) -> List[tuple]:
        items: List[tuple] = []
        if self.accept("begin"):
            if self.accept(":"):
                self.ident()
            while not self.at("end"):
                self.item(items, None)
            self.expect("end")
        else:
            self.item(items, None)
        return items

    def gen_if(self) -> tuple:
        line = self.expect("if").line
        self.expect("(")
        c = self.expr()
        self.expect(")")
        then = self.gen_block()
        other: List[tuple] = []
        if self.accept("else"):
            other = [self.gen_if()] if self.at("if") else self.gen_block()
        return ("genif", c, then, other, line)

    def gen_for(self) -> tuple:
        line = self.expect("for").line
        self.expect("(")
        self.accept("genvar")
        var = self.ident()
        self.expect("=")
        init = self.expr()
        self.expect(";")
        cond = self.expr()
        self.expect(";")
        sv = self.ident()
        self.expect("=")
        step = self.expr()
        self.expect(")")
        if sv != var:
            raise VerilogError("generate for must step its own genvar", line)
        return ("genfor", var, init, cond, step, self.gen_block(), line)

    def sensitivity(self, line: int):
        if self.accept("@*"):
            return ("comb", None)
        self.expect("@")
        if self.accept("*"):
            return ("comb", None)
        self.expect("(")
        if self.accept("*"):
            self.expect(")")
            return ("comb", None)
        edges = []
        while True:
            edge = None
            if self.peek().val in ("posedge", "negedge"):
                edge = self.next().val
            edges.append((edge, self.ident()))
            if not (self.accept("or") or self.accept(",")):
                break
        self.expect(")")
        if all(e is None for e, _ in edges):
            return ("comb", None)
        if len(edges) != 1 or edges[0][0] != "posedge":
            raise VerilogError("only single-clock `always @(posedge clk)` blocks are supported "
                               "(no asynchronous resets or negedge clocks)", line)
        return ("ff", edges[0][1])

    def instance(self) -> tuple:
        tk = self.peek()
        mod = self.ident()
        params: Dict[str, tuple] = {}
        pos_params: List[tuple] = []
        if self.accept("#"):
            self.expect("(")
            while not self.at(")"):
                if self.accept("."):
                    name = self.ident()
                    self.expect("(")
                    params[name] = self.expr()
                    self.expect(")")
                else:
                    pos_params.append(self.expr())
                if not self.accept(","):
                    break
            self.expect(")")
        inst = self.ident()
        self.expect("(")
        conns: Dict[str, Optional[tuple]] = {}
        pos: List[tuple] = []
        while not self.at(")"):
            if self.accept("."):
                name = self.ident()
                self.expect("(")
                conns[name] = None if self.at(")") else self.expr()
                self.expect(")")
            else:
                pos.append(self.expr())
            if not self.accept(","):
                break
        self.expect(")")
        self.expect(";")
        return ("inst", mod, params, pos_params, inst, conns, pos, tk.line)

    # --- statements
    def stmt(self) -> tuple:
        tk = self.peek()
        v = tk.val
        if tk.kind == "op" and v == ";":
            self.next()
            return ("nop",)
        if tk.kind == "sys":
            self.next()
            if self.accept("("):
                depth = 1
                while depth:
                    t = self.next()
                    if t.kind == "eof":
                        raise VerilogError("unterminated system task call", tk.line)
                    if t.kind == "op" and t.val == "(":
                        depth += 1
                    elif t.kind == "op" and t.val == ")":
                        depth -= 1
            self.expect(";")
            return ("nop",)
        if tk.kind == "id":
            if v == "begin":
                self.next()
                if self.accept(":"):
                    self.ident()
                body = []
                while not self.at("end"):
                    body.append(self.stmt())
                self.expect("end")
                return ("blk", body)
            if v == "if":
                self.next()
                self.expect("(")
                c = self.expr()
                self.expect(")")
                then = self.stmt()
                other = self.stmt() if self.accept("else") else None
                return ("if", c, then, other, tk.line)
            if v in ("case", "casez", "casex", "unique", "priority"):
                if v in ("unique", "priority"):
                    self.next()
                    v = self.peek().val
                self.next()
                self.expect("(")
                sel = self.expr()
                self.expect(")")
                arms = []
                while not self.at("endcase"):
                    if self.accept("default"):
                        self.accept(":")
                        arms.append((None, self.stmt()))
                        continue
                    labels = [self.expr()]
                    while self.accept(","):
                        labels.append(self.expr())
                    self.expect(":")
                    arms.append((labels, self.stmt()))
                self.expect("endcase")
                return ("case", v, sel, arms, tk.line)
            if v == "for":
                self.next()
                self.expect("(")
                init = self.simple_assign()
                self.expect(";")
                cond = self.expr()
                self.expect(";")
                step = self.simple_assign()
                self.expect(")")
                return ("for", init, cond, step, self.stmt(), tk.line)
            if v in ("while", "repeat", "forever", "wait", "fork", "disable"):
                raise VerilogError(f"`{v}` statements are not supported", tk.line)
            # task call: name ; or name (...) ;
            if self.peek(1).kind == "op" and self.peek(1).val == ";":
                self.next()
                self.next()
                return ("task", v, tk.line)
        lv = self.lvalue()
        if self.accept("<="):
            blocking = False
        else:
            self.expect("=")
            blocking = True
        x = self.expr()
        self.expect(";")
        return ("asg", blocking, lv, x, tk.line)

    def simple_assign(self) -> tuple:
        line = self.peek().line
        lv = self.lvalue()
        self.expect("=")
        return ("asg", True, lv, self.expr(), line)

    def lvalue(self) -> tuple:
        if self.accept("{"):
            parts = [self.lvalue()]
            while self.accept(","):
                parts.append(self.lvalue())
            self.expect("}")
            return ("cat", parts)
        return self.selects(("id", self.ident()))

    def selects(self, base: tuple) -> tuple:
        while self.at("["):
            self.next()
            a = self.expr()
            if self.accept(":"):
                b = self.expr()
                self.expect("]")
                base = ("sel", base, "part", a, b)
            elif self.accept("+:"):
                b = self.expr()
                self.expect("]")
                base = ("sel", base, "up", a, b)
            elif self.accept("-:"):
                b = self.expr()
                self.expect("]")
                base = ("sel", base, "down", a, b)
            else:
                self.expect("]")
                base = ("sel", base, "bit", a, None)
        return base

    # --- expressions (precedence climbing)
    _BIN = [("||",), ("&&",), ("|",), ("^", "~^", "^~"), ("&",), ("==", "!=", "===", "!=="),
            ("<", "<=", ">", ">="), ("<<", ">>", "<<<", ">>>"), ("+", "-"), ("*", "/", "%"), ("**",)]

    def expr(self) -> tuple:
        c = self.binary(0)
        if self.accept("?"):
            a = self.expr()
            self.expect(":")
            b = self.expr()
            return ("tern", c, a, b)
        return c

    def binary(self, level: int) -> tuple:
        if level == len(self._BIN):
            return self.unary()
        x = self.binary(level + 1)
        ops = self._BIN[level]
        while self.peek().kind == "op" and self.peek().val in ops:
            op = self.next().val
            y = self.binary(level + 1)
            x = ("bin", op, x, y)
        return x

    def unary(self) -> tuple:
        tk = self.peek()
        if tk.kind == "op" and tk.val in ("+", "-", "!", "~", "&", "~&", "|", "~|", "^", "~^", "^~"):
            self.next()
            return ("un", tk.val, self.unary())
        return self.primary()

    def primary(self) -> tuple:
        tk = self.next()
        if tk.kind == "num":
            return ("num", tk.val)
        if tk.kind == "str":
            return ("str", tk.val)
        if tk.kind == "sys":
            args = []
            if self.accept("("):
                while not self.at(")"):
                    args.append(self.expr())
                    if not self.accept(","):
                        break
                self.expect(")")
            return ("call", tk.val, args)
        if tk.kind == "id":
            return self.selects(("id", tk.val))
        if tk.kind == "op" and tk.val == "(":
            x = self.expr()
            self.expect(")")
            return x
        if tk.kind == "op" and tk.val == "{":
            first = self.expr()
            if self.at("{"):
                self.next()
                parts = [self.expr()]
                while self.accept(","):
                    parts.append(self.expr())
                self.expect("}")
                self.expect("}")
                return ("rep", first, parts)
            parts = [first]
            while self.accept(","):
                parts.append(self.expr())
            self.expect("}")
            return ("cat", parts)
        raise VerilogError(f"unexpected {tk.val!r} in expression", tk.line)


def parse(text: str, defines: Optional[Dict[str, str]] = None) -> List[ModuleAst]:
    return Parser(lex(preprocess(text, defines))).source()


# ============================================================ folding builders
def _lit(v: int, w: int) -> Expr:
    return Expr("lit", uint(w), (), v & ((1 << w) - 1))


def _is_lit(x: Expr) -> bool:
    return x.op == "lit"


def _ones(w: int) -> int:
    return (1 << w) - 1


def f_not(a: Expr) -> Expr:
    if _is_lit(a):
        return _lit(~a.attr, a.width)
    if a.op == "not":
        return a.args[0]
    return Expr("not", a.ty, (a,))


def _bitwise(op: str, a: Expr, b: Expr) -> Expr:
    w = a.width
    if _is_lit(a) and _is_lit(b):
        f = {"and": lambda x, y: x & y, "or": lambda x, y: x | y, "xor": lambda x, y: x ^ y}[op]
        return _lit(f(a.attr, b.attr), w)
    if _is_lit(a):
        a, b = b, a
    if _is_lit(b):
        v = b.attr
        if op == "and":
            if v == 0:
                return _lit(0, w)
            if v == _ones(w):
                return a
        elif op == "or":
            if v == 0:
                return a
            if v == _ones(w):
                return _lit(v, w)
        elif op == "xor":
            if v == 0:
                return a
            if v == _ones(w):
                return f_not(a)
    if a == b:
        return _lit(0, w) if op == "xor" else a
    return Expr(op, a.ty, (a, b))


def f_and(a, b): return _bitwise("and", a, b)
def f_or(a, b): return _bitwise("or", a, b)
def f_xor(a, b): return _bitwise("xor", a, b)


def f_add(a: Expr, b: Expr) -> Expr:
    w = a.width
    if _is_lit(a) and _is_lit(b):
        return _lit(a.attr + b.attr, w)
    if _is_lit(a) and a.attr == 0:
        return b
    if _is_lit(b) and b.attr == 0:
        return a
    return Expr("add", a.ty, (a, b))


def f_sub(a: Expr, b: Expr) -> Expr:
    w = a.width
    if _is_lit(a) and _is_lit(b):
        return _lit(a.attr - b.attr, w)
    if _is_lit(b) and b.attr == 0:
        return a
    if a == b:
        return _lit(0, w)
    return Expr("sub", a.ty, (a, b))


def f_mul(a: Expr, b: Expr) -> Expr:
    w = a.width
    if _is_lit(a) and _is_lit(b):
        return _lit(a.attr * b.attr, w)
    if _is_lit(a):
        a, b = b, a
    if _is_lit(b):
        if b.attr == 0:
            return _lit(0, w)
        if b.attr == 1:
            return a
    return Expr("mul", a.ty, (a, b))


def f_eq(a: Expr, b: Expr) -> Expr:
    if _is_lit(a) and _is_lit(b):
        return _lit(int(a.attr == b.attr), 1)
    if a == b:
        return _lit(1, 1)
    if a.width == 1:
        if _is_lit(a):
            a, b = b, a
        if _is_lit(b):
            return a if b.attr == 1 else f_not(a)
    return Expr("eq", uint(1), (a, b))


def f_ne(a: Expr, b: Expr) -> Expr:
    if _is_lit(a) and _is_lit(b):
        return _lit(int(a.attr != b.attr), 1)
    if a == b:
        return _lit(0, 1)
    if a.width == 1:
        if _is_lit(a):
            a, b = b, a
        if _is_lit(b):
            return f_not(a) if b.attr == 1 else a
    return Expr("ne", uint(1), (a, b))


def _sval(v: int, w: int) -> int:
    return v - (1 << w) if v >> (w - 1) & 1 else v


def f_lt(a: Expr, b: Expr, signed: bool) -> Expr:
    w = a.width
    if _is_lit(a) and _is_lit(b):
        x, y = (_sval(a.attr, w), _sval(b.attr, w)) if signed else (a.attr, b.attr)
        return _lit(int(x < y), 1)
    if a == b:
        return _lit(0, 1)
    if not signed and _is_lit(b) and b.attr == 0:
        return _lit(0, 1)
    if signed:
        return Expr("lt", uint(1), (Expr("cast", BitType(w, True), (a,), True),
                                    Expr("cast", BitType(w, True), (b,), True)))
    return Expr("lt", uint(1), (a, b))


def f_mux(c: Expr, a: Expr, b: Expr) -> Expr:
    if _is_lit(c):
        return a if c.attr & 1 else b
    if a == b:
        return a
    if a.width == 1 and _is_lit(a) and _is_lit(b):
        return c if a.attr == 1 else f_not(c)
    if c.op == "not":
        return f_mux(c.args[0], b, a)
    # mux(c, x, mux(c, y, z)) = mux(c, x, z); mux(c, mux(c, y, z), x) = mux(c, y, x)
    if b.op == "mux" and b.args[0] == c:
        b = b.args[2]
    if a.op == "mux" and a.args[0] == c:
        a = a.args[1]
    if a == b:
        return a
    return Expr("mux", a.ty, (c, a, b))


def f_slice(a: Expr, lsb: int, w: int) -> Expr:
    if lsb == 0 and w == a.width:
        return a
    if lsb >= a.width:
        return _lit(0, w)
    if lsb + w > a.width:  # reads above the value: zero bits
        return f_concat([_lit(0, lsb + w - a.width), f_slice(a, lsb, a.width - lsb)])
    if _is_lit(a):
        return _lit(a.attr >> lsb, w)
    if a.op == "slice":
        return f_slice(a.args[0], a.attr + lsb, w)
    if a.op == "zext":
        src = a.args[0]
        if lsb >= src.width:
            return _lit(0, w)
        if lsb + w <= src.width:
            return f_slice(src, lsb, w)
    if a.op == "concat":
        off = a.width
        for p in a.args:  # MSB first
            off -= p.width
            if off <= lsb and lsb + w <= off + p.width:
                return f_slice(p, lsb - off, w)
    return Expr("slice", uint(w), (a,), lsb)


def f_concat(parts: Sequence[Expr]) -> Expr:
    flat: List[Expr] = []
    for p in parts:
        flat.extend(p.args if p.op == "concat" else (p,))
    merged: List[Expr] = []
    for p in flat:
        if merged and _is_lit(p) and _is_lit(merged[-1]):
            q = merged.pop()
            merged.append(_lit((q.attr << p.width) | p.attr, q.width + p.width))
        else:
            merged.append(p)
    if len(merged) == 1:
        return merged[0]
    return Expr("concat", uint(sum(p.width for p in merged)), tuple(merged))


def f_zext(a: Expr, w: int) -> Expr:
    if w == a.width:
        return a
    if _is_lit(a):
        return _lit(a.attr, w)
    if a.op == "zext":
        return f_zext(a.args[0], w)
    return Expr("zext", uint(w), (a,), w)


def f_sext(a: Expr, w: int) -> Expr:
    if w == a.width:
        return a
    if _is_lit(a):
        return _lit(_sval(a.attr, a.width), w)
    return Expr("sext", uint(w), (a,), w)


def f_trunc(a: Expr, w: int) -> Expr:
    return f_slice(a, 0, w)


def f_bool(a: Expr) -> Expr:
    return a if a.width == 1 else f_ne(a, _lit(0, a.width))


def f_resize(a: Expr, w: int, signed: bool) -> Expr:
    if w < a.width:
        return f_trunc(a, w)
    return f_sext(a, w) if signed else f_zext(a, w)


def f_shift(a: Expr, amt: Expr, left: bool, arith: bool) -> Expr:
    """a shifted by the unsigned value amt (barrel of muxes for variable amounts)."""
    w = a.width
    if _is_lit(amt):
        k = amt.attr
        if k == 0:
            return a
        if left:
            return _lit(0, w) if k >= w else f_concat([f_slice(a, 0, w - k), _lit(0, k)])
        if k >= w:
            return f_sext(f_slice(a, w - 1, 1), w) if arith else _lit(0, w)
        fill = f_sext(f_slice(a, w - 1, 1), k) if arith else _lit(0, k)
        return f_concat([fill, f_slice(a, k, w - k)])
    stages = max(1, (w - 1).bit_length())
    v = a
    for s in range(min(stages, amt.width)):
        v = f_mux(f_slice(amt, s, 1), f_shift(v, _lit(1 << s, 32), left, arith), v)
    if amt.width > stages:
        big = f_ne(f_slice(amt, stages, amt.width - stages), _lit(0, amt.width - stages))
        v = f_mux(big, f_shift(a, _lit(w, 32), left, arith), v)
    return v


# ================================================================ elaboration
@dataclass(eq=False)
class Var:
    name: str               # hierarchical
    width: int
    signed: bool
    lo: int                 # bit index of the LSB as declared ([7:0] -> 0, [8:1] -> 1)
    desc: bool              # [msb:lsb] with msb >= lsb
    kind: str               # input output wire reg integer
    depth: int = 0          # memories: words
    mem_lo: int = 0         # memories: first index
    is_port: bool = False
    sig: Optional[ir.Signal] = None
    mem: Optional[ir.Memory] = None
    role: str = ""          # "in" | "comb" | "ff" after elaboration

    def bit_pos(self, idx: int) -> int:
        return idx - self.lo if self.desc else self.lo - idx


@dataclass
class TypedConst:
    value: Expr   # lit
    signed: bool


class Scope:
    def __init__(self, prefix: str, ast: ModuleAst):
        self.prefix = prefix
        self.ast = ast
        self.params: Dict[str, TypedConst] = {}
        self.vars: Dict[str, Var] = {}
        self.tasks: Dict[str, list] = {}
        self.consts: Dict[str, TypedConst] = {}  # loop and genvar values

    def hname(self, local: str) -> str:
        return self.prefix + local


class _Proc:
    """State of one always block during symbolic execution."""

    def __init__(self, ff: bool):
        self.ff = ff
        self.env: Dict[Var, Expr] = {}       # blocking values
        self.nba: Dict[Var, Expr] = {}       # pending non-blocking values
        self.mem: List[tuple] = []           # (var, cond, addr, data)
        self.ints: Dict[str, int] = {}       # loop integers (compile time)

    def copy(self) -> "_Proc":
        p = _Proc(self.ff)
        p.env = dict(self.env)
        p.nba = dict(self.nba)
        p.mem = []
        p.ints = dict(self.ints)
        return p


@dataclass
class Imported:
    module: ir.Module
    top: str
    inputs: List[str]
    outputs: List[str]
    stats: Dict[str, int]


class Elaborator:
    MAX_UNROLL = 1 << 16

    def __init__(self, mods: List[ModuleAst], top: str, overrides: Optional[Dict[str, int]] = None):
        self.mods = {m.name: m for m in mods}
        if top not in self.mods:
            raise VerilogError(f"top module `{top}` not found")
        self.top = top
        self.overrides = overrides or {}
        self.scratch = ir.Module("__sv_scratch")
        self.vars: List[Var] = []
        self.comb: Dict[Var, Expr] = {}       # comb var -> value
        self.comb_line: Dict[Var, int] = {}
        self.nexts: Dict[Var, Expr] = {}      # ff var -> next value
        self.mem_writes: List[tuple] = []     # (var, enable, addr, data)
        self.top_scope: Optional[Scope] = None

    # ------------------------------------------------------------ scopes
    def elaborate(self) -> None:
        ast = self.mods[self.top]
        sc = Scope("", ast)
        self.top_scope = sc
        self._instantiate(sc, {k: TypedConst(_lit(v, 32), True) for k, v in self.overrides.items()}, top=True)

    def _instantiate(self, sc: Scope, overrides: Dict[str, TypedConst], top: bool) -> None:
        ast = sc.ast
        known = {p[0] for p in ast.params}
        for k in overrides:
            if k not in known:
                raise VerilogError(f"module `{ast.name}` has no parameter `{k}`")
        for (name, local, rng, signed, kind, x) in ast.params:
            if not local and name in overrides:
                tc = overrides[name]
            else:
                tc = self._const_typed(sc, x)
            sc.params[name] = self._param_type(sc, tc, rng, signed, kind)
        # generate blocks resolve per instance (the parameters differ), on a copy
        # of the module's items; declarations first (items may use names declared later)
        items = self._flatten_gen(sc, list(ast.items))
        self._declare_items(sc, items, top)
        for p in ast.ports:
            if p not in sc.vars:
                raise VerilogError(f"port `{p}` of `{ast.name}` has no declaration")
        self._behaviour_items(sc, items)

    def _param_type(self, sc: Scope, tc: TypedConst, rng, signed: bool, kind) -> TypedConst:
        if kind == "integer":
            return TypedConst(f_resize(tc.value, 32, tc.signed), True)
        if rng is not None:
            msb, lsb = self._const_int(sc, rng[0]), self._const_int(sc, rng[1])
            w = abs(msb - lsb) + 1
            return TypedConst(f_resize(tc.value, w, tc.signed), signed)
        if signed:
            return TypedConst(tc.value, True)
        return tc

    def _flatten_gen(self, sc: Scope, items: List[tuple]) -> List[tuple]:
        out = []
        for it in items:
            if it[0] == "genif":
                _, c, then, other, line = it
                out.extend(self._flatten_gen(sc, then if self._const_int(sc, c) else other))
            elif it[0] == "genfor":
                _, var, init, cond, step, body, line = it
                v = self._const_int(sc, init)
                n = 0
                while True:
                    sc.consts[var] = TypedConst(_lit(v, 32), True)
                    if not self._const_int(sc, cond):
                        break
                    if body:
                        raise VerilogError("generate for loops with items are not supported", line)
                    v = self._const_int(sc, step)
                    n += 1
                    if n > self.MAX_UNROLL:
                        raise VerilogError("generate loop does not terminate", line)
                sc.consts.pop(var, None)
            else:
                out.append(it)
        return out

    def _declare_items(self, sc: Scope, items: List[tuple], top: bool) -> None:
        for it in items:
            if it[0] == "port":
                _, d, kind, signed, rng, name, line = it
                v = self._new_var(sc, name, rng, signed, "input" if d == "input" else
                                  ("reg" if kind == "reg" else "wire"), [], line, existing_ok=True)
                if d == "inout":
                    raise VerilogError("inout ports are not supported", line)
                v.is_port = True
                if d == "input":
                    v.kind = "input"
                elif d == "output" and kind == "reg":
                    v.kind = "reg"
            elif it[0] == "decl":
                _, kind, signed, rng, name, dims, init, line = it
                if kind == "genvar":
                    continue
                if name in sc.vars and sc.vars[name].is_port:
                    v = sc.vars[name]
                    if rng is not None:
                        v.width, v.lo, v.desc = self._range(sc, rng)
                    if signed:
                        v.signed = True
                    if kind in ("reg", "logic") and v.kind != "input":
                        v.kind = "reg"
                    continue
                self._new_var(sc, name, rng, signed, kind, dims, line)
            elif it[0] == "task":
                _, name, body, line = it
                sc.tasks[name] = body
        if top:
            for name in sc.ast.ports:
                v = sc.vars.get(name)
                if v is not None and v.kind == "input":
                    v.sig = self.scratch.input(v.name, v.width)

    def _range(self, sc: Scope, rng) -> Tuple[int, int, bool]:
        msb, lsb = self._const_int(sc, rng[0]), self._const_int(sc, rng[1])
        return abs(msb - lsb) + 1, lsb, msb >= lsb

    def _new_var(self, sc: Scope, name: str, rng, signed: bool, kind: str, dims, line: int,
                 existing_ok: bool = False) -> Var:
        if name in sc.vars:
            if existing_ok:
                return sc.vars[name]
            raise VerilogError(f"`{name}` declared twice", line)
        if kind == "integer":
            w, lo, desc = 32, 0, True
            signed = True
        elif rng is not None:
            w, lo, desc = self._range(sc, rng)
        else:
            w, lo, desc = 1, 0, True
        v = Var(sc.hname(name), w, signed, lo, desc, "reg" if kind in ("reg", "logic", "integer", "bit") else kind)
        if dims:
            if len(dims) > 1:
                raise VerilogError(f"memory `{name}` has more than one unpacked dimension", line)
            a, b = self._const_int(sc, dims[0][0]), self._const_int(sc, dims[0][1])
            v.depth, v.mem_lo = abs(a - b) + 1, min(a, b)
            v.mem = self.scratch.mem(v.name, v.width, v.depth)
        sc.vars[name] = v
        self.vars.append(v)
        return v

    # ------------------------------------------------------------ behaviour
    def _behaviour_items(self, sc: Scope, items: List[tuple]) -> None:
        for it in items:
            tag = it[0]
            if tag == "decl":
                _, kind, signed, rng, name, dims, init, line = it
                if init is not None and kind != "genvar":
                    v = sc.vars[name]
                    if v.kind in ("wire", "tri"):
                        self._drive_comb(sc, ("id", name), init, line)
                    # initial values of variables are not modelled (every register starts at 0)
            elif tag == "assign":
                _, lv, x, line = it
                self._drive_comb(sc, lv, x, line)
            elif tag == "always":
                _, (kind, clk), body, line = it
                self._always(sc, kind == "ff", body, line)
            elif tag == "inst":
                self._instance(sc, it)

    def _instance(self, sc: Scope, it: tuple) -> None:
        _, modname, params, pos_params, inst, conns, pos, line = it
        ast = self.mods.get(modname)
        if ast is None:
            raise VerilogError(f"module `{modname}` not found", line)
        if pos_params:
            names = [p[0] for p in ast.params if not p[1]]
            for k, x in enumerate(pos_params):
                params[names[k]] = x
        if pos:
            for k, x in enumerate(pos):
                conns[ast.ports[k]] = x
        child = Scope(sc.hname(inst) + ".", ast)
        ov = {k: self._const_typed(sc, x) for k, x in params.items()}
        self._instantiate(child, ov, top=False)
        for pname in ast.ports:
            pv = child.vars[pname]
            x = conns.get(pname)
            if pv.kind == "input":
                if x is None:
                    self._set_comb(pv, _lit(0, pv.width), line)
                else:
                    val = self._rvalue(sc, x, pv.width)
                    self._set_comb(pv, val, line)
            else:
                if x is None:
                    continue
                self._drive_comb_value(sc, x, self._read(pv), line)
        for k in conns:
            if k not in child.vars or not child.vars[k].is_port:
                raise VerilogError(f"`{modname}` has no port `{k}`", line)

    def _set_comb(self, v: Var, val: Expr, line: int) -> None:
        if v.depth:
            raise VerilogError(f"memory `{v.name}` cannot be driven continuously", line)
        if v in self.comb or v in self.nexts:
            raise VerilogError(f"`{v.name}` has more than one driver", line)
        if val.width != v.width:
            val = f_resize(val, v.width, False)
        self.comb[v] = val
        self.comb_line[v] = line
        v.role = "comb"

    def _drive_comb(self, sc: Scope, lv: tuple, x: tuple, line: int) -> None:
        w = self._lv_width(sc, lv)
        self._drive_comb_value(sc, lv, self._rvalue(sc, x, w), line)

    def _drive_comb_value(self, sc: Scope, lv: tuple, val: Expr, line: int) -> None:
        """Continuous assignment of val to lvalue lv (whole vars, selects, concats)."""
        targets = self._lv_split(sc, lv, val)
        for var, sel, part in targets:
            if sel is None:
                self._set_comb(var, part, line)
            else:
                # partial continuous drivers: merge into a pending per-bit table
                pend = self._partial.setdefault(var, {})
                lo, w = sel
                for b in range(w):
                    if lo + b in pend:
                        raise VerilogError(f"bit {lo + b} of `{var.name}` has more than one driver", line)
                    pend[lo + b] = f_slice(part, b, 1)
                self.comb_line[var] = line

    _partial: Dict[Var, Dict[int, Expr]]

    def _finish_partials(self) -> None:
        for var, bits in self._partial.items():
            parts = [bits.get(b, _lit(0, 1)) for b in reversed(range(var.width))]
            self._set_comb(var, f_concat(parts), self.comb_line.get(var, 0))

    def _always(self, sc: Scope, ff: bool, body: tuple, line: int) -> None:
        st = _Proc(ff)
        self._exec(sc, body, st, _lit(1, 1))
        if ff:
            for v, x in st.nba.items():
                self._set_next(v, x, line)
            for v, x in st.env.items():
                if v in st.nba:
                    raise VerilogError(f"`{v.name}` has blocking and non-blocking assignments", line)
                self._set_next(v, x, line)
        else:
            if st.nba:
                raise VerilogError("non-blocking assignment in a combinational block", line)
            for v, x in st.env.items():
                self._set_comb(v, x, line)
        for (v, en, addr, data) in st.mem:
            if not ff:
                raise VerilogError(f"memory `{v.name}` written in a combinational block", line)
            self.mem_writes.append((v, en, addr, data))

    def _set_next(self, v: Var, x: Expr, line: int) -> None:
        if v.depth:
            raise VerilogError(f"memory `{v.name}` assigned as a whole", line)
        if v.kind == "input":
            raise VerilogError(f"input `{v.name}` assigned", line)
        if v in self.nexts or v in self.comb:
            raise VerilogError(f"`{v.name}` is assigned in more than one block", line)
        self.nexts[v] = x
        v.role = "ff"

    # ------------------------------------------------------------ statements
    def _exec(self, sc: Scope, s: tuple, st: _Proc, path: Expr) -> None:
        tag = s[0]
        if tag == "blk":
            for x in s[1]:
                self._exec(sc, x, st, path)
        elif tag == "nop":
            return
        elif tag == "task":
            body = sc.tasks.get(s[1])
            if body is None:
                raise VerilogError(f"unknown task `{s[1]}`", s[2])
            if any(b not in (("nop",), ("blk", [])) for b in body):
                raise VerilogError(f"task `{s[1]}` has statements (only empty tasks are supported)", s[2])
        elif tag == "asg":
            _, blocking, lv, x, line = s
            self._assign(sc, st, blocking, lv, x, path, line)
        elif tag == "if":
            _, c, then, other, line = s
            cond = f_bool(self._self_value(sc, c, st)[0])
            self._branch(sc, st, cond, then, other, path)
        elif tag == "case":
            self._case(sc, s, st, path)
        elif tag == "for":
            _, init, cond, step, body, line = s
            name = self._loop_var(init, line)
            self._int_assign(sc, st, init, line)
            n = 0
            while True:
                c = self._self_value(sc, cond, st)[0]
                if not _is_lit(c):
                    raise VerilogError("for loop condition is not a compile-time constant", line)
                if not c.attr:
                    break
                self._exec(sc, body, st, path)
                if self._loop_var(step, line) != name:
                    raise VerilogError("for loop must step its own variable", line)
                self._int_assign(sc, st, step, line)
                n += 1
                if n > self.MAX_UNROLL:
                    raise VerilogError("for loop does not terminate", line)
        else:
            raise VerilogError(f"unsupported statement {tag}")

    @staticmethod
    def _loop_var(a: tuple, line: int) -> str:
        if a[0] != "asg" or a[2][0] != "id":
            raise VerilogError("for loop init/step must assign a plain variable", line)
        return a[2][1]

    def _int_assign(self, sc: Scope, st: _Proc, a: tuple, line: int) -> None:
        name = a[2][1]
        v = self._lookup_var(sc, name, line)
        val = self._rvalue(sc, a[3], v.width, st)
        if not _is_lit(val):
            raise VerilogError("for loop variable must take compile-time values", line)
        st.env[v] = val

    def _branch(self, sc: Scope, st: _Proc, cond: Expr, then: tuple, other: Optional[tuple], path: Expr) -> None:
        if _is_lit(cond):
            taken = then if cond.attr & 1 else other
            if taken is not None:
                self._exec(sc, taken, st, path)
            return
        a = st.copy()
        self._exec(sc, then, a, f_and(path, cond))
        b = st.copy()
        if other is not None:
            self._exec(sc, other, b, f_and(path, f_not(cond)))
        self._merge(st, cond, a, b)

    def _merge(self, st: _Proc, cond: Expr, a: _Proc, b: _Proc) -> None:
        for table in ("env", "nba"):
            ta, tb, base = getattr(a, table), getattr(b, table), getattr(st, table)
            for v in set(ta) | set(tb):
                va, vb = ta.get(v), tb.get(v)
                if va is vb:
                    base[v] = va
                    continue
                dflt = base.get(v)
                if dflt is None:
                    dflt = self._hold(st, v, table)
                base[v] = f_mux(cond, va if va is not None else dflt, vb if vb is not None else dflt)
        st.mem.extend(a.mem)
        st.mem.extend(b.mem)

    def _hold(self, st: _Proc, v: Var, table: str) -> Expr:
        """Value of v on a path that does not assign it."""
        if not st.ff:
            return _lit(0, v.width)   # combinational: don't-care, no latch
        return self._read(v)          # clocked: the register keeps its value

    def _case(self, sc: Scope, s: tuple, st: _Proc, path: Expr) -> None:
        _, kind, sel, arms, line = s
        sw, ss = self._typ(sc, sel, st)
        chain: List[Tuple[Optional[Expr], tuple]] = []
        for labels, body in arms:
            if labels is None:
                chain.append((None, body))
                continue
            conds = []
            for lab in labels:
                lw, ls = self._typ(sc, lab, st)
                w, sg = max(sw, lw), ss and ls
                a = self._gen(sc, sel, w, sg, st)
                b = self._gen(sc, lab, w, sg, st)
                care = _ones(w)
                if kind in ("casez", "casex") and lab[0] == "num" and lab[1].xz:
                    care &= ~lab[1].xz
                if care != _ones(w):
                    conds.append(f_eq(f_and(a, _lit(care, w)), f_and(b, _lit(care, w))))
                else:
                    conds.append(f_eq(a, b))
            c = conds[0]
            for x in conds[1:]:
                c = f_or(c, x)
            chain.append((c, body))
        # first match wins: build an if/else chain
        default = next((b for c, b in chain if c is None), None)
        items = [(c, b) for c, b in chain if c is not None]

        def run(k: int, st2: _Proc, path2: Expr) -> None:
            if k == len(items):
                if default is not None:
                    self._exec(sc, default, st2, path2)
                return
            c, body = items[k]
            if _is_lit(c):
                if c.attr & 1:
                    self._exec(sc, body, st2, path2)
                else:
                    run(k + 1, st2, path2)
                return
            a = st2.copy()
            self._exec(sc, body, a, f_and(path2, c))
            b = st2.copy()
            run(k + 1, b, f_and(path2, f_not(c)))
            self._merge(st2, c, a, b)

        run(0, st, path)

    def _assign(self, sc: Scope, st: _Proc, blocking: bool, lv: tuple, x: tuple, path: Expr, line: int) -> None:
        w = self._lv_width(sc, lv, st)
        val = self._rvalue(sc, x, w, st)
        for var, sel, part, addr in self._lv_split_proc(sc, lv, val, st, line):
            if var.depth:
                data = part
                if sel is not None:
                    lo, sw = sel
                    old = self._mem_read(var, addr)
                    data = self._splice(old, lo, sw, part)
                st.mem.append((var, path, addr, data))
                continue
            if var.kind == "input":
                raise VerilogError(f"input `{var.name}` assigned", line)
            table = st.env if (blocking or not st.ff) else st.nba
            if sel is None:
                table[var] = part
            else:
                cur = table.get(var)
                if cur is None:
                    cur = self._cur(st, var, blocking or not st.ff)
                table[var] = self._splice(cur, sel[0], sel[1], part)

    def _cur(self, st: _Proc, v: Var, blocking: bool) -> Expr:
        if not st.ff:
            return _lit(0, v.width)
        if blocking:
            return st.env.get(v, self._read(v))
        return self._read(v)

    @staticmethod
    def _splice(old: Expr, lo: int, w: int, part: Expr) -> Expr:
        parts = []
        if lo + w < old.width:
            parts.append(f_slice(old, lo + w, old.width - lo - w))
        parts.append(part)
        if lo > 0:
            parts.append(f_slice(old, 0, lo))
        return f_concat(parts)

    # ------------------------------------------------------------ lvalues
    def _lookup_var(self, sc: Scope, name: str, line: Optional[int] = None) -> Var:
        v = sc.vars.get(name)
        if v is None:
            raise VerilogError(f"`{name}` is not declared in `{sc.ast.name}`", line)
        return v

    def _lv_width(self, sc: Scope, lv: tuple, st: Optional[_Proc] = None) -> int:
        if lv[0] == "cat":
            return sum(self._lv_width(sc, p, st) for p in lv[1])
        var, sels = self._lv_base(sc, lv)
        if var.depth:
            if not sels:
                raise VerilogError(f"memory `{var.name}` needs an index")
            sels = sels[1:]
        if not sels:
            return var.width
        if len(sels) > 1:
            raise VerilogError(f"nested select of `{var.name}` is not supported")
        kind, a, b = sels[0]
        if kind == "bit":
            return 1
        if kind == "part":
            return abs(self._const_int(sc, a, st) - self._const_int(sc, b, st)) + 1
        return self._const_int(sc, b, st)

    def _lv_base(self, sc: Scope, lv: tuple) -> Tuple[Var, List[tuple]]:
        sels = []
        while lv[0] == "sel":
            sels.append((lv[2], lv[3], lv[4]))
            lv = lv[1]
        if lv[0] != "id":
            raise VerilogError("unsupported assignment target")
        return self._lookup_var(sc, lv[1]), list(reversed(sels))

    def _lv_split(self, sc: Scope, lv: tuple, val: Expr) -> List[tuple]:
        return [(v, sel, part) for v, sel, part, _ in self._lv_split_proc(sc, lv, val, None, None)]

    def _lv_split_proc(self, sc: Scope, lv: tuple, val: Expr, st: Optional[_Proc], line) -> List[tuple]:
        """(var, (lsb, width) | None, value, memory address | None) per target, MSB-first order."""
        if lv[0] == "cat":
            out, off = [], val.width
            for p in lv[1]:
                w = self._lv_width(sc, p, st)
                off -= w
                out.extend(self._lv_split_proc(sc, p, f_slice(val, off, w), st, line))
            return out
        var, sels = self._lv_base(sc, lv)
        addr = None
        if var.depth:
            if not sels or sels[0][0] != "bit":
                raise VerilogError(f"memory `{var.name}` needs an index", line)
            addr = self._mem_addr(sc, var, sels[0][1], st)
            sels = sels[1:]
        if not sels:
            return [(var, None, f_resize(val, var.width, False), addr)]
        kind, a, b = sels[0]
        if kind == "bit":
            idx = self._self_value(sc, a, st)[0]
            if not _is_lit(idx):
                raise VerilogError(f"variable bit select on the left of an assignment to `{var.name}`", line)
            pos = var.bit_pos(_sval(idx.attr, idx.width) if self._typ(sc, a, st)[1] else idx.attr)
            if not 0 <= pos < var.width:
                return []  # out of range: no effect
            return [(var, (pos, 1), val, addr)]
        if kind == "part":
            msb, lsb = self._const_int(sc, a, st), self._const_int(sc, b, st)
            lo = min(var.bit_pos(msb), var.bit_pos(lsb))
            w = abs(msb - lsb) + 1
        else:
            base, w = self._const_int(sc, a, st), self._const_int(sc, b, st)
            lo = var.bit_pos(base if kind == "up" else base - w + 1)
        if lo < 0 or lo + w > var.width:
            raise VerilogError(f"part select outside `{var.name}`", line)
        return [(var, (lo, w), val, addr)]

    # ------------------------------------------------------------ reading
    def _read(self, v: Var) -> Expr:
        if v.depth:
            raise VerilogError(f"memory `{v.name}` read without an index")
        if v.sig is None:
            v.sig = self.scratch.wire(v.name, v.width)
        return v.sig.value()

    def _mem_addr(self, sc: Scope, var: Var, x: tuple, st: Optional[_Proc]) -> Expr:
        a, w, s = self._self_value(sc, x, st)
        if var.mem_lo:
            a = f_sub(f_zext(a, max(w, 32)), _lit(var.mem_lo, max(w, 32)))
        return a

    def _mem_read(self, var: Var, addr: Expr) -> Expr:
        return Expr("memread", uint(var.width), (addr,), var.mem)

    # ------------------------------------------------------------ expressions
    def _const_typed(self, sc: Scope, x: tuple) -> TypedConst:
        v, w, s = self._self_value(sc, x, None)
        if not _is_lit(v):
            raise VerilogError("expression is not a compile-time constant")
        return TypedConst(v, s)

    def _const_int(self, sc: Scope, x: tuple, st: Optional[_Proc] = None) -> int:
        v, w, s = self._self_value(sc, x, st)
        if not _is_lit(v):
            raise VerilogError("expression is not a compile-time constant")
        return _sval(v.attr, w) if s else v.attr

    def _rvalue(self, sc: Scope, x: tuple, target_w: int, st: Optional[_Proc] = None) -> Expr:
        """x assigned to a target of target_w bits (context = max of both)."""
        w, s = self._typ(sc, x, st)
        cw = max(w, target_w)
        return f_trunc(self._gen(sc, x, cw, s, st), target_w)

    def _self_value(self, sc: Scope, x: tuple, st: Optional[_Proc]) -> Tuple[Expr, int, bool]:
        w, s = self._typ(sc, x, st)
        return self._gen(sc, x, w, s, st), w, s

    def _name(self, sc: Scope, name: str, st: Optional[_Proc]):
        """A name as a constant (TypedConst) or a Var."""
        if name in sc.consts:
            return sc.consts[name]
        if name in sc.params:
            return sc.params[name]
        v = sc.vars.get(name)
        if v is None:
            raise VerilogError(f"`{name}` is not declared in `{sc.ast.name}`")
        return v

    def _typ(self, sc: Scope, x: tuple, st: Optional[_Proc]) -> Tuple[int, bool]:
        tag = x[0]
        if tag == "num":
            n: Num = x[1]
            if n.width is None:
                return 32, n.signed and not n.based or (n.signed and n.based)
            return n.width, n.signed
        if tag == "str":
            return max(8, 8 * len(x[1])), False
        if tag == "id":
            t = self._name(sc, x[1], st)
            if isinstance(t, TypedConst):
                return t.value.width, t.signed
            if t.depth:
                raise VerilogError(f"memory `{t.name}` read without an index")
            return t.width, t.signed
        if tag == "sel":
            base, kind, a, b = x[1], x[2], x[3], x[4]
            if base[0] == "id":
                t = self._name(sc, base[1], st)
                if isinstance(t, Var) and t.depth:
                    if kind != "bit":
                        raise VerilogError(f"memory `{t.name}` needs a word index")
                    return t.width, t.signed
            if kind == "bit":
                return 1, False
            if kind == "part":
                return abs(self._const_int(sc, a, st) - self._const_int(sc, b, st)) + 1, False
            return self._const_int(sc, b, st), False
        if tag == "un":
            op = x[1]
            if op in ("~", "-", "+"):
                return self._typ(sc, x[2], st)
            return 1, False
        if tag == "bin":
            op = x[1]
            wa, sa = self._typ(sc, x[2], st)
            wb, sb = self._typ(sc, x[3], st)
            if op in ("==", "!=", "===", "!==", "<", "<=", ">", ">=", "&&", "||"):
                return 1, False
            if op in ("<<", ">>", "<<<", ">>>", "**"):
                return wa, sa
            return max(wa, wb), sa and sb
        if tag == "tern":
            wa, sa = self._typ(sc, x[2], st)
            wb, sb = self._typ(sc, x[3], st)
            return max(wa, wb), sa and sb
        if tag == "cat":
            return sum(self._typ(sc, p, st)[0] for p in x[1]), False
        if tag == "rep":
            n = self._const_int(sc, x[1], st)
            if n <= 0:
                raise VerilogError("replication count must be positive")
            return n * sum(self._typ(sc, p, st)[0] for p in x[2]), False
        if tag == "call":
            name, args = x[1], x[2]
            if name in ("$signed", "$unsigned"):
                if len(args) != 1:
                    raise VerilogError(f"{name} takes one argument")
                return self._typ(sc, args[0], st)[0], name == "$signed"
            if name == "$clog2":
                return 32, True
            raise VerilogError(f"unsupported system function {name}")
        raise VerilogError(f"unsupported expression {tag}")

    def _gen(self, sc: Scope, x: tuple, w: int, s: bool, st: Optional[_Proc]) -> Expr:
        """x evaluated in a context of w bits whose type is signed iff s."""
        tag = x[0]
        if tag == "un" and x[1] in ("~", "-", "+"):
            a = self._gen(sc, x[2], w, s, st)
            if x[1] == "~":
                return f_not(a)
            if x[1] == "-":
                return f_sub(_lit(0, w), a)
            return a
        if tag == "bin" and x[1] in ("+", "-", "*", "&", "|", "^", "~^", "^~", "/", "%", "**"):
            op = x[1]
            if op == "**":
                a = self._gen(sc, x[2], w, s, st)
                e = self._self_value(sc, x[3], st)[0]
                if not (_is_lit(a) and _is_lit(e)):
                    raise VerilogError("`**` needs constant operands")
                return _lit(a.attr ** e.attr, w)
            a = self._gen(sc, x[2], w, s, st)
            b = self._gen(sc, x[3], w, s, st)
            if op == "+":
                return f_add(a, b)
            if op == "-":
                return f_sub(a, b)
            if op == "*":
                return f_mul(a, b)
            if op == "&":
                return f_and(a, b)
            if op == "|":
                return f_or(a, b)
            if op == "^":
                return f_xor(a, b)
            if op in ("~^", "^~"):
                return f_not(f_xor(a, b))
            if not (_is_lit(a) and _is_lit(b)):
                raise VerilogError(f"`{op}` needs constant operands")
            if b.attr == 0:
                return _lit(0, w)
            av, bv = (_sval(a.attr, w), _sval(b.attr, w)) if s else (a.attr, b.attr)
            q = abs(av) // abs(bv) * (1 if (av < 0) == (bv < 0) else -1)
            return _lit(q if op == "/" else av - q * bv, w)
        if tag == "bin" and x[1] in ("<<", ">>", "<<<", ">>>"):
            a = self._gen(sc, x[2], w, s, st)
            amt = self._self_value(sc, x[3], st)[0]
            left = x[1] in ("<<", "<<<")
            return f_shift(a, amt, left, x[1] == ">>>" and s)
        if tag == "tern":
            c = f_bool(self._self_value(sc, x[1], st)[0])
            if _is_lit(c):
                return self._gen(sc, x[2] if c.attr else x[3], w, s, st)
            return f_mux(c, self._gen(sc, x[2], w, s, st), self._gen(sc, x[3], w, s, st))
        # self-determined operand: evaluate at its own width, then extend
        v, vw = self._leaf(sc, x, st)
        return f_resize(v, w, s) if w != vw else v

    def _leaf(self, sc: Scope, x: tuple, st: Optional[_Proc]) -> Tuple[Expr, int]:
        tag = x[0]
        if tag == "num":
            n: Num = x[1]
            w = self._typ(sc, x, st)[0]
            return _lit(n.value, w), w
        if tag == "str":
            w = max(8, 8 * len(x[1]))
            v = 0
            for ch in x[1].encode("latin-1"):
                v = (v << 8) | ch
            return _lit(v, w), w
        if tag == "id":
            t = self._name(sc, x[1], st)
            if isinstance(t, TypedConst):
                return t.value, t.value.width
            if st is not None and t in st.env:
                return st.env[t], t.width
            if st is not None and not st.ff and t.role != "comb" and self._assigned_in(t, st):
                raise VerilogError(f"`{t.name}` read before it is assigned in a combinational block")
            return self._read(t), t.width
        if tag == "sel":
            return self._select(sc, x, st)
        if tag == "un":
            op = x[1]
            a = self._self_value(sc, x[2], st)[0]
            aw = a.width
            if op == "!":
                return f_not(f_bool(a)), 1
            if op == "|":
                return f_bool(a), 1
            if op == "~|":
                return f_not(f_bool(a)), 1
            if op == "&":
                return f_eq(a, _lit(_ones(aw), aw)), 1
            if op == "~&":
                return f_ne(a, _lit(_ones(aw), aw)), 1
            if op in ("^", "~^", "^~"):
                r = f_slice(a, 0, 1)
                for i in range(1, aw):
                    r = f_xor(r, f_slice(a, i, 1))
                return (r if op == "^" else f_not(r)), 1
            raise VerilogError(f"unsupported unary `{op}`")
        if tag == "bin":
            op = x[1]
            if op in ("&&", "||"):
                a = f_bool(self._self_value(sc, x[2], st)[0])
                if _is_lit(a) and (a.attr == (0 if op == "&&" else 1)):
                    return a, 1  # short circuit (the other side may be a disabled feature)
                b = f_bool(self._self_value(sc, x[3], st)[0])
                return (f_and(a, b) if op == "&&" else f_or(a, b)), 1
            wa, sa = self._typ(sc, x[2], st)
            wb, sb = self._typ(sc, x[3], st)
            w, sg = max(wa, wb), sa and sb
            a = self._gen(sc, x[2], w, sg, st)
            b = self._gen(sc, x[3], w, sg, st)
            if op in ("==", "==="):
                return f_eq(a, b), 1
            if op in ("!=", "!=="):
                return f_ne(a, b), 1
            if op == "<":
                return f_lt(a, b, sg), 1
            if op == ">":
                return f_lt(b, a, sg), 1
            if op == "<=":
                return f_not(f_lt(b, a, sg)), 1
            if op == ">=":
                return f_not(f_lt(a, b, sg)), 1
            raise VerilogError(f"unsupported operator `{op}`")
        if tag == "cat":
            parts = [self._self_value(sc, p, st)[0] for p in x[1]]
            v = f_concat(parts)
            return v, v.width
        if tag == "rep":
            n = self._const_int(sc, x[1], st)
            parts = [self._self_value(sc, p, st)[0] for p in x[2]]
            v = f_concat(parts * n)
            return v, v.width
        if tag == "call":
            name, args = x[1], x[2]
            if name in ("$signed", "$unsigned"):
                v = self._self_value(sc, args[0], st)[0]
                return v, v.width
            if name == "$clog2":
                n = self._const_int(sc, args[0], st)
                return _lit(max(0, (n - 1).bit_length()), 32), 32
        raise VerilogError(f"unsupported expression {tag}")

    def _assigned_in(self, v: Var, st: _Proc) -> bool:
        return False

    def _select(self, sc: Scope, x: tuple, st: Optional[_Proc]) -> Tuple[Expr, int]:
        base, kind, a, b = x[1], x[2], x[3], x[4]
        if base[0] == "id":
            t = self._name(sc, base[1], st)
            if isinstance(t, Var) and t.depth:
                v = self._mem_read(t, self._mem_addr(sc, t, a, st))
                return v, t.width
        if base[0] == "id":
            t = self._name(sc, base[1], st)
            if isinstance(t, TypedConst):
                val, var = t.value, None
            else:
                val, var = self._leaf(sc, base, st)[0], t
        else:
            val, _ = self._leaf(sc, base, st)
            var = None

        def pos(i: int) -> int:
            return var.bit_pos(i) if var is not None else i

        if kind == "bit":
            iv, iw, isg = self._self_value(sc, a, st)
            if _is_lit(iv):
                p = pos(_sval(iv.attr, iw) if isg else iv.attr)
                if not 0 <= p < val.width:
                    return _lit(0, 1), 1
                return f_slice(val, p, 1), 1
            if var is not None and (var.lo != 0 or not var.desc):
                raise VerilogError(f"variable bit select of `{var.name}` with a non-zero base")
            return f_slice(f_shift(val, iv, False, False), 0, 1), 1
        if kind == "part":
            msb, lsb = self._const_int(sc, a, st), self._const_int(sc, b, st)
            lo, w = min(pos(msb), pos(lsb)), abs(msb - lsb) + 1
            if lo < 0:
                return f_concat([f_slice(val, 0, w + lo), _lit(0, -lo)]) if w + lo > 0 else _lit(0, w), w
            return f_slice(val, lo, w), w
        w = self._const_int(sc, b, st)
        iv, iw, isg = self._self_value(sc, a, st)
        if _is_lit(iv):
            bs = _sval(iv.attr, iw) if isg else iv.attr
            lo = pos(bs if kind == "up" else bs - w + 1)
            if lo < 0:
                return _lit(0, w), w
            return f_slice(val, lo, w), w
        if var is not None and (var.lo != 0 or not var.desc):
            raise VerilogError(f"indexed part select of `{var.name}` with a non-zero base")
        amt = iv if kind == "up" else f_sub(iv, _lit(w - 1, iv.width))
        return f_slice(f_shift(val, amt, False, False), 0, w), w


# ================================================================ final module
def _substitute(e: Expr, memo: Dict[int, Expr], leaf) -> Expr:
    """Rebuilds e bottom-up with folding; leaf(node) maps sig/memread nodes."""
    got = memo.get(id(e))
    if got is not None:
        return got
    stack = [(e, False)]
    while stack:
        node, done = stack.pop()
        if id(node) in memo:
            continue
        if not done:
            stack.append((node, True))
            for a in node.args:
                if id(a) not in memo:
                    stack.append((a, False))
            continue
        args = [memo[id(a)] for a in node.args]
        op = node.op
        if op in ("sig", "memread"):
            out = leaf(node, args)
        elif op == "lit":
            out = node
        elif op == "not":
            out = f_not(args[0])
        elif op in ("and", "or", "xor"):
            out = _bitwise(op, args[0], args[1])
        elif op == "add":
            out = f_add(*args)
        elif op == "sub":
            out = f_sub(*args)
        elif op == "mul":
            out = f_mul(*args)
        elif op == "eq":
            out = f_eq(*args)
        elif op == "ne":
            out = f_ne(*args)
        elif op == "lt":
            a, b = args
            if a.op == "cast" and b.op == "cast":
                out = f_lt(a.args[0], b.args[0], True)
            else:
                out = f_lt(a, b, node.args[0].ty.signed)
        elif op == "cast":
            out = args[0] if args[0].op == "lit" and False else Expr("cast", node.ty, (args[0],), node.attr)
        elif op == "mux":
            out = f_mux(*args)
        elif op == "slice":
            out = f_slice(args[0], node.attr, node.width)
        elif op == "concat":
            out = f_concat(args)
        elif op == "zext":
            out = f_zext(args[0], node.width)
        elif op == "sext":
            out = f_sext(args[0], node.width)
        elif op == "trunc":
            out = f_trunc(args[0], node.width)
        else:
            raise VerilogError(f"internal: unknown node {op}")
        memo[id(node)] = out
    return memo[id(e)]


def _walk(e: Expr, fn) -> None:
    seen, stack = set(), [e]
    while stack:
        x = stack.pop()
        if id(x) in seen:
            continue
        seen.add(id(x))
        fn(x)
        stack.extend(x.args)


def build_module(el: Elaborator, name: str, keep: Sequence[str] = ()) -> Imported:
    sc = el.top_scope
    by_sig: Dict[int, Var] = {}
    for v in el.vars:
        if v.sig is not None:
            by_sig[id(v.sig)] = v
    top_vars = {n: sc.vars[n] for n in sc.ast.ports}
    inputs = [n for n in sc.ast.ports if top_vars[n].kind == "input"]
    outputs = [n for n in sc.ast.ports if top_vars[n].kind != "input"]
    keep_vars = set()
    by_name = {v.name: v for v in el.vars}
    for k in keep:
        if k not in by_name:
            raise VerilogError(f"keep: no variable `{k}`")
        keep_vars.add(by_name[k])
    for v in el.vars:
        if v.sig is not None and v.role == "" and v.kind != "input" and not v.depth:
            # read but never driven: an undriven net is 0
            el.comb[v] = _lit(0, v.width)
            v.role = "comb"

    # which comb vars stay signals: top-level outputs and kept names; the rest inline
    materialize = {v for v in el.comb if (v.is_port and v.name in outputs) or v in keep_vars}

    # liveness from the roots, through comb definitions and register next values
    roots: List[Expr] = []
    live_vars = set()
    for n in outputs:
        live_vars.add(top_vars[n])
    live_vars |= keep_vars
    for (mv, en, addr, data) in el.mem_writes:
        roots.extend((en, addr, data))
    work = list(live_vars)
    live_mems = {mv for (mv, *_r) in el.mem_writes}

    def visit(e: Expr) -> None:
        def f(x: Expr) -> None:
            if x.op == "sig":
                v = by_sig.get(id(x.attr))
                if v is not None and v not in live_vars:
                    live_vars.add(v)
                    work.append(v)
            elif x.op == "memread":
                live_mems.add(x.attr)
        _walk(e, f)

    for r in roots:
        visit(r)
    while work:
        v = work.pop()
        if v in el.comb:
            visit(el.comb[v])
        elif v in el.nexts:
            visit(el.nexts[v])

    m = ir.Module(name)
    final_sig: Dict[Var, ir.Signal] = {}
    final_mem: Dict[int, ir.Memory] = {}
    for n in inputs:
        v = top_vars[n]
        final_sig[v] = m.input(n, v.width)
    for v in el.vars:
        if v in live_vars and v.kind != "input" and not v.depth and v in el.nexts:
            final_sig[v] = m.reg(v.name, v.width)
    for v in el.vars:
        if v in live_vars and v in materialize:
            kind_out = v.is_port and v.name in outputs
            final_sig[v] = m.output(v.name, v.width) if kind_out else m.wire(v.name, v.width)
    for v in el.vars:
        if v.depth and (v.mem in live_mems or any(mv is v for mv, *_r in el.mem_writes)):
            final_mem[id(v.mem)] = m.mem(v.name, v.width, v.depth)

    memo: Dict[int, Expr] = {}
    inlining: set = set()

    def leaf(node: Expr, args: List[Expr]) -> Expr:
        if node.op == "memread":
            fm = final_mem.get(id(node.attr))
            return Expr("memread", node.ty, (args[0],), fm)
        v = by_sig[id(node.attr)]
        s = final_sig.get(v)
        if s is not None:
            return s.value()
        if v in el.comb:
            if v in inlining:
                raise VerilogError(f"combinational loop through `{v.name}`")
            inlining.add(v)
            out = _substitute(el.comb[v], memo, leaf)
            inlining.discard(v)
            return out
        raise VerilogError(f"internal: `{v.name}` has no definition")

    def conv(e: Expr) -> Expr:
        return _substitute(e, memo, leaf)

    for v in el.vars:
        s = final_sig.get(v)
        if s is None:
            continue
        if v in el.nexts:
            m.next(s, conv(el.nexts[v]))
        elif v in el.comb:
            m.assign(s, conv(el.comb[v]))
    for (mv, en, addr, data) in el.mem_writes:
        en2 = conv(en)
        if _is_lit(en2) and en2.attr == 0:
            continue
        m.mem_write(final_mem[id(mv.mem)], en2, conv(addr), conv(data))
    stats = {"registers": len(m.registers), "wires": sum(1 for s in m.signals if s.kind == ir.WIRE),
             "memories": len(m.memories), "register_bits": sum(s.width for s in m.registers),
             "declared": len(el.vars)}
    return Imported(m, name, inputs, outputs, stats)


def import_verilog(text: str, top: str, params: Optional[Dict[str, int]] = None,
                   defines: Optional[Dict[str, str]] = None, keep: Sequence[str] = ()) -> Imported:
    """Parses ``text`` (one or more modules), elaborates ``top`` with parameter
    overrides ``params`` and returns the flat design (see the module docstring)."""
    mods = parse(text, defines)
    el = Elaborator(mods, top, params)
    el._partial = {}
    el.elaborate()
    el._finish_partials()
    return build_module(el, top, keep)
