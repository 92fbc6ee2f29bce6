"""Yosys JSON gate-netlist importer.

Restates ``import_yosys_netlist`` (crates/rrtl-sv-frontend/src/netlist.rs:49-188):
every net becomes a 1-bit signal (input-port bits named ``port`` or
``port_i``, flop outputs ``n{id}`` as registers, everything else ``n{id}``
wires); ``$_DFF*`` cells give ``next(Q) = D`` (``$_DFFE_*``: ``mux(E, D, Q)``);
gate cells ``$_NOT_ $_AND_ $_OR_ $_XOR_ $_NAND_ $_NOR_ $_XNOR_ $_ANDNOT_
$_ORNOT_ $_MUX_`` (Y = S ? B : A); constant bits ``"0"``/``"1"`` are literals
and ``"x"``/``"z"`` read as 0; output-port bits become outputs assigned from
their nets.  Nets are created in ascending id order and cells in name order so
signal indices are deterministic (the reference iterates a hash set).
"""
from __future__ import annotations

import json
from typing import Dict, List, Union

from . import ir
from .ir import lit, mux


def _net(b) -> Union[int, str]:
    if isinstance(b, int):
        return b
    if b in ("0", "1"):
        return b
    if b in ("x", "z"):
        return "0"
    raise ValueError(f"E_NETLIST: bad net reference {b!r}")


def import_yosys_netlist(text: str, top: str) -> ir.Module:
    v = json.loads(text)
    mod = v["modules"].get(top)
    if mod is None:
        raise ValueError(f"E_NETLIST: module `{top}` not found")
    cells = mod.get("cells", {})
    ports = mod.get("ports", {})
    input_nets: Dict[int, str] = {}
    reg_nets = set()
    all_nets = set()
    out_bits = []
    for pname, p in ports.items():
        bits = p.get("bits", [])
        multi = len(bits) > 1
        for i, b in enumerate(bits):
            n = _net(b)
            if isinstance(n, int):
                all_nets.add(n)
                if p.get("direction") == "input":
                    input_nets[n] = f"{pname}_{i}" if multi else pname
            if p.get("direction") == "output":
                out_bits.append((pname, i, n, multi))
    for cname in sorted(cells):
        cell = cells[cname]
        for arr in cell.get("connections", {}).values():
            for b in arr:
                n = _net(b)
                if isinstance(n, int):
                    all_nets.add(n)
        if cell["type"].startswith("$_DFF"):
            q = _net(cell["connections"]["Q"][0])
            if isinstance(q, int):
                reg_nets.add(q)

    m = ir.Module(top)
    sig: Dict[int, ir.Signal] = {}
    for n in sorted(all_nets):
        if n in input_nets:
            sig[n] = m.input(input_nets[n], 1)
        elif n in reg_nets:
            sig[n] = m.reg(f"n{n}", 1)
        else:
            sig[n] = m.wire(f"n{n}", 1)

    def val(n) -> ir.Expr:
        if isinstance(n, int):
            return sig[n].value()
        return lit(int(n), 1)

    def conn(cell, port) -> ir.Expr:
        arr = cell["connections"][port]
        if len(arr) != 1:
            raise ValueError(f"E_NETLIST: gate cell port `{port}` is not 1-bit")
        return val(_net(arr[0]))

    for cname in sorted(cells):
        cell = cells[cname]
        ty = cell["type"]
        if ty.startswith("$_DFF"):
            q = _net(cell["connections"]["Q"][0])
            if not isinstance(q, int):
                raise ValueError(f"E_NETLIST: DFF `{cname}` Q is constant")
            c = _net(cell["connections"]["C"][0])
            if isinstance(c, int):
                m.clock(sig[q], sig[c])
            d = conn(cell, "D")
            nxt = mux(conn(cell, "E"), d, sig[q].value()) if ty.startswith("$_DFFE") else d
            m.next(sig[q], nxt)
            continue
        y = _net(cell["connections"]["Y"][0])
        if not isinstance(y, int):
            raise ValueError(f"E_NETLIST: cell `{cname}` output is constant")
        a = lambda p: conn(cell, p)
        if ty == "$_NOT_":
            e = ~a("A")
        elif ty == "$_AND_":
            e = a("A") & a("B")
        elif ty == "$_OR_":
            e = a("A") | a("B")
        elif ty == "$_XOR_":
            e = a("A") ^ a("B")
        elif ty == "$_NAND_":
            e = ~(a("A") & a("B"))
        elif ty == "$_NOR_":
            e = ~(a("A") | a("B"))
        elif ty == "$_XNOR_":
            e = ~(a("A") ^ a("B"))
        elif ty == "$_ANDNOT_":
            e = a("A") & ~a("B")
        elif ty == "$_ORNOT_":
            e = a("A") | ~a("B")
        elif ty == "$_MUX_":
            e = mux(a("S"), a("B"), a("A"))
        else:
            raise ValueError(f"E_NETLIST: unsupported gate cell `{ty}` (cell `{cname}`)")
        m.assign(sig[y], e)
    for pname, i, n, multi in out_bits:
        o = m.output(f"{pname}_{i}" if multi else pname, 1)
        m.assign(o, val(n))
    return m
