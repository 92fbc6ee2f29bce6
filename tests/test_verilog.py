"""The Verilog-subset importer (rrtl_amd.verilog): language semantics against
closed forms written from the Verilog-2005 rules (IEEE 1364-2005 5.4-5.5 for
expression sizing, 9.2 for blocking/non-blocking assignments), run through
the lowering and the packed oracle; the reference's benchmark designs
(bench/sv/crc32.sv, workload.sv, picorv32.v + picorv32_soc.v) against their
closed forms and known answers when the reference tree is present."""
import os
import random

import numpy as np
import pytest

import oracle
from rrtl_amd import lower, verilog
from rrtl_amd.ir import DesignError

SV = "/root/reference/bench/sv"
need_ref = pytest.mark.skipif(not os.path.isdir(SV), reason="reference sources not present")


class Dut:
    """A lowered import on the packed oracle, one lane per stimulus vector."""

    def __init__(self, text, top, lanes, **kw):
        self.imp = verilog.import_verilog(text, top, **kw)
        self.p = lower.lower_module(self.imp.module)
        self.ref = oracle.PackedRef(self.p, lanes)
        self.lanes = lanes

    def set(self, name, values):
        self.ref.set_signal(self.p.signal_index(name), values, eval=False)

    def get(self, name):
        self.ref.eval_combinational()
        return self.ref.get_signal(self.p.signal_index(name))

    def tick(self):
        self.ref.eval_combinational()
        self.ref.tick_from_evaluated()


def sx(v, w):
    v &= (1 << w) - 1
    return v - (1 << w) if v >> (w - 1) else v


def test_preprocessor_macros_and_conditionals():
    src = """
`define W 6
`define ADD(a, b) ((a) + (b))
`define FEATURE
module m(input [`W-1:0] x, output [`W-1:0] y, output z);
`ifdef FEATURE
  assign y = `ADD(x, 6'd3); // comment
`else
  assign y = 0;
`endif
`ifndef FEATURE
  assign z = 1;
`elsif MISSING
  assign z = 1;
`else
  /* block
     comment */ assign z = 0;
`endif
endmodule"""
    d = Dut(src, "m", 4)
    d.set("x", [0, 1, 60, 63])
    assert d.get("y") == [3, 4, 63, 2]
    assert d.get("z") == [0] * 4


def test_expression_sizing_rules():
    """Context-determined widths: a 9-bit target keeps the carry of 8-bit
    operands; inside a concatenation the sum is self-determined (8 bits);
    comparisons size both operands to the wider one; an unsized ~0 is 32 bits
    then truncated; shifts keep the left operand's width; signed compare,
    arithmetic shift and sign extension follow $signed."""
    src = """
module m(input [7:0] a, input [7:0] b, input [3:0] s,
         output [8:0] sum9, output [8:0] cat9, output lt, output slt, output [7:0] sra, output [7:0] srl,
         output [15:0] sext, output [7:0] ones, output [7:0] shl, output eqw, output [3:0] red);
  assign sum9 = a + b;
  assign cat9 = {1'b0, a + b};
  assign lt = a < b;
  assign slt = $signed(a) < $signed(b);
  assign sra = $signed(a) >>> s;
  assign srl = a >> s;
  assign sext = $signed(a);
  assign ones = ~0;
  assign shl = a << s;
  assign eqw = (a + b) == 9'h100;
  assign red = {|a, &a, ^a, ~|b};
endmodule"""
    rng = random.Random(1)
    vec = [(rng.randrange(256), rng.randrange(256), rng.randrange(16)) for _ in range(200)]
    vec += [(255, 1, 0), (128, 127, 7), (0, 0, 15), (255, 255, 8)]
    d = Dut(src, "m", len(vec))
    d.set("a", [v[0] for v in vec])
    d.set("b", [v[1] for v in vec])
    d.set("s", [v[2] for v in vec])
    got = {n: d.get(n) for n in ("sum9", "cat9", "lt", "slt", "sra", "srl", "sext", "ones", "shl", "eqw", "red")}
    for k, (a, b, s) in enumerate(vec):
        assert got["sum9"][k] == a + b
        assert got["cat9"][k] == (a + b) & 0xFF
        assert got["lt"][k] == int(a < b)
        assert got["slt"][k] == int(sx(a, 8) < sx(b, 8))
        assert got["sra"][k] == (sx(a, 8) >> s) & 0xFF
        assert got["srl"][k] == a >> s
        assert got["sext"][k] == sx(a, 8) & 0xFFFF
        assert got["ones"][k] == 0xFF
        assert got["shl"][k] == (a << s) & 0xFF
        assert got["eqw"][k] == int(a + b == 0x100)
        red = (int(a != 0) << 3) | (int(a == 255) << 2) | ((bin(a).count("1") & 1) << 1) | int(b == 0)
        assert got["red"][k] == red


def test_case_casez_and_combinational_defaults():
    """case picks the first matching item (several labels per item, default);
    casez treats z/? label bits as don't-care; a variable left unassigned on a
    path of an always @* block is 0 (no latch)."""
    src = """
module m(input [3:0] x, output reg [3:0] y, output reg [1:0] z, output reg w);
  always @* begin
    case (x)
      4'd1, 4'd2: y = 4'd10;
      4'd3: y = 4'd11;
      4'd2: y = 4'd12;   // shadowed by the first item
      default: y = x;
    endcase
    casez (x)
      4'b1???: z = 2'd3;
      4'b01?1: z = 2'd2;
      default: z = 2'd0;
    endcase
    if (x == 4'd9) w = 1'b1;
  end
endmodule"""
    d = Dut(src, "m", 16)
    d.set("x", list(range(16)))
    y, z, w = d.get("y"), d.get("z"), d.get("w")
    for x in range(16):
        assert y[x] == (10 if x in (1, 2) else 11 if x == 3 else x)
        assert z[x] == (3 if x & 8 else 2 if (x & 0b1101) == 0b0101 else 0)
        assert w[x] == int(x == 9)


def test_blocking_and_nonblocking_semantics():
    """Non-blocking assignments read pre-edge values (a swap), blocking ones
    update at once (a chain), later assignments win, a register without an
    assignment on a path holds, part and concatenation targets update only
    their bits, a for loop over an integer unrolls."""
    src = """
module m(input clk, input en, input [7:0] d,
         output reg [7:0] p, output reg [7:0] q, output reg [7:0] r, output reg [7:0] h,
         output reg [7:0] part, output reg [3:0] hi, output reg [3:0] lo, output reg [7:0] rev);
  reg [7:0] t;
  integer i;
  always @(posedge clk) begin
    p <= q;              // swap
    q <= p ^ d;
    t = d + 8'd1;        // blocking chain
    r <= t + t;
    r <= r + 8'd1 + t;   // the later assignment wins
    if (en) h <= d;      // holds when en = 0
    part[3:0] <= d[7:4];
    {hi, lo} <= {d[3:0], d[7:4]};
    for (i = 0; i < 8; i = i + 1)
      rev[i] <= d[7 - i];
  end
endmodule"""
    lanes, cycles = 8, 12
    rng = random.Random(3)
    d = Dut(src, "m", lanes)
    d.set("clk", 1)
    st = [dict(p=0, q=0, r=0, h=0, part=0, hi=0, lo=0, rev=0) for _ in range(lanes)]
    for c in range(cycles):
        dv = [rng.randrange(256) for _ in range(lanes)]
        ev = [rng.randrange(2) for _ in range(lanes)]
        d.set("d", dv)
        d.set("en", ev)
        d.tick()
        for k in range(lanes):
            s, x = st[k], dv[k]
            t = (x + 1) & 0xFF
            n = dict(p=s["q"], q=s["p"] ^ x, r=(s["r"] + 1 + t) & 0xFF, h=x if ev[k] else s["h"],
                     part=(s["part"] & 0xF0) | (x >> 4), hi=x & 0xF, lo=x >> 4,
                     rev=int(f"{x:08b}"[::-1], 2))
            st[k] = n
        for name in st[0]:
            assert d.get(name) == [s[name] for s in st], (c, name)


def test_memory_write_after_read_in_one_edge():
    """A memory read in a clocked block returns the pre-edge word even when the
    same edge writes it (the reference SoC's registered read)."""
    src = """
module m(input clk, input we, input [2:0] addr, input [15:0] wd, output reg [15:0] rd);
  reg [15:0] mem [0:7];
  always @(posedge clk) begin
    rd <= mem[addr];
    if (we) mem[addr] <= wd;
  end
endmodule"""
    lanes = 4
    d = Dut(src, "m", lanes)
    d.set("clk", 1)
    model = [[0] * 8 for _ in range(lanes)]
    rd = [0] * lanes
    rng = random.Random(7)
    for c in range(40):
        we = [rng.randrange(2) for _ in range(lanes)]
        ad = [rng.randrange(8) for _ in range(lanes)]
        wd = [rng.randrange(1 << 16) for _ in range(lanes)]
        d.set("we", we)
        d.set("addr", ad)
        d.set("wd", wd)
        d.tick()
        for k in range(lanes):
            rd[k] = model[k][ad[k]]
            if we[k]:
                model[k][ad[k]] = wd[k]
        assert d.get("rd") == rd, c
    mem = d.ref.get_memory_limbs(d.p.memory_index("mem"))[:, :, 0]
    assert mem.tolist() == model


def test_instances_parameters_and_generate():
    src = """
module add #(parameter W = 4, parameter [0:0] INV = 0) (input [W-1:0] a, input [W-1:0] b, output [W-1:0] y);
  generate if (INV) begin
    assign y = ~(a + b);
  end else begin
    assign y = a + b;
  end endgenerate
endmodule
module top(input [7:0] x, output [7:0] s, output [7:0] n);
  add #(.W(8)) u0 (.a(x), .b(8'd5), .y(s));
  add #(.W(8), .INV(1)) u1 (.a(x), .b(8'd1), .y(n));
endmodule"""
    d = Dut(src, "top", 5)
    xs = [0, 1, 100, 250, 255]
    d.set("x", xs)
    assert d.get("s") == [(x + 5) & 0xFF for x in xs]
    assert d.get("n") == [~(x + 1) & 0xFF for x in xs]


def test_dead_logic_is_removed_and_keep_retains_it():
    src = """
module m(input clk, input [7:0] x, output reg [7:0] y);
  reg [7:0] dead;
  always @(posedge clk) begin
    y <= x;
    dead <= dead + x;
  end
endmodule"""
    a = verilog.import_verilog(src, "m")
    assert [s.name for s in a.module.registers] == ["y"]
    b = verilog.import_verilog(src, "m", keep=["dead"])
    assert sorted(s.name for s in b.module.registers) == ["dead", "y"]


@pytest.mark.parametrize("src,msg", [
    ("module m(output y); assign y = nope; endmodule", "not declared"),
    ("module m(input a, output y); assign y = a; assign y = ~a; endmodule", "more than one driver"),
    ("module m(input clk, input a, output reg y); always @(posedge clk or negedge a) y <= a; endmodule",
     "single-clock"),
    ("module m(input a, output y); function f; input x; f = x; endfunction assign y = a; endmodule", "functions"),
    ("module m(input a, output y); assign y = a; endmodule\nmodule n(); m u(.q(1'b0)); endmodule", "no port"),
])
def test_unsupported_or_invalid_input_is_rejected(src, msg):
    top = "n" if "module n" in src else "m"
    with pytest.raises(DesignError, match=msg):
        verilog.import_verilog(src, top)


def crc_step(crc: int, byte: int) -> int:
    """One byte of the MSB-first CRC-32 recurrence written in crc32.sv."""
    for i in range(8):
        fb = ((crc >> 31) ^ (byte >> i)) & 1
        crc = ((crc << 1) & 0xFFFFFFFF) ^ (0x04C11DB7 if fb else 0)
    return crc


@need_ref
def test_reference_crc32_sv_closed_form():
    """bench/sv/crc32.sv (blocking temporary and a for loop in a clocked
    block): one reset cycle, then 200 cycles of din = 0xAB, gives the closed
    form of its recurrence, 0x53535102."""
    d = Dut(open(os.path.join(SV, "crc32.sv")).read(), "crc32", 3)
    d.set("clk", 1)
    d.set("rst", 1)
    d.tick()
    d.set("rst", 0)
    d.set("din", [0xAB, 0x00, 0x5C])
    for _ in range(200):
        d.tick()
    want = []
    for b in (0xAB, 0x00, 0x5C):
        c = 0xFFFFFFFF
        for _ in range(200):
            c = crc_step(c, b)
        want.append(c)
    assert want[0] == 0x53535102
    assert d.get("crc") == want


@need_ref
def test_reference_workload_sv_halts_at_seed():
    """bench/sv/workload.sv: load a per-lane seed, then acc <- acc * din +
    counter until the counter reaches 0 (data-dependent halt)."""
    seeds = [0, 1, 5, 17]
    d = Dut(open(os.path.join(SV, "workload.sv")).read(), "workload", len(seeds))
    d.set("clk", 1)
    d.set("seed", seeds)
    d.set("din", 3)
    d.set("load", 1)
    d.tick()
    d.set("load", 0)
    acc = [0] * len(seeds)
    cnt = list(seeds)
    for _ in range(25):
        d.tick()
        for k in range(len(seeds)):
            if cnt[k]:
                acc[k] = (acc[k] * 3 + cnt[k]) & 0xFFFFFFFF
                cnt[k] -= 1
        assert d.get("acc") == acc


@need_ref
def test_committed_picorv32_program_is_the_import():
    """rrtl_amd/data/picorv32_soc.rrtl is exactly what the importer and the
    lowering make of the reference's picorv32 sources today."""
    from rrtl_amd import picorv32
    from rrtl_amd.data import make_picorv32
    p, _, _ = make_picorv32.build_program(SV)
    with open(picorv32.BLOB, "rb") as f:
        assert p.to_blob() == f.read()


def test_picorv32_sum_known_answer_on_packed_oracle():
    """picorv32_gpu.rs:24-36, :85-91 on the CPU oracle: result == 15,
    done == 1, trap == 0 after 8 reset cycles and 3000 cycles."""
    from rrtl_amd import picorv32 as pv
    p = pv.program()
    ref = oracle.PackedRef(p, 2)
    ref.set_memory_limbs(p.memory_index("mem"), pv.images([pv.sum_program()] * 2))
    ref.set_signal(p.signal_index("clk"), 1)
    ref.set_signal(p.signal_index("resetn"), 0)
    ref.tick_many(8)
    ref.set_signal(p.signal_index("resetn"), 1)
    ref.tick_many(3000)
    assert ref.get_signal(p.signal_index("result")) == [15, 15]
    assert ref.get_signal(p.signal_index("done")) == [1, 1]
    assert ref.get_signal(p.signal_index("trap")) == [0, 0]


def test_picorv32_random_programs_match_iss_and_host_engine():
    """Random programs on picorv32 (packed oracle): lanes that reach ecall hold
    the ISS's registers and memory; the host batch-SIMD engine's digest of
    every signal every cycle equals the packed oracle's."""
    from harness import packed_run
    from rrtl_amd import picorv32 as pv
    from rrtl_amd.program import KIND_INPUT
    p = pv.program()
    lanes, cycles = 12, 2500
    rng = random.Random(5)
    progs = [pv.random_program(rng, n=48, backward=0.15) for _ in range(lanes)]
    ins = pv.inputs(p)
    obs = [i for i, s in enumerate(p.signals) if s.kind != KIND_INPUT]
    pk = oracle.PackedRef(p, lanes)
    pk.set_memory_limbs(p.memory_index("mem"), pv.images(progs))
    pk, _, od = packed_run(p, lanes, 0, cycles, ins, obs, ref=pk)
    simd = oracle.SimdCpu(p, lanes, threads=2, shard_lanes=8)
    simd.set_memory_limbs(p.memory_index("mem"), pv.images(progs))
    dg, _ = simd.run(0, cycles, ins, obs)
    assert np.array_equal(dg, od)
    trap = pk.get_signal(p.signal_index("trap"))
    regs = pk.get_memory_limbs(p.memory_index("cpu.cpuregs"))[:, :, 0]
    mem = pk.get_memory_limbs(p.memory_index("mem"))[:, :, 0]
    halted = 0
    for l, pr in enumerate(progs):
        s = pv.iss_run(pr, 50_000)
        if trap[l]:
            assert s.trap
            assert [int(x) for x in regs[l]] == s.regs, l
            assert [int(x) for x in mem[l]] == s.mem, l
            halted += 1
    assert halted >= lanes // 3
