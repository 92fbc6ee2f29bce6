"""Instance folding (rrtl_amd/instance_fold.py): the structural analysis on
the cases of the reference's tests (crates/rrtl-sim-ir/src/instance_fold.rs:
detects independent instances; rejects inter-instance dataflow) and more, and
the folded child at N x L lanes against the flattened top at L lanes on the
packed oracle, cycle by cycle (the reference's fold_matches_naive_cpu,
crates/rrtl-gpu-sim/src/instance_fold.rs:140-200)."""
import random

import numpy as np
import pytest

import oracle
from rrtl_amd import lower, verilog
from rrtl_amd.instance_fold import InstanceFoldError, InstanceFoldPlan, InstanceFoldSimulator, analyze_instance_fold

CHILD = """
module M #(parameter K = 32'h9e3779b9) (input clk, input [31:0] din, input en, output [31:0] o, output odd);
    reg [31:0] acc;
    always @(posedge clk) if (en) acc <= acc * K + din;
    assign o = acc;
    assign odd = acc[0] ^ din[3];
endmodule
"""


def top(n, body_extra="", conn=None, child="M", params=""):
    ports = ["input clk"] + [f"input [31:0] din{i}, input en{i}, output [31:0] o{i}, output odd{i}" for i in range(n)]
    insts = []
    for i in range(n):
        c = conn(i) if conn else f".clk(clk), .din(din{i}), .en(en{i}), .o(o{i}), .odd(odd{i})"
        insts.append(f"    {child} {params} u{i} ({c});")
    return CHILD + "module Top (" + ", ".join(ports) + ");\n" + body_extra + "\n".join(insts) + "\nendmodule\n"


def test_detects_independent_instances():
    f = analyze_instance_fold(top(4), "Top")
    assert f is not None and f.child_module == "M" and f.num_instances == 4
    assert f.instance_names == ["u0", "u1", "u2", "u3"]
    assert {(c.top_signal, c.instance, c.port, c.direction) for c in f.connections if c.instance == 2} == \
        {("clk", 2, "clk", "input"), ("din2", 2, "din", "input"), ("en2", 2, "en", "input"),
         ("o2", 2, "o", "output"), ("odd2", 2, "odd", "output")}
    # positional connections and equal parameter overrides fold too
    pos = top(3, conn=lambda i: f"clk, din{i}, en{i}, o{i}, odd{i}", params="#(32'd5)")
    f = analyze_instance_fold(pos, "Top")
    assert f is not None and f.child_params == {"K": 5}


def test_rejects_what_is_not_a_fold():
    assert analyze_instance_fold(top(1), "Top") is None                      # one instance
    # an instance output feeding another instance's input
    chain = top(2, conn=lambda i: f".clk(clk), .din({'o0' if i else 'din0'}), .en(en{i}), .o(o{i}), .odd(odd{i})")
    assert analyze_instance_fold(chain, "Top") is None
    # two instances driving one signal
    shared = top(2, conn=lambda i: f".clk(clk), .din(din{i}), .en(en{i}), .o(o0), .odd(odd{i})")
    assert analyze_instance_fold(shared, "Top") is None
    # logic of the top's own
    assert analyze_instance_fold(top(2, body_extra="    wire x;\n    assign x = din0[0];\n"), "Top") is None
    # a connection that is an expression
    expr = top(2, conn=lambda i: f".clk(clk), .din(din{i} + 1), .en(en{i}), .o(o{i}), .odd(odd{i})")
    assert analyze_instance_fold(expr, "Top") is None
    # different parameters
    two = CHILD + """module Top (input clk, input [31:0] din0, din1, input en0, en1, output [31:0] o0, o1,
                     output odd0, odd1);
        M #(.K(3)) u0 (.clk(clk), .din(din0), .en(en0), .o(o0), .odd(odd0));
        M #(.K(5)) u1 (.clk(clk), .din(din1), .en(en1), .o(o1), .odd(odd1));
    endmodule"""
    assert analyze_instance_fold(two, "Top") is None
    with pytest.raises(InstanceFoldError):
        InstanceFoldPlan(top(1), "Top")


class PackedEngine:
    """The packed oracle behind the engine interface the fold simulator uses."""

    def __init__(self, program, lanes):
        self.r = oracle.PackedRef(program, lanes)

    def set_signal_limbs(self, sig, v, settle=True):
        self.r.set_signal_limbs(sig, v, eval=settle)

    def get_signal_limbs(self, sig):
        return self.r.get_signal_limbs(sig)

    def eval_combinational(self):
        self.r.eval_combinational()

    def tick(self):
        self.r.tick()

    def tick_many(self, n):
        for _ in range(n):
            self.r.tick()


def run_fold_vs_flat(n, base, cycles, make_engine, seed=3):
    text = top(n)
    fold = InstanceFoldSimulator(text, "Top", base, engine=make_engine)
    flat_p = lower.lower_module(verilog.import_verilog(text, "Top").module)
    flat = make_engine(flat_p, base)
    rng = random.Random(seed)
    clk = np.ones((1, base), np.uint32)
    fold.set_top_limbs("clk", clk)
    flat.set_signal_limbs(flat_p.signal_index("clk"), clk, settle=False)
    for c in range(cycles):
        for i in range(n):
            din = np.array([[rng.getrandbits(32) for _ in range(base)]], np.uint32)
            en = np.array([[rng.getrandbits(1) for _ in range(base)]], np.uint32)
            fold.set_top_limbs(f"din{i}", din)
            fold.set_top_limbs(f"en{i}", en)
            flat.set_signal_limbs(flat_p.signal_index(f"din{i}"), din, settle=False)
            flat.set_signal_limbs(flat_p.signal_index(f"en{i}"), en, settle=False)
        fold.tick()
        flat.tick()
        for i in range(n):
            for s in (f"o{i}", f"odd{i}"):
                assert np.array_equal(fold.get_top_limbs(s), flat.get_signal_limbs(flat_p.signal_index(s))), (c, s)
    return fold, flat_p


def test_fold_matches_flattened_on_the_packed_oracle():
    fold, flat_p = run_fold_vs_flat(4, 5, 12, PackedEngine)
    assert fold.total_lanes == 20 and fold.instance_lane(2, 3) == 13
    # the folded program is the child's: a quarter of the flattened one
    assert len(fold.plan.child.signals) * 3 < len(flat_p.signals)
