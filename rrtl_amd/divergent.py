"""A divergent batch: instances that halt at data-dependent cycles.

The reference's divergent-batch example (crates/rrtl-sim-ir/examples/
divergent_batch.rs) runs bench/sv/workload.sv, a counter loaded with a
per-instance run length that does a multiply-accumulate per cycle until the
counter reaches 0, with 1/8 of the instances 25x longer than the rest, and
compares a static schedule with work stealing.  This module's design is the
same shape written for halting on the device: ``countdown`` below adds a
``done`` output (counter at 0 after the load), which is the lane's halt
signal.  The design is imported with ``rrtl_amd.verilog``.
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np

from . import lower, verilog
from .program import Program

SOURCE = """
// Per-lane countdown with real work per cycle: load `seed` into the counter,
// then accumulate acc = acc * din + counter until the counter is 0.
module countdown (
    input             clk,
    input             load,
    input      [31:0] seed,
    input      [31:0] din,
    output reg [31:0] acc,
    output reg [31:0] left,
    output            done
);
    assign done = ~load & (left == 32'd0);
    always @(posedge clk) begin
        if (load) begin
            left <= seed;
            acc  <= 32'd0;
        end else if (left != 32'd0) begin
            acc  <= acc * din + left;
            left <= left - 32'd1;
        end
    end
endmodule
"""


@lru_cache(maxsize=None)
def program() -> Program:
    return lower.lower_module(verilog.import_verilog(SOURCE, "countdown").module)


def run_lengths(lanes: int, short: int, long: int, seed: int = 1) -> np.ndarray:
    """Run length per lane: 1/8 of the lanes (scattered by a seeded
    permutation) ``long``, the rest ``short``."""
    rng = np.random.default_rng(seed)
    n = np.full(lanes, short, dtype=np.uint32)
    n[rng.permutation(lanes)[: lanes // 8]] = long
    return n


def closed_form(seeds: np.ndarray, din: int) -> np.ndarray:
    """acc after the countdown of each seed: acc <- acc * din + k for k = seed .. 1 (mod 2^32)."""
    acc = np.zeros(seeds.size, dtype=np.uint64)
    left = seeds.astype(np.uint64)
    for _ in range(int(seeds.max())):
        live = left > 0
        acc = np.where(live, (acc * np.uint64(din) + left) & np.uint64(0xFFFFFFFF), acc)
        left = np.where(live, left - np.uint64(1), left)
    return acc.astype(np.uint32)


def run_inputs(p: Program):
    """The load pulse in cycle 0, the clock, and din held at its set value."""
    from .backend import IN_CONST, IN_PULSE
    from .workloads import pulse_arg
    return [(p.signal_index("clk"), IN_CONST, 1), (p.signal_index("load"), IN_PULSE, pulse_arg(1, 1, 0))]
