"""The picorv32 SoC workload: the reference's RV32I core running per-lane programs.

The design is the reference's ``picorv32_soc`` (bench/sv/picorv32_soc.v: the
picorv32 core of bench/sv/picorv32.v with ENABLE_COUNTERS=0, ENABLE_IRQ=0,
CATCH_MISALIGN=0, CATCH_ILLINSN=0, BARREL_SHIFTER=0, no MUL/DIV, no
compressed ISA, plus a 1024-word synchronous RAM and a store-to-0x100 result
latch), imported by ``rrtl_amd.verilog`` and lowered to the packed program
committed as ``data/picorv32_soc.rrtl`` (``data/make_picorv32.py`` rebuilds
it from the sources).

Per lane: the RAM holds the lane's program from address 0; ``resetn`` is
held low for the first ``RESET_CYCLES`` cycles; ``ecall`` moves the core to
its trap state and raises ``trap`` (CATCH_ILLINSN=0 turns ``ecall``/``ebreak``
into a trap, bench/sv/picorv32.v:1945-1947), which halts the lane.  The
reference's own check runs the sum 1..5 program and expects ``result == 15``,
``done == 1``, ``trap == 0`` (crates/rrtl-gpu-sim/examples/picorv32_gpu.rs:
24-36, :85-91).
"""
from __future__ import annotations

import os
import random
from functools import lru_cache
from typing import List, Optional, Sequence

import numpy as np

from . import riscv
from .program import Program

BLOB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "picorv32_soc.rrtl")
RESET_CYCLES = 4
MEM_WORDS = 1024
# random programs keep their data clear of the code: words 640..1023
DATA_BASE = 0xA00
DATA_SPAN = 0x5FC


@lru_cache(maxsize=None)
def _blob() -> bytes:
    with open(BLOB, "rb") as f:
        return f.read()


def program() -> Program:
    return Program.from_blob(_blob(), "picorv32_soc")


def sum_program() -> List[int]:
    """The reference example's program: sum 1..5 into x5, store it to 0x100
    (latched into ``result``/``done``), then spin (picorv32_gpu.rs:24-36)."""
    A = riscv.asm
    return [A("addi", 5, 0, 0), A("addi", 6, 0, 1), A("addi", 7, 0, 6),
            A("add", 5, 5, 6), A("addi", 6, 6, 1), A("blt", 6, 7, -8),
            A("addi", 8, 0, 0x100), A("sw", 5, 8, 0), A("jal", 0, 0)]


def random_program(rng: random.Random, n: int = 512, backward: Optional[float] = None) -> List[int]:
    """A random RV32I program (ALU 60 %, load/store 25 %, branch 15 %, in-range
    targets, backward with probability ``backward``, None: uniform over the
    program) with its data above the code, ending in ``nop; ecall``.

    The ``nop``: with CATCH_ILLINSN=0 picorv32 traps when it decodes
    ``ecall`` (bench/sv/picorv32.v:1945-1947), and it decodes the next
    instruction while a load or store is still in its memory state, so a load
    or store directly before ``ecall`` would be cut off before its register
    write.  Branches that targeted the ``ecall`` reach the ``nop`` instead."""
    prog = riscv.random_program(rng, n=n, backward=backward, data_base=DATA_BASE, data_span=DATA_SPAN)
    prog.insert(len(prog) - 1, riscv.asm("addi", 0, 0, 0))
    return prog


def inputs(p: Program, reset_cycles: int = RESET_CYCLES):
    """Run inputs: the clock constant, ``resetn`` low for ``reset_cycles``."""
    from .backend import IN_CONST, IN_PULSE
    from .workloads import pulse_arg
    return [(p.signal_index("clk"), IN_CONST, 1),
            (p.signal_index("resetn"), IN_PULSE, pulse_arg(reset_cycles, 0, 1))]


def images(progs: Sequence[Sequence[int]]) -> np.ndarray:
    """[lanes][1024][1] memory limbs of the programs."""
    return np.array([riscv.load_image(pr, MEM_WORDS) for pr in progs], dtype=np.uint32).reshape(len(progs), -1, 1)


def iss_run(prog: Sequence[int], max_insns: int) -> riscv.Iss:
    s = riscv.Iss(riscv.load_image(prog, MEM_WORDS))
    s.run(max_insns)
    return s
