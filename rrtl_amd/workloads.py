"""Multi-design batch workload: the four designs of the mixed configuration
(the README counter, the 20k-gate netlist, a 16-entry ready/valid FIFO with
32-bit data, picorv32 running per-lane random programs) with lanes per
design drawn uniformly from [1k, 64k] and cycles log-uniformly from [1e3, 1e5],
all from one seed.  Used by the batch benchmark and its tests."""
from __future__ import annotations

import math
import random
from dataclasses import dataclass
from typing import Callable, List, Optional, Tuple

import numpy as np

from . import designs, lower, picorv32, riscv
from .backend import IN_CONST, IN_PULSE, IN_RANDOM, HipSimulator
from .program import KIND_INPUT, KIND_OUTPUT, Program


def pulse_arg(n: int, during: int, after: int) -> int:
    """IN_PULSE argument: bit 32 is the value while cycle < n, bit 33 after."""
    return (n & 0xFFFFFFFF) | ((during & 1) << 32) | ((after & 1) << 33)


@dataclass
class Job:
    name: str
    program: Program
    lanes: int
    engine: str
    run: dict                                         # HipSimulator.run keyword arguments
    setup: Optional[Callable[[HipSimulator], None]] = None
    check_mems: Tuple[str, ...] = ()                 # state compared with the host engine after a run
    check_sigs: Tuple[str, ...] = ()

    def make(self, **opts) -> HipSimulator:
        sim = HipSimulator(self.program, self.lanes, engine=self.engine, **opts)
        if self.setup:
            self.setup(sim)
        return sim


def counter_job(lanes: int, cycles: int, seed: int) -> Job:
    p = lower.lower_module(designs.counter())
    return Job("counter", p, lanes, "word", dict(
        cycles=cycles, seed=seed, observe=[p.signal_index("out")],
        inputs=[(p.signal_index("clk"), IN_CONST, 1), (p.signal_index("rst"), IN_PULSE, pulse_arg(1, 1, 0)),
                (p.signal_index("en"), IN_RANDOM, 0)]))


def netlist_job(lanes: int, cycles: int, seed: int, **netlist) -> Job:
    p = lower.lower_module(designs.synthetic_netlist(**netlist))
    return Job("gates20k", p, lanes, "gate", dict(
        cycles=cycles, seed=seed,
        inputs=[(i, IN_RANDOM, 0) for i, s in enumerate(p.signals) if s.kind == KIND_INPUT],
        observe=[i for i, s in enumerate(p.signals) if s.kind == KIND_OUTPUT]))


def fifo_job(lanes: int, cycles: int, seed: int) -> Job:
    p = lower.lower_module(designs.ready_valid_fifo(16, 32))
    return Job("fifo", p, lanes, "word", dict(
        cycles=cycles, seed=seed,
        inputs=[(p.signal_index("clk"), IN_CONST, 1), (p.signal_index("rst"), IN_PULSE, pulse_arg(1, 1, 0)),
                (p.signal_index("in_valid"), IN_RANDOM, 0), (p.signal_index("out_ready"), IN_RANDOM, 0),
                (p.signal_index("in_data"), IN_RANDOM, 0)],
        observe=[p.signal_index(n) for n in ("out_valid", "out_data", "count")]))


def rv32i_job(lanes: int, cycles: int, seed: int, n_insns: int = 512, distinct: int = 256) -> Job:
    """Per-lane random RV32I programs (``distinct`` different programs dealt
    round-robin to the lanes), halting on ``ecall``."""
    p = lower.lower_module(riscv.rv32i_core())
    rng = random.Random(seed)
    images = np.array([riscv.load_image(riscv.random_program(rng, n=n_insns)) for _ in range(distinct)],
                      dtype=np.uint32)

    def setup(sim: HipSimulator) -> None:
        img = images[np.arange(sim.lanes) % distinct]
        sim.set_memory_limbs("mem", img.reshape(sim.lanes, -1, 1))

    return Job("rv32i", p, lanes, "word", dict(
        cycles=cycles, seed=seed, halt="trap", observe=[],
        inputs=[(p.signal_index("clk"), IN_CONST, 1), (p.signal_index("rst"), IN_CONST, 0)]), setup,
        check_mems=("regs", "mem"), check_sigs=("pc", "trap"))


def picorv32_job(lanes: int, cycles: int, seed: int, n_insns: int = 512, distinct: int = 256,
                 backward: Optional[float] = None) -> Job:
    """The reference's picorv32 SoC (imported from bench/sv) running per-lane
    random RV32I programs (``distinct`` programs dealt round-robin), halting
    on ``trap`` after ``ecall``; ``resetn`` low for the first cycles."""
    p = picorv32.program()
    rng = random.Random(seed)
    images = picorv32.images([picorv32.random_program(rng, n=n_insns, backward=backward) for _ in range(distinct)])

    def setup(sim: HipSimulator) -> None:
        sim.set_memory_limbs("mem", images[np.arange(sim.lanes) % distinct])

    return Job("picorv32", p, lanes, "word", dict(cycles=cycles, seed=seed, halt="trap", observe=[],
                                                   inputs=picorv32.inputs(p)), setup,
               check_mems=("cpu.cpuregs", "mem"), check_sigs=("cpu.reg_pc", "trap"))


def mixed_workload(seed: int = 1, lanes_range=(1000, 65536), cycles_range=(1e3, 1e5)) -> List[Job]:
    rng = random.Random(seed)

    def draw():
        lanes = rng.randint(*lanes_range)
        cycles = int(round(math.exp(rng.uniform(math.log(cycles_range[0]), math.log(cycles_range[1])))))
        return lanes, cycles

    jobs = []
    for k, make in enumerate((counter_job, netlist_job, fifo_job, picorv32_job)):
        lanes, cycles = draw()
        jobs.append(make(lanes, cycles, seed * 16 + k))
    return jobs
