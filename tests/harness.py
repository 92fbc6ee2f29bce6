"""Oracle-side harnesses that replay a device run's stimulus on the CPU oracles.

The per-cycle sequence matches the device runs: drive inputs, settle, observe
(PRE), edge (capture, memory writes, commit), then settle + observe (POST) or
nothing; a final settle after the last cycle when observing PRE.
"""
from __future__ import annotations

import numpy as np

import oracle
from oracle import stim
from rrtl_amd.backend import IN_CONST, IN_HOLD, IN_PULSE, IN_RANDOM


def input_limbs(program, sig: int, mode: int, arg: int, seed: int, cycle: int, lanes: int) -> np.ndarray:
    s = program.signals[sig]
    if mode == IN_RANDOM:
        return stim.random_input_lanes(seed, sig, s.width, cycle, lanes)
    v = np.zeros((s.limbs, lanes), dtype=np.uint32)
    if mode == IN_CONST:
        val = arg & ((1 << s.width) - 1)
        for l in range(s.limbs):
            v[l, :] = (val >> (32 * l)) & 0xFFFFFFFF
    elif mode == IN_PULSE:
        n, first, after = arg & 0xFFFFFFFF, (arg >> 32) & 1, (arg >> 33) & 1
        v[0, :] = first if cycle < n else after
    return v


def packed_run(program, lanes, seed, cycles, inputs, observe, observe_post=False, ref=None):
    """Replays a word-engine run on the packed oracle.  Returns (ref, trace, digest)
    with trace [cycles][observed limbs][lanes] and the per-lane u64 digest chain."""
    if ref is None:
        ref = oracle.PackedRef(program, lanes)
    trace = []
    d = np.zeros(lanes, dtype=np.uint64)

    def obs():
        rows = []
        nonlocal d
        for j, s in enumerate(observe):
            lm = ref.get_signal_limbs(s)
            for k in range(lm.shape[0]):
                rows.append(lm[k].copy())
                d = stim.lane_digest(d, lm[k], j, k)
        trace.append(np.stack(rows) if rows else np.zeros((0, lanes), np.uint32))

    for c in range(cycles):
        for (sig, mode, arg) in inputs:
            if mode == IN_HOLD:
                continue
            ref.set_signal_limbs(sig, input_limbs(program, sig, mode, arg, seed, c, lanes), eval=False)
        ref.eval_combinational()
        if not observe_post:
            obs()
        ref.tick_from_evaluated()
        if observe_post:
            ref.eval_combinational()
            obs()
    if not observe_post:
        ref.eval_combinational()
    return ref, np.stack(trace) if trace else None, d
