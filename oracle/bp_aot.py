"""ORACLE (test infrastructure only): the host AOT bit-parallel engine.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this module, as the checker or as the timed CPU baseline.

Restates ``BitParallelAot`` (crates/rrtl-sim-ir/src/bitparallel.rs:414-525):
the design's packed program is emitted as straight-line C, one statement per
gate instruction on u64 words (64 lanes each), in ``generate_c``'s order
(comb, tick_next captures, commits) inside a loop over lane groups and cycles;
the C is compiled once with ``gcc -O3 -march=native`` (the reference runs
``clang -O3``, absent here) and called through ctypes; groups are split over
threads with the cycle loop inside each group, as ``tick_many_parallel``
(:513-525) gives each core a range of groups for all cycles.

Differences from the reference's emitted C, each one in the baseline's favour:
* only the live cone is emitted: gates that reach a register capture or an
  observed signal (the same cone the GPU tile compiler keeps), so both
  engines evaluate the same gates;
* a group can be W u64 words (gcc vector extensions: W = 4 or 8 use AVX2 /
  AVX-512 registers), not only the reference's single word;
* the run loop drives the bench's stimulus on the fly (the counter hash of
  oracle/stim.h) and folds observed signals into the same per-lane-word
  signature as the GPU and the interpreted engines, so results can be compared.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess
from typing import Dict, List, Sequence, Tuple

import numpy as np

from rrtl_amd.program import (EF_CAPTURE, EF_STORE, KIND_INPUT, KIND_REG, OP_AND, OP_LIT, OP_MUX, OP_NOT, OP_OR,
                              OP_SIGNAL, OP_XOR, POL_LOW, RESET_ASYNC, Program)

HERE = os.path.dirname(os.path.abspath(__file__))
SANITIZE = os.environ.get("RRTL_ORACLE_SANITIZE") == "1"   # compile with ASan + UBSan (sanitizer test)
BUILD = os.path.join(HERE, "_build", "san") if SANITIZE else os.path.join(HERE, "_build")
IN_RANDOM, IN_CONST, IN_PULSE = 1, 2, 3
_GATE_OPS = (OP_NOT, OP_AND, OP_OR, OP_XOR, OP_MUX)


def _block_gates(block) -> List[Tuple[object, List]]:
    """(packet instrs, effects) per op-packet group: values are local to it."""
    out, cur = [], []
    for k, pk in enumerate(block.packets):
        cur.extend(pk.instrs)
        if pk.effects or k + 1 == len(block.packets):
            out.append((cur, pk.effects))
            cur = []
    return out


class _Gen:
    def __init__(self, p: Program, W: int, inputs: Sequence[Tuple], observe: Sequence[int]):
        if any(b.packets for b in (p.blocks[0], p.blocks[3])):
            raise ValueError("E_BITPARALLEL_AOT: async resets and memories are outside the AOT engine's scope")
        for b in p.blocks:
            for pk in b.packets:
                for i in pk.instrs:
                    if i.width != 1 or i.op not in (OP_LIT, OP_SIGNAL) + _GATE_OPS:
                        raise ValueError("E_BITPARALLEL_AOT: not a 1-bit gate-level program")
        self.p, self.W = p, W
        self.inputs, self.observe = list(inputs), list(observe)
        self.slots: Dict[int, int] = {}
        for i, s in enumerate(p.signals):
            if s.kind in (KIND_INPUT, KIND_REG):
                self.slots[i] = len(self.slots)
        self.gates = 0

    def s(self, sig: int) -> str:
        return f"s[{self.slots[sig]}]"

    def source(self) -> str:
        p, W = self.p, self.W
        comb, nxt = _block_gates(p.blocks[1]), _block_gates(p.blocks[2])
        # ---- liveness: from captures and observed signals backwards
        need_sig = set(self.observe)
        for instrs, effs in nxt:
            for e in effs:
                if e.kind == EF_CAPTURE and e.reset is not None:
                    need_sig.add(e.reset.signal)
        live_n: List[set] = []
        for instrs, effs in nxt:
            live = {e.value for e in effs if e.kind == EF_CAPTURE}
            for ins in reversed(instrs):
                if ins.dst in live:
                    if ins.op == OP_SIGNAL:
                        need_sig.add(ins.args[0])
                    elif ins.op != OP_LIT:
                        live.update(ins.args)
            live_n.append(live)
        live_c: List[set] = [set() for _ in comb]
        for gi in range(len(comb) - 1, -1, -1):
            instrs, effs = comb[gi]
            live = live_c[gi]
            for e in effs:
                if e.kind == EF_STORE and e.dst in need_sig:
                    live.add(e.value)
            for ins in reversed(instrs):
                if ins.dst in live:
                    if ins.op == OP_SIGNAL:
                        need_sig.add(ins.args[0])
                    elif ins.op != OP_LIT:
                        live.update(ins.args)
        # ---- emission
        L: List[str] = []
        cur: Dict[int, str] = {}  # signal -> C expression of its settled value

        def sigval(sig: int) -> str:
            if sig in cur:
                return cur[sig]
            if sig in self.slots:
                return self.s(sig)
            return "Z"  # a wire no live packet stored

        def emit(groups, lives, prefix):
            for gi, (instrs, effs) in enumerate(groups):
                live = lives[gi]
                name = {}
                for ins in instrs:
                    if ins.dst not in live:
                        continue
                    v = f"{prefix}{gi}_{ins.dst}"
                    name[ins.dst] = v
                    a = [name.get(x) for x in ins.args] if ins.op in _GATE_OPS else []
                    if ins.op == OP_LIT:
                        rhs = "O" if ins.args and ins.args[0] & 1 else "Z"
                    elif ins.op == OP_SIGNAL:
                        rhs = sigval(ins.args[0])
                    elif ins.op == OP_NOT:
                        rhs = f"~{a[0]}"
                    elif ins.op == OP_AND:
                        rhs = f"{a[0]} & {a[1]}"
                    elif ins.op == OP_OR:
                        rhs = f"{a[0]} | {a[1]}"
                    elif ins.op == OP_XOR:
                        rhs = f"{a[0]} ^ {a[1]}"
                    else:  # mux(cond, then, else)
                        rhs = f"({a[0]} & {a[1]}) | (~{a[0]} & {a[2]})"
                    if ins.op in _GATE_OPS:
                        self.gates += 1
                    L.append(f"   const vw {v} = {rhs};")
                yield gi, effs, name

        for gi, effs, name in emit(comb, live_c, "c"):
            for e in effs:
                if e.kind == EF_STORE and e.dst in need_sig and e.value in name:
                    cur[e.dst] = name[e.value]
        obs = [sigval(o) for o in self.observe]
        caps = []
        for gi, effs, name in emit(nxt, live_n, "n"):
            for e in effs:
                if e.kind != EF_CAPTURE:
                    continue
                v = name[e.value]
                if e.reset is not None:
                    a = sigval(e.reset.signal)
                    if e.reset.polarity == POL_LOW:
                        a = f"(~{a})"
                    rv = "O" if e.reset.value and e.reset.value[0] & 1 else "Z"
                    v = f"(({a} & {rv}) | (~{a} & {v}))"
                caps.append((e.dst, f"x{len(caps)}"))
                L.append(f"   const vw x{len(caps) - 1} = {v};")
        body = "\n".join(L)
        commit = "\n".join(f"   {self.s(d)} = {x};" for d, x in caps)
        ns = len(self.slots)
        nobs = max(1, len(obs))
        obs_store = "\n".join(f"    ob[{j}] = {o};" for j, o in enumerate(obs))
        # stimulus: random inputs from precomputed per-cycle keys, constants and pulses
        stim, nrand = [], 0
        for (sig, mode, arg) in self.inputs:
            slot = self.slots[sig]
            if mode == IN_RANDOM:
                stim.append(f"    {{ uint64_t* q = (uint64_t*)&s[{slot}]; const uint64_t k = kc[{nrand}];\n"
                            f"      for (int w = 0; w < W; w++) q[w] = stim_word_from_key(k, (uint64_t)(g * W + w)); }}")
                nrand += 1
            elif mode == IN_CONST:
                stim.append(f"    s[{slot}] = {'O' if arg & 1 else 'Z'};")
            elif mode == IN_PULSE:
                n, a, b = arg & 0xFFFFFFFF, (arg >> 32) & 1, (arg >> 33) & 1
                stim.append(f"    s[{slot}] = (c < {n}) ? {'O' if a else 'Z'} : {'O' if b else 'Z'};")
            else:
                raise ValueError("E_BITPARALLEL_AOT: input modes are random, const and pulse")
        self.nrand = nrand
        rand_sigs = [sig for (sig, mode, _a) in self.inputs if mode == IN_RANDOM]
        vt = "typedef uint64_t vw;" if W == 1 else f"typedef uint64_t vw __attribute__((vector_size({8 * W})));"
        return f"""/* generated by oracle/bp_aot.py: {p.top}, {self.gates} gates, W = {W} */
#include <stdint.h>
#include <stdlib.h>
#include <omp.h>
#include "stim.h"
#define W {W}
#define NS {ns}
#define NOBS {nobs}
#define NRAND {nrand}
{vt}
static const uint32_t rand_sig[NRAND + 1] = {{{", ".join(str(x) for x in rand_sigs) or "0"}}};

static void cycle_group(vw* restrict s, long g, long c, const uint64_t* kc, uint64_t* dg, long words) {{
    const vw Z = {{0}};
    const vw O = ~Z;
    (void)c; (void)kc; (void)O;
{chr(10).join(stim)}
{body}
    if (dg) {{
      vw ob[NOBS];
{obs_store}
      for (int w = 0; w < W; w++) {{
        const long gw = g * W + w;
        if (gw >= words) break;
        uint64_t v[NOBS];
        for (int j = 0; j < NOBS; j++) v[j] = ((const uint64_t*)&ob[j])[w];
        sig_digest_cycle(&dg[gw], (uint64_t)words, (uint64_t)c, v, {len(obs)}, 1);
      }}
    }}
{commit}
}}

/* st: [groups][NS] vectors; dg: [64][words] signature or NULL.  Runs cycles
   0 .. cycles - 1 of the stimulus (every run starts the stimulus and the
   signature at cycle 0, as the interpreted engines' runs do). */
void aot_run(uint64_t* st, long groups, long words, uint64_t seed, long cycles, uint64_t* dg, int threads) {{
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(cycles * NRAND + 1));
    for (long c = 0; c < cycles; c++)
        for (int i = 0; i < NRAND; i++)
            keys[c * NRAND + i] = stim_stream_key(seed, (uint64_t)rand_sig[i] << 8, (uint64_t)c);
    #pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (long g = 0; g < groups; g++) {{
        vw* s = (vw*)(st + (size_t)g * NS * W);
        for (long c = 0; c < cycles; c++) cycle_group(s, g, c, keys + c * NRAND, dg, words);
    }}
    free(keys);
}}
"""


class BitParallelAot:
    """The design compiled to straight-line C for a fixed stimulus and
    observation list; ``lanes`` lanes in groups of ``group_words`` u64 words."""

    def __init__(self, program: Program, lanes: int, inputs: Sequence[Tuple], observe: Sequence[int],
                 group_words: int = 1, threads: int = 1):
        gen = _Gen(program, group_words, inputs, observe)
        src = gen.source()
        self.gates = gen.gates
        self.p, self.W, self.threads = program, group_words, max(1, threads)
        self.words = (lanes + 63) // 64
        self.groups = (self.words + group_words - 1) // group_words
        self.slots = gen.slots
        self.observe = list(observe)
        os.makedirs(BUILD, exist_ok=True)
        tag = hashlib.sha256(src.encode()).hexdigest()[:16]
        so = os.path.join(BUILD, f"aot_{tag}.so")
        if not os.path.exists(so):
            c = os.path.join(BUILD, f"aot_{tag}.c")
            with open(c, "w") as f:
                f.write(src)
            tmp = so + f".tmp{os.getpid()}"
            san = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                   "-fno-sanitize-recover=undefined"] if SANITIZE else ["-O3"]
            r = subprocess.run(["gcc"] + san + ["-march=native", "-fopenmp", "-fPIC", "-shared", "-std=gnu11",
                                                "-I", HERE, "-o", tmp, c], capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError("E_BITPARALLEL_AOT: gcc failed: " + r.stderr[-2000:])
            os.replace(tmp, so)
        self.lib = ctypes.CDLL(so)
        self.lib.aot_run.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_long, ctypes.c_uint64, ctypes.c_long,
                                     ctypes.c_void_p, ctypes.c_int]
        # state vectors are aligned to 64 bytes (the vector loads assume their type's alignment)
        n = self.groups * len(self.slots) * group_words
        raw = np.zeros(n + 8, dtype=np.uint64)
        off = (-(raw.ctypes.data // 8)) % 8
        self.st = raw[off:off + n]
    def run(self, seed: int, cycles: int, digest: bool = True):
        """One run of ``cycles`` cycles continuing from the current state (the
        stimulus and signature start at cycle 0, as the interpreted engines'
        runs do); returns (signature [64][words] or None, seconds)."""
        import time
        dg = np.zeros((64, self.words), dtype=np.uint64) if digest else None
        t0 = time.perf_counter()
        self.lib.aot_run(self.st.ctypes.data, self.groups, self.words, seed, cycles,
                         dg.ctypes.data if dg is not None else None, self.threads)
        return dg, time.perf_counter() - t0

    def get_words(self, sig: int) -> np.ndarray:
        """The u64 lane words of an input or register."""
        k = self.slots[sig]
        v = self.st.reshape(self.groups, len(self.slots), self.W)[:, k, :].reshape(-1)
        return v[: self.words].copy()
