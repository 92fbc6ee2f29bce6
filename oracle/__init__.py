"""ORACLE -- test infrastructure, never part of the product.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import this package, and only as the checker (or, for ``bench.py``, as the
timed CPU baseline).  The product path (``rrtl_amd``) never imports it.

Contents:
  * ``scalar_sim.ScalarSimulator`` -- rrtl_core::Simulator restated (Python).
  * ``PackedRef`` -- rrtl_sim_ir::PackedSimulator restated (C, packed_ref.c).
  * ``BitParallelRef`` -- rrtl_sim_ir::bitparallel::BitParallelSimulator restated
    (C, bitparallel_ref.c, faithful engine).
  * ``BitParallelFast`` -- the group-major host engine timed as the CPU
    baseline (same file), validated against ``BitParallelRef``.

The reference implementation (Rust) cannot be built in this environment, so
there is no ``oracle/_ref``; parity is pinned by closed forms, brute force and
the reference's documented known answers (see tests/test_oracle_pins.py).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# RRTL_ORACLE_SANITIZE=1 loads the engines built with ASan + UBSan (make
# sanitize); the process must preload the ASan runtime.
SANITIZE = os.environ.get("RRTL_ORACLE_SANITIZE") == "1"
BUILD = os.path.join(HERE, "_build", "san") if SANITIZE else os.path.join(HERE, "_build")


def build() -> None:
    """Compile the C oracles with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE] + (["sanitize"] if SANITIZE else []), check=True)


def _lib(name: str) -> ctypes.CDLL:
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        build()
    return ctypes.CDLL(path)


_packed = None
_bp = None


def _packed_lib():
    global _packed
    if _packed is None:
        L = _lib("libpacked_ref.so")
        L.pref_create.restype = ctypes.c_void_p
        L.pref_create.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t]
        L.pref_destroy.argtypes = [ctypes.c_void_p]
        for fn in ("pref_eval_combinational", "pref_tick", "pref_tick_from_evaluated"):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
        L.pref_tick_many.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        for fn in ("pref_set_signal_limbs", "pref_set_memory_limbs"):
            getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int]
        for fn in ("pref_get_signal_limbs", "pref_get_memory_limbs"):
            getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.pref_snapshot.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.pref_restore.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _packed = L
    return _packed


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


class PackedRef:
    """ctypes wrapper of the packed per-lane oracle."""

    def __init__(self, program, lanes: int):
        L = _packed_lib()
        blob = program.to_blob()
        err = ctypes.create_string_buffer(512)
        h = L.pref_create(blob, len(blob), lanes, err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self.L, self.h, self.program, self.lanes = L, h, program, lanes

    def __del__(self):
        if getattr(self, "h", None):
            self.L.pref_destroy(self.h)
            self.h = None

    def set_signal_limbs(self, sig: int, vals: np.ndarray, eval: bool = True) -> None:
        v = np.ascontiguousarray(vals, dtype=np.uint32).reshape(self.program.signals[sig].limbs, self.lanes)
        self.L.pref_set_signal_limbs(self.h, sig, _ptr(v), int(eval))

    def set_signal(self, sig: int, lane_values, eval: bool = True) -> None:
        s = self.program.signals[sig]
        v = np.zeros((s.limbs, self.lanes), dtype=np.uint32)
        lv = np.asarray(lane_values, dtype=object) if not np.isscalar(lane_values) else None
        for l in range(s.limbs):
            if lv is None:
                v[l, :] = (int(lane_values) >> (32 * l)) & 0xFFFFFFFF
            else:
                v[l, :] = [(int(x) >> (32 * l)) & 0xFFFFFFFF for x in lv]
        self.set_signal_limbs(sig, v, eval)

    def get_signal_limbs(self, sig: int) -> np.ndarray:
        out = np.zeros((self.program.signals[sig].limbs, self.lanes), dtype=np.uint32)
        self.L.pref_get_signal_limbs(self.h, sig, _ptr(out))
        return out

    def get_signal(self, sig: int) -> list:
        lm = self.get_signal_limbs(sig)
        return [sum(int(lm[l, i]) << (32 * l) for l in range(lm.shape[0])) for i in range(self.lanes)]

    def set_memory_limbs(self, mem: int, vals: np.ndarray, eval: bool = True) -> None:
        m = self.program.memories[mem]
        v = np.ascontiguousarray(vals, dtype=np.uint32).reshape(self.lanes, m.depth, m.limbs)
        self.L.pref_set_memory_limbs(self.h, mem, _ptr(v), int(eval))

    def get_memory_limbs(self, mem: int) -> np.ndarray:
        m = self.program.memories[mem]
        out = np.zeros((self.lanes, m.depth, m.limbs), dtype=np.uint32)
        self.L.pref_get_memory_limbs(self.h, mem, _ptr(out))
        return out

    def eval_combinational(self): self.L.pref_eval_combinational(self.h)
    def tick(self): self.L.pref_tick(self.h)
    def tick_from_evaluated(self): self.L.pref_tick_from_evaluated(self.h)
    def tick_many(self, n: int): self.L.pref_tick_many(self.h, n)

    def snapshot(self):
        v = np.zeros(self.program.total_signal_words * self.lanes, dtype=np.uint32)
        m = np.zeros(max(1, self.program.total_memory_words_per_lane * self.lanes), dtype=np.uint32)
        self.L.pref_snapshot(self.h, _ptr(v), _ptr(m))
        return v, m[: self.program.total_memory_words_per_lane * self.lanes]

    def restore(self, values: np.ndarray, memories: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.uint32)
        m = np.ascontiguousarray(memories, dtype=np.uint32)
        if m.size == 0:
            m = np.zeros(1, dtype=np.uint32)
        self.L.pref_restore(self.h, _ptr(v), _ptr(m))


# ------------------------------------------------------------ bit-parallel
IN_HOLD, IN_RANDOM, IN_CONST, IN_PULSE, IN_GROUP_RANDOM = 0, 1, 2, 3, 5


class RunSpec(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64), ("cycles", ctypes.c_uint32), ("n_inputs", ctypes.c_uint32),
        ("input_sig", ctypes.c_void_p), ("input_mode", ctypes.c_void_p), ("input_arg", ctypes.c_void_p),
        ("n_obs", ctypes.c_uint32), ("obs_sig", ctypes.c_void_p), ("observe_post", ctypes.c_uint32),
        ("word_base", ctypes.c_uint64), ("digest", ctypes.c_void_p), ("trace", ctypes.c_void_p),
        ("group_key", ctypes.c_void_p),
    ]


def pulse_arg(n: int, first: int, after: int) -> int:
    return (n & 0xFFFFFFFF) | ((first & 1) << 32) | ((after & 1) << 33)


def _bp_lib():
    global _bp
    if _bp is None:
        L = _lib("libbitparallel_ref.so")
        L.bp_create.restype = ctypes.c_void_p
        L.bp_create.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t]
        L.bp_destroy.argtypes = [ctypes.c_void_p]
        L.bp_set_force.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.bp_clear_forces.argtypes = [ctypes.c_void_p]
        L.bp_set_signal_words.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.bp_get_signal_words.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        for fn in ("bp_settle", "bp_tick"):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
        L.bp_tick_many.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.bp_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(RunSpec)]
        L.bpg_create.restype = ctypes.c_void_p
        L.bpg_create.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        L.bpg_destroy.argtypes = [ctypes.c_void_p]
        L.bpg_set_force.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.bpg_set_signal_words.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.bpg_get_signal_words.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.bpg_run.restype = ctypes.c_double
        L.bpg_run.argtypes = [ctypes.c_void_p, ctypes.POINTER(RunSpec)]
        L.bpg_gate_count.restype = ctypes.c_uint32
        L.bpg_gate_count.argtypes = [ctypes.c_void_p]
        _bp = L
    return _bp


class _BPBase:
    words: int

    def _spec(self, *, seed, cycles, inputs, observe, observe_post=False, word_base=0,
              digest=True, trace=False, group_key=None):
        """``inputs``: list of (signal, mode, arg); ``observe``: signal list;
        ``group_key``: IN_GROUP_RANDOM key per u64 word of this engine."""
        self._keep = []
        isig = np.array([i[0] for i in inputs] or [0], dtype=np.uint32)
        imode = np.array([i[1] for i in inputs] or [0], dtype=np.uint32)
        iarg = np.array([i[2] for i in inputs] or [0], dtype=np.uint64)
        osig = np.array(list(observe) or [0], dtype=np.uint32)
        dg = np.zeros((64, self.words), dtype=np.uint64) if digest else None
        tr = np.zeros((cycles, len(observe), self.words), dtype=np.uint64) if trace else None
        gk = None if group_key is None else np.ascontiguousarray(group_key, dtype=np.uint64)
        if gk is not None and gk.size < self.words:
            raise ValueError("group_key needs one key per u64 word")
        self._keep += [isig, imode, iarg, osig, dg, tr, gk]
        spec = RunSpec(seed, cycles, len(inputs), isig.ctypes.data, imode.ctypes.data, iarg.ctypes.data,
                       len(observe), osig.ctypes.data, int(observe_post), word_base,
                       dg.ctypes.data if dg is not None else None, tr.ctypes.data if tr is not None else None,
                       gk.ctypes.data if gk is not None else None)
        return spec, dg, tr


class BitParallelRef(_BPBase):
    """Faithful ``BitParallelSimulator`` restatement."""

    def __init__(self, program, lanes: int):
        L = _bp_lib()
        blob = program.to_blob()
        err = ctypes.create_string_buffer(512)
        h = L.bp_create(blob, len(blob), lanes, err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self.L, self.h, self.program, self.lanes = L, h, program, lanes
        self.words = (lanes + 63) // 64

    def __del__(self):
        if getattr(self, "h", None):
            self.L.bp_destroy(self.h)
            self.h = None

    def set_force(self, sig, lane, value): self.L.bp_set_force(self.h, sig, lane, int(value))
    def clear_forces(self): self.L.bp_clear_forces(self.h)
    def settle(self): self.L.bp_settle(self.h)
    def tick(self): self.L.bp_tick(self.h)
    def tick_many(self, n): self.L.bp_tick_many(self.h, n)

    def set_words(self, sig, words):
        w = np.ascontiguousarray(words, dtype=np.uint64)
        self.L.bp_set_signal_words(self.h, sig, _ptr(w))

    def get_words(self, sig) -> np.ndarray:
        out = np.zeros(self.words, dtype=np.uint64)
        self.L.bp_get_signal_words(self.h, sig, _ptr(out))
        return out

    def set_signal_all(self, sig, bit):
        self.set_words(sig, np.full(self.words, 0xFFFFFFFFFFFFFFFF if bit else 0, dtype=np.uint64))

    def get_bit(self, sig, lane) -> int:
        return int(self.get_words(sig)[lane // 64] >> np.uint64(lane % 64)) & 1

    def run(self, **kw):
        spec, dg, tr = self._spec(**kw)
        self.L.bp_run(self.h, ctypes.byref(spec))
        return dg, tr


class BitParallelFast(_BPBase):
    """Group-major, threaded host engine (the timed CPU baseline)."""

    def __init__(self, program, lanes: int, threads: int = 0):
        L = _bp_lib()
        blob = program.to_blob()
        err = ctypes.create_string_buffer(512)
        h = L.bpg_create(blob, len(blob), lanes, threads, err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self.L, self.h, self.program, self.lanes = L, h, program, lanes
        self.words = (lanes + 63) // 64

    def __del__(self):
        if getattr(self, "h", None):
            self.L.bpg_destroy(self.h)
            self.h = None

    def set_force(self, sig, lane, value): self.L.bpg_set_force(self.h, sig, lane, int(value))

    def set_words(self, sig, words):
        w = np.ascontiguousarray(words, dtype=np.uint64)
        self.L.bpg_set_signal_words(self.h, sig, _ptr(w))

    def get_words(self, sig) -> np.ndarray:
        out = np.zeros(self.words, dtype=np.uint64)
        self.L.bpg_get_signal_words(self.h, sig, _ptr(out))
        return out

    @property
    def gate_count(self) -> int:
        return self.L.bpg_gate_count(self.h)

    def run(self, **kw):
        """Returns (digest, trace, seconds)."""
        spec, dg, tr = self._spec(**kw)
        secs = self.L.bpg_run(self.h, ctypes.byref(spec))
        return dg, tr, secs


# ------------------------------------------------------------ host batch-SIMD
_simd = None


def _simd_lib():
    global _simd
    if _simd is None:
        L = _lib("libsimd_cpu.so")
        L.simd_create.restype = ctypes.c_void_p
        L.simd_create.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32,
                                  ctypes.c_char_p, ctypes.c_size_t]
        L.simd_destroy.argtypes = [ctypes.c_void_p]
        L.simd_run.restype = ctypes.c_double
        L.simd_run.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int,
                               ctypes.c_void_p]
        L.simd_tick_many.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.simd_get_signal_limbs.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.simd_set_signal_limbs.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.simd_set_memory_limbs.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.simd_get_memory_limbs.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        _simd = L
    return _simd


class SimdCpu:
    """Host batch-SIMD engine (``SimdCpuSimulator`` restated), lane-sharded over threads."""

    def __init__(self, program, lanes: int, threads: int = 0, shard_lanes: int = 0):
        L = _simd_lib()
        blob = program.to_blob()
        err = ctypes.create_string_buffer(512)
        h = L.simd_create(blob, len(blob), lanes, threads, shard_lanes, err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self.L, self.h, self.program, self.lanes = L, h, program, lanes

    def __del__(self):
        if getattr(self, "h", None):
            self.L.simd_destroy(self.h)
            self.h = None

    def run(self, seed, cycles, inputs, observe, observe_post=False):
        """Returns (per-lane u64 digest, seconds)."""
        isig = np.array([i[0] for i in inputs] or [0], dtype=np.uint32)
        imode = np.array([i[1] for i in inputs] or [0], dtype=np.uint32)
        iarg = np.array([i[2] for i in inputs] or [0], dtype=np.uint64)
        osig = np.array(list(observe) or [0], dtype=np.uint32)
        dg = np.zeros(self.lanes, dtype=np.uint64)
        secs = self.L.simd_run(self.h, seed, cycles, len(inputs), _ptr(isig), _ptr(imode), _ptr(iarg), len(observe),
                               _ptr(osig), int(observe_post), _ptr(dg))
        return dg, secs

    def tick_many(self, n): self.L.simd_tick_many(self.h, n)

    def get_signal_limbs(self, sig) -> np.ndarray:
        out = np.zeros((self.program.signals[sig].limbs, self.lanes), dtype=np.uint32)
        self.L.simd_get_signal_limbs(self.h, sig, _ptr(out))
        return out

    def set_signal_limbs(self, sig, vals) -> None:
        v = np.ascontiguousarray(vals, dtype=np.uint32).reshape(self.program.signals[sig].limbs, self.lanes)
        self.L.simd_set_signal_limbs(self.h, sig, _ptr(v))

    def set_memory_limbs(self, mem, vals) -> None:
        """vals: [lanes][depth][limbs]; re-settles comb."""
        m = self.program.memories[mem]
        v = np.ascontiguousarray(vals, dtype=np.uint32).reshape(self.lanes, m.depth, m.limbs)
        self.L.simd_set_memory_limbs(self.h, mem, _ptr(v))

    def get_memory_limbs(self, mem) -> np.ndarray:
        m = self.program.memories[mem]
        out = np.zeros((self.lanes, m.depth, m.limbs), dtype=np.uint32)
        self.L.simd_get_memory_limbs(self.h, mem, _ptr(out))
        return out
