"""Python binding of the C ABI (``include/rrtl_hip.h``) via ctypes.

``HipSimulator`` mirrors the reference's backend trait ``SimBackend``
(crates/rrtl-sim-ir/src/lib.rs:5258-5270) and the ``RuntimeEngine`` surface
(crates/rrtl-runtime/src/lib.rs:7785-8100): signal and memory access in u32
limbs, ``eval_combinational``, ``tick``, ``tick_from_evaluated`` (the
reference's ``tick_from_evaluated_no_post_eval``), ``tick_many``,
``snapshot``/``restore`` in the ``PackedSimulatorStorage`` layout, stuck-at
forces, and the device-resident ``run``.  Errors raise ``HipError`` carrying
the library's ``E_*`` diagnostic code, as the reference returns an
``ErrorReport``.

There is no fallback: if ``librrtl_hip.so`` is missing the import fails.
"""
from __future__ import annotations

import ctypes
import json
import os
from typing import Iterable, List, Optional, Sequence, Tuple, Union

import numpy as np

from .program import Program

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "librrtl_hip.so")

RRTL_HIP_AUTO, RRTL_HIP_WORD, RRTL_HIP_GATE = 0, 1, 2
IN_HOLD, IN_RANDOM, IN_CONST, IN_PULSE, IN_TRACE, IN_GROUP_RANDOM, IN_TRACE_UNIFORM = 0, 1, 2, 3, 4, 5, 6
REPLAY_REPLICATED, REPLAY_INDEPENDENT = 0, 1   # ReplayLaneMode
REPLAY_LANE0, REPLAY_ALL_LANES = 0, 1          # ReplayCheckMode (Lane0Fast, AllLanes)
OBSERVE_PRE, OBSERVE_POST = 0, 1
NO_HALT = 0xFFFFFFFF
ENGINE_NAMES = {"auto": RRTL_HIP_AUTO, "word": RRTL_HIP_WORD, "gate": RRTL_HIP_GATE}
# word engine: generated per-design kernel ("auto", "on") or the record interpreter ("off")
CODEGEN_NAMES = {"auto": 0, "on": 1, "off": 2}

# every symbol the header declares
EXPORTS = [
    "rrtl_hip_create", "rrtl_hip_destroy", "rrtl_hip_last_error", "rrtl_hip_create_error", "rrtl_hip_kind",
    "rrtl_hip_lanes", "rrtl_hip_set_signal_limbs", "rrtl_hip_set_signal_limbs_raw", "rrtl_hip_get_signal_limbs",
    "rrtl_hip_get_signal_bits", "rrtl_hip_set_memory_limbs", "rrtl_hip_get_memory_limbs",
    "rrtl_hip_eval_combinational", "rrtl_hip_tick", "rrtl_hip_tick_from_evaluated", "rrtl_hip_tick_many",
    "rrtl_hip_snapshot", "rrtl_hip_restore", "rrtl_hip_set_force", "rrtl_hip_set_forces", "rrtl_hip_clear_forces",
    "rrtl_hip_synchronize", "rrtl_hip_run", "rrtl_hip_last_run_stats", "rrtl_hip_digest_bytes",
    "rrtl_hip_get_digest", "rrtl_hip_trace_bytes", "rrtl_hip_get_trace", "rrtl_hip_get_halt_cycles",
    "rrtl_hip_get_detect",
    "rrtl_hip_gate_tile_info", "rrtl_hip_word_launch_info", "rrtl_hip_batch_create", "rrtl_hip_batch_run", "rrtl_hip_batch_stats",
    "rrtl_hip_batch_describe", "rrtl_hip_batch_last_error", "rrtl_hip_batch_destroy", "rrtl_hip_emit_kernel", "rrtl_hip_emit_split_kernel", "rrtl_hip_compile_kernel",
    "rrtl_hip_word_program_info",
    "rrtl_hip_replay",
]
BATCH_MODES = {"persistent": 0, "per_design": 1, "concurrent": 2, "streams": 3}


class ReplaySpec(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("steps", ctypes.c_uint64), ("lane_mode", ctypes.c_uint32),
                ("check_mode", ctypes.c_uint32), ("n_inputs", ctypes.c_uint32), ("input_signal", ctypes.c_void_p),
                ("input_values", ctypes.c_void_p), ("n_checks", ctypes.c_uint32), ("check_signal", ctypes.c_void_p),
                ("check_index", ctypes.c_void_p), ("expected", ctypes.c_void_p), ("check_active", ctypes.c_void_p),
                ("max_mismatches", ctypes.c_uint32)]


class ReplayMismatch(ctypes.Structure):
    _fields_ = [("step", ctypes.c_uint64), ("check_index", ctypes.c_uint32), ("lane", ctypes.c_uint32),
                ("expected", ctypes.c_uint32 * 4), ("actual", ctypes.c_uint32 * 4)]


class HipError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status
        self.code = message.split(":", 1)[0] if message else ""


class Options(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("engine", ctypes.c_uint32), ("device", ctypes.c_int32),
                ("block_threads", ctypes.c_uint32), ("tile_words", ctypes.c_uint32),
                ("ticks_per_launch", ctypes.c_uint32), ("tile_chunks", ctypes.c_uint32),
                ("stage_gates", ctypes.c_uint32), ("codegen", ctypes.c_uint32), ("wave_lanes", ctypes.c_uint32),
                ("compact_lanes", ctypes.c_uint32), ("lane_waves", ctypes.c_uint32),
                ("interp_macros", ctypes.c_uint32)]


class RunSpec(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("cycles", ctypes.c_uint64),
                ("n_inputs", ctypes.c_uint32), ("input_signal", ctypes.c_void_p), ("input_mode", ctypes.c_void_p),
                ("input_arg", ctypes.c_void_p), ("trace_inputs", ctypes.c_void_p), ("n_observe", ctypes.c_uint32),
                ("observe_signal", ctypes.c_void_p), ("observe_point", ctypes.c_uint32),
                ("record_trace", ctypes.c_uint32), ("halt_signal", ctypes.c_uint32), ("detect", ctypes.c_uint32),
                ("group_key", ctypes.c_void_p)]


class RunStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("launches", ctypes.c_uint64), ("lane_cycles", ctypes.c_uint64)]


_lib = None


def load_library() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"E_HIP_LIBRARY: {LIB_PATH} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t
    L.rrtl_hip_create.argtypes = [ctypes.c_char_p, sz, u32, ctypes.POINTER(Options), ctypes.POINTER(vp)]
    L.rrtl_hip_destroy.argtypes = [vp]
    L.rrtl_hip_last_error.restype = ctypes.c_char_p
    L.rrtl_hip_last_error.argtypes = [vp]
    L.rrtl_hip_create_error.restype = ctypes.c_char_p
    L.rrtl_hip_kind.restype = u32
    L.rrtl_hip_kind.argtypes = [vp]
    L.rrtl_hip_lanes.restype = u32
    L.rrtl_hip_lanes.argtypes = [vp]
    for fn in ("rrtl_hip_set_signal_limbs", "rrtl_hip_set_signal_limbs_raw", "rrtl_hip_get_signal_limbs",
               "rrtl_hip_get_signal_bits", "rrtl_hip_set_memory_limbs", "rrtl_hip_get_memory_limbs"):
        getattr(L, fn).argtypes = [vp, u32, vp]
    for fn in ("rrtl_hip_eval_combinational", "rrtl_hip_tick", "rrtl_hip_tick_from_evaluated",
               "rrtl_hip_clear_forces", "rrtl_hip_synchronize"):
        getattr(L, fn).argtypes = [vp]
    L.rrtl_hip_tick_many.argtypes = [vp, u64]
    L.rrtl_hip_snapshot.argtypes = [vp, vp, vp]
    L.rrtl_hip_restore.argtypes = [vp, vp, vp]
    L.rrtl_hip_set_force.argtypes = [vp, u32, u32, ctypes.c_int]
    L.rrtl_hip_set_forces.argtypes = [vp, vp, vp, vp, sz]
    L.rrtl_hip_get_detect.argtypes = [vp, vp]
    L.rrtl_hip_run.argtypes = [vp, ctypes.POINTER(RunSpec)]
    L.rrtl_hip_last_run_stats.argtypes = [vp, ctypes.POINTER(RunStats)]
    L.rrtl_hip_digest_bytes.restype = sz
    L.rrtl_hip_digest_bytes.argtypes = [vp]
    L.rrtl_hip_trace_bytes.restype = sz
    L.rrtl_hip_trace_bytes.argtypes = [vp]
    L.rrtl_hip_get_digest.argtypes = [vp, vp]
    L.rrtl_hip_get_trace.argtypes = [vp, vp]
    L.rrtl_hip_get_halt_cycles.argtypes = [vp, vp]
    L.rrtl_hip_gate_tile_info.argtypes = [vp, vp]
    L.rrtl_hip_word_launch_info.argtypes = [vp, vp]
    L.rrtl_hip_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(RunSpec), u32, ctypes.POINTER(vp)]
    L.rrtl_hip_batch_run.argtypes = [vp, u32]
    L.rrtl_hip_batch_stats.argtypes = [vp, ctypes.POINTER(RunStats)]
    L.rrtl_hip_batch_describe.argtypes = [vp, ctypes.c_char_p, sz, ctypes.POINTER(sz)]
    L.rrtl_hip_batch_last_error.restype = ctypes.c_char_p
    L.rrtl_hip_batch_last_error.argtypes = [vp]
    L.rrtl_hip_batch_destroy.argtypes = [vp]
    L.rrtl_hip_emit_kernel.argtypes = [ctypes.c_char_p, sz, u32, u32, ctypes.c_char_p, sz, ctypes.POINTER(sz)]
    L.rrtl_hip_emit_split_kernel.argtypes = [ctypes.c_char_p, sz, u32, ctypes.c_char_p, sz, ctypes.POINTER(sz)]
    L.rrtl_hip_word_program_info.argtypes = [ctypes.c_char_p, sz, u32, ctypes.POINTER(u32)]
    L.rrtl_hip_compile_kernel.argtypes = [ctypes.c_char_p, ctypes.POINTER(sz)]
    L.rrtl_hip_replay.argtypes = [vp, ctypes.POINTER(ReplaySpec), ctypes.POINTER(u64), ctypes.POINTER(ReplayMismatch),
                                  ctypes.POINTER(u32)]
    for fn in EXPORTS:
        if fn != "rrtl_hip_destroy" and getattr(L, fn).restype is ctypes.c_int:
            getattr(L, fn).restype = ctypes.c_int
    _lib = L
    return L


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


SignalRef = Union[int, str]


STORAGE_HBM, STORAGE_SCRATCH_LDS, STORAGE_ALL_LDS = 0, 1, 2


def emit_kernel(program: Program, storage: int = STORAGE_ALL_LDS, block_threads: int = 0) -> str:
    """The word engine's generated HIP kernel for ``program`` (the counterpart
    of gpu_codegen::emit_kernel, crates/rrtl-sim-ir/src/gpu_codegen.rs:503);
    with ``block_threads``, small memories staged in LDS as the engine stages
    them for that workgroup size.  Needs no GPU."""
    L = load_library()
    blob = program.to_blob()
    n = ctypes.c_size_t()
    st = L.rrtl_hip_emit_kernel(blob, len(blob), storage, block_threads, None, 0, ctypes.byref(n))
    if st != 0:
        raise HipError(st, L.rrtl_hip_create_error().decode())
    buf = ctypes.create_string_buffer(n.value + 1)
    st = L.rrtl_hip_emit_kernel(blob, len(blob), storage, block_threads, buf, n.value + 1, ctypes.byref(n))
    if st != 0:
        raise HipError(st, L.rrtl_hip_create_error().decode())
    return buf.value.decode()


def emit_split_kernel(program: Program, lane_waves: int) -> str:
    """The generated kernel split between ``lane_waves`` waves per 64 lanes
    (first line: estimated speed-up, barrier phases, exchange rows).  Needs no
    GPU."""
    L = load_library()
    blob = program.to_blob()
    n = ctypes.c_size_t()
    st = L.rrtl_hip_emit_split_kernel(blob, len(blob), lane_waves, None, 0, ctypes.byref(n))
    if st != 0:
        raise HipError(st, L.rrtl_hip_create_error().decode())
    buf = ctypes.create_string_buffer(n.value + 1)
    st = L.rrtl_hip_emit_split_kernel(blob, len(blob), lane_waves, buf, n.value + 1, ctypes.byref(n))
    if st != 0:
        raise HipError(st, L.rrtl_hip_create_error().decode())
    return buf.value.decode()


def word_program_info(program: Program, macros: bool = True) -> dict:
    """Counts of the word engine's interpreter encoding of ``program`` (with
    macro records or with the fusion menu).  Needs no GPU."""
    L = load_library()
    blob = program.to_blob()
    out = (ctypes.c_uint32 * 4)()
    st = L.rrtl_hip_word_program_info(blob, len(blob), 1 if macros else 2, out)
    if st != 0:
        raise HipError(st, L.rrtl_hip_create_error().decode())
    return {"records": out[0], "macro_records": out[1], "macro_ops": out[2], "scratch_slots": out[3]}


def compile_kernel(source: str) -> int:
    """Compiles HIP source for gfx950 with the library's run-time compiler;
    returns the code object's size.  Needs no GPU."""
    L = load_library()
    n = ctypes.c_size_t()
    st = L.rrtl_hip_compile_kernel(source.encode(), ctypes.byref(n))
    if st != 0:
        raise HipError(st, L.rrtl_hip_create_error().decode())
    return n.value


class HipSimulator:
    """A batch of ``lanes`` independent instances of one program on one GPU."""

    def __init__(self, program: Program, lanes: int, engine: str = "auto", device: int = 0,
                 tile_words: int = 0, ticks_per_launch: int = 0, block_threads: int = 0, tile_chunks: int = 0,
                 stage_gates: int = 0, codegen: str = "auto", wave_lanes: int = 0, compact_lanes: bool = False,
                 lane_waves: int = 0, interp_macros: str = "auto"):
        self.L = load_library()
        self.program = program
        blob = program.to_blob()
        opt = Options(ctypes.sizeof(Options), ENGINE_NAMES[engine], device, block_threads, tile_words,
                      ticks_per_launch, tile_chunks, stage_gates, CODEGEN_NAMES[codegen], wave_lanes,
                      int(bool(compact_lanes)), lane_waves, {"auto": 0, "on": 1, "off": 2}[interp_macros])
        h = ctypes.c_void_p()
        st = self.L.rrtl_hip_create(blob, len(blob), lanes, ctypes.byref(opt), ctypes.byref(h))
        if st != 0:
            raise HipError(st, self.L.rrtl_hip_create_error().decode())
        self.h = h
        self.lanes = lanes
        self.words32 = (lanes + 31) // 32

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.rrtl_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ helpers
    def _check(self, st: int) -> None:
        if st != 0:
            raise HipError(st, self.L.rrtl_hip_last_error(self.h).decode())

    def sig(self, s: SignalRef) -> int:
        i = self.program.signal_index(s) if isinstance(s, str) else int(s)
        if not 0 <= i < len(self.program.signals):
            raise HipError(1, f"E_SIM_IR_SIGNAL: signal index {i} is out of range")
        return i

    def mem(self, m: Union[int, str]) -> int:
        i = self.program.memory_index(m) if isinstance(m, str) else int(m)
        if not 0 <= i < len(self.program.memories):
            raise HipError(1, f"E_SIM_IR_MEMORY: memory index {i} is out of range")
        return i

    @property
    def kind(self) -> str:
        return {1: "word", 2: "gate"}[self.L.rrtl_hip_kind(self.h)]

    # ------------------------------------------------------------ signals
    def set_signal_limbs(self, s: SignalRef, limbs: np.ndarray, settle: bool = True) -> None:
        i = self.sig(s)
        n = self.program.signals[i].limbs
        a = np.ascontiguousarray(np.asarray(limbs, dtype=np.uint32).reshape(n, self.lanes))
        fn = self.L.rrtl_hip_set_signal_limbs if settle else self.L.rrtl_hip_set_signal_limbs_raw
        self._check(fn(self.h, i, _ptr(a)))

    def set_signal(self, s: SignalRef, values, settle: bool = True) -> None:
        i = self.sig(s)
        n = self.program.signals[i].limbs
        a = np.zeros((n, self.lanes), dtype=np.uint32)
        if np.isscalar(values):
            for l in range(n):
                a[l, :] = (int(values) >> (32 * l)) & 0xFFFFFFFF
        else:
            vals = list(values)
            for l in range(n):
                a[l, :] = [(int(v) >> (32 * l)) & 0xFFFFFFFF for v in vals]
        self.set_signal_limbs(i, a, settle)

    def get_signal_limbs(self, s: SignalRef) -> np.ndarray:
        i = self.sig(s)
        out = np.zeros((self.program.signals[i].limbs, self.lanes), dtype=np.uint32)
        self._check(self.L.rrtl_hip_get_signal_limbs(self.h, i, _ptr(out)))
        return out

    def get_signal(self, s: SignalRef) -> List[int]:
        a = self.get_signal_limbs(s)
        if a.shape[0] == 1:
            return [int(x) for x in a[0]]
        return [sum(int(a[l, k]) << (32 * l) for l in range(a.shape[0])) for k in range(self.lanes)]

    def get_signal_bits(self, s: SignalRef) -> np.ndarray:
        out = np.zeros(self.words32, dtype=np.uint32)
        self._check(self.L.rrtl_hip_get_signal_bits(self.h, self.sig(s), _ptr(out)))
        return out

    def set_memory_limbs(self, m, values: np.ndarray) -> None:
        i = self.mem(m)
        pm = self.program.memories[i]
        a = np.ascontiguousarray(np.asarray(values, dtype=np.uint32).reshape(self.lanes, pm.depth, pm.limbs))
        self._check(self.L.rrtl_hip_set_memory_limbs(self.h, i, _ptr(a)))

    def get_memory_limbs(self, m) -> np.ndarray:
        i = self.mem(m)
        pm = self.program.memories[i]
        out = np.zeros((self.lanes, pm.depth, pm.limbs), dtype=np.uint32)
        self._check(self.L.rrtl_hip_get_memory_limbs(self.h, i, _ptr(out)))
        return out

    # ------------------------------------------------------------ ticking
    def eval_combinational(self) -> None:
        self._check(self.L.rrtl_hip_eval_combinational(self.h))

    def tick(self) -> None:
        self._check(self.L.rrtl_hip_tick(self.h))

    def tick_from_evaluated(self) -> None:
        self._check(self.L.rrtl_hip_tick_from_evaluated(self.h))

    def tick_many(self, steps: int) -> None:
        self._check(self.L.rrtl_hip_tick_many(self.h, steps))

    def synchronize(self) -> None:
        self._check(self.L.rrtl_hip_synchronize(self.h))

    def snapshot(self) -> Tuple[np.ndarray, np.ndarray]:
        v = np.zeros(self.program.total_signal_words * self.lanes, dtype=np.uint32)
        m = np.zeros(max(1, self.program.total_memory_words_per_lane * self.lanes), dtype=np.uint32)
        self._check(self.L.rrtl_hip_snapshot(self.h, _ptr(v), _ptr(m)))
        return v, m[: self.program.total_memory_words_per_lane * self.lanes]

    def restore(self, values: np.ndarray, memories: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.uint32)
        m = np.ascontiguousarray(memories, dtype=np.uint32)
        if m.size == 0:
            m = np.zeros(1, dtype=np.uint32)
        self._check(self.L.rrtl_hip_restore(self.h, _ptr(v), _ptr(m)))

    def set_force(self, s: SignalRef, lane: int, value: int) -> None:
        self._check(self.L.rrtl_hip_set_force(self.h, self.sig(s), lane, int(value)))

    def set_forces(self, signals, lanes, values) -> None:
        """Many forces at once: arrays of signal index, lane and stuck value."""
        s = np.ascontiguousarray(signals, dtype=np.uint32)
        ln = np.ascontiguousarray(lanes, dtype=np.uint32)
        v = np.ascontiguousarray(values, dtype=np.uint8)
        if not (s.size == ln.size == v.size):
            raise ValueError("set_forces: arrays differ in length")
        self._check(self.L.rrtl_hip_set_forces(self.h, _ptr(s), _ptr(ln), _ptr(v), s.size))

    def clear_forces(self) -> None:
        self._check(self.L.rrtl_hip_clear_forces(self.h))

    # ------------------------------------------------------------ runs
    def run(self, cycles: int, seed: int = 0, inputs: Sequence[Tuple] = (), observe: Sequence[SignalRef] = (),
            observe_post: bool = False, trace: bool = False, halt: Optional[SignalRef] = None,
            trace_inputs: Optional[np.ndarray] = None, detect: bool = False,
            group_key: Optional[np.ndarray] = None) -> None:
        """``inputs``: (signal, mode, arg) triples with modes IN_*; ``observe``: signals
        folded into the digest (and the trace when ``trace``).  ``detect``: flag
        lanes whose observed values leave lane 0 of their 64-lane group (read
        with ``detect_flags``); ``group_key``: IN_GROUP_RANDOM key per group."""
        spec = self.run_spec(cycles, seed, inputs, observe, observe_post, trace, halt, trace_inputs, detect, group_key)
        self._check(self.L.rrtl_hip_run(self.h, ctypes.byref(spec)))

    def run_spec(self, cycles: int, seed: int = 0, inputs: Sequence[Tuple] = (), observe: Sequence[SignalRef] = (),
                 observe_post: bool = False, trace: bool = False, halt: Optional[SignalRef] = None,
                 trace_inputs: Optional[np.ndarray] = None, detect: bool = False,
                 group_key: Optional[np.ndarray] = None) -> RunSpec:
        """Builds the C run spec; its arrays stay alive on this object until the next call."""
        isig = np.array([self.sig(i[0]) for i in inputs] or [0], dtype=np.uint32)
        imode = np.array([i[1] for i in inputs] or [0], dtype=np.uint32)
        iarg = np.array([int(i[2]) if len(i) > 2 else 0 for i in inputs] or [0], dtype=np.uint64)
        osig = np.array([self.sig(o) for o in observe] or [0], dtype=np.uint32)
        tin = None if trace_inputs is None else np.ascontiguousarray(trace_inputs, dtype=np.uint32)
        gk = None
        if group_key is not None:
            gk = np.ascontiguousarray(group_key, dtype=np.uint64)
            if gk.size != (self.lanes + 63) // 64:
                raise ValueError("group_key needs one key per 64-lane group")
        self._keep = (isig, imode, iarg, osig, tin, gk)
        spec = RunSpec(ctypes.sizeof(RunSpec), seed, cycles, len(inputs), isig.ctypes.data, imode.ctypes.data,
                       iarg.ctypes.data, tin.ctypes.data if tin is not None else None, len(observe),
                       osig.ctypes.data, OBSERVE_POST if observe_post else OBSERVE_PRE, int(trace),
                       NO_HALT if halt is None else self.sig(halt), int(bool(detect)),
                       gk.ctypes.data if gk is not None else None)
        return spec

    def stats(self) -> dict:
        s = RunStats()
        self._check(self.L.rrtl_hip_last_run_stats(self.h, ctypes.byref(s)))
        return {"kernel_ms": s.kernel_ms, "launches": s.launches, "lane_cycles": s.lane_cycles}

    def digest(self) -> np.ndarray:
        n = self.L.rrtl_hip_digest_bytes(self.h)
        if self.kind == "gate":
            out = np.zeros((64, self.words32), dtype=np.uint32)
        else:
            out = np.zeros(self.lanes, dtype=np.uint64)
        assert out.nbytes == n, (out.nbytes, n)
        self._check(self.L.rrtl_hip_get_digest(self.h, _ptr(out)))
        return out

    def trace(self) -> np.ndarray:
        n = self.L.rrtl_hip_trace_bytes(self.h)
        out = np.zeros(n // 4, dtype=np.uint32)
        self._check(self.L.rrtl_hip_get_trace(self.h, _ptr(out)))
        return out

    def replay(self, inputs: Sequence[SignalRef], input_values: np.ndarray, checks: Sequence[SignalRef],
               expected: np.ndarray, lane_mode: int = REPLAY_REPLICATED, check_mode: int = REPLAY_LANE0,
               check_active: Optional[np.ndarray] = None, check_index: Optional[Sequence[int]] = None,
               max_mismatches: int = 16) -> dict:
        """Trace replay on the device (``SimBackend::replay_trace``,
        crates/rrtl-sim-ir/src/lib.rs:5270).  ``input_values``: replicated
        [steps][input limbs] or independent [steps][input limbs][lanes] u32;
        ``expected`` likewise over the checked signals' limbs; ``check_active``
        [steps][ceil(checks/32)] u32 bits or None.  Returns the mismatch count
        and the first ``max_mismatches`` as (step, check_index, lane, expected,
        actual) with values as Python ints."""
        isig = np.array([self.sig(x) for x in inputs] or [0], dtype=np.uint32)
        csig = np.array([self.sig(x) for x in checks] or [0], dtype=np.uint32)
        iv = np.ascontiguousarray(input_values, dtype=np.uint32)
        ev = np.ascontiguousarray(expected, dtype=np.uint32)
        steps = int(iv.shape[0]) if len(inputs) else int(ev.shape[0])
        ca = None if check_active is None else np.ascontiguousarray(check_active, dtype=np.uint32)
        ci = None if check_index is None else np.ascontiguousarray(check_index, dtype=np.uint32)
        sp = ReplaySpec(ctypes.sizeof(ReplaySpec), steps, lane_mode, check_mode, len(inputs), isig.ctypes.data,
                        iv.ctypes.data, len(checks), csig.ctypes.data, None if ci is None else ci.ctypes.data,
                        ev.ctypes.data, None if ca is None else ca.ctypes.data, max_mismatches)
        out = (ReplayMismatch * max(1, max_mismatches))()
        count, n = ctypes.c_uint64(), ctypes.c_uint32()
        self._check(self.L.rrtl_hip_replay(self.h, ctypes.byref(sp), ctypes.byref(count), out, ctypes.byref(n)))

        def u128(v):
            return sum(int(v[k]) << (32 * k) for k in range(4))
        mism = [(int(m.step), int(m.check_index), int(m.lane), u128(m.expected), u128(m.actual))
                for m in out[: n.value]]
        return {"mismatch_count": int(count.value), "mismatches": mism}

    def detect_flags(self) -> np.ndarray:
        """Per-lane detection flags of the last detect run (bool array)."""
        out = np.zeros(self.words32, dtype=np.uint32)
        self._check(self.L.rrtl_hip_get_detect(self.h, _ptr(out)))
        bits = np.unpackbits(out.view(np.uint8), bitorder="little")
        return bits[: self.lanes].astype(bool)

    def halt_cycles(self) -> np.ndarray:
        out = np.zeros(self.lanes, dtype=np.uint64)
        self._check(self.L.rrtl_hip_get_halt_cycles(self.h, _ptr(out)))
        return out

    def launch_info(self) -> dict:
        """Word engine: the launch shape it chose."""
        out = np.zeros(4, dtype=np.uint32)
        self._check(self.L.rrtl_hip_word_launch_info(self.h, _ptr(out)))
        return {"generated": bool(out[0]), "lane_waves": int(out[1]), "wave_lanes": int(out[2]),
                "block_threads": int(out[3])}

    def tile_info(self) -> dict:
        out = np.zeros(4, dtype=np.uint32)
        self._check(self.L.rrtl_hip_gate_tile_info(self.h, _ptr(out)))
        return {"tile_words": int(out[0]), "live_slots": int(out[1]), "levels": int(out[2]), "gates": int(out[3])}


class Batch:
    """Several engines (designs) run together (``rrtl_hip_batch_*``).  ``runs``
    pairs each simulator with its ``run`` keyword arguments; results land in
    each simulator as after ``HipSimulator.run``.  Modes: ``persistent`` (one
    graph launch of the shared work-queue kernel), ``concurrent`` (each design
    on its own kernels and stream, the fork/join captured into one graph),
    ``streams`` (the same without the graph) and ``per_design`` (one design
    after another, the comparison)."""

    def __init__(self, runs: Sequence[Tuple["HipSimulator", dict]]):
        self.L = load_library()
        self.sims = [r[0] for r in runs]
        # every spec needs its own arrays alive for the batch's lifetime
        self._specs = []
        for sim, kw in runs:
            holder = HipSimulator.__new__(HipSimulator)
            holder.program, holder.lanes = sim.program, sim.lanes
            spec = HipSimulator.run_spec(holder, **kw)
            self._specs.append((holder, spec))
        arr_h = (ctypes.c_void_p * len(runs))(*[s.h for s in self.sims])
        arr_s = (RunSpec * len(runs))(*[sp for _, sp in self._specs])
        self._arrays = (arr_h, arr_s)
        h = ctypes.c_void_p()
        st = self.L.rrtl_hip_batch_create(arr_h, arr_s, len(runs), ctypes.byref(h))
        if st != 0:
            raise HipError(st, self.L.rrtl_hip_create_error().decode())
        self.h = h

    def run(self, mode: str = "persistent") -> dict:
        st = self.L.rrtl_hip_batch_run(self.h, BATCH_MODES[mode])
        if st != 0:
            raise HipError(st, self.L.rrtl_hip_batch_last_error(self.h).decode())
        s = RunStats()
        self.L.rrtl_hip_batch_stats(self.h, ctypes.byref(s))
        return {"kernel_ms": s.kernel_ms, "launches": s.launches, "lane_cycles": s.lane_cycles}

    def describe(self) -> dict:
        """Shape of the last persistent launch: grid, workgroups per CU and
        how each design's items run inside the generated kernel."""
        n = ctypes.c_size_t()
        self.L.rrtl_hip_batch_describe(self.h, None, 0, ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value + 1)
        self.L.rrtl_hip_batch_describe(self.h, buf, n.value + 1, ctypes.byref(n))
        return json.loads(buf.value.decode())

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.rrtl_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
