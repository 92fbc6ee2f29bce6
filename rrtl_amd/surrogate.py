"""Python side of the triage surrogate (include/rrtl_mlp.h): a ReLU MLP with
one logit, trained and evaluated on the GPU by the library's BF16
matrix-core kernels.  No CPU path: every call goes through the HIP library and
raises when it is missing."""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from .backend import HipError, load_library

MAX_LAYERS = 8


class MlpConfig(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("n_layers", ctypes.c_uint32),
                ("dims", ctypes.c_uint32 * (MAX_LAYERS + 1)), ("max_batch", ctypes.c_uint32),
                ("device", ctypes.c_int32), ("seed", ctypes.c_uint64), ("lr", ctypes.c_float),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
                ("weight_decay", ctypes.c_float)]


class MlpStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("gemm_ms", ctypes.c_double), ("gemm_flop", ctypes.c_double),
                ("steps", ctypes.c_uint64)]


EXPORTS = ["rrtl_mlp_create", "rrtl_mlp_destroy", "rrtl_mlp_last_error", "rrtl_mlp_create_error",
           "rrtl_mlp_param_count", "rrtl_mlp_get_params", "rrtl_mlp_set_params", "rrtl_mlp_forward_backward",
           "rrtl_mlp_get_grads", "rrtl_mlp_forward", "rrtl_mlp_set_tables", "rrtl_mlp_train_pairs",
           "rrtl_mlp_predict_pairs", "rrtl_mlp_last_stats", "rrtl_mlp_gemm_nt", "rrtl_mlp_gemm_nt_noalloc"]

_bound = False


def _lib():
    global _bound
    L = load_library()
    if not _bound:
        vp, u32, u64, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t
        L.rrtl_mlp_create.argtypes = [ctypes.POINTER(MlpConfig), ctypes.POINTER(vp)]
        L.rrtl_mlp_destroy.argtypes = [vp]
        L.rrtl_mlp_last_error.restype = ctypes.c_char_p
        L.rrtl_mlp_last_error.argtypes = [vp]
        L.rrtl_mlp_create_error.restype = ctypes.c_char_p
        L.rrtl_mlp_param_count.restype = sz
        L.rrtl_mlp_param_count.argtypes = [vp]
        L.rrtl_mlp_get_params.argtypes = [vp, vp]
        L.rrtl_mlp_set_params.argtypes = [vp, vp]
        L.rrtl_mlp_forward_backward.argtypes = [vp, vp, vp, u32, ctypes.POINTER(ctypes.c_float)]
        L.rrtl_mlp_get_grads.argtypes = [vp, vp]
        L.rrtl_mlp_forward.argtypes = [vp, vp, u64, vp]
        L.rrtl_mlp_set_tables.argtypes = [vp, vp, u32, u32, vp, u32, u32, u32]
        L.rrtl_mlp_train_pairs.argtypes = [vp, vp, vp, vp, vp, u64, u32, u32, u64, vp]
        L.rrtl_mlp_predict_pairs.argtypes = [vp, vp, vp, vp, u64, vp]
        L.rrtl_mlp_last_stats.argtypes = [vp, ctypes.POINTER(MlpStats)]
        L.rrtl_mlp_gemm_nt.argtypes = [vp, vp, vp, u32, u32, u32, u32, vp, u32, ctypes.POINTER(ctypes.c_double)]
        L.rrtl_mlp_gemm_nt_noalloc.argtypes = [vp, vp, u32, u32, u32, u32, vp, vp]
        _bound = True
    return L


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def gemm_nt_noalloc(a_dev: int, b_dev: int, M: int, N: int, K: int, c_dev: int, tile: int = 0,
                    stream: int = 0) -> None:
    """C = A . B^T on caller-owned device buffers (addresses as integers:
    A [M][K] and B [N][K] bf16 patterns, C [M][N] fp32), enqueued on `stream`
    (a hipStream_t as an integer, 0 = the default stream) without allocating
    device memory; split-K partials use the library's preallocated region."""
    L = _lib()
    st = L.rrtl_mlp_gemm_nt_noalloc(ctypes.c_void_p(a_dev), ctypes.c_void_p(b_dev), M, N, K, tile,
                                    ctypes.c_void_p(c_dev), ctypes.c_void_p(stream))
    if st != 0:
        raise HipError(st, L.rrtl_mlp_create_error().decode())


class Surrogate:
    """``dims``: input width, hidden widths, 1 (e.g. 512, 1024, 1024, 1024, 1)."""

    def __init__(self, dims: Sequence[int], max_batch: int = 8192, seed: int = 1, lr: float = 1e-3,
                 beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, weight_decay: float = 0.0,
                 device: int = 0):
        self.L = _lib()
        self.dims = list(dims)
        c = MlpConfig()
        c.struct_size = ctypes.sizeof(MlpConfig)
        c.n_layers = len(self.dims) - 1
        for i, d in enumerate(self.dims):
            c.dims[i] = d
        c.max_batch, c.device, c.seed = max_batch, device, seed
        c.lr, c.beta1, c.beta2, c.eps, c.weight_decay = lr, beta1, beta2, eps, weight_decay
        h = ctypes.c_void_p()
        st = self.L.rrtl_mlp_create(ctypes.byref(c), ctypes.byref(h))
        if st != 0:
            raise HipError(st, self.L.rrtl_mlp_create_error().decode())
        self.h = h
        self.max_batch = max_batch
        self.n_params = int(self.L.rrtl_mlp_param_count(self.h))

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.rrtl_mlp_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, st: int) -> None:
        if st != 0:
            raise HipError(st, self.L.rrtl_mlp_last_error(self.h).decode())

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        self._check(self.L.rrtl_mlp_get_params(self.h, _ptr(out)))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError("parameter count mismatch")
        self._check(self.L.rrtl_mlp_set_params(self.h, _ptr(p)))

    def grads(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        self._check(self.L.rrtl_mlp_get_grads(self.h, _ptr(out)))
        return out

    def forward_backward(self, x_bits: np.ndarray, y: np.ndarray) -> float:
        """x_bits: [n][dims[0]] uint16 bf16 patterns; y: [n] labels.  Returns the mean loss."""
        x = np.ascontiguousarray(x_bits, dtype=np.uint16)
        yv = np.ascontiguousarray(y, dtype=np.float32)
        loss = ctypes.c_float()
        self._check(self.L.rrtl_mlp_forward_backward(self.h, _ptr(x), _ptr(yv), x.shape[0], ctypes.byref(loss)))
        return float(loss.value)

    def forward(self, x_bits: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x_bits, dtype=np.uint16)
        out = np.zeros(x.shape[0], dtype=np.float32)
        self._check(self.L.rrtl_mlp_forward(self.h, _ptr(x), x.shape[0], _ptr(out)))
        return out

    def set_tables(self, table_a: np.ndarray, table_b: np.ndarray, extra_width: int = 0) -> None:
        """Feature rows are concat(table_a[ia], table_b[ib], extra[row]) (uint16
        bf16 patterns); ``extra_width`` columns come with each row."""
        ta = np.ascontiguousarray(table_a, dtype=np.uint16)
        tb = np.ascontiguousarray(table_b, dtype=np.uint16)
        self.extra_width = extra_width
        self._check(self.L.rrtl_mlp_set_tables(self.h, _ptr(ta), ta.shape[0], ta.shape[1], _ptr(tb), tb.shape[0],
                                               tb.shape[1], extra_width))

    def _extra(self, extra, n):
        if not getattr(self, "extra_width", 0):
            return None, None
        ex = np.ascontiguousarray(extra, dtype=np.uint16)
        if ex.shape != (n, self.extra_width):
            raise ValueError("extra features must be [rows][extra_width]")
        return ex, _ptr(ex)

    def train_pairs(self, ia, ib, y, epochs: int = 1, batch: int = 8192, seed: int = 1, extra=None) -> np.ndarray:
        ia = np.ascontiguousarray(ia, dtype=np.uint32)
        ib = np.ascontiguousarray(ib, dtype=np.uint32)
        yv = np.ascontiguousarray(y, dtype=np.float32)
        ex, exp = self._extra(extra, ia.size)
        losses = np.zeros(epochs, dtype=np.float32)
        self._check(self.L.rrtl_mlp_train_pairs(self.h, _ptr(ia), _ptr(ib), exp, _ptr(yv), ia.size, epochs, batch,
                                                seed, _ptr(losses)))
        return losses

    def predict_pairs(self, ia, ib, extra=None) -> np.ndarray:
        ia = np.ascontiguousarray(ia, dtype=np.uint32)
        ib = np.ascontiguousarray(ib, dtype=np.uint32)
        ex, exp = self._extra(extra, ia.size)
        out = np.zeros(ia.size, dtype=np.float32)
        self._check(self.L.rrtl_mlp_predict_pairs(self.h, _ptr(ia), _ptr(ib), exp, ia.size, _ptr(out)))
        return out

    def gemm_nt(self, a_bits: np.ndarray, b_bits: np.ndarray, iters: int = 1, want_c: bool = True, tile: int = 0):
        """C = A . B^T on the matrix cores (A [M][K], B [N][K] bf16 patterns).
        tile: 0 automatic, 1 = 128 x 128, 2 = 256 x 128, 3 = 256 x 256 workgroup tiles.
        Returns (C or None, mean ms per launch)."""
        a = np.ascontiguousarray(a_bits, dtype=np.uint16)
        b = np.ascontiguousarray(b_bits, dtype=np.uint16)
        M, K = a.shape
        N = b.shape[0]
        if b.shape[1] != K:
            raise ValueError("inner dimensions differ")
        c = np.zeros((M, N), dtype=np.float32) if want_c else None
        ms = ctypes.c_double()
        self._check(self.L.rrtl_mlp_gemm_nt(self.h, _ptr(a), _ptr(b), M, N, K, tile, _ptr(c) if want_c else None,
                                            iters, ctypes.byref(ms)))
        return c, ms.value

    def stats(self) -> dict:
        s = MlpStats()
        self._check(self.L.rrtl_mlp_last_stats(self.h, ctypes.byref(s)))
        return {"kernel_ms": s.kernel_ms, "gemm_ms": s.gemm_ms, "gemm_flop": s.gemm_flop, "steps": s.steps}
