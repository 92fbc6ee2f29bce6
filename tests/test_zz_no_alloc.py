"""The allocation-free GEMM entry point (rrtl_mlp_gemm_nt_noalloc) against the
allocating one (rrtl_mlp_gemm_nt): same kernels, same K split, so every
element must be identical.  Two calls in a row, the second smaller, both with
split-K partials in the preallocated region; a product whose partials exceed
the region is rejected before anything is launched.

Device buffers come from the HIP runtime the library itself is linked
against (libamdhip64.so.7, already loaded with it), not from PyTorch: a
PyTorch-ROCm wheel carries its own copy of the HIP runtime, and importing it
after this library has started the GPU puts a second runtime in the process,
whose initialisation then finds no GPU."""
import ctypes

import numpy as np
import pytest

from conftest import flush_print
from oracle import mlp_ref
from rrtl_amd.backend import HipError
from rrtl_amd.surrogate import Surrogate, gemm_nt_noalloc

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree / hipStreamCreate of the library's runtime."""

    def __init__(self):
        Surrogate  # the library (and with it its runtime) is loaded by the import above
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        self.check(self.L.hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        self.bufs.append(p.value)
        return p.value

    def upload(self, arr: np.ndarray) -> int:
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        self.check(self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p: int, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        self.check(self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def stream(self) -> int:
        s = ctypes.c_void_p()
        self.check(self.L.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s.value

    def close(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def test_noalloc_gemm_matches_allocating_entry_point_twice():
    s = Surrogate([128, 128, 1], max_batch=128)
    hip = Hip()
    stream = hip.stream()
    # (512, 512, 4096) splits K 16 ways (16 MiB of partials), (256, 128, 1024) 4 ways
    for M, N, K in ((512, 512, 4096), (256, 128, 1024)):
        flush_print(f"case: noalloc gemm {M}x{N}x{K}")
        rng = np.random.default_rng(M * 7 + N * 3 + K)
        a = mlp_ref.bf16_bits(rng.normal(size=(M, K)).astype(np.float32))
        b = mlp_ref.bf16_bits(rng.normal(size=(N, K)).astype(np.float32))
        want, _ = s.gemm_nt(a, b)
        da, db = hip.upload(a), hip.upload(b)
        dc = hip.upload(np.full((M, N), np.nan, dtype=np.float32))
        gemm_nt_noalloc(da, db, M, N, K, dc, stream=stream)
        hip.check(hip.L.hipStreamSynchronize(stream), "hipStreamSynchronize")
        got = hip.download(dc, (M, N), np.float32)
        assert np.array_equal(got, want), float(np.nanmax(np.abs(got - want)))
    hip.L.hipStreamDestroy(stream)
    hip.close()
    s.close()


def test_noalloc_gemm_rejects_partials_beyond_the_region():
    M, N, K = 3072, 2048, 1024  # 384 tiles -> 2 K slices -> 48 MiB of partials > 32 MiB
    flush_print(f"case: noalloc gemm rejection {M}x{N}x{K}")
    hip = Hip()
    da = hip.upload(np.zeros((M, K), dtype=np.uint16))
    db = hip.upload(np.zeros((N, K), dtype=np.uint16))
    dc = hip.upload(np.zeros((M, N), dtype=np.float32))
    with pytest.raises(HipError, match="E_MLP_NOALLOC"):
        gemm_nt_noalloc(da, db, M, N, K, dc)
    assert float(np.abs(hip.download(dc, (M, N), np.float32)).max()) == 0.0
    hip.close()
