"""The allocation-free GEMM entry point (rrtl_mlp_gemm_nt_noalloc) against the
allocating one (rrtl_mlp_gemm_nt): same kernels, same K split, so every
element must be identical.  Two calls in a row, the second smaller, both with
split-K partials in the preallocated region; a product whose partials exceed
the region is rejected before anything is launched."""
import numpy as np
import pytest

from conftest import flush_print
from oracle import mlp_ref
from rrtl_amd.backend import HipError
from rrtl_amd.surrogate import Surrogate, gemm_nt_noalloc

pytestmark = pytest.mark.gpu


def _device_copy(torch, bits: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(bits).view(np.int16)).to("cuda")


def test_noalloc_gemm_matches_allocating_entry_point_twice():
    import torch
    s = Surrogate([128, 128, 1], max_batch=128)
    stream = torch.cuda.current_stream()
    # (512, 512, 4096) splits K 16 ways (16 MiB of partials), (256, 128, 1024) 4 ways
    for M, N, K in ((512, 512, 4096), (256, 128, 1024)):
        flush_print(f"case: noalloc gemm {M}x{N}x{K}")
        rng = np.random.default_rng(M * 7 + N * 3 + K)
        a = mlp_ref.bf16_bits(rng.normal(size=(M, K)).astype(np.float32))
        b = mlp_ref.bf16_bits(rng.normal(size=(N, K)).astype(np.float32))
        want, _ = s.gemm_nt(a, b)
        da, db = _device_copy(torch, a), _device_copy(torch, b)
        dc = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        gemm_nt_noalloc(da.data_ptr(), db.data_ptr(), M, N, K, dc.data_ptr(), stream=stream.cuda_stream)
        torch.cuda.synchronize()
        got = dc.cpu().numpy()
        assert np.array_equal(got, want), float(np.max(np.abs(got - want)))
    s.close()


def test_noalloc_gemm_rejects_partials_beyond_the_region():
    import torch
    M, N, K = 3072, 2048, 1024  # 384 tiles -> 2 K slices -> 48 MiB of partials > 32 MiB
    flush_print(f"case: noalloc gemm rejection {M}x{N}x{K}")
    da = torch.zeros((M, K), dtype=torch.int16, device="cuda")
    db = torch.zeros((N, K), dtype=torch.int16, device="cuda")
    dc = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(HipError, match="E_MLP_NOALLOC"):
        gemm_nt_noalloc(da.data_ptr(), db.data_ptr(), M, N, K, dc.data_ptr())
    torch.cuda.synchronize()
    assert float(dc.abs().max()) == 0.0
