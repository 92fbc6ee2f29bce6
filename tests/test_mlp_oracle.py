"""Pins the surrogate's CPU oracle (oracle/mlp_ref.py): analytic gradients
against central finite differences in float64, and its bf16 rounding against
torch's float32 -> bfloat16 conversion."""
import numpy as np
import pytest
import torch

from oracle import mlp_ref


@pytest.mark.parametrize("dims", [[6, 5, 4, 1], [3, 7, 1], [4, 3, 3, 3, 1]])
def test_gradient_matches_finite_differences(dims):
    rng = np.random.default_rng(len(dims))
    flat = mlp_ref.init_params(dims, rng).astype(np.float64) * 3
    x = rng.normal(size=(9, dims[0]))
    y = rng.integers(0, 2, 9).astype(np.float64)
    loss, g, _ = mlp_ref.forward_backward(flat, dims, x, y, bf16=False, dtype=np.float64)
    h = 1e-6
    fd = np.zeros_like(flat)
    for i in range(flat.size):
        e = np.zeros_like(flat)
        e[i] = h
        lp, _, _ = mlp_ref.forward_backward(flat + e, dims, x, y, bf16=False, dtype=np.float64)
        lm, _, _ = mlp_ref.forward_backward(flat - e, dims, x, y, bf16=False, dtype=np.float64)
        fd[i] = (lp - lm) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-8), np.abs(g - fd).max()
    assert np.abs(g).max() > 1e-3


def test_loss_closed_form_at_zero_weights():
    """All-zero weights: logit = 0 for every row, loss = log 2, output-bias
    gradient = mean(0.5 - y)."""
    dims = [4, 3, 1]
    flat = np.zeros(3 * 4 + 3 + 3 + 1)
    x = np.arange(8.0).reshape(2, 4)
    y = np.array([1.0, 0.0])
    loss, g, logit = mlp_ref.forward_backward(flat, dims, x, y, bf16=False, dtype=np.float64)
    assert np.allclose(logit, 0) and abs(loss - np.log(2)) < 1e-12
    assert abs(g[-1] - np.mean(0.5 - y)) < 1e-12


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=10000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 10000),
                        np.array([1.0, 1.00390625, 1.01171875, -2.5, 0.0, 65504.0], dtype=np.float32)])
    # exact ties: halfway patterns with both parities of the kept bit
    ties = (np.arange(1, 2000, dtype=np.uint32) << np.uint32(16)) | np.uint32(0x8000)
    x = np.concatenate([x, ties.view(np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mlp_ref.bf16_round(x), want)
    assert np.array_equal(mlp_ref.from_bf16_bits(mlp_ref.bf16_bits(x)), want)


def test_bf16_recipe_close_to_float64():
    dims = [128, 128, 128, 1]
    rng = np.random.default_rng(5)
    flat = mlp_ref.init_params(dims, rng)
    x = rng.normal(size=(64, 128)).astype(np.float32)
    y = rng.integers(0, 2, 64).astype(np.float32)
    l1, g1, _ = mlp_ref.forward_backward(flat, dims, x, y, bf16=True)
    l2, g2, _ = mlp_ref.forward_backward(flat.astype(np.float64), dims, x.astype(np.float64), y, bf16=False,
                                         dtype=np.float64)
    assert abs(l1 - l2) < 1e-2
    assert np.linalg.norm(g1 - g2) / np.linalg.norm(g2) < 3e-2
