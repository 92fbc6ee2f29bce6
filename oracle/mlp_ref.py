"""ORACLE (test infrastructure only): CPU restatement of the triage surrogate's
forward and backward passes (include/rrtl_mlp.h, rrtl_amd/csrc/mlp.hip).

There is no reference counterpart for this network.  The restatement is
pinned by central finite differences of the loss in float64 (tests/
test_mlp_oracle.py); with ``bf16=True`` it rounds at the same points as the
device recipe (bf16 operands of every product, bf16 activations and
pre-activation gradients, fp32 accumulation), which is what the device
results are compared against.

Parameters use the library's flat layout: for each layer l the weight
[dims[l+1]][dims[l]] (row-major) then the bias [dims[l+1]].
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << np.uint32(16)).view(np.float32)


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """bfloat16 bit patterns (uint16) of float32 values, round to nearest even."""
    return (bf16_round(x).view(np.uint32) >> np.uint32(16)).astype(np.uint16)


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << np.uint32(16)).view(np.float32)


def split_params(flat: np.ndarray, dims: Sequence[int]) -> List[Tuple[np.ndarray, np.ndarray]]:
    out, off = [], 0
    for l in range(len(dims) - 1):
        n, k = dims[l + 1], dims[l]
        W = flat[off: off + n * k].reshape(n, k)
        off += n * k
        b = flat[off: off + n]
        off += n
        out.append((W, b))
    assert off == flat.size
    return out


def init_params(dims: Sequence[int], rng: np.random.Generator) -> np.ndarray:
    parts = []
    for l in range(len(dims) - 1):
        bound = 1.0 / np.sqrt(dims[l])
        parts.append(rng.uniform(-bound, bound, dims[l + 1] * dims[l] + dims[l + 1]))
    return np.concatenate(parts).astype(np.float32)


def forward_backward(flat: np.ndarray, dims: Sequence[int], x: np.ndarray, y: np.ndarray, bf16: bool = True,
                     dtype=np.float32):
    """Mean binary cross-entropy of logit = MLP(x) and its gradient.

    Layers l < L-1: H[l+1] = relu(H[l] W[l]^T + b[l]); output: logit =
    H[L-1] w + b.  Returns (loss, flat gradient, logits).
    """
    P = split_params(np.asarray(flat, dtype=dtype), dims)
    L = len(P)
    r = bf16_round if bf16 else (lambda a: a)
    Wq = [r(W.astype(np.float32)).astype(dtype) if bf16 else W for W, _ in P]
    H = [r(np.asarray(x, dtype=np.float32)).astype(dtype) if bf16 else np.asarray(x, dtype=dtype)]
    for l in range(L - 1):
        z = H[l] @ Wq[l].T + P[l][1]
        H.append(r(np.maximum(z, 0).astype(np.float32)).astype(dtype) if bf16 else np.maximum(z, 0))
    logit = H[-1] @ Wq[-1][0] + P[-1][1][0]
    n = x.shape[0]
    yv = np.asarray(y, dtype=dtype)
    loss = np.mean(np.maximum(logit, 0) - yv * logit + np.log1p(np.exp(-np.abs(logit))))
    p = 1.0 / (1.0 + np.exp(-logit))
    dlogit = (p - yv) / n
    grads = [None] * L
    # output layer
    grads[L - 1] = (H[-1].T @ dlogit, np.array([dlogit.sum()], dtype=dtype))
    dZ = np.outer(dlogit, Wq[-1][0]) * (H[-1] > 0)
    if bf16:
        dZ = r(dZ.astype(np.float32)).astype(dtype)
    for l in range(L - 2, -1, -1):
        grads[l] = (dZ.T @ H[l], dZ.sum(axis=0))
        if l > 0:
            dZ = (dZ @ Wq[l]) * (H[l] > 0)
            if bf16:
                dZ = r(dZ.astype(np.float32)).astype(dtype)
    flat_g = np.concatenate([np.concatenate([g[0].ravel(), g[1].ravel()]) for g in grads])
    return float(loss), flat_g, logit
