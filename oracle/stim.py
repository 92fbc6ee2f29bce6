"""ORACLE (test infrastructure only): numpy restatement of oracle/stim.h.

Stimulus: bit b of input signal s at cycle c for lane l is bit (l mod 64) of
``stim_word(seed, (s << 8) | b, c, l // 64)`` (SplitMix64 finaliser used as a
counter hash).  Digests: the per-lane 64-bit chain of word-level runs and the
bit-sliced 64-word signature of gate-level runs.
"""
from __future__ import annotations

import numpy as np

U64 = np.uint64
K1 = U64(0x9E3779B97F4A7C15)
K2 = U64(0xD1B54A32D192ED03)


def mix64(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=U64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> U64(30))) * U64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> U64(27))) * U64(0x94D049BB133111EB)
    return z ^ (z >> U64(31))


def stim_key(seed: int, stream: int, cycle: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = np.array([seed], dtype=U64) ^ (np.array([stream + 1], dtype=U64) * K1) ^ (np.array([cycle + 1], dtype=U64) * K2)
    return mix64(x)


def stim_words(seed: int, stream: int, cycle: int, words: np.ndarray) -> np.ndarray:
    key = stim_key(seed, stream, cycle)
    with np.errstate(over="ignore"):
        return mix64(key + (np.asarray(words, dtype=U64) + U64(1)) * K1)


def random_input_lanes(seed: int, sig_index: int, width: int, cycle: int, lanes: int) -> np.ndarray:
    """[limbs, lanes] u32 values of a RANDOM-mode input for one cycle."""
    limbs = max(1, (width + 31) // 32)
    out = np.zeros((limbs, lanes), dtype=np.uint32)
    lane = np.arange(lanes, dtype=U64)
    wordidx = lane >> U64(6)
    for b in range(width):
        w = stim_words(seed, (sig_index << 8) | b, cycle, wordidx)
        bit = ((w >> (lane & U64(63))) & U64(1)).astype(np.uint32)
        out[b // 32] |= bit << np.uint32(b % 32)
    return out


def random_input_words(seed: int, sig_index: int, cycle: int, words64: int, base: int = 0) -> np.ndarray:
    """u64 lane words of a 1-bit RANDOM-mode input for one cycle."""
    return stim_words(seed, sig_index << 8, cycle, np.arange(base, base + words64, dtype=U64))


def lane_digest(d: np.ndarray, limb: np.ndarray, j: int, k: int) -> np.ndarray:
    tag = U64(((j * 64 + k + 1) & 0xFFFFFFFF) << 32)
    return mix64(np.asarray(d, dtype=U64) ^ np.asarray(limb, dtype=U64) ^ tag)


def sig_digest_cycle(S: np.ndarray, cycle: int, v: np.ndarray) -> np.ndarray:
    """S: [64, words] u64 (or u32); v: [n_obs, words] of the same dtype."""
    S = S.copy()
    for j in range(v.shape[0]):
        S[(j + cycle) & 63] ^= v[j]
    p = np.arange(64)
    return S ^ (S[(p + 17) & 63] & ~S[(p + 41) & 63])


def u32_rows_to_u64(a: np.ndarray) -> np.ndarray:
    """[..., 2n] u32 lane words (32 lanes each) -> [..., n] u64 (64 lanes each)."""
    a = np.asarray(a, dtype=np.uint32)
    if a.shape[-1] % 2:
        pad = np.zeros(a.shape[:-1] + (1,), dtype=np.uint32)
        a = np.concatenate([a, pad], axis=-1)
    lo = a[..., 0::2].astype(U64)
    hi = a[..., 1::2].astype(U64)
    return lo | (hi << U64(32))


def group_random_words(seed: int, sig_index: int, cycle: int, keys: np.ndarray) -> np.ndarray:
    """u64 lane words of a 1-bit GROUP_RANDOM input for one cycle: bit 0 of
    ``stim_word(seed, s << 8, c, key)`` for each 64-lane group, copied to all of
    its lanes."""
    w = stim_words(seed, sig_index << 8, cycle, np.asarray(keys, dtype=U64))
    return np.where((w & U64(1)) != 0, U64(0xFFFFFFFFFFFFFFFF), U64(0))


def detect_from_trace(trace: np.ndarray) -> np.ndarray:
    """Fault-detection flags from a bit-sliced observation trace
    [cycles][n_obs][words] u64: lane l of word w is flagged when any observed
    bit differs from lane 0 of the same word at any cycle (the golden-lane
    comparison of crates/rrtl-sim-ir/examples/fault_sim.rs:86-95).  Returns a
    bool array [words * 64]."""
    t = np.asarray(trace, dtype=U64)
    gold = np.where((t & U64(1)) != 0, U64(0xFFFFFFFFFFFFFFFF), U64(0))
    d = np.bitwise_or.reduce((t ^ gold).reshape(-1, t.shape[-1]), axis=0)
    bits = np.unpackbits(d.view(np.uint8), bitorder="little")
    return bits.astype(bool)
