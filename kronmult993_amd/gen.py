"""Deterministic counter-based random numbers for the synthetic batches.

The reference seeds `std::default_random_engine` from `std::random_device` (tests/utils/utils_cpu.h:23-24,
tests/utils/utils_gpu.h:55-56), so its test data can never be reproduced. Here every value is a pure
function of (seed, stream, index): SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014) applied to a Weyl sequence, keyed per (seed, stream). Uniforms use the top 53
bits; normals use the Box-Muller transform on two consecutive counters.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_MIX1 = np.uint64(0xBF58476D1CE4E5B9)
_MIX2 = np.uint64(0x94D049BB133111EB)
_TWO_M53 = 1.0 / 9007199254740992.0  # 2^-53


def _mix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser on a uint64 array (wrapping arithmetic)."""
    z = (z ^ (z >> np.uint64(30))) * _MIX1
    z = (z ^ (z >> np.uint64(27))) * _MIX2
    return z ^ (z >> np.uint64(31))


def _key(seed: int, stream: int) -> np.ndarray:
    s = np.array([seed & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
    t = np.array([stream & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix64(_mix64(s * _GOLDEN + t) ^ (t * _MIX2))


def raw64(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    """`count` uint64 words: word i is mix64(key + (offset + i + 1) * golden)."""
    idx = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix64(_key(seed, stream) + idx * _GOLDEN)


def uniform(seed: int, stream: int, count: int, lo: float = -1.0, hi: float = 1.0,
            offset: int = 0) -> np.ndarray:
    """float64 samples of U[lo, hi)."""
    u = (raw64(seed, stream, count, offset) >> np.uint64(11)).astype(np.float64) * _TWO_M53
    return lo + (hi - lo) * u


def normal(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    """float64 samples of N(0, 1) by Box-Muller on counters (2i, 2i+1)."""
    w = raw64(seed, stream, 2 * count, 2 * offset)
    u1 = ((w[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * _TWO_M53  # (0, 1]
    u2 = (w[1::2] >> np.uint64(11)).astype(np.float64) * _TWO_M53          # [0, 1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
