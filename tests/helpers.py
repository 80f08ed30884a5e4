"""Shared checks for the test suite."""
from __future__ import annotations

from functools import reduce

import numpy as np


def normwise(y: np.ndarray, ref: np.ndarray) -> float:
    """max|y - ref| / max|ref| (the normwise relative error used throughout)."""
    scale = float(np.max(np.abs(ref)))
    if scale == 0.0:
        return float(np.max(np.abs(y)))
    return float(np.max(np.abs(y - ref))) / scale


def factors(b, k):
    """The d factor matrices of entry k as dense n x n arrays (column-major strided storage)."""
    out = []
    for t in range(b.d):
        off = int(b.mat_off[k * b.d + t])
        cols = [b.mat_pool[off + c * b.stride: off + c * b.stride + b.n] for c in range(b.n)]
        out.append(np.stack(cols, axis=1))
    return out


def kron_dense(mats):
    return reduce(np.kron, mats)


def numpy_kronmult(b) -> np.ndarray:
    """Brute force: output[k] += (A_0 ⊗ ... ⊗ A_{d-1}) @ input[k] with np.kron (A_0 slowest, as np.kron)."""
    out = b.output_pool_init().astype(np.float64)
    size = b.size
    for k in range(b.nb):
        K = kron_dense([f.astype(np.float64) for f in factors(b, k)])
        x = b.in_pool[int(b.in_off[k]): int(b.in_off[k]) + size].astype(np.float64)
        o = int(b.out_off[k])
        out[o: o + size] += K @ x
    return out


def accumulation_bound(b, ulp: float) -> np.ndarray:
    """Componentwise error bound for the accumulation into shared outputs: reordering the sum of the E_o
    contributions into output o changes it by at most ~E_o * ulp * sum_k |contribution_k|, which is
    bounded here by E_o * ulp * sum_k |K_k| |x_k| via the oracle on |A|, |x|."""
    import oracle
    absb = b.astype(np.float64)
    absb = type(absb)(**{**absb.__dict__, "mat_pool": np.abs(absb.mat_pool), "in_pool": np.abs(absb.in_pool),
                         "out_init": None if absb.out_init is None else np.abs(absb.out_init)})
    mag = oracle.kronmult(absb)
    counts = np.bincount(b.out_index, minlength=b.out_count).astype(np.float64)
    per_elem = np.repeat(counts, b.size)
    return (per_elem + 2.0) * ulp * mag
