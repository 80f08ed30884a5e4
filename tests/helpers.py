"""Shared checks for the test suite."""
from __future__ import annotations

from functools import reduce

import numpy as np

from kronmult993_amd import workloads as W


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


def grouped_by_output(b):
    """The same entries reordered (stably) so that entries sharing an output are consecutive."""
    from kronmult993_amd.workloads import HostBatch
    p = np.argsort(b.out_index, kind="stable")
    moff = b.mat_off.reshape(b.nb, b.d)[p].reshape(-1)
    return HostBatch(**{**b.__dict__, "name": b.name + "_grouped", "mat_off": moff.copy(),
                        "in_off": b.in_off[p].copy(), "out_index": b.out_index[p].copy()})


def ordered_reference(b, epw: int) -> np.ndarray:
    """The summation order of the ordered-accumulation call, from the oracle's per-entry contributions
    (bit-identical to the device's): per output segment, pieces = its parts in ranges of `epw` entries,
    P = ((0 + y_a) + y_a+1) + ...; out + P for a one-piece segment, else S = P_0 + P_1 + ... (left to
    right) and out + S. Arithmetic in the batch's own precision."""
    import oracle
    from kronmult993_amd.workloads import HostBatch
    u = HostBatch(**{**b.__dict__, "out_count": b.nb, "out_index": np.arange(b.nb, dtype=np.int64),
                     "out_init": None})
    y = oracle.kronmult(u).reshape(b.nb, b.size)
    out = b.output_pool_init().reshape(b.out_count, b.size)
    k = 0
    while k < b.nb:
        o = int(b.out_index[k])
        s1 = k
        while s1 < b.nb and int(b.out_index[s1]) == o:
            s1 += 1
        pieces = []
        ps = k
        while ps < s1:
            pe = min(s1, (ps // epw + 1) * epw)
            acc = np.zeros(b.size, dtype=b.dtype)
            for e in range(ps, pe):
                acc = acc + y[e]
            pieces.append(acc)
            ps = pe
        total = pieces[0]
        for p in pieces[1:]:
            total = total + p
        out[o] = out[o] + total
        k = s1
    return out.reshape(-1)


def terms_reference(b, fused: bool) -> np.ndarray:
    """Several-term call on a batch in expanded form (workloads.expand_terms) with every base entry writing
    its own output: from the oracle's per-entry contributions y_{t,k}, out_k + (((0 + y_0k) + y_1k) + ...)
    for the fused kernel (terms summed in registers), ((out_k + y_0k) + y_1k) + ... for one launch per term."""
    import oracle
    from kronmult993_amd.workloads import HostBatch
    terms, nb = int(b.meta["terms"]), int(b.meta["term_nb"])
    assert np.array_equal(np.sort(b.out_index[:nb]), np.arange(nb)), "base entries must have distinct outputs"
    u = HostBatch(**{**b.__dict__, "out_count": b.nb, "out_index": np.arange(b.nb, dtype=np.int64),
                     "out_init": None})
    y = oracle.kronmult(u).reshape(terms, nb, b.size)
    out = b.output_pool_init().reshape(b.out_count, b.size)
    for k in range(nb):
        o = int(b.out_index[k])
        if fused:
            acc = np.zeros(b.size, dtype=b.dtype)
            for t in range(terms):
                acc = acc + y[t, k]
            out[o] = out[o] + acc
        else:
            for t in range(terms):
                out[o] = out[o] + y[t, k]
    return out.reshape(-1)


def shift_odd_inputs(b):
    """Each entry's input copied into its own slot of size+1 elements, odd entries shifted by one element:
    the input pointers alternate between 16-byte and element alignment along every run of entries."""
    size = b.size
    pool = np.zeros(b.nb * (size + 1), dtype=b.dtype)
    in_off = np.arange(b.nb, dtype=np.int64) * (size + 1) + (np.arange(b.nb) % 2)
    for k in range(b.nb):
        pool[in_off[k]: in_off[k] + size] = b.in_pool[b.in_off[k]: b.in_off[k] + size]
    return W.HostBatch(**{**b.__dict__, "in_pool": pool, "in_off": in_off})
