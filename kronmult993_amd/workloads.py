"""Synthetic batches for `kronmult_batched`, built on the host as flat pools plus per-entry offsets.

A batch is described the way the reference's API sees it (kronmult_gpu/kronmult.cuh:12-32): entry k has
d factor matrices (n x n, column-major, leading dimension `stride`), one input vector and one output
vector of n^d elements; output vectors may be shared between entries. Here the storage is a few flat
pools and the pointer arrays are offsets into them, so the same description can be turned into host
pointers (for the CPU oracle) or device pointers (for the GPU library).

Builders:
  * `c1` .. `c5`: the five workloads the benchmark and the parity tests use (d=3 n=4 CPU-sized; the
    d=6 n=4 ASGarD-like operator with digit-indexed 4x4 blocks of six 64x64 matrices; five mixed (d, n)
    groups; d=4 n=8 with a wide dynamic range; d=3 n=8 with shared factor triples and a cotangent);
  * `reference_case`: the reference test-harness cases (tests/kronmult_test.cpp:15-53) with its
    batch-size model (tests/utils/batch_size.h:8-21) and output alias map (tests/utils/utils_cpu.h:79-83),
    seeded instead of drawn from std::random_device.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np

from . import gen


def pow_int(number: int, power: int) -> int:
    """n^d in integers (the reference's pow_int, kronmult_gpu/kronmult.cu:11-15)."""
    r = 1
    for _ in range(power):
        r *= number
    return r


def flops_per_entry(d: int, n: int) -> int:
    """2*d*n^(d+1) flops for the d contraction steps plus n^d adds into the output."""
    return 2 * d * n ** (d + 1) + n ** d


def bytes_per_entry(d: int, n: int, itemsize: int) -> int:
    """Per-entry contract bytes: d factors, the input, the output read and written, d+2 pointers."""
    return itemsize * (d * n * n + 3 * n ** d) + 8 * (d + 2)


@dataclass
class HostBatch:
    """One batch of entries sharing (d, n, stride), with flat host pools."""

    name: str
    d: int
    n: int
    stride: int
    mat_pool: np.ndarray            # all factor storage
    mat_off: np.ndarray             # int64 [nb*d]: element offset of factor t of entry k at [k*d + t]
    in_pool: np.ndarray
    in_off: np.ndarray              # int64 [nb]
    out_count: int                  # number of distinct output vectors
    out_index: np.ndarray           # int64 [nb]: which output vector entry k accumulates into
    out_init: Optional[np.ndarray] = None   # initial output pool (out_count*size); zeros if None
    grad_out_pool: Optional[np.ndarray] = None  # cotangent per output vector (adjoint workloads)
    meta: dict = field(default_factory=dict)

    @property
    def nb(self) -> int:
        return int(self.in_off.shape[0])

    @property
    def size(self) -> int:
        return pow_int(self.n, self.d)

    @property
    def out_off(self) -> np.ndarray:
        return self.out_index * self.size

    @property
    def dtype(self):
        return self.mat_pool.dtype

    def output_pool_init(self) -> np.ndarray:
        if self.out_init is not None:
            return self.out_init.copy()
        return np.zeros(self.out_count * self.size, dtype=self.dtype)

    def astype(self, dtype) -> "HostBatch":
        conv = lambda a: None if a is None else np.ascontiguousarray(a, dtype=dtype)
        return replace(self, mat_pool=conv(self.mat_pool), in_pool=conv(self.in_pool),
                       out_init=conv(self.out_init), grad_out_pool=conv(self.grad_out_pool))

    def subset(self, count: int) -> "HostBatch":
        """The first `count` entries (pools kept whole)."""
        return replace(self, mat_off=self.mat_off[: count * self.d].copy(), in_off=self.in_off[:count].copy(),
                       out_index=self.out_index[:count].copy())

    def flops(self) -> int:
        return self.nb * flops_per_entry(self.d, self.n)

    def contract_bytes(self, itemsize: Optional[int] = None) -> int:
        s = self.dtype.itemsize if itemsize is None else itemsize
        return self.nb * bytes_per_entry(self.d, self.n, s)


def _per_entry_factors(seed: int, stream: int, nb: int, d: int, n: int, scale_stream: Optional[int] = None):
    """nb*d contiguous n x n factors (stride n), entries U[-1,1] (optionally times 10^u per matrix)."""
    pool = gen.uniform(seed, stream, nb * d * n * n)
    if scale_stream is not None:
        u = gen.uniform(seed, scale_stream, nb * d, -2.0, 2.0)
        pool = (pool.reshape(nb * d, n * n) * (10.0 ** u)[:, None]).reshape(-1)
    off = np.arange(nb * d, dtype=np.int64) * (n * n)
    return pool, off


def c1(seed: int = 0, nb: int = 4096) -> HostBatch:
    """d=3, n=4, stride 4; each entry owns 3 factors U[-1,1] and a 64-long input U[-1,1];
    entry k accumulates into output k/8 (512 outputs for 4096 entries)."""
    d, n = 3, 4
    size = n ** d
    mats, moff = _per_entry_factors(seed, 1, nb, d, n)
    inp = gen.uniform(seed, 2, nb * size)
    return HostBatch("c1_d3n4", d, n, n, mats, moff, inp, np.arange(nb, dtype=np.int64) * size,
                     (nb + 7) // 8, np.arange(nb, dtype=np.int64) // 8)


def c2(seed: int = 0, rows: int = 128, cols: int = 64, order: str = "output") -> HostBatch:
    """ASGarD-like 6D operator apply, d=6, n=4, stride 64.

    For each dimension t a 64x64 column-major coefficient matrix C_t ~ N(0,1)/8. Entry (i, j) (row element
    i < rows, column element j < cols) uses, in dimension t, the 4x4 block of C_t at (4*r, 4*c) where r and
    c are digit t of i and j in base 16 (digit t = floor(id / 16^t) mod 16). Entry (i, j) reads input[j]
    (4096 values U[-1,1]) and accumulates into output[i]. order "output": entry k = i*cols + j (the
    forward's order, entries of one output consecutive); "input": k = j*rows + i (the same entries with
    those of one input consecutive, the adjoint's order). Same data either way.
    """
    if order not in ("output", "input"):
        raise ValueError("order is 'output' or 'input'")
    d, n, ld = 6, 4, 64
    size = n ** d
    coef = np.concatenate([gen.normal(seed, 10 + t, ld * ld) / 8.0 for t in range(d)])
    if order == "output":
        i = np.repeat(np.arange(rows, dtype=np.int64), cols)
        j = np.tile(np.arange(cols, dtype=np.int64), rows)
    else:
        j = np.repeat(np.arange(cols, dtype=np.int64), rows)
        i = np.tile(np.arange(rows, dtype=np.int64), cols)
    moff = np.empty((rows * cols, d), dtype=np.int64)
    for t in range(d):
        r = (i >> (4 * t)) & 15
        c = (j >> (4 * t)) & 15
        moff[:, t] = t * ld * ld + 4 * r + 4 * c * ld
    inp = gen.uniform(seed, 2, cols * size)
    name = "c2_d6n4_asgard" if order == "output" else "c2_d6n4_asgard_input_major"
    return HostBatch(name, d, n, ld, coef, moff.reshape(-1), inp, j * size, rows, i,
                     meta={"rows": rows, "cols": cols, "order": order})


def expand_terms(b: HostBatch, factor_pools) -> HostBatch:
    """A several-term batch in expanded form: entry t*nb + k is term t of entry k of `b` (same input and
    output), its factors at b's offsets in factor_pools[t] (laid out like b.mat_pool). The first nb input
    and output pointers with all nb*terms*d factor pointers are the term-major call's arguments."""
    terms = len(factor_pools)
    pool = np.concatenate([np.asarray(f, dtype=b.dtype) for f in factor_pools])
    moff = np.concatenate([b.mat_off + t * len(b.mat_pool) for t in range(terms)])
    return replace(b, name=f"{b.name}_x{terms}terms", mat_pool=pool, mat_off=moff,
                   in_off=np.tile(b.in_off, terms), out_index=np.tile(b.out_index, terms),
                   meta=dict(b.meta, terms=terms, term_nb=b.nb))


def c2_terms(seed: int = 0, terms: int = 3, rows: int = 128, cols: int = 64) -> HostBatch:
    """The d=6 n=4 operator with `terms` Kronecker terms over the same element pairs: term tau draws its own
    six 64x64 coefficient matrices (term 0 is the one-term workload's)."""
    b = c2(seed, rows, cols)
    pools = [np.concatenate([gen.normal(seed, 10 + t + 16 * tau, 64 * 64) / 8.0 for t in range(6)])
             for tau in range(terms)]
    return expand_terms(b, pools)


C3_GROUPS = ((2, 8, 20000), (3, 6, 6000), (4, 4, 3000), (5, 3, 1500), (6, 2, 800))


def c3(seed: int = 0, groups=C3_GROUPS, scale: float = 1.0) -> List[HostBatch]:
    """Mixed (d, n, count) groups; contiguous factors (stride n) U[-1,1], inputs U[-1,1], four consecutive
    entries per output. `scale` shrinks the counts for quick checks."""
    out = []
    for g, (d, n, count) in enumerate(groups):
        count = max(4, int(count * scale) // 4 * 4)
        size = n ** d
        mats, moff = _per_entry_factors(seed, 100 + 2 * g, count, d, n)
        inp = gen.uniform(seed, 101 + 2 * g, count * size)
        out.append(HostBatch(f"c3_g{g}_d{d}n{n}", d, n, n, mats, moff, inp,
                             np.arange(count, dtype=np.int64) * size, count // 4,
                             np.arange(count, dtype=np.int64) // 4))
    return out


def c4(seed: int = 0, nb: int = 16384) -> HostBatch:
    """d=4, n=8, 4096-long inputs N(0,1) (one per entry); factors U[-1,1] * 10^u, u ~ U[-2,2] per matrix;
    entry k accumulates into output k/16."""
    d, n = 4, 8
    size = n ** d
    mats, moff = _per_entry_factors(seed, 3, nb, d, n, scale_stream=4)
    inp = gen.normal(seed, 5, nb * size)
    return HostBatch("c4_d4n8_range", d, n, n, mats, moff, inp, np.arange(nb, dtype=np.int64) * size,
                     (nb + 15) // 16, np.arange(nb, dtype=np.int64) // 16)


def c5(seed: int = 0, nb: int = 8192, triples: int = 256) -> HostBatch:
    """d=3, n=8; 256 distinct factor triples U[-1,1], entry k uses triple k mod 256; 512-long inputs U[-1,1]
    (one per entry); entry k accumulates into output k/4; a cotangent N(0,1) per output."""
    d, n = 3, 8
    size = n ** d
    mats = gen.uniform(seed, 6, triples * d * n * n)
    k = np.arange(nb, dtype=np.int64)
    moff = ((k % triples)[:, None] * d + np.arange(d, dtype=np.int64)[None, :]) * (n * n)
    inp = gen.uniform(seed, 7, nb * size)
    nout = (nb + 3) // 4
    g = gen.normal(seed, 8, nout * size)
    return HostBatch("c5_d3n8_adjoint", d, n, n, mats, moff.reshape(-1), inp, k * size, nout, k // 4,
                     grad_out_pool=g, meta={"triples": triples})


def compute_batch_size(degree: int, dimension: int, grid_level: int, nb_distinct_outputs: int) -> int:
    """The reference's ASGarD-like batch count (tests/utils/batch_size.h:8-21)."""
    size = pow_int(degree, dimension)
    max_elements = 395000000000
    max_batch = (max_elements - nb_distinct_outputs * size) // (size * (2 + dimension * degree * degree))
    formula = pow_int(2, grid_level) * pow_int(grid_level, min(1, dimension - 1))
    return int(min(max_batch, formula))


def reference_case(degree: int, dimension: int, grid_level: int, nb_distinct_outputs: int = 5,
                   stride: int = 67, seed: int = 0, nb: Optional[int] = None) -> HostBatch:
    """A case of the reference's own test harness (tests/kronmult_test.cpp:15-53): every factor is the
    top-left n x n corner of its own n x 67 column-major array (stride 67), all data N(0,1), outputs start
    from N(0,1) values, and entry i >= nb_distinct_outputs accumulates into output (i*distinct)/nb
    (tests/utils/utils_cpu.h:79-83)."""
    n, d = degree, dimension
    size = pow_int(n, d)
    if nb is None:
        nb = compute_batch_size(degree, dimension, grid_level, nb_distinct_outputs)
    mats = gen.normal(seed, 20, nb * d * n * stride)
    moff = np.arange(nb * d, dtype=np.int64) * (n * stride)
    inp = gen.normal(seed, 21, nb * size)
    distinct = min(nb_distinct_outputs, nb)
    idx = np.arange(nb, dtype=np.int64)
    out_index = np.where(idx < distinct, idx, (idx * distinct) // max(nb, 1))
    out_init = gen.normal(seed, 22, distinct * size)
    return HostBatch(f"ref_deg{degree}_dim{dimension}_lvl{grid_level}", d, n, stride, mats, moff, inp,
                     idx * size, distinct, out_index, out_init=out_init,
                     meta={"grid_level": grid_level, "distinct": distinct})


def random_case(seed: int, d: int, n: int, nb: int, stride: Optional[int] = None,
                distinct: Optional[int] = None, shared_factors: Optional[int] = None,
                dtype=np.float64) -> HostBatch:
    """A small irregular batch for parity tests: padded strides, aliased outputs in a scattered (not
    sorted) order, optionally factor sets shared between entries, non-zero initial outputs and a
    cotangent."""
    size = pow_int(n, d)
    stride = n + (seed % 3) if stride is None else stride
    sets = nb if shared_factors is None else shared_factors
    mats = gen.uniform(seed, 31, sets * d * n * stride)
    k = np.arange(nb, dtype=np.int64)
    moff = (((k % sets) * d)[:, None] + np.arange(d, dtype=np.int64)[None, :]) * (n * stride)
    inp = gen.uniform(seed, 32, nb * size)
    distinct = max(1, nb // 3) if distinct is None else distinct
    perm = np.argsort(gen.raw64(seed, 33, nb))
    out_index = (perm * distinct) // nb
    out_init = gen.uniform(seed, 34, distinct * size)
    g = gen.normal(seed, 35, distinct * size)
    hb = HostBatch(f"rand_s{seed}_d{d}n{n}", d, n, stride, mats, moff.reshape(-1), inp, k * size, distinct,
                   out_index.astype(np.int64), out_init=out_init, grad_out_pool=g)
    return hb.astype(dtype)


def random_terms(seed: int, d: int, n: int, nb: int, terms: int, **kw) -> HostBatch:
    """random_case with `terms` factor sets per entry (expanded form, see expand_terms)."""
    b = random_case(seed, d, n, nb, **kw)
    pools = [b.mat_pool] + [gen.uniform(seed, 31 + 100 * t, len(b.mat_pool)).astype(b.dtype)
                            for t in range(1, terms)]
    return expand_terms(b, pools)


def ref_large(seed: int = 0) -> HostBatch:
    """The reference benchmark's "large" case (tests/kronmult_bench_gpu.cpp: degree 8, dimension 6, level 7):
    896 entries of d=6, n=8 (n^d = 262144), stride 67, 5 distinct outputs."""
    return reference_case(8, 6, 7, seed=seed)


def by_name(name: str, seed: int = 0):
    table = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c5": c5, "ref_large": ref_large, "c2_terms": c2_terms,
             "c2_input": lambda seed: c2(seed=seed, order="input")}
    return table[name](seed=seed)
