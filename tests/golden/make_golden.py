"""Writes the golden fixtures of this directory: small seeded batches with the outputs of the reference's
own kronmult_omp (compiled from /root/reference by oracle/Makefile into oracle/_ref, one OpenMP thread) in
its three builds: as shipped (its own flags), every term fused, every product rounded.

Run from the repository root after `make -C oracle all`:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from kronmult993_amd import workloads as W  # noqa: E402


def nan_padding(b):
    """Factor storage the algorithm never reads (rows n..stride-1 of each column, and gaps) set to NaN, so a
    fixture also catches reads outside the n x n matrices."""
    used = np.zeros(b.mat_pool.shape[0], dtype=bool)
    for off in np.unique(b.mat_off):
        for c in range(b.n):
            used[off + c * b.stride: off + c * b.stride + b.n] = True
    pool = b.mat_pool.copy()
    pool[~used] = np.nan
    return W.HostBatch(**{**b.__dict__, "mat_pool": pool})


def cases():
    yield "c1_d3n4", W.c1(nb=128)
    yield "c2_d6n4", W.c2(rows=2, cols=2)
    yield "ref_toy", W.reference_case(4, 1, 2)
    yield "ref_small", W.reference_case(4, 2, 4)
    yield "ref_medium64", W.reference_case(6, 3, 6, nb=64)
    for g in W.c3(scale=0.002):
        yield g.name, g
    yield "c4_d4n8", W.c4(nb=4)
    yield "c5_d3n8", W.c5(nb=16, triples=4)
    yield "rand_shared", W.random_case(7, 4, 3, 30, shared_factors=5)


def main():
    for name, b in cases():
        b = nan_padding(b)
        for dt in (np.float64, np.float32):
            bb = b.astype(dt)
            out = {v: oracle.ref_kronmult(bb, nthreads=1, variant=v) for v in ("shipped", "fused", "unfused")}
            path = os.path.join(HERE, f"{name}_{np.dtype(dt).name}.npz")
            np.savez_compressed(path, d=bb.d, n=bb.n, stride=bb.stride, mat_pool=bb.mat_pool, mat_off=bb.mat_off,
                                in_pool=bb.in_pool, in_off=bb.in_off, out_count=bb.out_count,
                                out_index=bb.out_index, out_init=bb.output_pool_init(),
                                out_shipped=out["shipped"], out_fused=out["fused"], out_unfused=out["unfused"])
            print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
