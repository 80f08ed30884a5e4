"""Precision study on the GPU: FP32, and FP64 data through the FP32 matrix cores (kronmult_batched_split_f64),
against the oracle, for every benchmark workload.

Errors are reported as max relative (max|y - ref| / max|ref|, the normwise error used throughout) and RMS
relative (||y - ref||_2 / ||ref||_2). The reference is the oracle's FP64 sweep (kronmult_omp's arithmetic);
small random cases use its extended-precision (long double) sweep, and for the d=4, n=8 wide-dynamic-range
workload the FP64 dense Kronecker oracle is checked on a subset too.

Bars: error-free splitting (fixed3) <= 1e-10 (the target of the split path; it lands near 1e-14);
FP32 data and FP32 / FP32-pair products <= 1e-5; BF16 pairs <= 1e-3; BF16 <= 1e-1.

Run as a script to write the report (markdown) with timings:
    python tests/test_precision_gpu.py --report <path.md>
"""
import functools
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle  # noqa: E402
from kronmult993_amd import api, workloads as W  # noqa: E402
from kronmult993_amd.device import DeviceBatch  # noqa: E402

pytestmark = pytest.mark.gpu

SPLIT_BOUND = {"fixed3": 1e-10, "fp32": 1e-5, "pair32": 1e-5, "pairbf16": 1e-3, "bf16": 1e-1}
FP32_BOUND = 1e-5


def errors(y, ref):
    y = y.astype(np.float64)
    d = y - ref
    return float(np.max(np.abs(d)) / np.max(np.abs(ref))), float(np.linalg.norm(d) / np.linalg.norm(ref))


@functools.lru_cache(maxsize=None)
def _ref(config):
    """(batch, FP64 oracle output) of a single-batch workload, computed once per session."""
    b = W.by_name(config)
    print(f"[oracle] {config} ...", flush=True)
    return b, oracle.kronmult(b)


def _split(b, gpu, mode):
    print(f"[split] {b.name} d={b.d} n={b.n} nb={b.nb} mode={mode}", flush=True)
    db = DeviceBatch(b, device=str(gpu), workspace="none")
    db.run_split(mode)
    return db.output()


def _fp32(b, gpu):
    print(f"[fp32] {b.name} d={b.d} n={b.n} nb={b.nb}", flush=True)
    db = DeviceBatch(b.astype(np.float32), device=str(gpu))
    db.run()
    return db.output()


@pytest.mark.parametrize("d,n", [(1, 3), (2, 8), (3, 5), (3, 8), (4, 4), (5, 3), (5, 4), (6, 2), (6, 4), (4, 7)])
def test_split_fixed3_random(gpu, d, n):
    b = W.random_case(seed=50 + d * n, d=d, n=n, nb=23)
    mx, rms = errors(_split(b, gpu, "fixed3"), oracle.kronmult_hp(b))
    assert mx <= 1e-12, (mx, rms)


@pytest.mark.parametrize("mode", ["fixed3", "fp32", "pair32", "bf16", "pairbf16"])
@pytest.mark.parametrize("config", ["c2", "c4"])
def test_split_modes(gpu, config, mode):
    b, ref = _ref(config)
    mx, rms = errors(_split(b, gpu, mode), ref)
    print(f"[split] {config} {mode}: max {mx:.3e} rms {rms:.3e}")
    assert mx <= SPLIT_BOUND[mode], (mx, rms)


@pytest.mark.parametrize("d", [4, 5, 6])
def test_split_register_sweep_n4(gpu, d):
    """The register-resident split sweep (the n = 4 plans): runs of entries into scattered shared outputs,
    padded strides, factors spanning 10^-3 .. 10^3, an all-zero factor row and an all-zero input vector,
    against the extended-precision oracle."""
    from helpers import factors
    b = W.random_case(seed=70 + d, d=d, n=4, nb=300, distinct=17)
    pool = b.mat_pool.copy()
    scale = 10.0 ** np.linspace(-3, 3, b.nb * d)
    for k in range(b.nb):
        for t in range(d):
            off = int(b.mat_off[k * d + t])
            for c in range(4):
                pool[off + c * b.stride: off + c * b.stride + 4] = factors(b, k)[t][:, c] * scale[k * d + t]
    off = int(b.mat_off[5 * d + 1])
    for c in range(4):
        pool[off + 2 + c * b.stride] = 0.0  # row 2 of factor 1 of entry 5
    inp = b.in_pool.copy()
    inp[int(b.in_off[7]): int(b.in_off[7]) + b.size] = 0.0
    b = W.HostBatch(**{**b.__dict__, "mat_pool": pool, "in_pool": inp})
    assert api.kernel_name(d, 4, 8, b.nb).startswith("reg_sweep")
    mx, rms = errors(_split(b, gpu, "fixed3"), oracle.kronmult_hp(b))
    assert mx <= 1e-12, (mx, rms)


def test_split_register_sweep_misaligned_inputs(gpu):
    """d=6 n=4 with input pointers alternating between 16-byte and element alignment along every run: the
    LDS-DMA staging and its register fallback switch from one entry to the next."""
    from helpers import shift_odd_inputs
    b = shift_odd_inputs(W.c2(rows=3, cols=64))
    mx, rms = errors(_split(b, gpu, "fixed3"), oracle.kronmult(b))
    assert mx <= 1e-12, (mx, rms)


def test_split_c4_dense_subset(gpu):
    """64 entries (four whole outputs) of the wide-dynamic-range workload against the dense FP64 oracle."""
    b = W.c4().subset(64)
    b = W.HostBatch(b.name, b.d, b.n, b.stride, b.mat_pool, b.mat_off, b.in_pool, b.in_off, 4, b.out_index)
    dense = oracle.dense(b)
    mx, _ = errors(dense, oracle.kronmult(b))
    assert mx <= 1e-13  # two FP64 evaluations in different orders (4096-term dense dots vs the sweep)
    mx, _ = errors(_split(b, gpu, "fixed3"), dense)
    assert mx <= 1e-10


@pytest.mark.parametrize("config", ["c1", "c2", "c3", "c4", "c5"])
def test_fp32_every_config(gpu, config):
    if config == "c3":
        pairs = [(b, oracle.kronmult(b)) for b in W.c3()]
    else:
        pairs = [_ref(config)]
    for b, ref in pairs:
        mx, rms = errors(_fp32(b, gpu), ref)
        print(f"[fp32] {b.name}: max {mx:.3e} rms {rms:.3e}")
        assert mx <= FP32_BOUND, (b.name, mx, rms)


def test_split_invalid(gpu):
    b = W.random_case(seed=1, d=2, n=9, nb=2)  # n > 8
    db = DeviceBatch(b, device=str(gpu), workspace="none")
    with pytest.raises(api.KronmultError):
        db.run_split("fixed3")
    b = W.random_case(seed=1, d=5, n=8, nb=2)  # n^d * 8 bytes beyond the on-chip vector
    db = DeviceBatch(b, device=str(gpu), workspace="none")
    with pytest.raises(api.KronmultError):
        db.run_split("fixed3")


# ---- report ----------------------------------------------------------------------------------------------

def _time_us(fn, reps=20):
    import torch
    fn()
    torch.cuda.synchronize()
    a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return a.elapsed_time(e) / reps * 1e3


def report(path):
    import torch
    dev = "cuda:0"
    lines = ["# Precision of the FP32 and split paths against the FP64 oracle", "",
             "Command: `python tests/test_precision_gpu.py --report <path>` (MI355X). Errors: max relative = "
             "max|y - ref| / max|ref|, RMS relative = ||y - ref|| / ||ref|| over all outputs; ref = the oracle's "
             "FP64 sweep (kronmult_omp's arithmetic). Times: mean kernel time per launch over 20 launches (HIP events).", "",
             "| workload | path | max rel | RMS rel | us/launch |", "|---|---|---|---|---|"]
    for config in ("c1", "c2", "c3", "c4", "c5"):
        batches = W.by_name(config)
        for b in (batches if isinstance(batches, list) else [batches]):
            print(f"[report] {b.name}", flush=True)
            ref = oracle.kronmult(b)
            rows = []
            db64 = DeviceBatch(b, device=dev)
            db64.run()
            rows.append(("FP64 (" + api.kernel_name(b.d, b.n, 8, b.nb) + ")", db64.output(),
                         _time_us(lambda: db64.run(sync=False))))
            db32 = DeviceBatch(b.astype(np.float32), device=dev)
            db32.run()
            rows.append(("FP32 data (" + api.kernel_name(b.d, b.n, 4, b.nb) + ")", db32.output(),
                         _time_us(lambda: db32.run(sync=False))))
            if b.n <= 8 and b.size * 8 <= 60000:
                for mode in ("fp32", "pair32", "bf16", "pairbf16", "fixed3"):
                    dbs = DeviceBatch(b, device=dev, workspace="none")
                    dbs.run_split(mode)
                    out = dbs.output()
                    rows.append((f"FP64 data, FP32 matrix cores: {mode}", out,
                                 _time_us(lambda: dbs.run_split(mode, sync=False))))
            for name, out, us in rows:
                mx, rms = errors(out, ref)
                lines.append(f"| {b.name} (d={b.d} n={b.n} nb={b.nb}) | {name} | {mx:.2e} | {rms:.2e} | {us:.1f} |")
                print(lines[-1], flush=True)
            torch.cuda.empty_cache()
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--report", required=True)
    report(ap.parse_args().report)
