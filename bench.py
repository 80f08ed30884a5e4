#!/usr/bin/env python3
"""Benchmark of the batched Kronecker mat-vec on one MI355X.

A step is one call of the library over the whole batch (one kernel launch). The default workload is the
d=6, n=4 ASGarD-like operator apply in FP64 (8192 entries: 128 output elements x 64 input elements; six
64x64 coefficient matrices whose 4x4 blocks are the factors, stride 64). Inputs are resident in HBM before
the timed region; the outputs accumulate across steps as the operation specifies.

Timing: the K steps are captured once as a HIP graph of K launches on the bench stream and replayed; the
timed region (bracketed by device synchronisation) is one replay, so `ms_per_step` is steady-state
throughput without host launch gaps. A separate eager pass records HIP events around every launch on the
same stream; their mean is the kernel duration used for the roofline.

Printed: one JSON line with GFLOP/s (value), the binding roofline of the kernel (`roofline`: compute --
flops per launch over the kernel duration against the peak of the pipe that does them -- when the flops at
that peak take longer than the measured bytes at 8 TB/s, else HBM -- the HBM bytes per launch measured with
rocprofv3 counters (profiles/traffic.json) over the kernel duration), both rooflines (`roofline_compute`,
`roofline_hbm`), the per-entry contract bytes over the kernel duration (`contract_gbs`, not a fraction of
anything), the split-precision leg (`split`) and the reference's CPU implementation timed on this host.

    python bench.py [--steps K] [--warmup W] [--config c2] [--dtype f64] [--op forward|adjoint|split|ordered|terms]
                    [--intermediates auto|recompute|store] [--cpu-baseline auto|off] [--calibrate-copy]
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# CPUs this process may run on, read before any OpenMP runtime starts: with threads bound to places, the
# runtime pins the main thread to one core when it initialises, and the mask read afterwards holds one CPU
try:
    HOST_CPUS = len(os.sched_getaffinity(0))
except AttributeError:
    HOST_CPUS = os.cpu_count() or 1

# OpenMP placement for the CPU baseline (read when the reference's OpenMP runtime starts): threads pinned to
# cores, packed; the caller's settings win
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3

WORKLOADS = {
    "c1": "d=3 n=4 stride=4, nb=4096 entries with own factors, 512 outputs x 8 contributions",
    "c2": ("ASGarD-like 6D operator apply: d=6 n=4 stride=64, 128 output x 64 input elements (nb=8192), "
           "factors = digit-indexed 4x4 blocks of six 64x64 N(0,1)/8 matrices, 64 shared inputs of 4096, "
           "128 outputs x 64 contributions"),
    "c3": ("mixed (d, n, count) groups (2,8,20000) (3,6,6000) (4,4,3000) (5,3,1500) (6,2,800), contiguous "
           "factors, 4 consecutive entries per output, one grouped call (strategy and shapes autotuned)"),
    "c4": "d=4 n=8 stride=8, nb=16384, factors U[-1,1]*10^U[-2,2], 1024 outputs x 16 contributions",
    "c5": "d=3 n=8 stride=8, nb=8192, 256 shared factor triples, 2048 outputs x 4 contributions",
    "ref_large": ("the reference benchmark's large case: d=6 n=8 (n^d = 262144) stride=67, nb=896, 5 outputs, "
                  "N(0,1) data; multi-pass sweep through the workspace"),
    "c2_terms": ("the d=6 n=4 operator with 3 Kronecker terms per element pair (each term its own six 64x64 "
                 "coefficient matrices), 8192 entries x 3 terms, 128 outputs; with --op terms"),
    "c2_input": ("the d=6 n=4 batch's 8192 entries in input-major order (the 128 entries of one input "
                 "consecutive): the adjoint's order, d(input) summed per run of equal input pointers"),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_threads() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return min(int(env), HOST_CPUS)
    return HOST_CPUS


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(hbs, runs: int):
    """The reference's kronmult_omp (built from its header, its own flags) on all host threads, one call per
    batch (a grouped workload is one call per group, back to back), full batches."""
    import oracle
    threads = cpu_threads()
    kind = "reference" if oracle.have_ref("shipped") else "port"
    # the reference builds with -march=native (tests/CMakeLists.txt:5): on an AVX-512 host its x86-64-v4
    # build is the nearest one made in the build container, else the x86-64-v3 build
    variant = "avx512" if kind == "reference" and oracle.have_ref("avx512") and oracle.host_has_avx512() \
        else "shipped"
    calls, resets = [], []
    for hb in hbs:
        if kind == "reference":
            runner = oracle.RefRunner(hb, variant)
            calls.append(lambda r=runner: r.run(threads))
            resets.append(runner.reset)
        else:  # the restatement, same arithmetic and threading scheme
            out = hb.output_pool_init()
            calls.append(lambda hb=hb, out=out: oracle.kronmult_omp(hb, threads, out_pool=out))
            resets.append(lambda: None)
    times = []
    for r in range(runs + 1):
        for reset in resets:
            reset()
        t0 = time.perf_counter()
        for call in calls:
            call()
        t1 = time.perf_counter()
        if r > 0:
            times.append(t1 - t0)
        log(f"[cpu] run {r} {'(warm-up) ' if r == 0 else ''}{(t1 - t0) * 1e3:.1f} ms on {threads} threads")
    med = statistics.median(times)
    flops = sum(hb.flops() for hb in hbs)
    build = ({"avx512": "reference kronmult_omp, g++ -O3 -march=x86-64-v4 -fopenmp (AVX-512)",
              "shipped": "reference kronmult_omp, g++ -O3 -march=x86-64-v3 -fopenmp (AVX2+FMA)"}[variant]
             if kind == "reference" else "the oracle's OpenMP restatement")
    placement = (f"{build}; {cpu_model()}; "
                 f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')} OMP_PLACES={os.environ.get('OMP_PLACES')} "
                 f"OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS')}, {HOST_CPUS} CPUs in this process's "
                 f"affinity mask at start, {os.cpu_count()} online")
    names = "+".join(hb.name for hb in hbs)
    nb = sum(hb.nb for hb in hbs)
    return {"value": round(flops / med / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": kind,
            "sample": (f"full {names} batch ({nb} entries, {flops / 1e9:.3f} GFLOP), private input "
                       f"copies, 1 warm-up + {runs} runs, median {med * 1e3:.1f} ms, min {min(times) * 1e3:.1f} ms, "
                       f"{placement}")}


def load_traffic(key: str, kernel: str):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (profiles/traffic.json), if
    recorded for this workload and kernel family."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    entry = table.get(key)
    if entry is None or entry.get("kernel") != kernel:
        return None
    return entry.get("bytes_per_launch")


def time_steps(step, stream, dev, steps: int, warmup: int):
    """W untimed warm-up steps; an eager pass with HIP events around every launch on the launch stream (the
    kernel duration); then the timed region: one replay of a graph holding K launches, bracketed by device
    synchronisation. Returns (ms per step of the replay, per-launch eager ms, wall seconds of the replay)."""
    import torch
    for _ in range(warmup):
        step(stream)
    torch.cuda.synchronize(dev)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    for i in range(steps):
        starts[i].record(stream)
        step(stream)
        ends[i].record(stream)
    torch.cuda.synchronize(dev)
    kernel_ms = [a.elapsed_time(b) for a, b in zip(starts, ends)]
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream(device=dev)
    cap.wait_stream(stream)
    with torch.cuda.stream(cap):
        with torch.cuda.graph(graph, stream=cap):
            for _ in range(steps):
                step(cap)
    torch.cuda.synchronize(dev)
    graph.replay()  # warm replay
    torch.cuda.synchronize(dev)
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t_start.record(stream)
    graph.replay()
    t_end.record(stream)
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - wall0
    return t_start.elapsed_time(t_end) / steps, kernel_ms, wall


def split_leg(db, stream, dev, steps: int, warmup: int, fp64_ms_per_step: float, flops: int):
    """The split-precision half of the metric on the same batch: FP64 data through the FP32 matrix cores
    (kronmult_batched_split_f64, error-free fixed-point splitting), timed like the FP64 step, with its error
    against the FP64 kernel's result on one clean call (that kernel is pinned to the oracle by the tests:
    within 1e-13 normwise)."""
    import numpy as np
    import torch
    step = lambda s: db.run_split("fixed3", sync=False, stream=s)
    ms, kernel_ms, _ = time_steps(step, stream, dev, steps, warmup)
    db.reset_output()
    db.run(sync=True)
    ref = db.output() - db.hb.output_pool_init()
    db.reset_output()
    db.run_split("fixed3", sync=True)
    got = db.output() - db.hb.output_pool_init()
    torch.cuda.synchronize(dev)
    err = np.abs(got - ref)
    scale = float(np.max(np.abs(ref)))
    k_avg = statistics.mean(kernel_ms)
    log(f"[bench] split fixed3: {ms * 1e3:.2f} us/step (graph replay), kernel avg {k_avg * 1e3:.2f} us, "
        f"max rel err {float(np.max(err)) / scale:.3e}")
    return {"kernel": "split_fixed3", "ms_per_step": round(ms, 6), "value": round(flops / (ms * 1e-3) / 1e9, 2),
            "unit": "GFLOP/s (FP64-equivalent)", "kernel_us": round(k_avg * 1e3, 3),
            "max_rel_err": float(np.max(err)) / scale,
            "rms_rel_err": float(np.sqrt(np.mean(err ** 2)) / np.sqrt(np.mean(ref ** 2))),
            "error_reference": "this batch through kronmult_batched_f64 (FP64 kernel, within 1e-13 of the oracle)",
            "speedup_vs_fp64": round(fp64_ms_per_step / ms, 4), "target_rel_err": 1e-10}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    ap.add_argument("--op", default="forward", choices=["forward", "adjoint", "split", "ordered", "terms"],
                    help="forward: kronmult_batched; adjoint: kronmult_batched_adjoint (input and factor "
                         "gradients); split: FP64 data through the FP32 matrix cores (error-free splitting); "
                         "ordered: kronmult_batched_ordered (reproducible sums into shared outputs); terms: "
                         "kronmult_batched_terms on a several-term workload (--config c2_terms)")
    ap.add_argument("--intermediates", default="auto", choices=["auto", "recompute", "store"],
                    help="adjoint: how the forward intermediates are obtained")
    ap.add_argument("--cpu-baseline", default="auto", choices=["auto", "off"])
    ap.add_argument("--split", default="on", choices=["on", "off"],
                    help="forward FP64 runs: also time the split-precision path on the same batch")
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after timing, copy a 256 MiB device buffer (reference kernel for byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        log("bench.py measures one GPU: this library runs the batch on a single MI355X and has no "
            "multi-GPU path (the operator would shard by output rows, with no exchange between GPUs).")
        return 2

    import numpy as np
    import torch

    from kronmult993_amd import api, workloads as W
    from kronmult993_amd.device import DeviceBatch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    np_dtype = np.float64 if args.dtype == "f64" else np.float32
    batches = W.by_name(args.config)
    batches = [b.astype(np_dtype) for b in (batches if isinstance(batches, list) else [batches])]
    dbs = [DeviceBatch(hb, device=str(dev), workspace="per_entry" if args.op == "ordered" else "auto")
           for hb in batches]
    stream = torch.cuda.current_stream(dev)
    extra_config = {}
    if args.op != "forward" and len(dbs) != 1:
        log("--op adjoint / split take a single-batch workload")
        return 2
    if len(dbs) == 1:
        hb, db = batches[0], dbs[0]
        kernel = api.kernel_name(hb.d, hb.n, hb.dtype.itemsize, hb.nb)
        step = lambda s: db.run(sync=False, stream=s)
        extra_config = {"d": hb.d, "n": hb.n, "matrix_stride": hb.stride, "nb_batch": hb.nb,
                        "distinct_outputs": hb.out_count, "op": args.op}
        if args.op == "adjoint":
            if hb.grad_out_pool is None:  # a cotangent N(0,1) per output, seeded
                from kronmult993_amd import gen
                hb.grad_out_pool = gen.normal(0, 90, hb.out_count * hb.size).astype(hb.dtype)
            db.adjoint_setup(intermediates=args.intermediates)
            kernel = f"adjoint ({args.intermediates})"
            step = lambda s: db.run_adjoint(sync=False, stream=s)
            extra_config["intermediates"] = args.intermediates
        elif args.op == "split":
            if args.dtype != "f64":
                log("--op split takes FP64 data")
                return 2
            kernel = "split_fixed3"
            step = lambda s: db.run_split("fixed3", sync=False, stream=s)
        elif args.op == "ordered":
            kernel = kernel + " (ordered) + reduce_pieces"
            step = lambda s: db.run_ordered(sync=False, stream=s)
        elif args.op == "terms":
            if "terms" not in hb.meta:
                log("--op terms takes a several-term workload (--config c2_terms)")
                return 2
            base_nb = int(hb.meta["term_nb"])
            fused = api.kernel_name(hb.d, hb.n, hb.dtype.itemsize, base_nb).startswith("reg_sweep")
            kernel = ("reg_terms" if fused else "per-term launches") + f" ({hb.meta['terms']} terms)"
            step = lambda s: db.run_terms(sync=False, stream=s)
            extra_config.update({"terms": int(hb.meta["terms"]), "nb_batch": base_nb})
    else:
        from kronmult993_amd.autotune import Autotuner, grouped_call_from_choice
        choice = Autotuner().tune_grouped(dbs)
        for db in dbs:
            db.reset_output()
        groups, structs, kw = grouped_call_from_choice(dbs, choice)
        kernel = "group_persistent" if choice["strategy"] == "persistent" else "separate launches"
        step = lambda s: api.kronmult_batched_grouped(groups, dtype=args.dtype, stream=s, structs=structs, **kw)
        extra_config = {"groups": [[hb.d, hb.n, hb.nb] for hb in batches], "autotuned": choice}
    log(f"[bench] {'+'.join(hb.name for hb in batches)} {args.dtype}: kernel={kernel} on "
        f"{torch.cuda.get_device_name(dev)}")

    ms_per_step, kernel_ms, wall = time_steps(step, stream, dev, args.steps, args.warmup)
    k_avg = statistics.mean(kernel_ms)
    log(f"[bench] {args.steps} steps: {ms_per_step * 1e3:.2f} us/step (graph replay), kernel avg "
        f"{k_avg * 1e3:.2f} us (eager, min {min(kernel_ms) * 1e3:.2f}), wall {wall * 1e6 / args.steps:.2f} us/step")

    flops = sum(hb.flops() for hb in batches)
    contract = sum(hb.contract_bytes() for hb in batches)
    if args.op == "adjoint":
        # per entry: transposed sweep + factor-gradient contractions + forward chain, 2 flops per term;
        # bytes: input, cotangent, input gradient read and written, factors, factor gradients read and written
        s_ = hb.dtype.itemsize
        per_flop = (6 * hb.d - 2) * hb.n ** (hb.d + 1)
        per_byte = s_ * (4 * hb.size + 3 * hb.d * hb.n * hb.n) + 8 * (2 + 2 * hb.d)
        flops, contract = hb.nb * per_flop, hb.nb * per_byte
    gflops = flops / (ms_per_step * 1e-3) / 1e9
    contract_gbs = contract / (k_avg * 1e-3) / 1e9
    peak_tf = FP64_PEAK_TFLOPS if args.dtype == "f64" else FP32_PEAK_TFLOPS
    achieved_tf = flops / (k_avg * 1e-3) / 1e12
    traffic_key = f"{args.config}_{args.dtype}" + ("" if args.op == "forward" else f"_{args.op}")
    traffic = load_traffic(traffic_key, kernel)
    # the arithmetic runs on the vector FMA units (the adjoint's factor gradients partly on the FP64 matrix
    # cores, same peak on MI355X); the split path's on the FP32 matrix cores
    pipe = "fp32_mfma" if args.op == "split" else ("fp64_valu" if args.dtype == "f64" else "fp32_valu")
    if args.op == "split":
        peak_tf = FP32_PEAK_TFLOPS
    # HBM: the bytes that actually move (rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size)
    # over the kernel duration; the per-entry contract bytes are reported beside it, not as a fraction
    hbm_gbs = traffic / (k_avg * 1e-3) / 1e9 if traffic is not None else None
    # binding roofline: the longer of the measured bytes at 8 TB/s (contract bytes when unmeasured) and the
    # flops at the pipe's peak
    t_hbm = (traffic if traffic is not None else contract) / (HBM_PEAK_GBS * 1e9)
    t_cmp = flops / (peak_tf * 1e12)
    result = {
        "metric": "batched kron-matvec GFLOP/s & %roofline, d=6 n=4 FP64 vs FP32-split at 1e-10",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded counter-based generator, seed 0)",
        "config": dict({"workload": f"{args.config}: {WORKLOADS[args.config]}", "kernel": kernel},
                       **extra_config),
        "roofline": None,  # the binding one of the two below
        "roofline_hbm": {"bound": "hbm", "achieved": None if hbm_gbs is None else round(hbm_gbs, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": None if hbm_gbs is None else round(hbm_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                         "note": "measured HBM bytes per launch over the kernel duration"},
        "roofline_compute": {"bound": pipe, "achieved": round(achieved_tf, 3), "peak": peak_tf,
                             "unit": "TFLOP/s", "frac": round(achieved_tf / peak_tf, 4), "traffic": traffic,
                             "note": ("FP64 vector peak of MI355X (the FP64 matrix peak is the same, 78.6 "
                                      "TFLOP/s); the sweep runs on the vector FMA units" if pipe == "fp64_valu"
                                      else "FP32 peak 157.3 TFLOP/s")},
        "contract_gbs": round(contract_gbs, 1),
        "kernel_us": {"avg": round(k_avg * 1e3, 3), "min": round(min(kernel_ms) * 1e3, 3),
                      "max": round(max(kernel_ms) * 1e3, 3)},
        "contract_bytes_per_launch": contract,
        "flops_per_launch": flops,
    }
    result["roofline"] = result["roofline_compute"] if t_cmp > t_hbm else result["roofline_hbm"]
    if args.op == "forward" and args.split == "on" and args.dtype == "f64" and len(dbs) == 1:
        result["split"] = split_leg(dbs[0], stream, dev, args.steps, args.warmup, ms_per_step, flops)
    if args.cpu_baseline == "auto" and args.op == "forward":
        result["cpu_baseline"] = cpu_baseline(batches, args.cpu_runs)
    if args.calibrate_copy:
        src = torch.ones(256 * 1024 * 1024 // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize(dev)
        dst.copy_(src)
        torch.cuda.synchronize(dev)
        log("[bench] calibration copy: 268435456 bytes read, 268435456 bytes written")
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
