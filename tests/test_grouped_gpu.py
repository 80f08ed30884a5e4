"""GPU parity of the grouped call (kronmult_batched_grouped_f64/_f32): several batches of different (d, n)
in one persistent launch, or one launch per group, through the C ABI.

Per group the bar is the one of test_kronmult_gpu.py: entries that each own an output are bit-identical to
the oracle's fused sweep; aliased outputs stay within the componentwise reordering bound and the normwise
tolerance (1e-13 FP64, 1e-5 FP32). The persistent launch sums consecutive same-output entries of a pass
before its atomic add, which is one more reordering of the same sum, covered by the same bound.
"""
import numpy as np
import pytest

import oracle
from helpers import accumulation_bound, normwise
from kronmult993_amd import api, workloads as W
from kronmult993_amd.device import DeviceBatch

pytestmark = pytest.mark.gpu

EPS = {np.dtype(np.float64): np.finfo(np.float64).eps / 2, np.dtype(np.float32): np.finfo(np.float32).eps / 2}
NORM_TOL = {np.dtype(np.float64): 1e-13, np.dtype(np.float32): 1e-5}


def _check(b, out):
    ref = oracle.kronmult(b)
    bound = accumulation_bound(b, EPS[b.dtype])
    err = np.abs(out.astype(np.float64) - ref.astype(np.float64))
    worst = float(np.max(err - bound))
    assert worst <= 0.0, f"{b.name}: componentwise bound exceeded by {worst:.3e}"
    assert normwise(out, ref) <= NORM_TOL[b.dtype]


def _run(batches, gpu, strategy, wpc=0, epw=None, tunings=None):
    import torch
    print(f"[grouped] strategy={strategy} wpc={wpc} groups="
          + ",".join(f"({b.d},{b.n},{b.nb},{b.dtype})" for b in batches), flush=True)
    dbs = [DeviceBatch(b, device=str(gpu)) for b in batches]
    groups = [db.group(tuning=None if tunings is None else tunings[i],
                       entries_per_tile=0 if epw is None else epw[i]) for i, db in enumerate(dbs)]
    api.kronmult_batched_grouped(groups, dtype=dbs[0].dtype_name, strategy=strategy, workgroups_per_cu=wpc)
    torch.cuda.synchronize()
    return dbs


@pytest.mark.parametrize("strategy", ["persistent", "separate", "auto"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_c3_full(gpu, dtype, strategy):
    batches = [b.astype(dtype) for b in W.c3()]
    for b, db in zip(batches, _run(batches, gpu, strategy)):
        _check(b, db.output())


def test_c3_bitexact_per_entry(gpu):
    """Every entry into its own output: the persistent launch reproduces the oracle's fused sweep bits."""
    batches = []
    for b in W.c3(scale=0.05):
        batches.append(W.HostBatch(b.name + "_own", b.d, b.n, b.stride, b.mat_pool, b.mat_off, b.in_pool,
                                   b.in_off, b.nb, np.arange(b.nb, dtype=np.int64)))
    for b, db in zip(batches, _run(batches, gpu, "persistent")):
        np.testing.assert_array_equal(db.output(), oracle.kronmult(b))


@pytest.mark.parametrize("wpc,epw", [(1, [1, 1, 1, 1, 1]), (4, [64, 21, 16, 9, 8]), (8, [999, 7, 5, 3, 2])])
def test_c3_launch_shapes(gpu, wpc, epw):
    batches = W.c3(scale=0.1)
    for b, db in zip(batches, _run(batches, gpu, "persistent", wpc=wpc, epw=epw)):
        _check(b, db.output())


def test_outputs_shared_across_groups(gpu):
    """(2,8) and (6,2) both have 64-element vectors: both groups accumulate into one output pool."""
    import torch
    a = W.random_case(seed=21, d=2, n=8, nb=50, distinct=6)
    b = W.random_case(seed=22, d=6, n=2, nb=40, distinct=6)
    b = W.HostBatch(b.name, b.d, b.n, b.stride, b.mat_pool, b.mat_off, b.in_pool, b.in_off, b.out_count,
                    b.out_index, out_init=np.zeros_like(a.out_init))
    da, db_ = DeviceBatch(a, device=str(gpu)), DeviceBatch(b, device=str(gpu))
    db_.out_ptrs = da._ptrs(da.out, b.out_off)  # group b writes into a's output pool
    api.kronmult_batched_grouped([da.group(), db_.group()], dtype="f64", strategy="persistent")
    torch.cuda.synchronize()
    ref = oracle.kronmult(b, out_pool=oracle.kronmult(a))
    assert normwise(da.output(), ref) <= 1e-13


def test_mixed_persistent_and_separate(gpu):
    """Groups the persistent launch does not take (n > 8; a vector beyond 64 KiB needing the workspace) run
    as launches of their own inside the same call."""
    batches = [W.random_case(seed=31, d=3, n=4, nb=64), W.random_case(seed=32, d=2, n=12, nb=20),
               W.random_case(seed=33, d=6, n=8, nb=3), W.random_case(seed=34, d=2, n=5, nb=33)]
    for b, db in zip(batches, _run(batches, gpu, "persistent")):
        _check(b, db.output())


def test_graph_replays(gpu):
    """The grouped call captured in a graph: each replay (with eager resets and synchronisation between
    replays) matches the eager result."""
    import torch
    batches = W.c3(scale=0.2)
    dbs = [DeviceBatch(b, device=str(gpu)) for b in batches]
    groups = [db.group() for db in dbs]
    structs = api.grouped_structs(groups, "f64")
    api.kronmult_batched_grouped(groups, dtype="f64", strategy="persistent", structs=structs)
    torch.cuda.synchronize()
    eager = [db.output() for db in dbs]
    for b, e in zip(batches, eager):
        _check(b, e)
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        for db in dbs:
            db.reset_output()
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            api.kronmult_batched_grouped(groups, dtype="f64", strategy="persistent", stream=s, structs=structs)
    for replay in range(3):
        for db in dbs:
            db.reset_output()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for b, db, e in zip(batches, dbs, eager):
            out = db.output()
            assert normwise(out, e) <= 1e-13, f"replay {replay}: {b.name}"


def test_grouped_invalid(gpu):
    b = W.random_case(seed=41, d=2, n=3, nb=8)
    db = DeviceBatch(b, device=str(gpu))
    with pytest.raises(api.KronmultError):
        api.kronmult_batched_grouped([db.group(entries_per_tile=-1)], dtype="f64")
    with pytest.raises(api.KronmultError):
        api.kronmult_batched_grouped([db.group()], dtype="f64", strategy=7)
    bad = api.Group(0, 3, db.mat_ptrs, 3, db.in_ptrs, db.out_ptrs, None, 8)
    with pytest.raises(api.KronmultError):
        api.kronmult_batched_grouped([db.group(), bad], dtype="f64")
    # nothing was launched by the rejected calls
    np.testing.assert_array_equal(db.output(), b.output_pool_init())
    api.kronmult_batched_grouped([], dtype="f64")


def test_autotuner_grouped(gpu, tmp_path):
    import json
    import torch
    from kronmult993_amd.autotune import Autotuner, grouped_call_from_choice
    batches = W.c3(scale=0.05)
    dbs = [DeviceBatch(b, device=str(gpu)) for b in batches]
    path = str(tmp_path / "tune.json")
    tuner = Autotuner(cache_path=path, reps=2)
    choice = tuner.tune_grouped(dbs)
    print(f"[autotune] {choice}", flush=True)
    assert choice["strategy"] in ("persistent", "separate")
    for db in dbs:
        db.reset_output()
    groups, structs, kw = grouped_call_from_choice(dbs, choice)
    api.kronmult_batched_grouped(groups, dtype="f64", structs=structs, **kw)
    torch.cuda.synchronize()
    for b, db in zip(batches, dbs):
        _check(b, db.output())
    with open(path) as f:
        saved = json.load(f)
    assert any(k.startswith("grouped|") for k in saved)
    assert sum(k.count("|nb") for k in saved) >= len(batches)  # every group's own shape is cached too
    assert Autotuner(cache_path=path).tune_grouped(dbs) == choice  # served from the cache
