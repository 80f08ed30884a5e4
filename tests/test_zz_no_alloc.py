"""GPU tests of kronmult_batched_ordered_f64_noalloc, the ordered accumulation that takes no workspace and
allocates nothing (its partial sums live in the library's own device array), through the C ABI.

Every call is compared bit for bit with kronmult_batched_ordered_f64_async on the same batch and launch shape
(per-entry workspaces from the caller) and with the documented summation order applied to the oracle's
per-entry contributions. Two calls in a row, the second on a smaller batch, check that the array's leftover
contents from one call do not leak into the next.
"""
import numpy as np
import pytest

from helpers import grouped_by_output, ordered_reference
from kronmult993_amd import api, workloads as W
from kronmult993_amd.device import DeviceBatch

pytestmark = pytest.mark.gpu


def _noalloc(db, tuning=None, stream=None):
    api.kronmult_batched_ordered_noalloc(db.d, db.n, db.mat_ptrs, db.stride, db.in_ptrs, db.out_ptrs, db.nb,
                                         tuning=tuning, stream=stream)


def _compare(b, gpu, tuning=None):
    """noalloc result == workspace-taking ordered result == ordered reference, bit for bit."""
    import torch
    info = api.ordered_launch_info(b.d, b.n, 8, b.nb, tuning)
    print(f"[case] noalloc {b.name} d={b.d} n={b.n} nb={b.nb} tuning={tuning} family={info['family']} "
          f"epw={info['entries_per_workgroup']}", flush=True)
    with_ws = DeviceBatch(b, device=str(gpu), workspace="per_entry")
    with_ws.run_ordered(sync=False, tuning=tuning)
    no_ws = DeviceBatch(b, device=str(gpu), workspace="none")
    _noalloc(no_ws, tuning)
    torch.cuda.synchronize()
    want, got = with_ws.output(), no_ws.output()
    if not np.array_equal(got, want):
        bad = np.flatnonzero(got != want)
        pytest.fail(f"{b.name}: {bad.size} elements differ from the workspace-taking call, first at {bad[0]}: "
                    f"{got[bad[0]]!r} vs {want[bad[0]]!r}")
    assert np.array_equal(got, ordered_reference(b, info["entries_per_workgroup"]))


@pytest.mark.parametrize("tuning", [None, {"entries_per_workgroup": 5}])
def test_noalloc_matches_ordered_call_twice(gpu, tuning):
    """Register family on part of the d=6 n=4 operator (multi-range segments), then a smaller irregular
    batch on the one-entry-per-workgroup fiber kernel: both equal the workspace-taking call."""
    _compare(W.c2(rows=6, cols=64), gpu, tuning)
    _compare(grouped_by_output(W.random_case(301, 3, 4, 90, shared_factors=5)), gpu)


def test_noalloc_full_operator_then_smaller(gpu):
    """The whole d=6 n=4 operator batch (8192 entries, 64 MiB of the array), then 192 of its entries."""
    _compare(W.c2(), gpu)
    _compare(W.c2(rows=3, cols=64), gpu)


def test_noalloc_large_vectors(gpu):
    """Multi-pass sweep (every entry's vector kept in the array, reduced from there) and the workspace fiber
    sweep (n > 8: the array is the sweep's working storage)."""
    _compare(grouped_by_output(W.random_case(303, 6, 8, 4, distinct=2)), gpu)
    _compare(grouped_by_output(W.random_case(304, 3, 21, 6, distinct=2)), gpu)


def test_noalloc_rejects_batches_beyond_the_array(gpu):
    """d=6 n=8 vectors are 2 MiB; 64 one-entry ranges need 128 MiB: refused before anything runs."""
    import torch
    b = W.random_case(305, 6, 8, 64, distinct=4)
    need = 64 * 8 + 64 * 8 ** 6 * 8
    assert need > api.NOALLOC_SCRATCH_BYTES
    db = DeviceBatch(b, device=str(gpu), workspace="none")
    before = db.output()
    with pytest.raises(api.KronmultError) as e:
        _noalloc(db)
    assert e.value.code == api.HIP_ERROR_OUT_OF_MEMORY
    torch.cuda.synchronize()
    assert np.array_equal(db.output(), before)
    api.kronmult_batched_ordered_noalloc(db.d, db.n, db.mat_ptrs, db.stride, db.in_ptrs, db.out_ptrs, 0)  # nb = 0


def test_noalloc_graph_replays(gpu):
    """Captured in a HIP graph (the pointer-array fill, the sweep and the reduction kernels; no allocation,
    no memset), every replay, with eager output resets between replays, equals the eager call."""
    import torch
    b = W.c2(rows=4, cols=64)
    db = DeviceBatch(b, device=str(gpu), workspace="none")
    _noalloc(db)
    torch.cuda.synchronize()
    eager = db.output()
    assert np.array_equal(eager, ordered_reference(b, api.ordered_launch_info(6, 4, 8, b.nb)["entries_per_workgroup"]))
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        db.reset_output()
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            _noalloc(db, stream=s)
    for replay in range(3):
        db.reset_output()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(db.output(), eager), f"replay {replay} differs"
