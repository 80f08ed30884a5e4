"""GPU tests of the ordered-accumulation entry points (kronmult_batched_ordered_f64/_f32), through the C ABI.

The device result must equal, bit for bit, the documented summation order applied to the oracle's per-entry
contributions (helpers.ordered_reference), and repeat calls must reproduce it exactly. Covers the register
family (its range length from the launch), the one-entry-per-workgroup fiber kernels (LDS and workspace),
the multi-pass sweep, explicit launch shapes, and the layout checks of the synchronous entry point.
"""
import numpy as np
import pytest

import oracle
from helpers import grouped_by_output, normwise, ordered_reference, shift_odd_inputs
from kronmult993_amd import api, workloads as W
from kronmult993_amd.device import DeviceBatch

pytestmark = pytest.mark.gpu


def _ordered(b, gpu, tuning=None, repeats=2):
    info = api.ordered_launch_info(b.d, b.n, b.dtype.itemsize, b.nb, tuning)
    print(f"[case] {b.name} d={b.d} n={b.n} nb={b.nb} dtype={b.dtype} tuning={tuning} "
          f"family={info['family']} epw={info['entries_per_workgroup']}", flush=True)
    db = DeviceBatch(b, device=str(gpu), workspace="per_entry")
    outs = []
    for _ in range(repeats):
        db.reset_output()
        db.run_ordered(tuning=tuning)
        outs.append(db.output())
    return outs, info["entries_per_workgroup"]


def _check(b, gpu, tuning=None):
    outs, epw = _ordered(b, gpu, tuning)
    ref = ordered_reference(b, epw)
    for out in outs:
        if not np.array_equal(out, ref):
            bad = np.flatnonzero(out != ref)
            pytest.fail(f"{b.name}: {bad.size} elements differ from the ordered sum, first at {bad[0]}: "
                        f"{out[bad[0]]!r} vs {ref[bad[0]]!r}")
    tol = 1e-13 if b.dtype == np.float64 else 1e-5
    assert normwise(outs[0], oracle.kronmult(b)) <= tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ordered_c2_register_family(gpu, dtype):
    b = W.c2().astype(dtype)
    assert api.ordered_launch_info(6, 4, b.dtype.itemsize, b.nb)["family"] == api.FAMILY_REGISTER
    _check(b, gpu)


@pytest.mark.parametrize("epw", [1, 5, 64, 200])
def test_ordered_register_range_lengths(gpu, epw):
    """Ranges shorter than, misaligned with and longer than the 64-entry output segments."""
    _check(W.c2(rows=5, cols=64), gpu, tuning={"entries_per_workgroup": epw})


@pytest.mark.parametrize("epw", [3, 64])
def test_ordered_register_mixed_alignment(gpu, epw):
    """FP64 d=6 n=4 with input pointers alternating between 16-byte and element alignment inside every
    range: the LDS-DMA input staging and its register fallback switch from one entry to the next."""
    _check(shift_odd_inputs(W.c2(rows=5, cols=64)), gpu, tuning={"entries_per_workgroup": epw})


@pytest.mark.parametrize("d,n,nb,dtype", [(3, 4, 300, np.float64), (3, 4, 300, np.float32), (5, 3, 90, np.float64),
                                          (4, 8, 70, np.float64), (3, 8, 70, np.float64), (2, 20, 40, np.float64)])
def test_ordered_irregular_batches(gpu, d, n, nb, dtype):
    """Scattered aliasing regrouped by output, padded strides, shared factors, non-zero initial outputs."""
    b = grouped_by_output(W.random_case(90 + d + n, d, n, nb, shared_factors=7, dtype=dtype))
    _check(b, gpu)


@pytest.mark.parametrize("d,n,nb", [(3, 21, 6), (6, 8, 4), (6, 7, 3), (2, 70, 5), (1, 100, 7)])
def test_ordered_large_vectors(gpu, d, n, nb):
    """Workspace fiber sweep (n > 8), the multi-pass sweep (every entry reduced from its workspace) and the
    wide-fiber sweep (n > 64, workspace and LDS)."""
    b = grouped_by_output(W.random_case(95 + d + n, d, n, nb, distinct=2))
    _check(b, gpu)


def test_ordered_fiber_family_on_a_register_shape(gpu):
    _check(W.c2(rows=2, cols=64), gpu, tuning={"family": api.FAMILY_FIBER})


def test_ordered_rejects_bad_layouts(gpu):
    b = W.random_case(97, 3, 4, 60, distinct=10)     # outputs scattered, not grouped
    db = DeviceBatch(b, device=str(gpu), workspace="per_entry")
    before = db.output()
    with pytest.raises(api.KronmultError) as e:
        db.run_ordered()
    assert e.value.code == api.HIP_ERROR_INVALID_VALUE
    g = grouped_by_output(b)
    dg = DeviceBatch(g, device=str(gpu), workspace="per_entry")
    args = (g.d, g.n, dg.mat_ptrs, g.stride, dg.in_ptrs, dg.out_ptrs)
    with pytest.raises(api.KronmultError):             # no workspace
        api.kronmult_batched_ordered(*args, None, g.nb, dtype="f64")
    import torch
    shared_ws = dg.ws_ptrs.clone()
    shared_ws[1] = shared_ws[0]                         # two entries on one workspace
    with pytest.raises(api.KronmultError):
        api.kronmult_batched_ordered(*args, shared_ws, g.nb, dtype="f64")
    overlap = dg.ws_ptrs.clone()
    overlap[3] = dg.out_ptrs[0]                         # a workspace on an output
    with pytest.raises(api.KronmultError):
        api.kronmult_batched_ordered(*args, overlap, g.nb, dtype="f64")
    with pytest.raises(api.KronmultError):             # no matrix-core variant
        api.ordered_launch_info(6, 4, 8, 100, {"family": api.FAMILY_MATRIX})
    torch.cuda.synchronize()
    assert np.array_equal(db.output(), before)
    api.kronmult_batched_ordered(*args, dg.ws_ptrs, 0, dtype="f64")   # nb = 0: nothing to do


def test_ordered_graph_replays_are_bitwise_identical(gpu):
    """The stream-ordered ordered call captured in a graph (sweep + reduction kernels, nothing allocated):
    every replay, with eager output resets between replays, reproduces the eager result bit for bit."""
    import torch
    b = W.c2(rows=6, cols=64)
    db = DeviceBatch(b, device=str(gpu), workspace="per_entry")
    db.run_ordered(sync=False)
    torch.cuda.synchronize()
    eager = db.output()
    assert np.array_equal(eager, ordered_reference(b, api.ordered_launch_info(6, 4, 8, b.nb)["entries_per_workgroup"]))
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        db.reset_output()
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            db.run_ordered(sync=False, stream=s)
    for replay in range(3):
        db.reset_output()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(db.output(), eager), f"replay {replay} differs"
