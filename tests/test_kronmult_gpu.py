"""GPU parity of libkronmult_hip.so against the oracle, through the C ABI.

Bar (FP64 and FP32 alike):
  * an entry's contribution is bit-identical to the oracle's fused-multiply-add sweep, so wherever every
    output receives one contribution the result is compared with np.array_equal;
  * with aliased outputs only the order of the atomic additions differs: the error must stay within the
    componentwise bound (E_o + 2) * u * (|K| |x| summed over contributions), and the normwise relative error
    within 1e-13 (FP64) / 1e-5 (FP32).
"""
import numpy as np
import pytest

import oracle
from helpers import grouped_by_output, accumulation_bound, normwise, shift_odd_inputs
from kronmult993_amd import api, workloads as W
from kronmult993_amd.device import DeviceBatch

pytestmark = pytest.mark.gpu

EPS = {np.dtype(np.float64): np.finfo(np.float64).eps / 2, np.dtype(np.float32): np.finfo(np.float32).eps / 2}
NORM_TOL = {np.dtype(np.float64): 1e-13, np.dtype(np.float32): 1e-5}


def _unaliased(b):
    """Same entries, each into its own zero-initialised output."""
    return W.HostBatch(b.name + "_unaliased", b.d, b.n, b.stride, b.mat_pool, b.mat_off, b.in_pool, b.in_off,
                       b.nb, np.arange(b.nb, dtype=np.int64))


def _run(b, gpu, **kw):
    print(f"[case] {b.name} d={b.d} n={b.n} nb={b.nb} stride={b.stride} dtype={b.dtype} "
          f"kernel={api.kernel_name(b.d, b.n, b.dtype.itemsize, b.nb)}", flush=True)
    db = DeviceBatch(b, device=str(gpu), **kw)
    inputs_before = db.inp.clone()
    db.run()
    out = db.output()
    assert bool((db.inp == inputs_before).all()), "inputs must be left unmodified"
    return out


def _check_aliased(b, out):
    ref = oracle.kronmult(b)
    bound = accumulation_bound(b, EPS[b.dtype])
    err = np.abs(out.astype(np.float64) - ref.astype(np.float64))
    worst = float(np.max(err - bound))
    assert worst <= 0.0, f"{b.name}: componentwise bound exceeded by {worst:.3e}"
    assert normwise(out, ref) <= NORM_TOL[b.dtype]


def _check_bitexact(b, out):
    ref = oracle.kronmult(b)
    if not np.array_equal(out, ref):
        bad = np.flatnonzero(out != ref)
        pytest.fail(f"{b.name}: {bad.size} elements differ, first at {bad[0]}: {out[bad[0]]!r} vs {ref[bad[0]]!r}, "
                    f"normwise {normwise(out, ref):.3e}")


DTYPES = [np.float64, np.float32]


@pytest.mark.parametrize("dtype", DTYPES)
def test_c1_full(gpu, dtype):
    b = W.c1().astype(dtype)
    _check_aliased(b, _run(b, gpu))
    u = _unaliased(b)
    _check_bitexact(u, _run(u, gpu))


@pytest.mark.parametrize("dtype", DTYPES)
def test_c2_asgard_full(gpu, dtype):
    b = W.c2().astype(dtype)
    _check_aliased(b, _run(b, gpu))


@pytest.mark.parametrize("dtype", DTYPES)
def test_c2_entries_bitexact(gpu, dtype):
    u = _unaliased(W.c2(rows=3, cols=64).astype(dtype))
    _check_bitexact(u, _run(u, gpu))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("group", range(5))
def test_c3_groups(gpu, dtype, group):
    b = W.c3()[group].astype(dtype)
    _check_aliased(b, _run(b, gpu))
    u = _unaliased(b.subset(min(b.nb, 2000)))
    _check_bitexact(u, _run(u, gpu))


@pytest.mark.parametrize("dtype", DTYPES)
def test_c4_range(gpu, dtype):
    """The precision workload at its full size (16384 entries, runs of 16 entries per output)."""
    b = W.c4().astype(dtype)
    _check_aliased(b, _run(b, gpu))
    u = _unaliased(W.c4(nb=256).astype(dtype))
    _check_bitexact(u, _run(u, gpu))


@pytest.mark.parametrize("dtype", DTYPES)
def test_c5_forward(gpu, dtype):
    b = W.c5().astype(dtype)
    _check_aliased(b, _run(b, gpu))


@pytest.mark.parametrize("deg,dim,lvl", [(4, 1, 2), (4, 2, 4), (6, 3, 6)])
def test_reference_harness_cases(gpu, deg, dim, lvl):
    """The reference's own test cases (tests/kronmult_test_gpu.cpp:74-75 and the bench's medium case):
    stride 67, N(0,1) data, non-zero outputs aliased by tests/utils/utils_cpu.h:79-83."""
    b = W.reference_case(deg, dim, lvl)
    _check_aliased(b, _run(b, gpu))


@pytest.mark.parametrize("d,n,nb,seed", [
    (1, 1, 7, 1), (1, 4, 33, 2), (2, 3, 50, 3), (2, 7, 65, 4), (3, 5, 40, 5), (4, 6, 9, 6), (5, 2, 100, 7),
    (6, 3, 20, 8), (2, 9, 30, 9), (3, 10, 12, 10), (2, 16, 8, 11), (1, 64, 5, 12), (7, 2, 33, 13),
    (8, 3, 3, 14)])
def test_random_shapes(gpu, d, n, nb, seed):
    for dtype in DTYPES:
        b = W.random_case(seed, d, n, nb, shared_factors=max(1, nb // 4)).astype(dtype)
        _check_aliased(b, _run(b, gpu))
        u = _unaliased(b)
        _check_bitexact(u, _run(u, gpu))


@pytest.mark.parametrize("d,n,nb,seed", [(1, 65, 9, 21), (1, 300, 5, 22), (2, 70, 12, 23), (2, 129, 6, 24),
                                         (3, 65, 3, 25), (2, 66, 40, 26)])
def test_large_n_wide_fibers(gpu, d, n, nb, seed):
    """n > 64 (the reference takes any n, kronmult_gpu/kronmult.cu:54-78): the wide-fiber kernel, vector in
    LDS (d=1) or in the caller's workspace, padded strides, shared factors, aliased outputs; bit-exact per
    entry, in FP64 and FP32."""
    for dtype in DTYPES:
        b = W.random_case(seed, d, n, nb, shared_factors=max(1, nb // 3)).astype(dtype)
        assert api.kernel_name(d, n, b.dtype.itemsize, b.nb).startswith("wide_")
        _check_aliased(b, _run(b, gpu))
        u = _unaliased(b)
        _check_bitexact(u, _run(u, gpu))


def test_large_n_limit(gpu):
    """A fiber must fit the 32 KiB staging chunk: n = 4097 in FP64 is refused, n = 1500 (d = 1) runs."""
    with pytest.raises(api.KronmultError):
        api.launch_info(1, 4097, 8, 1)
    u = _unaliased(W.random_case(27, 1, 1500, 2))
    _check_bitexact(u, _run(u, gpu))


@pytest.mark.parametrize("d,n", [(6, 4), (5, 4), (4, 4), (4, 8), (3, 8)])
@pytest.mark.parametrize("dtype", DTYPES)
def test_register_plans_irregular(gpu, d, n, dtype):
    """The register-resident plans with padded strides, scrambled aliased outputs, shared factors and inputs
    whose addresses are only element-aligned (odd entries shifted by one element)."""
    b = W.random_case(60 + d * n, d, n, 97, stride=n + 3, shared_factors=13).astype(dtype)
    assert api.kernel_name(d, n, b.dtype.itemsize, b.nb).startswith("reg_sweep")
    size = b.size
    pool = np.zeros(b.nb * (size + 1), dtype=b.dtype)
    in_off = np.arange(b.nb, dtype=np.int64) * (size + 1) + (np.arange(b.nb) % 2)
    for k in range(b.nb):
        pool[in_off[k]: in_off[k] + size] = b.in_pool[b.in_off[k]: b.in_off[k] + size]
    b = W.HostBatch(**{**b.__dict__, "in_pool": pool, "in_off": in_off})
    _check_aliased(b, _run(b, gpu))
    u = _unaliased(b)
    _check_bitexact(u, _run(u, gpu))


@pytest.mark.parametrize("case", ["d6n4_runs", "d5n4_scattered"])
def test_register_dma_staging_mixed_alignment(gpu, case):
    """The FP64 d=5, 6 n=4 sweeps load inputs straight into LDS when the pointer is 16-byte aligned and
    through registers otherwise: runs of two or more entries per workgroup that switch between the two
    paths from one entry to the next, with same-output runs (d=6) and scattered outputs (d=5)."""
    if case == "d6n4_runs":
        b = W.c2(rows=32, cols=64)  # 2048 entries, runs of 64 entries per output
    else:
        b = W.random_case(91, 5, 4, 4500, stride=6, distinct=300)
    info = api.launch_info(b.d, b.n, 8, b.nb)
    assert info["family"] == api.FAMILY_REGISTER and info["entries_per_workgroup"] >= 2, info
    b = shift_odd_inputs(b)
    _check_aliased(b, _run(b, gpu))
    u = _unaliased(b)
    _check_bitexact(u, _run(u, gpu))


@pytest.mark.parametrize("aliased", [False, True])
def test_register_graph_replays(gpu, aliased):
    """The stream-ordered call captured in a graph (FP64 d=6 n=4: LDS-DMA input staging, run sums, atomics),
    replayed several times with eager output resets in between: unaliased outputs equal the oracle bit for
    bit on every replay, aliased ones stay within the accumulation bound."""
    import torch
    b = shift_odd_inputs(W.c2(rows=6, cols=64))
    if not aliased:
        b = _unaliased(b)
    print(f"[case] graph {b.name} nb={b.nb}", flush=True)
    db = DeviceBatch(b, device=str(gpu))
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        db.reset_output()
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            db.run(sync=False, stream=s)
    for replay in range(3):
        db.reset_output()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        if aliased:
            _check_aliased(b, db.output())
        else:
            _check_bitexact(b, db.output())


@pytest.mark.parametrize("d,n,nb", [(3, 21, 6), (4, 12, 4)])
def test_large_vectors_use_workspace(gpu, d, n, nb):
    """n^d * 8 bytes above the on-chip size, n > 8: the fiber sweep runs from the caller's workspace."""
    b = W.random_case(40 + d, d, n, nb)
    assert api.kernel_name(d, n, 8, nb).startswith("fiber_global")
    _check_aliased(b, _run(b, gpu))
    u = _unaliased(b)
    _check_bitexact(u, _run(u, gpu))


@pytest.mark.parametrize("d,n,nb,dtype", [(6, 8, 3, np.float64), (6, 7, 3, np.float64), (5, 8, 4, np.float64),
                                          (7, 5, 2, np.float64), (6, 6, 2, np.float64), (8, 4, 2, np.float64),
                                          (9, 3, 2, np.float64), (15, 3, 1, np.float64), (6, 8, 3, np.float32),
                                          (5, 8, 5, np.float32)])
def test_large_vectors_multipass(gpu, d, n, nb, dtype):
    """n^d * sizeof(T) above the on-chip size, n <= 8: groups of modes in passes over the workspace (two
    passes; three for d=15 n=3). Per entry the result is bit-identical to the oracle's sweep."""
    b = W.random_case(60 + d + n, d, n, nb, dtype=dtype)
    assert api.kernel_name(d, n, b.dtype.itemsize, nb).startswith("multipass")
    _check_aliased(b, _run(b, gpu))
    u = _unaliased(b)
    _check_bitexact(u, _run(u, gpu))


@pytest.mark.parametrize("d", [3, 4])
@pytest.mark.parametrize("epw", [1, 3, 7])
def test_matrix_core_n8_run_lengths(gpu, d, epw):
    """The n = 8 FP64 matrix-core sweep with runs that start at odd entries (its alternating factor buffers)
    and cut output segments at every place: grouped outputs of 1..5 entries, padded strides, element-aligned
    odd inputs; bit-exact per entry."""
    tu = {"family": api.FAMILY_REGISTER, "entries_per_workgroup": epw}
    b = grouped_by_output(W.random_case(140 + d + epw, d, 8, 61, stride=11, distinct=17, shared_factors=5))
    assert api.kernel_name(d, 8, 8, b.nb) == "reg_sweep_mfma_f64"
    b = shift_odd_inputs(b)
    _check_aliased(b, _run_tuned(b, gpu, tu))
    u = _unaliased(b)
    _check_bitexact(u, _run_tuned(u, gpu, tu))


def test_empty_batch_and_invalid_arguments(gpu):
    b = W.random_case(50, 3, 4, 8)
    db = DeviceBatch(b, device=str(gpu))
    before = db.output()
    api.kronmult_batched(3, 4, db.mat_ptrs, b.stride, db.in_ptrs, db.out_ptrs, None, 0, dtype="f64")
    assert np.array_equal(db.output(), before)
    for args in ((0, 4, b.stride), (3, 0, b.stride), (3, 4, 3)):
        with pytest.raises(api.KronmultError) as e:
            api.kronmult_batched(args[0], args[1], db.mat_ptrs, args[2], db.in_ptrs, db.out_ptrs, None, 8,
                                 dtype="f64")
        assert e.value.code == api.HIP_ERROR_INVALID_VALUE
    with pytest.raises(api.KronmultError):
        api.kronmult_batched(3, 4, None, b.stride, db.in_ptrs, db.out_ptrs, None, 8, dtype="f64")
    assert np.array_equal(db.output(), before)


def test_async_on_a_stream_and_repeat_accumulates(gpu):
    import torch
    b = W.random_case(51, 4, 4, 64)
    db = DeviceBatch(b, device=str(gpu))
    s = torch.cuda.Stream(device=gpu)
    db.run(sync=False, stream=s)
    db.run(sync=False, stream=s)
    s.synchronize()
    once = oracle.kronmult(b)
    twice = oracle.kronmult(b, out_pool=once.copy())
    assert normwise(db.output(), twice) <= 1e-13


def _run_tuned(b, gpu, tuning):
    print(f"[case] {b.name} d={b.d} n={b.n} nb={b.nb} stride={b.stride} dtype={b.dtype} tuning={tuning}", flush=True)
    db = DeviceBatch(b, device=str(gpu))
    db.run(tuning=tuning)
    return db.output()


@pytest.mark.parametrize("epw", [0, 1, 3])
def test_matrix_core_family_d6n4(gpu, epw):
    """The opt-in matrix-core sweep (d=6, n=4, FP64): per entry bit-identical to the oracle, aliased outputs
    within the bound, also with a padded stride, element-aligned inputs and short entry runs."""
    tu = {"family": api.FAMILY_MATRIX, "entries_per_workgroup": epw}
    u = _unaliased(W.c2(rows=3, cols=64))
    _check_bitexact(u, _run_tuned(u, gpu, tu))
    b = W.c2()
    _check_aliased(b, _run_tuned(b, gpu, tu))
    r = W.random_case(77, 6, 4, 97, stride=7, shared_factors=13)
    size = r.size
    pool = np.zeros(r.nb * (size + 1))
    in_off = np.arange(r.nb, dtype=np.int64) * (size + 1) + (np.arange(r.nb) % 2)
    for k in range(r.nb):
        pool[in_off[k]: in_off[k] + size] = r.in_pool[r.in_off[k]: r.in_off[k] + size]
    r = W.HostBatch(**{**r.__dict__, "in_pool": pool, "in_off": in_off})
    _check_aliased(r, _run_tuned(r, gpu, tu))
    ru = _unaliased(r)
    _check_bitexact(ru, _run_tuned(ru, gpu, tu))


def test_matrix_core_family_rejects_other_shapes(gpu):
    tu = {"family": api.FAMILY_MATRIX}
    assert api.launch_info(6, 4, 8, 100, tu)["family"] == api.FAMILY_MATRIX
    for d, n, size in ((6, 4, 4), (5, 4, 8), (6, 8, 8), (3, 4, 8)):
        with pytest.raises(api.KronmultError) as e:
            api.launch_info(d, n, size, 100, tu)
        assert e.value.code == api.HIP_ERROR_INVALID_VALUE
    assert api.kernel_name(6, 4, 8, 100) == "reg_sweep_f64"


@pytest.mark.parametrize("d,n", [(3, 4), (2, 8), (3, 6), (5, 2), (6, 2), (4, 3), (1, 8), (2, 2), (3, 7)])
@pytest.mark.parametrize("dtype", DTYPES)
def test_packed_family(gpu, d, n, dtype):
    """Several small vectors per workgroup (the default for n^(d-1) <= 64 and large batches of shapes above 64
    elements; forced here for the wave shapes too): padded strides, shared factors, scattered aliasing;
    bit-identical per entry."""
    b = W.random_case(120 + d * n, d, n, 6000, stride=n + 1, shared_factors=37).astype(dtype)
    wave = (d, n) in WAVE_SHAPES and not ((d, n) == (3, 6) and dtype == np.float64)
    assert api.kernel_name(d, n, b.dtype.itemsize, b.nb).startswith("wave" if wave else "packed")
    tu = {"family": api.FAMILY_PACKED}
    _check_aliased(b, _run_tuned(b, gpu, tu))
    u = _unaliased(b)
    _check_bitexact(u, _run_tuned(u, gpu, tu))


WAVE_SHAPES = [(1, 2), (2, 2), (3, 2), (4, 2), (5, 2), (6, 2), (1, 3), (2, 3), (3, 3), (1, 4), (2, 4), (3, 4),
               (1, 5), (2, 5), (1, 6), (2, 6), (1, 7), (2, 7), (1, 8), (2, 8),
               (7, 2), (8, 2), (4, 3), (5, 3), (3, 5), (3, 6), (2, 9), (2, 10)]  # 2 to 4 elements per lane


@pytest.mark.parametrize("d,n", WAVE_SHAPES)
@pytest.mark.parametrize("dtype", DTYPES)
def test_wave_family(gpu, d, n, dtype):
    """One wavefront per entry (n^d <= 256, up to 4 elements per lane): padded strides, shared factors, scattered aliasing and runs of
    equal outputs cut at every place (grouped outputs, 5 entries per wavefront); bit-identical per entry."""
    tu = {"family": api.FAMILY_WAVE, "entries_per_workgroup": 5}
    b = W.random_case(160 + d * n, d, n, 203, stride=n + 2, shared_factors=29).astype(dtype)
    print(f"[case] wave d={d} n={n} {np.dtype(dtype).name} nb={b.nb}", flush=True)
    _check_aliased(b, _run_tuned(b, gpu, tu))
    g = shift_odd_inputs(grouped_by_output(W.random_case(170 + d * n, d, n, 97, distinct=23).astype(dtype)))
    _check_aliased(g, _run_tuned(g, gpu, tu))
    u = _unaliased(b)
    _check_bitexact(u, _run_tuned(u, gpu, tu))


@pytest.mark.parametrize("epw,wpc", [(1, 0), (3, 2), (0, 32), (64, 0)])
def test_wave_family_launch_shapes(gpu, epw, wpc):
    """Forced wave launches with explicit run lengths and grids; shapes outside the family are refused."""
    tu = {"family": api.FAMILY_WAVE, "entries_per_workgroup": epw, "workgroups_per_cu": wpc}
    b = W.random_case(181, 2, 8, 300, distinct=41)
    _check_aliased(b, _run_tuned(b, gpu, tu))
    u = _unaliased(b)
    _check_bitexact(u, _run_tuned(u, gpu, tu))
    for d, n in ((3, 7), (4, 4), (1, 9), (2, 1), (6, 3)):
        with pytest.raises(api.KronmultError):
            api.launch_info(d, n, 8, 100, {"family": api.FAMILY_WAVE})


@pytest.mark.parametrize("d,n", [(3, 4), (2, 8), (6, 2), (1, 5), (5, 3)])
@pytest.mark.parametrize("dtype", DTYPES)
def test_fiber_family_small_shapes(gpu, d, n, dtype):
    """The one-workgroup-per-entry fiber sweep, forced on shapes whose default is now the wave family."""
    tu = {"family": api.FAMILY_FIBER}
    b = W.random_case(190 + d * n, d, n, 150, stride=n + 1, shared_factors=11).astype(dtype)
    assert api.launch_info(d, n, b.dtype.itemsize, b.nb, tu)["family"] == api.FAMILY_FIBER
    _check_aliased(b, _run_tuned(b, gpu, tu))
    u = _unaliased(b)
    _check_bitexact(u, _run_tuned(u, gpu, tu))


@pytest.mark.parametrize("ept,wpc", [(1, 0), (3, 2), (0, 64)])
def test_packed_family_launch_shapes(gpu, ept, wpc):
    """Forced packed launches on a small batch with explicit tile lengths and grid sizes."""
    tu = {"family": api.FAMILY_PACKED, "entries_per_workgroup": ept, "workgroups_per_cu": wpc}
    b = W.random_case(131, 3, 4, 77, distinct=9)
    _check_aliased(b, _run_tuned(b, gpu, tu))
    u = _unaliased(b)
    _check_bitexact(u, _run_tuned(u, gpu, tu))
    with pytest.raises(api.KronmultError):
        api.launch_info(4, 5, 8, 100, {"family": api.FAMILY_PACKED})   # 125 fibers: not packed
