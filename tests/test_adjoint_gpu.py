"""GPU parity of the adjoint (kronmult_batched_adjoint_f64/_f32) and of the check-only autograd Function.

The adjoint has no counterpart in the reference. It is pinned three ways:
  * against the oracle's adjoint (oracle/kronmult_oracle_body.h, itself pinned in test_oracle.py against
    the dense Kronecker gradient and central differences) on seeded batches with shared factors, shared
    inputs and aliased outputs, for both ways of obtaining the intermediates (recompute / store) and both
    places the scratch can live (on chip / caller workspace);
  * through the autograd Function against central finite differences (h = 1e-6) of the GPU forward: the
    map is affine in every single scalar, so the difference quotient is exact up to rounding,
    |error| <~ eps * |L| / h;
  * through the autograd Function against the dense Kronecker gradient computed with numpy.
Tolerances (normwise, max|err| / max|ref|): FP64 1e-12 (only summation order differs: the factor
gradient's contraction runs on the matrix cores and shared gradients are summed by atomics); FP32 against
the FP64 oracle 1e-4.
"""
import numpy as np
import pytest

import oracle
from helpers import factors, kron_dense, normwise
from kronmult993_amd import api, workloads as W
from kronmult993_amd.device import DeviceBatch

pytestmark = pytest.mark.gpu

TOL = {np.dtype(np.float64): 1e-12, np.dtype(np.float32): 1e-4}


def _gpu_adjoint(b, gpu, intermediates, want_input=True, want_mats=True):
    ws = api.adjoint_workspace(b.d, b.n, b.dtype.itemsize, intermediates)
    print(f"[adjoint] {b.name} d={b.d} n={b.n} nb={b.nb} dtype={b.dtype} intermediates={intermediates} "
          f"workspace/entry={ws}", flush=True)
    db = DeviceBatch(b, device=str(gpu), workspace="none")
    db.adjoint_setup(want_input=want_input, want_mats=want_mats, intermediates=intermediates)
    inp_before, gout_before = db.inp.clone(), db.gout.clone()
    db.run_adjoint()
    assert bool((db.inp == inp_before).all()) and bool((db.gout == gout_before).all())
    return db.gradients()


def _check(b, gi, gm, want_input=True, want_mats=True):
    ref_b = b.astype(np.float64)
    ri, rm = oracle.adjoint(ref_b, want_input=want_input, want_mats=want_mats)
    tol = TOL[b.dtype]
    if want_input:
        e = normwise(gi.astype(np.float64), ri)
        assert e <= tol, f"{b.name}: d(input) normwise error {e:.3e}"
    else:
        assert gi is None
    if want_mats:
        e = normwise(gm.astype(np.float64), rm)
        assert e <= tol, f"{b.name}: d(factors) normwise error {e:.3e}"
    else:
        assert gm is None


SHAPES = [(1, 1), (1, 5), (2, 3), (2, 8), (3, 5), (3, 8), (4, 4), (4, 8), (5, 3), (5, 4), (6, 2), (6, 4), (3, 7)]


@pytest.mark.parametrize("intermediates", ["auto", "recompute", "store"])
@pytest.mark.parametrize("d,n", SHAPES)
def test_adjoint_random_f64(gpu, d, n, intermediates):
    nb = 37 if n ** d <= 4096 else 9
    b = W.random_case(seed=100 * d + n, d=d, n=n, nb=nb, shared_factors=5)
    gi, gm = _gpu_adjoint(b, gpu, intermediates)
    _check(b, gi, gm)


@pytest.mark.parametrize("d,n", [(2, 3), (3, 8), (4, 4), (5, 4), (6, 4)])
def test_adjoint_random_f32(gpu, d, n):
    b = W.random_case(seed=7 * d + n, d=d, n=n, nb=29, shared_factors=4, dtype=np.float32)
    gi, gm = _gpu_adjoint(b, gpu, "auto")
    _check(b, gi, gm)


@pytest.mark.parametrize("want_input,want_mats", [(True, False), (False, True)])
def test_adjoint_one_gradient(gpu, want_input, want_mats):
    b = W.random_case(seed=5, d=3, n=4, nb=20, shared_factors=3)
    gi, gm = _gpu_adjoint(b, gpu, "auto", want_input, want_mats)
    _check(b, gi, gm, want_input, want_mats)


@pytest.mark.parametrize("intermediates", ["recompute", "store"])
def test_adjoint_workspace_path(gpu, intermediates):
    """n^d = 32768: the scratch vectors do not fit on chip and live in the caller's workspace."""
    b = W.random_case(seed=3, d=5, n=8, nb=4, shared_factors=2)
    assert api.adjoint_workspace(5, 8, 8, intermediates) > 0
    gi, gm = _gpu_adjoint(b, gpu, intermediates)
    _check(b, gi, gm)


@pytest.mark.parametrize("intermediates", ["auto", "recompute", "store"])
def test_adjoint_c5_full(gpu, intermediates):
    """The differentiation workload: d=3, n=8, 8192 entries, 256 shared factor triples (32 entries each),
    outputs k/4, cotangent N(0,1)."""
    b = W.c5()
    gi, gm = _gpu_adjoint(b, gpu, intermediates)
    _check(b, gi, gm)


def test_adjoint_c5_f32(gpu):
    b = W.c5().astype(np.float32)
    gi, gm = _gpu_adjoint(b, gpu, "auto")
    _check(b, gi, gm)


@pytest.mark.parametrize("order", ["output", "input"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_adjoint_c2_run_sums(gpu, order, dtype):
    """d=6 n=4 operator subset (24 output x 16 input elements, 384 entries, factor blocks of modes 2..5 shared
    by every entry): in input-major order the kernel sums d(input) over each run of 24 entries with one
    input before its atomics, and in both orders the factor-gradient partials over the entries that share a
    gradient pointer in a row; both must equal the oracle adjoint."""
    from kronmult993_amd import gen
    b = W.c2(rows=24, cols=16, order=order).astype(dtype)
    b.grad_out_pool = gen.normal(0, 90, b.out_count * b.size).astype(dtype)
    gi, gm = _gpu_adjoint(b, gpu, "auto")
    _check(b, gi, gm)


def test_adjoint_invalid(gpu):
    import torch
    b = W.random_case(seed=1, d=2, n=9, nb=3)  # n > 8 is not supported by the adjoint
    with pytest.raises(api.KronmultError):
        api.adjoint_workspace(2, 9, 8)
    db = DeviceBatch(b, device=str(gpu), workspace="none")
    with pytest.raises(api.KronmultError):
        api.kronmult_batched_adjoint(b.d, b.n, db.mat_ptrs, b.stride, db.in_ptrs, db.out_ptrs, None, None, None,
                                     b.nb, dtype="f64")
    # a workspace is required for n^d = 32768 but none is given
    b = W.random_case(seed=1, d=5, n=8, nb=2)
    db = DeviceBatch(b, device=str(gpu), workspace="none")
    db.adjoint_setup()
    with pytest.raises(api.KronmultError):
        api.kronmult_batched_adjoint(b.d, b.n, db.mat_ptrs, b.stride, db.in_ptrs, db.gout_ptrs, db.gin_ptrs,
                                     db.gmat_ptrs, None, b.nb, dtype="f64")
    torch.cuda.synchronize()


# ---- autograd Function -------------------------------------------------------------------------------

def _autograd_setup(b, gpu):
    import torch
    from kronmult993_amd.autograd import Layout, kronmult
    dev = torch.device(str(gpu))
    mat = torch.tensor(b.mat_pool, device=dev, requires_grad=True)
    inp = torch.tensor(b.in_pool, device=dev, requires_grad=True)
    out0 = torch.tensor(b.output_pool_init(), device=dev, requires_grad=True)
    g = torch.tensor(b.grad_out_pool, device=dev)
    return kronmult, Layout.of(b), mat, inp, out0, g


def test_autograd_central_differences(gpu):
    import torch
    b = W.random_case(seed=11, d=3, n=4, nb=12, shared_factors=3)
    kronmult, layout, mat, inp, out0, g = _autograd_setup(b, gpu)
    out = kronmult(mat, inp, out0, layout)
    (out * g).sum().backward()
    gm, gi, go = mat.grad.cpu().numpy(), inp.grad.cpu().numpy(), out0.grad.cpu().numpy()
    np.testing.assert_array_equal(go, b.grad_out_pool)

    def loss(m, x):
        with torch.no_grad():
            return float((kronmult(m, x, out0.detach(), layout) * g).sum())

    h = 1e-6
    rng = np.random.default_rng(0)
    used_m = np.unique((b.mat_off[:, None] + (np.arange(b.n)[:, None] * b.stride + np.arange(b.n)[None, :]).reshape(1, -1)).reshape(-1))
    used_i = np.unique((b.in_off[:, None] + np.arange(b.size)[None, :]).reshape(-1))
    with torch.no_grad():
        L_abs = float((kronmult(mat.detach().abs(), inp.detach().abs(), out0.detach().abs(), layout) * g.abs()).sum())
    for which, idxs in (("mat", rng.choice(used_m, 12, replace=False)), ("in", rng.choice(used_i, 12, replace=False))):
        for i in idxs:
            m, x = mat.detach().clone(), inp.detach().clone()
            t = m if which == "mat" else x
            t[i] += h
            lp = loss(m, x)
            t[i] -= 2 * h
            lm = loss(m, x)
            fd = (lp - lm) / (2 * h)
            ad = (gm if which == "mat" else gi)[i]
            # affine in each scalar: only rounding of L (eps |L| / h) separates fd from the derivative
            assert abs(fd - ad) <= 1e-8 * max(1.0, L_abs), f"{which}[{i}]: fd {fd!r} vs adjoint {ad!r}"


def test_autograd_dense_kronecker_gradient(gpu):
    b = W.random_case(seed=12, d=3, n=3, nb=10, shared_factors=4)
    kronmult, layout, mat, inp, out0, g = _autograd_setup(b, gpu)
    (kronmult(mat, inp, out0, layout) * g).sum().backward()
    gm, gi = mat.grad.cpu().numpy(), inp.grad.cpu().numpy()
    rm, ri = np.zeros_like(b.mat_pool), np.zeros_like(b.in_pool)
    size, n = b.size, b.n
    for k in range(b.nb):
        A = factors(b, k)
        gk = b.grad_out_pool[int(b.out_off[k]): int(b.out_off[k]) + size]
        xk = b.in_pool[int(b.in_off[k]): int(b.in_off[k]) + size]
        ri[int(b.in_off[k]): int(b.in_off[k]) + size] += kron_dense(A).T @ gk
        for t in range(b.d):
            off = int(b.mat_off[k * b.d + t])
            for r in range(n):
                for c in range(n):
                    E = np.zeros((n, n))
                    E[r, c] = 1.0
                    rm[off + r + c * b.stride] += gk @ (kron_dense(A[:t] + [E] + A[t + 1:]) @ xk)
    assert normwise(gi, ri) <= 1e-13
    assert normwise(gm, rm) <= 1e-13


def test_autograd_gradcheck(gpu):
    import torch
    b = W.random_case(seed=13, d=2, n=3, nb=5, shared_factors=2)
    kronmult, layout, mat, inp, out0, g = _autograd_setup(b, gpu)
    # shared factors and inputs are summed by atomic adds in no fixed order: repeated backward passes may
    # differ in the last bits, hence a nondeterminism tolerance
    assert torch.autograd.gradcheck(lambda m, x, o: kronmult(m, x, o, layout), (mat, inp, out0),
                                    eps=1e-6, atol=1e-7, rtol=1e-6, nondet_tol=1e-13)


# ---- the differentiation workload's own checks, on 64 entries of its shape -----------------------------

def _c5_subset():
    """64 entries of the d=3 n=8 differentiation batch's shape: 8 shared factor triples (8 entries each,
    entry k uses triple k mod 8), outputs k/4, inputs U[-1,1], cotangent N(0,1)."""
    return W.c5(nb=64, triples=8)


def _dense_gradients(b):
    """The dense-Kronecker gradient of L = sum_o <g_o, y_o>: d(input_k) = K_k^T g_o(k) with K_k the explicit
    n^d x n^d Kronecker matrix, and d(A_t)[r, c] = sum over the entries of K_k's elements whose mode-t indices
    are (r, c), of (g x^T)[i, j] times the other factors' elements: the gradient of <g, K x> with respect to
    one factor of the product K = A_0 (x) A_1 (x) A_2, read off the dense outer product g x^T."""
    n, d, size = b.n, b.d, b.size
    assert d == 3
    ri, rm = np.zeros_like(b.in_pool), np.zeros_like(b.mat_pool)
    for k in range(b.nb):
        A = factors(b, k)
        gk = b.grad_out_pool[int(b.out_off[k]): int(b.out_off[k]) + size]
        xk = b.in_pool[int(b.in_off[k]): int(b.in_off[k]) + size]
        ri[int(b.in_off[k]): int(b.in_off[k]) + size] += kron_dense(A).T @ gk
        GX = np.outer(gk, xk).reshape(n, n, n, n, n, n)  # (i0, i1, i2, j0, j1, j2)
        grads = (np.einsum("abcdef,be,cf->ad", GX, A[1], A[2]),
                 np.einsum("abcdef,ad,cf->be", GX, A[0], A[2]),
                 np.einsum("abcdef,ad,be->cf", GX, A[0], A[1]))
        for t in range(d):
            off = int(b.mat_off[k * d + t])
            for c in range(n):
                rm[off + c * b.stride: off + c * b.stride + n] += grads[t][:, c]
    return ri, rm


@pytest.mark.parametrize("intermediates", ["recompute", "store"])
def test_adjoint_c5_subset_dense_gradient(gpu, intermediates):
    b = _c5_subset()
    gi, gm = _gpu_adjoint(b, gpu, intermediates)
    ri, rm = _dense_gradients(b)
    assert normwise(gi, ri) <= 1e-13
    assert normwise(gm, rm) <= 1e-13
    # the dense gradient itself against the oracle's adjoint (pinned on the CPU)
    oi, om = oracle.adjoint(b)
    assert normwise(oi, ri) <= 1e-13 and normwise(om, rm) <= 1e-13


def test_adjoint_c5_subset_central_differences(gpu):
    """Central differences (h = 1e-6) of the GPU forward, through the autograd Function, at factor and input
    elements drawn from every shared triple and many entries."""
    import torch
    b = _c5_subset()
    kronmult, layout, mat, inp, out0, g = _autograd_setup(b, gpu)
    out = kronmult(mat, inp, out0, layout)
    (out * g).sum().backward()
    gm, gi = mat.grad.cpu().numpy(), inp.grad.cpu().numpy()

    def loss(m, x):
        with torch.no_grad():
            return float((kronmult(m, x, out0.detach(), layout) * g).sum())

    h = 1e-6
    rng = np.random.default_rng(5)
    with torch.no_grad():
        L_abs = float((kronmult(mat.detach().abs(), inp.detach().abs(), out0.detach().abs(), layout) * g.abs()).sum())
    idx_m = rng.choice(b.mat_pool.size, 24, replace=False)
    used_i = np.unique((b.in_off[:, None] + np.arange(b.size)[None, :]).reshape(-1))
    idx_i = rng.choice(used_i, 24, replace=False)
    for which, idxs in (("mat", idx_m), ("in", idx_i)):
        for i in idxs:
            m, x = mat.detach().clone(), inp.detach().clone()
            t = m if which == "mat" else x
            t[i] += h
            lp = loss(m, x)
            t[i] -= 2 * h
            lm = loss(m, x)
            fd = (lp - lm) / (2 * h)
            ad = (gm if which == "mat" else gi)[i]
            assert abs(fd - ad) <= 1e-8 * max(1.0, L_abs), f"{which}[{i}]: fd {fd!r} vs adjoint {ad!r}"
