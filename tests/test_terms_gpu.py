"""GPU tests of the several-term entry points (kronmult_batched_terms_f64/_f32), through the C ABI.

output[k] += sum_t kron(A_{t,k,*}) input[k], factor pointers term-major. With one output per entry the
result must equal, bit for bit, the oracle's per-term contributions summed in the documented order (terms
summed in registers for the register-resident plans, one launch per term otherwise); with aliased outputs it
must stay within the accumulation bound of the expanded batch.
"""
import numpy as np
import pytest

import oracle
from helpers import accumulation_bound, normwise, terms_reference
from kronmult993_amd import api, workloads as W
from kronmult993_amd.device import DeviceBatch

pytestmark = pytest.mark.gpu

EPS = {np.dtype(np.float64): np.finfo(np.float64).eps / 2, np.dtype(np.float32): np.finfo(np.float32).eps / 2}
NORM_TOL = {np.dtype(np.float64): 1e-13, np.dtype(np.float32): 1e-5}


def _base_unaliased(b):
    """Each base entry into its own output (the terms of one entry still share it)."""
    nb, terms = int(b.meta["term_nb"]), int(b.meta["terms"])
    idx = np.tile(np.arange(nb, dtype=np.int64), terms)
    init = None if b.out_init is None else np.resize(b.out_init, nb * b.size)
    return W.HostBatch(**{**b.__dict__, "name": b.name + "_unaliased", "out_count": nb, "out_index": idx,
                          "out_init": init})


def _run(b, gpu):
    fused = api.kernel_name(b.d, b.n, b.dtype.itemsize, int(b.meta["term_nb"])).startswith("reg_sweep")
    print(f"[case] {b.name} d={b.d} n={b.n} terms={b.meta['terms']} nb={b.meta['term_nb']} dtype={b.dtype} "
          f"fused={fused}", flush=True)
    db = DeviceBatch(b, device=str(gpu))
    db.run_terms()
    return db.output(), fused


def _check(b, gpu):
    out, _ = _run(b, gpu)
    ref = oracle.kronmult(b)
    err = np.abs(out.astype(np.float64) - ref.astype(np.float64))
    worst = float(np.max(err - accumulation_bound(b, EPS[b.dtype])))
    assert worst <= 0.0, f"{b.name}: componentwise bound exceeded by {worst:.3e}"
    assert normwise(out, ref) <= NORM_TOL[b.dtype]
    u = _base_unaliased(b)
    out, fused = _run(u, gpu)
    ref = terms_reference(u, fused)
    if not np.array_equal(out, ref):
        bad = np.flatnonzero(out != ref)
        pytest.fail(f"{u.name}: {bad.size} elements differ, first at {bad[0]}: {out[bad[0]]!r} vs {ref[bad[0]]!r}")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_terms_c2(gpu, dtype):
    _check(W.c2_terms(terms=3).astype(dtype), gpu)


@pytest.mark.parametrize("d,n", [(6, 4), (5, 4), (4, 4), (4, 8), (3, 8)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_terms_register_plans(gpu, d, n, dtype):
    """Fused kernel: padded strides, shared factor sets, scattered aliasing, 2 and 4 terms."""
    for terms in (2, 4):
        _check(W.random_terms(70 + d * n + terms, d, n, 61, terms, stride=n + 2, shared_factors=9,
                              dtype=dtype), gpu)


@pytest.mark.parametrize("d,n,nb", [(3, 4, 50), (2, 20, 12), (6, 8, 3)])
def test_terms_other_shapes(gpu, d, n, nb):
    """One launch of the ordinary path per term (fiber sweep in LDS, runtime n, multi-pass)."""
    _check(W.random_terms(80 + d + n, d, n, nb, 3), gpu)


def test_terms_one_term_is_the_plain_call(gpu):
    c = W.c2(rows=2, cols=64)
    c = W.HostBatch(**{**c.__dict__, "out_count": c.nb, "out_index": np.arange(c.nb, dtype=np.int64)})
    b = W.expand_terms(c, [c.mat_pool])
    out, _ = _run(b, gpu)
    db = DeviceBatch(b, device=str(gpu))
    db.run()
    assert np.array_equal(out, db.output())


def test_terms_invalid_arguments(gpu):
    b = W.c2_terms(terms=2, rows=1, cols=8)
    db = DeviceBatch(b, device=str(gpu))
    nb = int(b.meta["term_nb"])
    with pytest.raises(api.KronmultError) as e:
        api.kronmult_batched_terms(6, 4, 0, db.mat_ptrs, 64, db.in_ptrs[:nb], db.out_ptrs[:nb], None, nb, dtype="f64")
    assert e.value.code == api.HIP_ERROR_INVALID_VALUE
    with pytest.raises(api.KronmultError):
        api.kronmult_batched_terms(6, 4, 2, None, 64, db.in_ptrs[:nb], db.out_ptrs[:nb], None, nb, dtype="f64")
    before = db.output()
    api.kronmult_batched_terms(6, 4, 2, db.mat_ptrs, 64, db.in_ptrs[:nb], db.out_ptrs[:nb], None, 0, dtype="f64")
    assert np.array_equal(db.output(), before)


def test_terms_graph_replays(gpu):
    """The stream-ordered several-term call captured in a graph matches its eager result on every replay."""
    import torch
    b = W.c2_terms(terms=2, rows=4, cols=64)
    db = DeviceBatch(b, device=str(gpu))
    db.run_terms(sync=False)
    torch.cuda.synchronize()
    eager = db.output()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        db.reset_output()
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            db.run_terms(sync=False, stream=s)
    for replay in range(3):
        db.reset_output()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert normwise(db.output(), eager) <= 1e-13, f"replay {replay}"
