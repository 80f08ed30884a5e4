"""GPU check of the device-side batch builder (kronmult_build_element_batch_f64/_f32): the pointer arrays it
writes for the ASGarD-like d=6, n=4 operator are exactly the ones the host builds for that workload
(kronmult993_amd.workloads.c2: digit t of an element id, base 16, picks the 4x4 block of coefficient
matrix t), and an apply through them matches the apply through the host-built arrays. Also a second
geometry (d=3, n=5, base 3, scattered element ids) against a numpy restatement of the rule."""
import numpy as np
import pytest

from kronmult993_amd import api, workloads as W
from kronmult993_amd.device import DeviceBatch

pytestmark = pytest.mark.gpu


def _addrs(t, offs, itemsize):
    return np.int64(t.data_ptr()) + np.asarray(offs, dtype=np.int64) * np.int64(itemsize)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_builder_reproduces_the_c2_batch(gpu, dtype):
    import torch
    hb = W.c2().astype(dtype)
    db = DeviceBatch(hb, device=str(gpu))
    rows, cols, d, n, ld = 128, 64, hb.d, hb.n, hb.stride
    dev = torch.device(str(gpu))
    it = hb.dtype.itemsize
    coeff = torch.from_numpy(_addrs(db.mat, np.arange(d) * ld * ld, it)).to(dev)
    inputs = torch.from_numpy(_addrs(db.inp, np.arange(cols) * hb.size, it)).to(dev)
    outputs = torch.from_numpy(_addrs(db.out, np.arange(rows) * hb.size, it)).to(dev)
    row_ids = torch.arange(rows, dtype=torch.int32, device=dev)
    col_ids = torch.arange(cols, dtype=torch.int32, device=dev)
    mats = torch.zeros(hb.nb * d, dtype=torch.int64, device=dev)
    ins = torch.zeros(hb.nb, dtype=torch.int64, device=dev)
    outs = torch.zeros(hb.nb, dtype=torch.int64, device=dev)
    api.build_element_batch(d, n, 16, coeff, ld, row_ids, rows, col_ids, cols, inputs, outputs, mats, ins, outs,
                            dtype=db.dtype_name)
    torch.cuda.synchronize()
    assert torch.equal(mats, db.mat_ptrs) and torch.equal(ins, db.in_ptrs) and torch.equal(outs, db.out_ptrs)
    # apply through the device-built arrays, then through the host-built ones
    api.kronmult_batched(d, n, mats, ld, ins, outs, None, hb.nb, dtype=db.dtype_name)
    first = db.output()
    db.reset_output()
    db.run()
    second = db.output()
    scale = np.max(np.abs(second))
    assert np.max(np.abs(first - second)) <= (1e-13 if dtype == np.float64 else 1e-5) * scale


def test_builder_scattered_ids(gpu):
    import torch
    dev = torch.device(str(gpu))
    d, n, base, ld = 3, 5, 3, 17
    coeff_pool = torch.zeros(d * ld * ld, dtype=torch.float64, device=dev)
    vec_pool = torch.zeros(16, dtype=torch.float64, device=dev)
    rng = np.random.default_rng(3)
    row_ids = rng.integers(0, base ** d, 7).astype(np.int32)
    col_ids = rng.integers(0, base ** d, 5).astype(np.int32)
    coeff = torch.from_numpy(_addrs(coeff_pool, np.arange(d) * ld * ld, 8)).to(dev)
    inputs = torch.from_numpy(_addrs(vec_pool, np.arange(5), 8)).to(dev)
    outputs = torch.from_numpy(_addrs(vec_pool, 8 + np.arange(7), 8)).to(dev)
    nb = 7 * 5
    mats = torch.zeros(nb * d, dtype=torch.int64, device=dev)
    ins = torch.zeros(nb, dtype=torch.int64, device=dev)
    outs = torch.zeros(nb, dtype=torch.int64, device=dev)
    api.build_element_batch(d, n, base, coeff, ld, torch.from_numpy(row_ids).to(dev), 7,
                            torch.from_numpy(col_ids).to(dev), 5, inputs, outputs, mats, ins, outs, dtype="f64")
    torch.cuda.synchronize()
    base_addr = coeff_pool.data_ptr()
    want = []
    for a in range(7):
        for b in range(5):
            for t in range(d):
                r = (int(row_ids[a]) // base ** t) % base
                c = (int(col_ids[b]) // base ** t) % base
                want.append(base_addr + 8 * (t * ld * ld + n * r + n * c * ld))
    assert mats.cpu().numpy().tolist() == want
    assert ins.cpu().numpy().tolist() == [vec_pool.data_ptr() + 8 * b for a in range(7) for b in range(5)]
    assert outs.cpu().numpy().tolist() == [vec_pool.data_ptr() + 8 * (8 + a) for a in range(7) for b in range(5)]
