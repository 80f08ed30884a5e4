"""CPU tests of the host side: generator, workload builders, the C-ABI library's exports and its argument
checks (which return before touching a device)."""
import ctypes
import os
import re

import numpy as np
import pytest

from kronmult993_amd import _lib, gen, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---- generator --------------------------------------------------------------------------------------

def test_generator_is_deterministic_and_counter_based():
    a = gen.uniform(0, 1, 1000)
    assert np.array_equal(a, gen.uniform(0, 1, 1000))
    assert np.array_equal(a[500:], gen.uniform(0, 1, 500, offset=500))
    assert not np.array_equal(a, gen.uniform(0, 2, 1000))
    assert not np.array_equal(a, gen.uniform(1, 1, 1000))


def test_splitmix64_known_answer():
    """SplitMix64's published finaliser: mix64(0x9E3779B97F4A7C15) (the first output of the generator seeded
    with 0) is 0xE220A8397B1DCDAF."""
    z = gen._mix64(np.array([0x9E3779B97F4A7C15], dtype=np.uint64))
    assert int(z[0]) == 0xE220A8397B1DCDAF


def test_generator_distributions():
    u = gen.uniform(3, 4, 200000)
    assert u.min() >= -1.0 and u.max() < 1.0
    assert abs(u.mean()) < 0.01 and abs(u.var() - 1.0 / 3.0) < 0.01
    z = gen.normal(3, 5, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01


# ---- workloads ----------------------------------------------------------------------------------------

def test_c2_digit_map():
    b = W.c2()
    assert (b.d, b.n, b.stride, b.nb, b.out_count) == (6, 4, 64, 8192, 128)
    moff = b.mat_off.reshape(b.nb, b.d)
    for k in (0, 1, 63, 64, 1000, 8191):
        i, j = divmod(k, 64)
        assert b.in_off[k] == j * 4096 and b.out_index[k] == i
        for t in range(6):
            r, c = (i >> (4 * t)) & 15, (j >> (4 * t)) & 15
            assert moff[k, t] == t * 4096 + 4 * r + 4 * c * 64
    assert np.all(moff[:, 2:] % 4096 == 0)  # digits 2..5 of ids < 256 are zero


def test_config_shapes():
    c1 = W.c1()
    assert (c1.d, c1.n, c1.nb, c1.out_count) == (3, 4, 4096, 512)
    assert np.array_equal(np.bincount(c1.out_index), np.full(512, 8))
    c4 = W.c4(nb=64)
    assert (c4.d, c4.n, c4.size) == (4, 8, 4096)
    scale = np.abs(c4.mat_pool.reshape(-1, 64)).max(axis=1)
    assert scale.min() < 0.1 and scale.max() > 10.0
    c5 = W.c5(nb=1024)
    assert np.array_equal(c5.mat_off[:3], c5.mat_off[256 * 3: 256 * 3 + 3])
    assert c5.grad_out_pool.shape == (256 * 512,)
    groups = W.c3()
    assert [(g.d, g.n, g.nb) for g in groups] == [(2, 8, 20000), (3, 6, 6000), (4, 4, 3000), (5, 3, 1500),
                                                (6, 2, 800)]
    assert all(np.array_equal(np.bincount(g.out_index), np.full(g.nb // 4, 4)) for g in groups)


def test_reference_batch_size_model():
    """tests/utils/batch_size.h:8-21 for the reference harness cases."""
    assert W.compute_batch_size(4, 1, 2, 5) == 4
    assert W.compute_batch_size(4, 2, 4, 5) == 64
    assert W.compute_batch_size(6, 3, 6, 5) == 384
    assert W.compute_batch_size(8, 6, 7, 5) == 896
    assert W.compute_batch_size(8, 6, 9, 5) == 3903  # the published "large"/"realistic" batch


def test_reference_alias_map():
    """Entry i >= distinct goes to output (i*distinct)/N (tests/utils/utils_cpu.h:79-83)."""
    b = W.reference_case(4, 2, 4)
    assert b.nb == 64 and b.out_count == 5
    expect = [i if i < 5 else (i * 5) // 64 for i in range(64)]
    assert list(b.out_index) == expect


def test_flop_and_byte_model():
    assert W.flops_per_entry(6, 4) == 200704
    assert W.bytes_per_entry(6, 4, 8) == 99136
    assert W.flops_per_entry(6, 8) == 25427968
    assert W.c2().flops() == 8192 * 200704


# ---- C ABI ----------------------------------------------------------------------------------------------

def _declared_functions():
    text = open(os.path.join(ROOT, "include", "kronmult_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kronmult_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    handle = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared_functions()
    assert "kronmult_batched_f64" in names and "kronmult_batched_f32" in names
    for name in names:
        assert hasattr(handle, name), name
    assert set(names) <= set(_lib.SIGNATURES), "ctypes signatures out of date"


def test_invalid_arguments_are_rejected_before_any_device_call():
    lib = _lib.lib()
    dummy = ctypes.c_void_p(16)
    for d, n, stride, nb in ((0, 4, 4, 1), (2, 0, 4, 1), (2, 4, 3, 1), (2, 4, 4, -1), (40, 2, 2, 1)):
        assert lib.kronmult_batched_f64(d, n, dummy, stride, dummy, dummy, dummy, nb) == 1
        assert lib.kronmult_batched_f32_async(d, n, dummy, stride, dummy, dummy, dummy, nb, None) == 1
    # nb_batch == 0: success, nothing launched
    assert lib.kronmult_batched_f64(3, 4, None, 4, None, None, None, 0) == 0
    assert lib.kronmult_batched_f64_async(3, 4, None, 4, None, None, None, 0, None) == 0
    # null pointer arrays with work to do
    assert lib.kronmult_batched_f64_async(3, 4, None, 4, dummy, dummy, None, 5, None) == 1


def test_kernel_selection():
    from kronmult993_amd import api
    assert api.kernel_name(6, 4, 8, 8192) == "reg_sweep_f64"
    assert api.kernel_name(4, 8, 4, 16384) == "reg_sweep_f32"
    assert api.kernel_name(4, 8, 8, 16384) == "reg_sweep_mfma_f64"  # n = 8 FP64: matrix-core sweep
    assert api.kernel_name(3, 8, 8, 8192) == "reg_sweep_mfma_f64"
    assert api.kernel_name(3, 8, 4, 8192) == "reg_sweep_f32"
    assert api.kernel_name(3, 7, 8, 6000) == "packed_f64"      # n^(d-1) = 49 <= 64 fibers, large batch
    assert api.kernel_name(3, 6, 4, 6000) == "wave_f32"        # 216 elements: 4 per lane of one wavefront
    assert api.kernel_name(3, 6, 8, 6000) == "packed_f64"      # ... except in FP64 (152 VGPRs), measured slower
    assert api.kernel_name(3, 4, 8, 4096) == "wave_f64"        # n^d <= 64: one wavefront per entry
    assert api.kernel_name(2, 8, 8, 20000) == "wave_f64"
    assert api.kernel_name(6, 2, 4, 800) == "wave_f32"
    assert api.kernel_name(3, 7, 8, 100).startswith("fiber_lds")  # small batch: one workgroup per entry
    assert api.kernel_name(4, 5, 4, 6000).startswith("fiber_lds")  # 125 fibers: fills a workgroup alone
    assert api.kernel_name(6, 8, 8, 3903) == "multipass_f64"
    assert api.kernel_name(3, 21, 8, 6) == "fiber_global_f64"
    assert api.kernel_name(2, 100, 8, 1) == "wide_global_f64"  # n > 64: wide fibers, vector in the workspace
    assert api.kernel_name(1, 100, 8, 1) == "wide_lds_f64"
    assert api.kernel_name(1, 4097, 8, 1) == "unsupported"    # a fiber above the 32 KiB staging chunk
    assert api.kernel_name(1, 8192, 4, 1) == "wide_lds_f32" or api.kernel_name(1, 8192, 4, 1) == "wide_global_f32"
    assert "gfx950" in api.version()


def test_extension_entry_points_reject_invalid_arguments_before_any_device_call():
    lib = _lib.lib()
    dummy = ctypes.c_void_p(16)
    # adjoint: n > 8 is unsupported, intermediates out of range, shape errors
    assert lib.kronmult_batched_adjoint_f64(2, 9, dummy, 9, dummy, dummy, dummy, dummy, None, 3, 0) == 1
    assert lib.kronmult_batched_adjoint_f64(2, 4, dummy, 4, dummy, dummy, dummy, dummy, None, 3, 7) == 1
    assert lib.kronmult_batched_adjoint_f32_async(0, 4, dummy, 4, dummy, dummy, dummy, dummy, None, 3, 0,
                                                  None) == 1
    assert lib.kronmult_batched_adjoint_workspace(2, 9, 8, 0) == -1
    assert lib.kronmult_batched_adjoint_workspace(2, 4, 2, 0) == -1
    assert lib.kronmult_batched_adjoint_f64(3, 4, None, 4, None, None, None, None, None, 0, 0) == 0
    # FP64 through FP32 matrix cores: n > 8, mode out of range, vector beyond the on-chip size
    assert lib.kronmult_batched_split_f64(2, 9, dummy, 9, dummy, dummy, 3, 4, None) == 1
    assert lib.kronmult_batched_split_f64(2, 4, dummy, 4, dummy, dummy, 3, 9, None) == 1
    assert lib.kronmult_batched_split_f64(5, 8, dummy, 8, dummy, dummy, 3, 4, None) == 1
    # grouped: strategy out of range, an invalid group, a negative tile length
    from kronmult993_amd import api
    good = api.Group(2, 3, 16, 3, 16, 16, None, 4)
    for groups, opt in (([good], _lib.GroupedOptions(7, 0)),
                        ([good, api.Group(0, 3, 16, 3, 16, 16, None, 4)], _lib.GroupedOptions(1, 0)),
                        ([api.Group(2, 3, 16, 3, 16, 16, None, 4, entries_per_tile=-1)], _lib.GroupedOptions(1, 0))):
        arr = api.grouped_structs(groups, "f64")
        assert lib.kronmult_batched_grouped_f64(arr, len(groups), ctypes.byref(opt), None) == 1
    assert lib.kronmult_batched_grouped_f64(api.grouped_structs([], "f64"), 0, None, None) == 0


def test_autotune_keys_and_pass_rule():
    from kronmult993_amd import autotune
    assert [autotune.bucket(v) for v in (0, 1, 3, 4, 5, 8192, 8193)] == [1, 1, 4, 4, 8, 8192, 16384]
    assert autotune.key_single("gfx950", "f64", 6, 4, 8000) == "gfx950|f64|d6n4|nb8192"
    # entries per pass of the persistent grouped launch: 256 threads / n^(d-1) fibers, within 64 KiB of LDS
    assert autotune.entries_per_pass(2, 8, 8) == 32
    assert autotune.entries_per_pass(3, 6, 8) == 7
    assert autotune.entries_per_pass(4, 4, 8) == 4
    assert autotune.entries_per_pass(6, 4, 8) == 1


def test_builder_rejects_invalid_arguments_before_any_device_call():
    lib = _lib.lib()
    p = ctypes.c_void_p(16)
    args = lambda d, n, base, ld, nr, nc: (d, n, base, p, ld, p, nr, p, nc, p, p, p, p, p, None)
    assert lib.kronmult_build_element_batch_f64(*args(0, 4, 16, 64, 2, 2)) == 1
    assert lib.kronmult_build_element_batch_f64(*args(6, 4, 16, 63, 2, 2)) == 1  # ld < n * base
    assert lib.kronmult_build_element_batch_f64(*args(6, 4, 0, 64, 2, 2)) == 1
    assert lib.kronmult_build_element_batch_f32(*args(6, 4, 16, 64, -1, 2)) == 1
    assert lib.kronmult_build_element_batch_f64(*args(6, 4, 16, 64, 0, 5)) == 0  # empty batch


def test_ordered_reference_order_and_grouping():
    """The ordered-sum model used by the ordered-accumulation tests: regrouping entries by output and any
    range length give the oracle's result up to rounding; with one range per entry on integer-valued data
    it is exact."""
    import oracle
    from helpers import grouped_by_output, normwise, ordered_reference
    b = W.random_case(71, 3, 4, 50, distinct=6)
    g = grouped_by_output(b)
    assert np.all(np.diff(g.out_index) >= 0)
    ref = oracle.kronmult(b)
    assert normwise(oracle.kronmult(g), ref) <= 1e-14
    for epw in (1, 3, 7, 50):
        assert normwise(ordered_reference(g, epw), ref) <= 1e-14
    ints = W.HostBatch(**{**g.__dict__, "mat_pool": np.round(g.mat_pool * 4), "in_pool": np.round(g.in_pool * 4),
                          "out_init": np.round(g.out_init * 4)})
    assert np.array_equal(ordered_reference(ints, 4), oracle.kronmult(ints))
