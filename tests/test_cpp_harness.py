"""The rewritten reference harnesses (tests/kronmult_test_gpu.cpp, tests/kronmult_bench_gpu.cpp of the
reference) as executables under tests/cpp/bin, run on the GPU.

kronmult_test_gpu compares the C++ drop-in template against the oracle library on the reference's toy and
small cases (plus medium and d=6 n=4) and exits non-zero on failure; kronmult_bench_gpu is run on its
first three cases only (the d=6 n=8 cases need ~20 GB of device memory and tens of seconds of host-side
data generation, they belong to a benchmark run, not a test).
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "bin")


def _run(args, timeout):
    exe = os.path.join(BIN, args[0])
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (make -C tests/cpp)")
    p = subprocess.run([exe] + list(args[1:]), capture_output=True, text=True, timeout=timeout)
    print(p.stdout)
    print(p.stderr)
    return p


@pytest.mark.gpu
def test_cpp_kronmult_test_gpu(gpu):
    p = _run(["kronmult_test_gpu"], 120)
    assert p.returncode == 0, p.stderr
    assert "all cases passed" in p.stdout
    for name in ("toy", "small", "medium", "d6n4"):
        assert f"{name}: normwise relative error" in p.stdout


@pytest.mark.gpu
def test_cpp_kronmult_bench_gpu(gpu):
    p = _run(["kronmult_bench_gpu", "5", "3"], 120)
    assert p.returncode == 0, p.stderr
    for name in ("toy", "small", "medium"):
        assert name in p.stdout
    assert "GFLOP/s" in p.stdout


@pytest.mark.gpu
def test_cpp_kronmult_fullbench_gpu(gpu):
    """The full-grid benchmark on its small corner (degrees 2-3, all dimensions, level 2)."""
    p = _run(["kronmult_fullbench_gpu", "2", "2", "2", "3"], 120)
    assert p.returncode == 0, p.stderr
    assert p.stdout.count("GFLOP/s") == 12 and "fullbench done" in p.stdout


def test_cpp_harness_built():
    """The executables link against the product library (no GPU needed to check the link)."""
    for exe in ("kronmult_test_gpu", "kronmult_bench_gpu", "kronmult_fullbench_gpu"):
        path = os.path.join(BIN, exe)
        if not os.path.exists(path):
            pytest.skip("tests/cpp not built")
        out = subprocess.run(["ldd", path], capture_output=True, text=True).stdout
        assert "libkronmult_hip.so" in out and "not found" not in out.split("libkronmult_hip.so")[1].split("\n")[0]
