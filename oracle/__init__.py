"""CPU oracle for `kronmult_batched` -- TEST INFRASTRUCTURE ONLY.

Python bindings over two CPU libraries built by oracle/Makefile:
  * liboracle.so: the project's C restatement of the reference arithmetic (kronmult_oracle.c), which
    follows kronmult_omp/kronmult.hpp:32-104 and kronmult_omp/linear_algebra.hpp:78-121 (sweep) and
    tests/utils/kronmult_naive.h:31-121 (dense oracle), plus an adjoint and an extended-precision sweep;
  * _ref/libkronmult_ref.so: the reference's own kronmult_omp header and dense test oracle compiled from
    their sources (present only where it was built).

Only the test suite, the smoke check and the benchmark's CPU-baseline leg import this package, and only
as the checker / the timed CPU baseline. The GPU library never depends on it.

Batches are `kronmult993_amd.workloads.HostBatch` objects: flat pools plus element offsets. Each call
turns them into host pointer arrays.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
# reference builds: "shipped" = the reference's own flags; "fused" / "unfused" = the two arithmetic
# readings of `dotprod += x*a`; "avx512" = the reference's flags for x86-64-v4 (see oracle/Makefile)
REF_VARIANTS = {"shipped": "libkronmult_ref.so", "fused": "libkronmult_ref_fused.so",
                "unfused": "libkronmult_ref_unfused.so", "avx512": "libkronmult_ref_avx512.so"}
REF_SO = os.path.join(REF_DIR, REF_VARIANTS["shipped"])

_lib = None
_ref = {}


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref when the reference tree is present) with make."""
    import subprocess
    args = ["make", "-C", HERE, "all"]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _lib = ctypes.CDLL(ORACLE_SO)
        base = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                ctypes.c_int]
        for name in ("oracle_kronmult_batched_f64", "oracle_kronmult_batched_f32",
                     "oracle_dense_kronmult_batched_f64", "oracle_dense_kronmult_batched_f32"):
            f = getattr(_lib, name)
            f.restype = ctypes.c_int
            f.argtypes = base + [ctypes.c_int]
        _lib.oracle_kronmult_batched_hp.restype = ctypes.c_int
        _lib.oracle_kronmult_batched_hp.argtypes = base
        for name in ("oracle_kronmult_batched_omp_f64", "oracle_kronmult_batched_omp_f32"):
            f = getattr(_lib, name)
            f.restype = ctypes.c_int
            f.argtypes = base + [ctypes.c_int, ctypes.c_int]
        for name in ("oracle_kronmult_adjoint_f64", "oracle_kronmult_adjoint_f32"):
            f = getattr(_lib, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    return _lib


def have_ref(variant: str = "shipped") -> bool:
    return os.path.exists(os.path.join(REF_DIR, REF_VARIANTS[variant]))


def host_has_avx512() -> bool:
    """True when this host's CPUs report AVX-512F (the avx512 build's baseline instruction set)."""
    try:
        with open("/proc/cpuinfo") as f:
            return any(line.startswith("flags") and " avx512f" in line for line in f)
    except OSError:
        return False


def ref_lib(variant: str = "shipped") -> ctypes.CDLL:
    if variant not in _ref:
        path = os.path.join(REF_DIR, REF_VARIANTS[variant])
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (needs the reference tree at build time)")
        lib_ = ctypes.CDLL(path)
        for name in ("ref_kronmult_batched_f64", "ref_kronmult_batched_f32"):
            f = getattr(lib_, name)
            f.restype = None
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        for name in ("ref_kronmult_batched_naive_f64", "ref_kronmult_batched_naive_f32"):
            f = getattr(lib_, name)
            f.restype = None
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                          ctypes.c_void_p, ctypes.c_int]
        lib_.ref_omp_max_threads.restype = ctypes.c_int
        _ref[variant] = lib_
    return _ref[variant]


def ptrs(pool: np.ndarray, off: np.ndarray) -> np.ndarray:
    """uint64 array of host addresses pool + off (elements)."""
    assert pool.flags.c_contiguous
    return (np.uint64(pool.ctypes.data) + off.astype(np.uint64) * np.uint64(pool.itemsize)).astype(np.uint64)


def _addr(a: Optional[np.ndarray]):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _sfx(dtype) -> str:
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "f64"
    if dtype == np.float32:
        return "f32"
    raise TypeError(f"unsupported dtype {dtype}")


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise ValueError(f"{what}: invalid arguments (rc={rc})")


def kronmult(b, out_pool: Optional[np.ndarray] = None, fused: bool = True) -> np.ndarray:
    """Sequential restatement (== the reference on one thread); returns the output pool.
    `fused` selects the arithmetic reading of each dot-product term (fma vs rounded product + add)."""
    out = b.output_pool_init() if out_pool is None else out_pool
    m, i, o = ptrs(b.mat_pool, b.mat_off), ptrs(b.in_pool, b.in_off), ptrs(out, b.out_off)
    f = getattr(lib(), "oracle_kronmult_batched_" + _sfx(b.dtype))
    _check(f(b.d, b.n, _addr(m), b.stride, _addr(i), _addr(o), b.nb, int(fused)), "oracle_kronmult_batched")
    return out


def kronmult_omp(b, nthreads: int, out_pool: Optional[np.ndarray] = None, fused: bool = True) -> np.ndarray:
    out = b.output_pool_init() if out_pool is None else out_pool
    m, i, o = ptrs(b.mat_pool, b.mat_off), ptrs(b.in_pool, b.in_off), ptrs(out, b.out_off)
    f = getattr(lib(), "oracle_kronmult_batched_omp_" + _sfx(b.dtype))
    _check(f(b.d, b.n, _addr(m), b.stride, _addr(i), _addr(o), b.nb, nthreads, int(fused)),
           "oracle_kronmult_batched_omp")
    return out


def dense(b, out_pool: Optional[np.ndarray] = None, fused: bool = True) -> np.ndarray:
    """Dense Kronecker oracle (n^d <= 4096)."""
    out = b.output_pool_init() if out_pool is None else out_pool
    m, i, o = ptrs(b.mat_pool, b.mat_off), ptrs(b.in_pool, b.in_off), ptrs(out, b.out_off)
    f = getattr(lib(), "oracle_dense_kronmult_batched_" + _sfx(b.dtype))
    _check(f(b.d, b.n, _addr(m), b.stride, _addr(i), _addr(o), b.nb, int(fused)), "oracle_dense_kronmult_batched")
    return out


def kronmult_hp(b, out_pool: Optional[np.ndarray] = None) -> np.ndarray:
    """Extended-precision (long double) sweep on float64 data."""
    b64 = b.astype(np.float64)
    out = b64.output_pool_init() if out_pool is None else out_pool
    m, i, o = ptrs(b64.mat_pool, b64.mat_off), ptrs(b64.in_pool, b64.in_off), ptrs(out, b64.out_off)
    _check(lib().oracle_kronmult_batched_hp(b.d, b.n, _addr(m), b.stride, _addr(i), _addr(o), b.nb),
           "oracle_kronmult_batched_hp")
    return out


def adjoint(b, grad_out_pool: Optional[np.ndarray] = None, want_input: bool = True,
            want_mats: bool = True, fused: bool = True) -> Tuple[Optional[np.ndarray], Optional[np.ndarray]]:
    """Gradients of sum(grad_out * output) w.r.t. the input pool and the factor pool (same layouts as
    `in_pool` / `mat_pool`; shared storage accumulates)."""
    g = b.grad_out_pool if grad_out_pool is None else grad_out_pool
    g = np.ascontiguousarray(g, dtype=b.dtype)
    gi = np.zeros_like(b.in_pool) if want_input else None
    gm = np.zeros_like(b.mat_pool) if want_mats else None
    m, i, go = ptrs(b.mat_pool, b.mat_off), ptrs(b.in_pool, b.in_off), ptrs(g, b.out_off)
    gip = ptrs(gi, b.in_off) if want_input else None
    gmp = ptrs(gm, b.mat_off) if want_mats else None
    f = getattr(lib(), "oracle_kronmult_adjoint_" + _sfx(b.dtype))
    _check(f(b.d, b.n, _addr(m), b.stride, _addr(i), _addr(go), _addr(gip), _addr(gmp), b.nb, int(fused)),
           "oracle_kronmult_adjoint")
    return gi, gm


def ref_kronmult(b, nthreads: int = 1, out_pool: Optional[np.ndarray] = None,
                 variant: str = "shipped") -> np.ndarray:
    """The reference's kronmult_omp (compiled from its header). It clobbers inputs and workspaces
    (kronmult_omp/README.md), and entries that share an input would race, so every entry gets private
    copies here."""
    out = b.output_pool_init() if out_pool is None else out_pool
    size = b.size
    priv_in = np.ascontiguousarray(b.in_pool[(b.in_off[:, None] + np.arange(size)[None, :])].reshape(-1))
    ws = np.zeros(b.nb * size, dtype=b.dtype)
    offs = np.arange(b.nb, dtype=np.int64) * size
    m, i, o, w = ptrs(b.mat_pool, b.mat_off), ptrs(priv_in, offs), ptrs(out, b.out_off), ptrs(ws, offs)
    f = getattr(ref_lib(variant), "ref_kronmult_batched_" + _sfx(b.dtype))
    f(b.d, b.n, _addr(m), b.stride, _addr(i), _addr(o), _addr(w), b.nb, nthreads)
    return out


class RefRunner:
    """Repeated timed calls of the reference's kronmult_omp on one batch. The reference clobbers its
    inputs and workspaces, so every entry gets a private input copy and workspace, restored from pristine
    copies by `reset()` (outside any timer); `run()` is the bare reference call."""

    def __init__(self, b, variant: str = "shipped"):
        self.b = b
        size = b.size
        self.pristine = np.ascontiguousarray(b.in_pool[(b.in_off[:, None] + np.arange(size)[None, :])].reshape(-1))
        self.work = self.pristine.copy()
        self.ws = np.zeros(b.nb * size, dtype=b.dtype)
        self.out = b.output_pool_init()
        offs = np.arange(b.nb, dtype=np.int64) * size
        self._m = ptrs(b.mat_pool, b.mat_off)
        self._i = ptrs(self.work, offs)
        self._o = ptrs(self.out, b.out_off)
        self._w = ptrs(self.ws, offs)
        self._f = getattr(ref_lib(variant), "ref_kronmult_batched_" + _sfx(b.dtype))

    def reset(self) -> None:
        np.copyto(self.work, self.pristine)
        self.out[:] = self.b.output_pool_init()

    def run(self, nthreads: int) -> None:
        b = self.b
        self._f(b.d, b.n, _addr(self._m), b.stride, _addr(self._i), _addr(self._o), _addr(self._w), b.nb, nthreads)


def ref_naive(b, out_pool: Optional[np.ndarray] = None, variant: str = "shipped") -> np.ndarray:
    """The reference's dense test oracle (kronmult_batched_naive)."""
    out = b.output_pool_init() if out_pool is None else out_pool
    m, i, o = ptrs(b.mat_pool, b.mat_off), ptrs(b.in_pool, b.in_off), ptrs(out, b.out_off)
    f = getattr(ref_lib(variant), "ref_kronmult_batched_naive_" + _sfx(b.dtype))
    f(b.d, b.n, _addr(m), b.stride, _addr(i), _addr(o), b.nb)
    return out
