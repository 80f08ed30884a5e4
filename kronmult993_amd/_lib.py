"""ctypes loader for libkronmult_hip.so (the C ABI declared in include/kronmult_hip.h).

The library is built in-tree by kronmult993_amd/csrc/Makefile. There is no fallback: if it is missing the
import of any compute entry point raises, so nothing can silently run on a CPU path.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkronmult_hip.so")
INCLUDE_DIR = os.path.join(os.path.dirname(HERE), "include")

_lib = None
_lock = threading.Lock()

_P = ctypes.c_void_p
_I = ctypes.c_int


class Tuning(ctypes.Structure):
    """kronmult_tuning_t: family (0 auto, 1 register, 2 fiber), workgroups per CU, entries per workgroup;
    zero fields keep the library default."""
    _fields_ = [("family", _I), ("workgroups_per_cu", _I), ("entries_per_workgroup", _I)]


class LaunchInfo(ctypes.Structure):
    """kronmult_launch_info_t."""
    _fields_ = [("family", _I), ("grid", _I), ("block", _I), ("entries_per_workgroup", _I), ("lds_bytes", _I),
                ("resident_per_cu", _I), ("uses_workspace", _I)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}

class GroupedOptions(ctypes.Structure):
    """kronmult_grouped_options_t: strategy (0 auto, 1 persistent, 2 separate), workgroups per CU."""
    _fields_ = [("strategy", _I), ("workgroups_per_cu", _I)]


def _group_struct(name):
    return type(name, (ctypes.Structure,), {"_fields_": [
        ("matrix_count", _I), ("matrix_size", _I), ("matrix_stride", _I), ("nb_batch", _I),
        ("matrix_list_batched", _P), ("input_batched", _P), ("output_batched", _P), ("workspace_batched", _P),
        ("tuning", Tuning), ("entries_per_tile", _I)]})


GroupF64 = _group_struct("GroupF64")  # kronmult_group_f64_t
GroupF32 = _group_struct("GroupF32")  # kronmult_group_f32_t

# name -> (restype, argtypes); the order of arguments is the one of include/kronmult_hip.h
SIGNATURES = {
    "kronmult_batched_f64": (_I, [_I, _I, _P, _I, _P, _P, _P, _I]),
    "kronmult_batched_f32": (_I, [_I, _I, _P, _I, _P, _P, _P, _I]),
    "kronmult_batched_f64_async": (_I, [_I, _I, _P, _I, _P, _P, _P, _I, _P]),
    "kronmult_batched_f32_async": (_I, [_I, _I, _P, _I, _P, _P, _P, _I, _P]),
    "kronmult_batched_f64_tuned": (_I, [_I, _I, _P, _I, _P, _P, _P, _I, ctypes.POINTER(Tuning), _P]),
    "kronmult_batched_f32_tuned": (_I, [_I, _I, _P, _I, _P, _P, _P, _I, ctypes.POINTER(Tuning), _P]),
    "kronmult_hip_launch_info": (_I, [_I, _I, _I, _I, ctypes.POINTER(Tuning), ctypes.POINTER(LaunchInfo)]),
    "kronmult_batched_ordered_f64": (_I, [_I, _I, _P, _I, _P, _P, _P, _I]),
    "kronmult_batched_ordered_f32": (_I, [_I, _I, _P, _I, _P, _P, _P, _I]),
    "kronmult_batched_ordered_f64_async": (_I, [_I, _I, _P, _I, _P, _P, _P, _I, ctypes.POINTER(Tuning), _P]),
    "kronmult_batched_ordered_f32_async": (_I, [_I, _I, _P, _I, _P, _P, _P, _I, ctypes.POINTER(Tuning), _P]),
    "kronmult_batched_ordered_f64_noalloc": (_I, [_I, _I, _P, _I, _P, _P, _I, ctypes.POINTER(Tuning), _P]),
    "kronmult_hip_ordered_launch_info": (_I, [_I, _I, _I, _I, ctypes.POINTER(Tuning), ctypes.POINTER(LaunchInfo)]),
    "kronmult_batched_terms_f64": (_I, [_I, _I, _I, _P, _I, _P, _P, _P, _I]),
    "kronmult_batched_terms_f32": (_I, [_I, _I, _I, _P, _I, _P, _P, _P, _I]),
    "kronmult_batched_terms_f64_async": (_I, [_I, _I, _I, _P, _I, _P, _P, _P, _I, _P]),
    "kronmult_batched_terms_f32_async": (_I, [_I, _I, _I, _P, _I, _P, _P, _P, _I, _P]),
    "kronmult_batched_adjoint_f64": (_I, [_I, _I, _P, _I, _P, _P, _P, _P, _P, _I, _I]),
    "kronmult_batched_adjoint_f32": (_I, [_I, _I, _P, _I, _P, _P, _P, _P, _P, _I, _I]),
    "kronmult_batched_adjoint_f64_async": (_I, [_I, _I, _P, _I, _P, _P, _P, _P, _P, _I, _I, _P]),
    "kronmult_batched_adjoint_f32_async": (_I, [_I, _I, _P, _I, _P, _P, _P, _P, _P, _I, _I, _P]),
    "kronmult_batched_adjoint_workspace": (ctypes.c_longlong, [_I, _I, _I, _I]),
    "kronmult_batched_grouped_f64": (_I, [ctypes.POINTER(GroupF64), _I, ctypes.POINTER(GroupedOptions), _P]),
    "kronmult_batched_grouped_f32": (_I, [ctypes.POINTER(GroupF32), _I, ctypes.POINTER(GroupedOptions), _P]),
    "kronmult_batched_split_f64": (_I, [_I, _I, _P, _I, _P, _P, _I, _I, _P]),
    "kronmult_build_element_batch_f64": (_I, [_I, _I, _I, _P, _I, _P, _I, _P, _I, _P, _P, _P, _P, _P, _P]),
    "kronmult_build_element_batch_f32": (_I, [_I, _I, _I, _P, _I, _P, _I, _P, _I, _P, _P, _P, _P, _P, _P]),
    "kronmult_hip_version": (ctypes.c_char_p, []),
    "kronmult_hip_kernel_name": (ctypes.c_char_p, [_I, _I, _I, _I]),
}


class LibraryMissing(RuntimeError):
    pass


def build() -> None:
    """Compile the library with hipcc for gfx950 (in-tree)."""
    import subprocess
    subprocess.run(["make", "-C", os.path.join(HERE, "csrc"), "-B"], check=True)


def lib() -> ctypes.CDLL:
    """The loaded library; raises LibraryMissing if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise LibraryMissing(f"{LIB_PATH} is not built: run `make -C kronmult993_amd/csrc` "
                                     "(or __graft_entry__.build()) first")
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(handle, name)
                f.restype = res
                f.argtypes = args
            _lib = handle
    return _lib
