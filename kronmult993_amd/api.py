"""Host-side mirror of the reference's `kronmult_batched` interface, over the C ABI.

`kronmult_batched(matrix_count, matrix_size, matrix_list_batched, matrix_stride, input_batched,
output_batched, workspace_batched, nb_batch)` takes the same arguments, in the same order and with the same
meaning, as the reference (kronmult_gpu/kronmult.cuh:28-32): pointer arrays of device addresses (here a
torch int64 tensor on the device, or an integer address of such an array), column-major factors with
leading dimension `matrix_stride`, outputs accumulated and possibly aliased. The element type, a template
parameter in C++, is the keyword `dtype`.

Errors: the reference returns the cudaError of cudaDeviceSynchronize and its harness throws on failure
(tests/utils/utils_gpu.h:11-17). Here a non-zero HIP status raises `KronmultError` carrying the code.
"""
from __future__ import annotations

from typing import Any, Optional

import ctypes

from ._lib import GroupedOptions, GroupF32, GroupF64, LaunchInfo, Tuning, lib

HIP_SUCCESS = 0
HIP_ERROR_INVALID_VALUE = 1
HIP_ERROR_OUT_OF_MEMORY = 2
NOALLOC_SCRATCH_BYTES = 96 * 1024 * 1024  # KRONMULT_NOALLOC_SCRATCH_BYTES


class KronmultError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: hipError {code}")
        self.code = code


def _addr(x: Any) -> Optional[int]:
    """Device address of a pointer array: torch tensor -> data_ptr(), int -> itself, None -> NULL."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    ptr = getattr(x, "data_ptr", None)
    if ptr is not None:
        return int(ptr())
    raise TypeError(f"cannot take the address of {type(x)!r}")


def _suffix(dtype: Any) -> str:
    s = str(dtype).replace("torch.", "").replace("numpy.", "")
    if s in ("float64", "double", "f64", "<class 'numpy.float64'>"):
        return "f64"
    if s in ("float32", "float", "f32", "<class 'numpy.float32'>"):
        return "f32"
    raise TypeError(f"kronmult_batched is provided for float32 and float64 only, not {dtype}")


def _stream_handle(stream: Any) -> Optional[int]:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream or None
    h = getattr(stream, "cuda_stream", None)  # torch.cuda.Stream
    if h is not None:
        return int(h) or None
    raise TypeError(f"not a stream: {type(stream)!r}")


def kronmult_batched(matrix_count: int, matrix_size: int, matrix_list_batched, matrix_stride: int,
                     input_batched, output_batched, workspace_batched, nb_batch: int, *, dtype) -> None:
    """Synchronous drop-in call (returns once the device is idle, like the reference)."""
    f = getattr(lib(), "kronmult_batched_" + _suffix(dtype))
    rc = f(matrix_count, matrix_size, _addr(matrix_list_batched), matrix_stride, _addr(input_batched),
           _addr(output_batched), _addr(workspace_batched), nb_batch)
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_batched")


def kronmult_batched_async(matrix_count: int, matrix_size: int, matrix_list_batched, matrix_stride: int,
                           input_batched, output_batched, workspace_batched, nb_batch: int, *, dtype,
                           stream=None) -> None:
    """Stream-ordered call; `stream` is a torch.cuda.Stream, a raw hipStream_t address, or None."""
    f = getattr(lib(), "kronmult_batched_" + _suffix(dtype) + "_async")
    rc = f(matrix_count, matrix_size, _addr(matrix_list_batched), matrix_stride, _addr(input_batched),
           _addr(output_batched), _addr(workspace_batched), nb_batch, _stream_handle(stream))
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_batched_async")


FAMILY_AUTO, FAMILY_REGISTER, FAMILY_FIBER, FAMILY_MATRIX, FAMILY_PACKED, FAMILY_WAVE = 0, 1, 2, 3, 4, 5


def _tuning(tuning) -> Optional[Tuning]:
    if tuning is None:
        return None
    if isinstance(tuning, Tuning):
        return tuning
    return Tuning(int(tuning.get("family", 0)), int(tuning.get("workgroups_per_cu", 0)),
                  int(tuning.get("entries_per_workgroup", 0)))


def kronmult_batched_tuned(matrix_count: int, matrix_size: int, matrix_list_batched, matrix_stride: int,
                           input_batched, output_batched, workspace_batched, nb_batch: int, *, dtype,
                           tuning=None, stream=None) -> None:
    """Stream-ordered call with an explicit launch shape (dict with family / workgroups_per_cu /
    entries_per_workgroup, or a Tuning)."""
    f = getattr(lib(), "kronmult_batched_" + _suffix(dtype) + "_tuned")
    tu = _tuning(tuning)
    rc = f(matrix_count, matrix_size, _addr(matrix_list_batched), matrix_stride, _addr(input_batched),
           _addr(output_batched), _addr(workspace_batched), nb_batch, ctypes.byref(tu) if tu else None,
           _stream_handle(stream))
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_batched_tuned")


def launch_info(matrix_count: int, matrix_size: int, element_size: int, nb_batch: int, tuning=None) -> dict:
    """The launch the library would make (needs the HIP runtime)."""
    info = LaunchInfo()
    tu = _tuning(tuning)
    rc = lib().kronmult_hip_launch_info(matrix_count, matrix_size, element_size, nb_batch,
                                        ctypes.byref(tu) if tu else None, ctypes.byref(info))
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_hip_launch_info")
    return info.as_dict()


def kronmult_batched_ordered(matrix_count: int, matrix_size: int, matrix_list_batched, matrix_stride: int,
                             input_batched, output_batched, workspace_batched, nb_batch: int, *, dtype,
                             tuning=None, stream=None, sync: bool = True) -> None:
    """Ordered accumulation (bitwise reproducible sums into shared outputs; entries of one output consecutive,
    a workspace per entry). sync=True: the synchronous entry point, which checks the layout on the host;
    sync=False: stream-ordered with an optional launch shape, layout trusted."""
    sfx = _suffix(dtype)
    if sync:
        if tuning is not None:
            raise ValueError("tuning needs sync=False")
        rc = getattr(lib(), "kronmult_batched_ordered_" + sfx)(
            matrix_count, matrix_size, _addr(matrix_list_batched), matrix_stride, _addr(input_batched),
            _addr(output_batched), _addr(workspace_batched), nb_batch)
    else:
        tu = _tuning(tuning)
        rc = getattr(lib(), "kronmult_batched_ordered_" + sfx + "_async")(
            matrix_count, matrix_size, _addr(matrix_list_batched), matrix_stride, _addr(input_batched),
            _addr(output_batched), _addr(workspace_batched), nb_batch, ctypes.byref(tu) if tu else None,
            _stream_handle(stream))
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_batched_ordered")


def kronmult_batched_ordered_noalloc(matrix_count: int, matrix_size: int, matrix_list_batched, matrix_stride: int,
                                     input_batched, output_batched, nb_batch: int, *, tuning=None,
                                     stream=None) -> None:
    """Ordered accumulation (FP64) with no workspace and no allocation: the partial sums live in the library's
    own device array (NOALLOC_SCRATCH_BYTES per GPU). Stream-ordered, layout trusted; raises KronmultError
    with code HIP_ERROR_OUT_OF_MEMORY when the batch needs more than that array."""
    tu = _tuning(tuning)
    rc = lib().kronmult_batched_ordered_f64_noalloc(
        matrix_count, matrix_size, _addr(matrix_list_batched), matrix_stride, _addr(input_batched),
        _addr(output_batched), nb_batch, ctypes.byref(tu) if tu else None, _stream_handle(stream))
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_batched_ordered_f64_noalloc")


def ordered_launch_info(matrix_count: int, matrix_size: int, element_size: int, nb_batch: int,
                        tuning=None) -> dict:
    """The launch of an ordered call (entries_per_workgroup: the range length of its summation order)."""
    info = LaunchInfo()
    tu = _tuning(tuning)
    rc = lib().kronmult_hip_ordered_launch_info(matrix_count, matrix_size, element_size, nb_batch,
                                                ctypes.byref(tu) if tu else None, ctypes.byref(info))
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_hip_ordered_launch_info")
    return info.as_dict()


def kronmult_batched_terms(matrix_count: int, matrix_size: int, term_count: int, matrix_list_batched,
                           matrix_stride: int, input_batched, output_batched, workspace_batched, nb_batch: int, *,
                           dtype, stream=None, sync: bool = True) -> None:
    """output[k] += sum_t kron(A_{t,k,*}) input[k]; factor pointers term-major (term t of entry k at
    (t*nb_batch + k)*matrix_count). Synchronous, or stream-ordered with sync=False."""
    sfx = _suffix(dtype)
    args = (matrix_count, matrix_size, term_count, _addr(matrix_list_batched), matrix_stride, _addr(input_batched),
            _addr(output_batched), _addr(workspace_batched), nb_batch)
    if sync:
        rc = getattr(lib(), "kronmult_batched_terms_" + sfx)(*args)
    else:
        rc = getattr(lib(), "kronmult_batched_terms_" + sfx + "_async")(*args, _stream_handle(stream))
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_batched_terms")


ADJOINT_AUTO, ADJOINT_RECOMPUTE, ADJOINT_STORE = 0, 1, 2
_INTERMEDIATES = {"auto": ADJOINT_AUTO, "recompute": ADJOINT_RECOMPUTE, "store": ADJOINT_STORE}


def _intermediates(v) -> int:
    return _INTERMEDIATES[v] if isinstance(v, str) else int(v)


def kronmult_batched_adjoint(matrix_count: int, matrix_size: int, matrix_list_batched, matrix_stride: int,
                             input_batched, grad_output_batched, grad_input_batched, grad_matrix_batched,
                             workspace_batched, nb_batch: int, *, dtype, intermediates="auto", stream=None,
                             sync: bool = True) -> None:
    """Adjoint of kronmult_batched (no reference counterpart): accumulates K^T g into grad_input_batched[k]
    and dL/dA into grad_matrix_batched[k*d+t] (either may be None). `intermediates`: "auto", "recompute"
    or "store". Synchronous unless sync=False (then stream-ordered on `stream`)."""
    sfx = _suffix(dtype)
    args = [matrix_count, matrix_size, _addr(matrix_list_batched), matrix_stride, _addr(input_batched),
            _addr(grad_output_batched), _addr(grad_input_batched), _addr(grad_matrix_batched),
            _addr(workspace_batched), nb_batch, _intermediates(intermediates)]
    if sync:
        rc = getattr(lib(), "kronmult_batched_adjoint_" + sfx)(*args)
    else:
        rc = getattr(lib(), "kronmult_batched_adjoint_" + sfx + "_async")(*args, _stream_handle(stream))
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_batched_adjoint")


def adjoint_workspace(matrix_count: int, matrix_size: int, element_size: int, intermediates="auto") -> int:
    """Workspace elements per entry the adjoint reads (0: none); raises for unsupported arguments."""
    v = lib().kronmult_batched_adjoint_workspace(matrix_count, matrix_size, element_size,
                                                  _intermediates(intermediates))
    if v < 0:
        raise KronmultError(HIP_ERROR_INVALID_VALUE, "kronmult_batched_adjoint_workspace")
    return int(v)


GROUPED_AUTO, GROUPED_PERSISTENT, GROUPED_SEPARATE = 0, 1, 2
_STRATEGIES = {"auto": GROUPED_AUTO, "persistent": GROUPED_PERSISTENT, "separate": GROUPED_SEPARATE}


class Group:
    """Arguments of one kronmult_batched call inside a grouped call, plus launch hints (`tuning` for a
    launch of its own, `entries_per_tile` for the persistent launch)."""

    def __init__(self, matrix_count, matrix_size, matrix_list_batched, matrix_stride, input_batched,
                 output_batched, workspace_batched, nb_batch, tuning=None, entries_per_tile=0):
        self.args = (matrix_count, matrix_size, matrix_stride, nb_batch, _addr(matrix_list_batched),
                     _addr(input_batched), _addr(output_batched), _addr(workspace_batched))
        self.tuning = _tuning(tuning) or Tuning(0, 0, 0)
        self.entries_per_tile = int(entries_per_tile)


def grouped_structs(groups, dtype):
    """The ctypes array of kronmult_group_*_t for `groups` (kept by the caller while it is in use)."""
    S = GroupF64 if _suffix(dtype) == "f64" else GroupF32
    arr = (S * max(1, len(groups)))()
    for i, g in enumerate(groups):
        d, n, stride, nb, m, x, y, w = g.args
        arr[i] = S(d, n, stride, nb, m, x, y, w, g.tuning, g.entries_per_tile)
    return arr


def kronmult_batched_grouped(groups, *, dtype, strategy="auto", workgroups_per_cu: int = 0, stream=None,
                             structs=None) -> None:
    """Stream-ordered grouped call (no synchronisation). `structs`: a prebuilt grouped_structs() array
    (saves rebuilding it on every call)."""
    arr = structs if structs is not None else grouped_structs(groups, dtype)
    st = _STRATEGIES[strategy] if isinstance(strategy, str) else int(strategy)
    opt = GroupedOptions(st, int(workgroups_per_cu))
    f = getattr(lib(), "kronmult_batched_grouped_" + _suffix(dtype))
    rc = f(arr, len(groups), ctypes.byref(opt), _stream_handle(stream))
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_batched_grouped")


SPLIT_MODES = {"fp32": 0, "pair32": 1, "bf16": 2, "pairbf16": 3, "fixed3": 4}


def kronmult_batched_split(matrix_count: int, matrix_size: int, matrix_list_batched, matrix_stride: int,
                           input_batched, output_batched, nb_batch: int, *, mode="fixed3", stream=None) -> None:
    """FP64 data, contractions on the FP32 matrix cores (`mode`: fp32, pair32, bf16, pairbf16, fixed3);
    stream-ordered."""
    m = SPLIT_MODES[mode] if isinstance(mode, str) else int(mode)
    rc = lib().kronmult_batched_split_f64(matrix_count, matrix_size, _addr(matrix_list_batched), matrix_stride,
                                          _addr(input_batched), _addr(output_batched), nb_batch, m,
                                          _stream_handle(stream))
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_batched_split_f64")


def build_element_batch(matrix_count: int, matrix_size: int, base: int, coeff, coeff_ld: int, row_ids, nrows: int,
                        col_ids, ncols: int, inputs, outputs, matrix_list_batched, input_batched, output_batched,
                        *, dtype, stream=None) -> None:
    """Device-side construction of an element-pair operator batch's pointer arrays (stream-ordered)."""
    f = getattr(lib(), "kronmult_build_element_batch_" + _suffix(dtype))
    rc = f(matrix_count, matrix_size, base, _addr(coeff), coeff_ld, _addr(row_ids), nrows, _addr(col_ids), ncols,
           _addr(inputs), _addr(outputs), _addr(matrix_list_batched), _addr(input_batched), _addr(output_batched),
           _stream_handle(stream))
    if rc != HIP_SUCCESS:
        raise KronmultError(rc, "kronmult_build_element_batch")


def kernel_name(matrix_count: int, matrix_size: int, element_size: int, nb_batch: int) -> str:
    return lib().kronmult_hip_kernel_name(matrix_count, matrix_size, element_size, nb_batch).decode()


def version() -> str:
    return lib().kronmult_hip_version().decode()
