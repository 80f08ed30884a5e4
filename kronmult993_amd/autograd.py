"""Check-only torch.autograd.Function over the forward and adjoint C ABI entry points.

    out = kronmult(mat_pool, in_pool, out_init, layout)
        = out_init + scatter_k( kron(A_{k,0}, ..., A_{k,d-1}) @ in_k )

`mat_pool`, `in_pool` and `out_init` are flat device tensors (float64 or float32); `layout` says where each
entry's factors, input and output sit in them (element offsets, the same layout as `HostBatch`). The
backward pass calls kronmult_batched_adjoint for d(in_pool) and d(mat_pool); d(out_init) is the incoming
cotangent. Offsets may repeat (shared factors, inputs, outputs): gradients of shared storage are summed.

This exists to check the adjoint against central finite differences and the dense Kronecker gradient; it
is not a training API (no double backward, no graph capture, every call synchronises the device).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import api


def _torch():
    import torch
    return torch


@dataclass(frozen=True)
class Layout:
    d: int
    n: int
    stride: int
    mat_off: np.ndarray   # int64 [nb*d]
    in_off: np.ndarray    # int64 [nb]
    out_off: np.ndarray   # int64 [nb]
    intermediates: str = "auto"

    @property
    def nb(self) -> int:
        return int(self.in_off.shape[0])

    @property
    def size(self) -> int:
        return self.n ** self.d

    @staticmethod
    def of(hb, intermediates: str = "auto") -> "Layout":
        return Layout(hb.d, hb.n, hb.stride, hb.mat_off, hb.in_off, hb.out_off, intermediates)


def _ptrs(t, off: np.ndarray):
    torch = _torch()
    addr = np.int64(t.data_ptr()) + off.astype(np.int64) * np.int64(t.element_size())
    return torch.from_numpy(addr).to(t.device)


def _dtype_name(t) -> str:
    torch = _torch()
    if t.dtype == torch.float64:
        return "f64"
    if t.dtype == torch.float32:
        return "f32"
    raise TypeError(f"unsupported dtype {t.dtype}")


def _workspace(t, per_entry: int, nb: int):
    torch = _torch()
    if per_entry <= 0:
        return None, None
    ws = torch.zeros(nb * per_entry, dtype=t.dtype, device=t.device)
    return ws, _ptrs(ws, np.arange(nb, dtype=np.int64) * per_entry)


class KronmultFunction(_torch().autograd.Function):
    @staticmethod
    def forward(ctx, mat_pool, in_pool, out_init, layout: Layout):
        torch = _torch()
        L = layout
        mat_pool, in_pool = mat_pool.contiguous(), in_pool.contiguous()
        out = out_init.detach().clone().contiguous()
        dt = _dtype_name(out)
        uses_ws = api.kernel_name(L.d, L.n, out.element_size(), L.nb).startswith(("fiber_global", "multipass"))
        ws, ws_ptrs = _workspace(out, L.size if uses_ws else 0, L.nb)  # kept alive across the call
        api.kronmult_batched(L.d, L.n, _ptrs(mat_pool, L.mat_off), L.stride, _ptrs(in_pool, L.in_off),
                             _ptrs(out, L.out_off), ws_ptrs, L.nb, dtype=dt)
        del ws
        ctx.save_for_backward(mat_pool, in_pool)
        ctx.layout = L
        return out

    @staticmethod
    def backward(ctx, grad_out):
        torch = _torch()
        mat_pool, in_pool = ctx.saved_tensors
        L = ctx.layout
        g = grad_out.contiguous()
        want_mats, want_input = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        gm = torch.zeros_like(mat_pool) if want_mats else None
        gi = torch.zeros_like(in_pool) if want_input else None
        if want_mats or want_input:
            per = api.adjoint_workspace(L.d, L.n, g.element_size(), L.intermediates)
            ws, ws_ptrs = _workspace(g, per, L.nb)  # kept alive across the call
            api.kronmult_batched_adjoint(
                L.d, L.n, _ptrs(mat_pool, L.mat_off), L.stride, _ptrs(in_pool, L.in_off), _ptrs(g, L.out_off),
                _ptrs(gi, L.in_off) if want_input else None, _ptrs(gm, L.mat_off) if want_mats else None,
                ws_ptrs, L.nb, dtype=_dtype_name(g), intermediates=L.intermediates)
            del ws
        return gm, gi, (grad_out if ctx.needs_input_grad[2] else None), None


def kronmult(mat_pool, in_pool, out_init, layout: Layout):
    """out_init + the batched Kronecker products, differentiable in all three tensors."""
    return KronmultFunction.apply(mat_pool, in_pool, out_init, layout)
