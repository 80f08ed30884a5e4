"""A `HostBatch` placed in device memory, with the device pointer arrays the C ABI takes.

torch is used only to own device memory and to name streams; every computation goes through
libkronmult_hip.so.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import api
from .workloads import HostBatch



def _torch():
    import torch
    return torch


class DeviceBatch:
    """Device pools + pointer arrays for one HostBatch.

    workspace: "auto" allocates per-entry workspaces only when the library needs them (n^d * itemsize
    above the on-chip size), "per_entry" always, "none" never.
    """

    def __init__(self, hb: HostBatch, device: str = "cuda", workspace: str = "auto"):
        torch = _torch()
        self.hb = hb
        self.device = torch.device(device)
        self.d, self.n, self.stride, self.nb, self.size = hb.d, hb.n, hb.stride, hb.nb, hb.size
        self.np_dtype = hb.dtype
        self.itemsize = hb.dtype.itemsize
        self.mat = torch.from_numpy(np.ascontiguousarray(hb.mat_pool)).to(self.device)
        self.inp = torch.from_numpy(np.ascontiguousarray(hb.in_pool)).to(self.device)
        self._out_init = torch.from_numpy(hb.output_pool_init()).to(self.device)
        self.out = self._out_init.clone()
        self.mat_ptrs = self._ptrs(self.mat, hb.mat_off)
        self.in_ptrs = self._ptrs(self.inp, hb.in_off)
        self.out_ptrs = self._ptrs(self.out, hb.out_off)
        uses_ws = api.kernel_name(self.d, self.n, self.itemsize, self.nb).startswith(("fiber_global", "multipass", "wide_global"))
        need_ws = workspace == "per_entry" or (workspace == "auto" and uses_ws)
        if need_ws:
            self.ws = torch.zeros(self.nb * self.size, dtype=self.out.dtype, device=self.device)
            self.ws_ptrs = self._ptrs(self.ws, np.arange(self.nb, dtype=np.int64) * self.size)
        else:
            self.ws = None
            self.ws_ptrs = None

    def _ptrs(self, t, off: np.ndarray):
        torch = _torch()
        addr = np.int64(t.data_ptr()) + off.astype(np.int64) * np.int64(self.itemsize)
        return torch.from_numpy(addr).to(self.device)

    @property
    def dtype_name(self) -> str:
        return "f64" if self.itemsize == 8 else "f32"

    def reset_output(self) -> None:
        self.out.copy_(self._out_init)

    def run(self, sync: bool = True, stream=None, tuning=None) -> None:
        """One call of the library over the whole batch (`tuning`: explicit launch shape, stream-ordered)."""
        args = (self.d, self.n, self.mat_ptrs, self.stride, self.in_ptrs, self.out_ptrs, self.ws_ptrs, self.nb)
        if tuning is not None:
            api.kronmult_batched_tuned(*args, dtype=self.dtype_name, tuning=tuning, stream=stream)
            if sync:
                _torch().cuda.synchronize(self.device)
        elif sync:
            api.kronmult_batched(*args, dtype=self.dtype_name)
        else:
            api.kronmult_batched_async(*args, dtype=self.dtype_name, stream=stream)

    def run_ordered(self, sync: bool = True, stream=None, tuning=None) -> None:
        """The ordered-accumulation call (needs per-entry workspaces: workspace="per_entry")."""
        if self.ws_ptrs is None:
            raise ValueError("the ordered call needs per-entry workspaces (workspace='per_entry')")
        api.kronmult_batched_ordered(self.d, self.n, self.mat_ptrs, self.stride, self.in_ptrs, self.out_ptrs,
                                     self.ws_ptrs, self.nb, dtype=self.dtype_name, tuning=tuning, stream=stream,
                                     sync=sync and tuning is None)
        if sync and tuning is not None:
            _torch().cuda.synchronize(self.device)

    def run_terms(self, sync: bool = True, stream=None) -> None:
        """The several-term call for a batch in expanded form (workloads.expand_terms): all term-major factor
        pointers, the first term_nb input and output pointers."""
        terms, nb = int(self.hb.meta["terms"]), int(self.hb.meta["term_nb"])
        api.kronmult_batched_terms(self.d, self.n, terms, self.mat_ptrs, self.stride, self.in_ptrs[:nb],
                                   self.out_ptrs[:nb], None if self.ws_ptrs is None else self.ws_ptrs[:nb], nb,
                                   dtype=self.dtype_name, stream=stream, sync=sync)

    # ---- adjoint -----------------------------------------------------------------------------------------
    def adjoint_setup(self, grad_out: Optional[np.ndarray] = None, want_input: bool = True,
                      want_mats: bool = True, intermediates="auto") -> None:
        """Device buffers for the adjoint: the cotangent pool (one vector per output, default the batch's
        own), gradient pools shaped like the input and factor pools, and the workspace it needs."""
        torch = _torch()
        hb = self.hb
        g = hb.grad_out_pool if grad_out is None else grad_out
        if g is None:
            raise ValueError("no cotangent for this batch")
        self.gout = torch.from_numpy(np.ascontiguousarray(g, dtype=hb.dtype)).to(self.device)
        self.gout_ptrs = self._ptrs(self.gout, hb.out_off)
        self.gin = torch.zeros_like(self.inp) if want_input else None
        self.gin_ptrs = self._ptrs(self.gin, hb.in_off) if want_input else None
        self.gmat = torch.zeros_like(self.mat) if want_mats else None
        self.gmat_ptrs = self._ptrs(self.gmat, hb.mat_off) if want_mats else None
        self.intermediates = intermediates
        ws = api.adjoint_workspace(self.d, self.n, self.itemsize, intermediates)
        if ws > 0:
            self.adj_ws = torch.zeros(self.nb * ws, dtype=self.out.dtype, device=self.device)
            self.adj_ws_ptrs = self._ptrs(self.adj_ws, np.arange(self.nb, dtype=np.int64) * ws)
        else:
            self.adj_ws = None
            self.adj_ws_ptrs = None

    def reset_gradients(self) -> None:
        for t in (self.gin, self.gmat):
            if t is not None:
                t.zero_()

    def run_adjoint(self, sync: bool = True, stream=None) -> None:
        api.kronmult_batched_adjoint(self.d, self.n, self.mat_ptrs, self.stride, self.in_ptrs, self.gout_ptrs,
                                     self.gin_ptrs, self.gmat_ptrs, self.adj_ws_ptrs, self.nb,
                                     dtype=self.dtype_name, intermediates=self.intermediates, stream=stream,
                                     sync=sync)

    def gradients(self):
        gi = None if self.gin is None else self.gin.detach().cpu().numpy()
        gm = None if self.gmat is None else self.gmat.detach().cpu().numpy()
        return gi, gm

    def run_split(self, mode: str = "fixed3", sync: bool = True, stream=None) -> None:
        """FP64 batch through the FP32 matrix cores (kronmult_batched_split_f64)."""
        if self.itemsize != 8:
            raise TypeError("the split path takes FP64 batches")
        api.kronmult_batched_split(self.d, self.n, self.mat_ptrs, self.stride, self.in_ptrs, self.out_ptrs,
                                   self.nb, mode=mode, stream=stream)
        if sync:
            _torch().cuda.synchronize(self.device)

    def group(self, tuning=None, entries_per_tile: int = 0) -> "api.Group":
        """This batch as one group of a grouped call."""
        return api.Group(self.d, self.n, self.mat_ptrs, self.stride, self.in_ptrs, self.out_ptrs, self.ws_ptrs,
                         self.nb, tuning=tuning, entries_per_tile=entries_per_tile)

    def output(self) -> np.ndarray:
        return self.out.detach().cpu().numpy()

    def inputs(self) -> np.ndarray:
        return self.inp.detach().cpu().numpy()
