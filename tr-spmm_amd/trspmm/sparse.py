"""RoutedSparse: a routed sparse matrix kept resident on the GPU.

Preprocess once (device router -> device plan), multiply many times:

    A = RoutedSparse(rowptr, col, val, ncols)      # CSR on CPU or GPU
    C = A @ B                                      # B: [ncols, K] fp16/bf16 on the GPU, C fp32

For the backward pass the handle also builds, once and on first use, the
routed plan of A^T (`transpose()`), and computes the value gradient over A's
pattern (`sddmm()`); `trspmm.autograd.spmm` wires both into torch.autograd.

The router records the CSR edge behind every value slot, so new values
(`set_values()`) refresh the plans of A and A^T and the fused-backward plan
in place, without routing again.

With `reorder` (the default for tuned routing where the plan's kernels
support it: window 16, the deep tile kernel with the matrix-core residual
engine) the router also tries the rows in (first column, degree) order and
keeps that routing when it gathers at least a tenth fewer B rows; the plan then writes each
row to its own index, so results are unchanged.  Exact routing keeps the
reference's tiles unless reorder=True is passed.
"""
from __future__ import annotations

import ctypes
import os
import weakref

import torch

from . import _lib
from ._lib import check, lib, ptr
from .core import Plan, dtype_code, route


class _BwdPlan:
    """Owns a trspmm_bwd_plan handle."""

    def __init__(self, handle):
        self.handle = handle

    def __del__(self):
        h, self.handle = getattr(self, "handle", None), None
        if h:
            try:
                lib().trspmm_bwd_plan_free(h)
            except Exception:
                pass


def reorder_supported(mode: str, window: int, engine: str | None, plan_opts: dict) -> bool:
    """Whether a reordered routing can run with these options: exact or tuned
    routing at window 16, on the deep tile kernel with the matrix-core
    residual engine (the kernel that writes rows through the plan's row map)."""
    if engine is None:
        engine = "valu" if os.environ.get("TRSPMM_RESIDUAL") == "valu" else "mfma"
    return (mode in ("exact", "tuned") and window == 16 and engine == "mfma"
            and int(plan_opts.get("tile_kernel", 0)) == 0 and int(plan_opts.get("item_order", 0)) != 2)


class RoutedSparse:
    def __init__(self, rowptr, col, val, ncols: int, m: int | None = None, *, mode: str = "exact",
                 dtype: torch.dtype = torch.float16, engine: str | None = None, slab: int | None = None,
                 plan_opts: dict | None = None, reorder: bool | None = None, **router_kwargs):
        rowptr = torch.as_tensor(rowptr)
        col = torch.as_tensor(col)
        val = torch.as_tensor(val)
        self.rows = int(rowptr.numel() - 1)
        self.m = int(m if m is not None else self.rows)
        if not 0 <= self.m <= self.rows:
            raise ValueError(f"m={self.m} must be within [0, rows={self.rows}]")
        self.ncols = int(ncols)
        self.nnz = int(col.numel())
        self.dtype = dtype
        self.mode = mode
        self.router_kwargs = dict(router_kwargs)
        self.window = int(router_kwargs.get("window", 16))
        self.engine, self.slab = engine, slab
        self.plan_opts = dict(plan_opts or {})      # tile_cap / tc_cap / res_cap / item_order
        if reorder is None:
            reorder = mode == "tuned" and reorder_supported(mode, self.window, engine, self.plan_opts)
        self.reorder = bool(reorder)
        self.meta = route(rowptr, col, val, self.ncols, mode=mode, edges=True, reorder=self.reorder, **router_kwargs)
        self.plan = Plan.from_meta(self.meta, self.m, self.ncols, dtype, engine=engine, slab=slab, **self.plan_opts)
        # device CSR (validated by the router): the SDDMM pattern and the
        # source of the transposed plan
        dev = torch.device("cuda", torch.cuda.current_device())
        self.rowptr = rowptr.to(dev, torch.int32).contiguous()
        self.col = col.to(dev, torch.int32).contiguous()
        self.values = val.to(dev, torch.float32).contiguous()
        self._t = None
        self._tcsr = None
        self._bwd = None
        self._synced = (weakref.ref(self.values), self.values._version)   # values the plans hold
        self.values_epoch = 0                                               # counts set_values calls

    def set_values(self, values: torch.Tensor):
        """New values of A's nonzeros ([nnz], A's edge order, any float dtype):
        the plans of A (and of A^T and the fused backward, once built) take them
        in place; the pattern and routes stay."""
        v = values.detach()
        if v.shape != (self.nnz,):
            raise ValueError(f"values must have shape ({self.nnz},), got {tuple(v.shape)}")
        self.values = v.to(self.values.device, torch.float32).contiguous()
        self.plan.set_values(self.values)
        if self._tcsr is not None:
            t_val, perm = self._tcsr[2], self._tcsr[3]
            torch.index_select(self.values, 0, perm, out=t_val)
            if self._t is not None:
                self._t.set_values(t_val)
            if self._bwd is not None:
                check(lib().trspmm_bwd_plan_set_values(self._bwd.handle, ptr(self.values), self.nnz,
                                                       _lib.current_stream()), "trspmm_bwd_plan_set_values")
        self._synced = (weakref.ref(values), values._version)
        self.values_epoch += 1

    def sync_values(self, values: torch.Tensor):
        """set_values(values) unless the plans already hold this tensor at its
        current version (the same object, not modified in place since)."""
        ref, version = self._synced
        if ref() is not values or values._version != version:
            self.set_values(values)

    @property
    def stats(self):
        return dict(self.plan.stats, nnz=self.nnz, router_seconds=self.meta.seconds)

    def matmul(self, B: torch.Tensor, out: torch.Tensor | None = None, routes: int = _lib.ROUTE_ALL):
        return self.plan.matmul(B, out, routes)

    __matmul__ = matmul

    def timed(self, B, out=None, routes=_lib.ROUTE_ALL, warmup=10, iters=20) -> float:
        if out is None:
            out = torch.empty((self.m, B.shape[1]), device=B.device, dtype=torch.float32)
        return self.plan.timed(B, out, routes, warmup, iters)

    # ------------------------------------------------------------ backward
    def _transposed_csr(self):
        """A^T as a canonical device CSR (t_rowptr, t_col, t_val) plus perm (A's
        edge of each A^T entry), computed on the device once."""
        if self._tcsr is None:
            if self.m < self.rows and int(self.rowptr[self.m].item()) != self.nnz:
                raise ValueError("rows beyond m hold nonzeros: A^T would reference rows outside dY")
            dev = self.rowptr.device
            n1 = max(self.nnz, 1)
            t_rowptr = torch.empty(self.ncols + 1, dtype=torch.int32, device=dev)
            t_col = torch.empty(n1, dtype=torch.int32, device=dev)
            t_val = torch.empty(n1, dtype=torch.float32, device=dev)
            perm = torch.empty(n1, dtype=torch.int32, device=dev)
            check(lib().trspmm_csr_transpose(ptr(self.rowptr), ptr(self.col), ptr(self.values), self.rows, self.nnz,
                                             self.ncols, ptr(t_rowptr), ptr(t_col), ptr(t_val), ptr(perm),
                                             _lib.current_stream()), "trspmm_csr_transpose")
            n = self.nnz
            self._tcsr = (t_rowptr, t_col[:n], t_val[:n], perm[:n])
        return self._tcsr

    @property
    def _t_perm(self):
        return self._transposed_csr()[3]

    def transpose(self) -> "RoutedSparse":
        """A^T (ncols x m) as a routed matrix with the same routing mode and
        operand type, built on the device on first use and cached."""
        if self._t is None:
            t_rowptr, t_col, t_val, _ = self._transposed_csr()
            self._t = RoutedSparse(t_rowptr, t_col, t_val, self.m, self.ncols, mode=self.mode, dtype=self.dtype,
                                   reorder=self.reorder, **self.router_kwargs)
        return self._t

    def _backward_plan(self):
        if self._bwd is None:
            t_rowptr, t_col, t_val, perm = self._transposed_csr()
            h = ctypes.c_void_p()
            check(lib().trspmm_bwd_plan_create(ptr(t_rowptr), ptr(t_col), ptr(t_val), ptr(perm), self.ncols, self.nnz,
                                               self.m, dtype_code(self.dtype), _lib.current_stream(),
                                               ctypes.byref(h)), "trspmm_bwd_plan_create")
            self._bwd = _BwdPlan(h)
        return self._bwd

    @property
    def backward_stats(self) -> dict:
        """What the fused backward runs on (builds its plan on first use): the routed
        fused kernel or the lane-group kernels, A^T's TC sub-blocks / TC nonzeros /
        residual records, and whether A^T's rows were reordered."""
        c = (ctypes.c_int64 * 5)()
        check(lib().trspmm_bwd_plan_stats(self._backward_plan().handle, c), "trspmm_bwd_plan_stats")
        return {"routed": bool(c[0]), "tc_subblocks": int(c[1]), "tc_nnz": int(c[2]), "residual_nnz": int(c[3]),
                "reordered": bool(c[4])}

    def spmm_t_sddmm(self, dY: torch.Tensor, B: torch.Tensor, dB: torch.Tensor | None = None,
                     dvals: torch.Tensor | None = None):
        """Fused backward: (dB = A^T · dY  [ncols, K] fp32, dvals = SDDMM(dY, B)  [nnz]
        fp32 in A's edge order) in one pass over A^T's rows (K <= 512)."""
        for name, t, rows in (("dY", dY, self.m), ("B", B, self.ncols)):
            if not t.is_cuda or t.dtype != self.dtype or t.dim() != 2 or t.shape[0] != rows:
                raise ValueError(f"{name} must be a [{rows}, K] {self.dtype} device tensor, got "
                                 f"{tuple(t.shape)} {t.dtype} on {t.device}")
            if t.stride(1) != 1 or t.stride(0) % 8 or t.data_ptr() % 16:
                raise ValueError(f"{name} must be row-major with a row stride multiple of 8, 16-byte aligned")
        K = dY.shape[1]
        if B.shape[1] != K:
            raise ValueError("dY and B need the same K")
        if dB is None:
            dB = torch.empty((self.ncols, K), dtype=torch.float32, device=dY.device)
        if dvals is None:
            dvals = torch.empty(self.nnz, dtype=torch.float32, device=dY.device)
        plan = self._backward_plan()
        check(lib().trspmm_spmm_t_sddmm(plan.handle, ptr(dY), dY.stride(0), ptr(B), B.stride(0), K, ptr(dB),
                                        dB.stride(0), ptr(dvals), _lib.current_stream()), "trspmm_spmm_t_sddmm")
        return dB, dvals

    def sddmm(self, X: torch.Tensor, Y: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """dvals[e] = <X[row_e], Y[col_e]> over A's edges (fp32, A's edge order):
        the gradient of sum(X * (A @ Y)) with respect to A's values."""
        for name, t, rows in (("X", X, self.m), ("Y", Y, self.ncols)):
            if not t.is_cuda or t.dtype != self.dtype or t.dim() != 2 or t.shape[0] != rows:
                raise ValueError(f"{name} must be a [{rows}, K] {self.dtype} device tensor, got "
                                 f"{tuple(t.shape)} {t.dtype} on {t.device}")
            if t.stride(1) != 1 or t.stride(0) % 8 or t.data_ptr() % 16:
                raise ValueError(f"{name} must be row-major with a row stride multiple of 8, 16-byte aligned")
        if X.shape[1] != Y.shape[1] or X.shape[1] % 8:
            raise ValueError("X and Y need the same K, a multiple of 8")
        if out is None:
            out = torch.empty(self.nnz, dtype=torch.float32, device=X.device)
        elif out.dtype != torch.float32 or out.shape != (self.nnz,) or not out.is_contiguous() or not out.is_cuda:
            raise ValueError(f"out must be a contiguous float32 device tensor of shape ({self.nnz},)")
        check(lib().trspmm_sddmm(ptr(self.rowptr), self.rows, ptr(self.col), self.nnz, self.ncols, ptr(X), X.stride(0),
                                 ptr(Y), Y.stride(0), X.shape[1], dtype_code(self.dtype), ptr(out),
                                 _lib.current_stream()), "trspmm_sddmm")
        return out
