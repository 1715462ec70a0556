"""RoutedSparse: a routed sparse matrix kept resident on the GPU.

Preprocess once (device router -> device plan), multiply many times:

    A = RoutedSparse(rowptr, col, val, ncols)      # CSR on CPU or GPU
    C = A @ B                                      # B: [ncols, K] fp16/bf16 on the GPU, C fp32
"""
from __future__ import annotations

import torch

from . import _lib
from .core import Plan, route


class RoutedSparse:
    def __init__(self, rowptr, col, val, ncols: int, m: int | None = None, *, mode: str = "exact",
                 dtype: torch.dtype = torch.float16, **router_kwargs):
        rowptr = torch.as_tensor(rowptr)
        col = torch.as_tensor(col)
        val = torch.as_tensor(val)
        self.rows = int(rowptr.numel() - 1)
        self.m = int(m if m is not None else self.rows)
        self.ncols = int(ncols)
        self.nnz = int(col.numel())
        self.dtype = dtype
        self.mode = mode
        self.window = int(router_kwargs.get("window", 16))
        self.meta = route(rowptr, col, val, self.ncols, mode=mode, **router_kwargs)
        self.plan = Plan.from_meta(self.meta, self.m, self.ncols, dtype)

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
