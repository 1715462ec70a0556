"""Differentiable routed SpMM: C = A(values) · B with a hand-written backward.

    A = RoutedSparse(rowptr, col, val, ncols, dtype=torch.float16)
    vals = A.values.clone().requires_grad_()       # [nnz] fp32, A's edge order
    B = torch.randn(ncols, K, device="cuda", dtype=torch.float16, requires_grad=True)
    C = trspmm.autograd.spmm(A, B, vals)            # fp32 [m, K]
    C.backward(dY)

Backward (dY = dL/dC, cast to the operand type of the routed kernels):
    dB    = A^T · dY       the routed multiply on A^T's cached plan
    dvals = SDDMM(dY, B)   dvals[e] = <dY[row_e], B[col_e]> over A's pattern

The forward multiplies with the plan built from A's values: `vals` carries the
gradient and must hold those values (the plan is not rebuilt per call).
"""
from __future__ import annotations

import torch

from .sparse import RoutedSparse


class RoutedSpMM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A: RoutedSparse, values: torch.Tensor, B: torch.Tensor):
        ctx.A = A
        ctx.save_for_backward(B)
        return A.matmul(B)

    @staticmethod
    def backward(ctx, dC: torch.Tensor):
        A: RoutedSparse = ctx.A
        (B,) = ctx.saved_tensors
        dY = dC.to(A.dtype).contiguous()
        d_values = d_B = None
        if ctx.needs_input_grad[2]:
            d_B = A.transpose().matmul(dY).to(B.dtype)
        if ctx.needs_input_grad[1]:
            d_values = A.sddmm(dY, B)
        return None, d_values, d_B


def spmm(A: RoutedSparse, B: torch.Tensor, values: torch.Tensor | None = None) -> torch.Tensor:
    """Differentiable C = A · B (fp32).  `values` ([nnz], A's edge order) receives
    the value gradient; by default A's own value tensor is used."""
    if values is None:
        values = A.values
    if values.shape != (A.nnz,):
        raise ValueError(f"values must have shape ({A.nnz},), got {tuple(values.shape)}")
    return RoutedSpMM.apply(A, values, B)
