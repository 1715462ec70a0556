"""Differentiable routed SpMM: C = A(values) · B with a hand-written backward.

    A = RoutedSparse(rowptr, col, val, ncols, dtype=torch.float16)
    vals = A.values.clone().requires_grad_()       # [nnz] fp32, A's edge order
    B = torch.randn(ncols, K, device="cuda", dtype=torch.float16, requires_grad=True)
    C = trspmm.autograd.spmm(A, B, vals)            # fp32 [m, K]
    C.backward(dY)

Backward (dY = dL/dC, cast to the operand type of the routed kernels):
    dB    = A^T · dY       dvals = SDDMM(dY, B), dvals[e] = <dY[row_e], B[col_e]>
When both are needed (and K <= 512) one fused pass over A^T's rows computes
them together, gathering each dY row once ("fused", the default); otherwise,
or with backward="routed", dB runs on the routed kernels over A^T's cached
plan and dvals on the SDDMM kernel.

The forward multiplies with the values in `vals`: when they are not the ones
A's plans hold (another tensor, or the same one modified in place since), the
plans of A, A^T and the fused backward take them first (RoutedSparse.set_values,
a few value-refresh kernels; the routes are not rebuilt).
"""
from __future__ import annotations

import torch

from .sparse import RoutedSparse


class RoutedSpMM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A: RoutedSparse, values: torch.Tensor, B: torch.Tensor, backward: str = "fused"):
        if backward not in ("fused", "routed"):
            raise ValueError("backward must be 'fused' or 'routed'")
        A.sync_values(values)
        ctx.A = A
        ctx.epoch = A.values_epoch
        ctx.mode = backward
        ctx.save_for_backward(B, values)
        return A.matmul(B)

    @staticmethod
    def backward(ctx, dC: torch.Tensor):
        A: RoutedSparse = ctx.A
        B, values = ctx.saved_tensors
        dY = dC.to(A.dtype).contiguous()
        if A.values_epoch != ctx.epoch:      # a later forward set other values: back to this one's
            A.set_values(values)
        d_values = d_B = None
        want_v, want_b = ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        if want_v and want_b and ctx.mode == "fused" and dY.shape[1] <= 512:
            d_B, d_values = A.spmm_t_sddmm(dY, B)
            return None, d_values, d_B.to(B.dtype), None
        if want_b:
            d_B = A.transpose().matmul(dY).to(B.dtype)
        if want_v:
            d_values = A.sddmm(dY, B)
        return None, d_values, d_B, None


def spmm(A: RoutedSparse, B: torch.Tensor, values: torch.Tensor | None = None,
         backward: str = "fused") -> torch.Tensor:
    """Differentiable C = A · B (fp32).  `values` ([nnz], A's edge order) receives
    the value gradient; by default A's own value tensor is used."""
    if values is None:
        values = A.values
    if values.shape != (A.nnz,):
        raise ValueError(f"values must have shape ({A.nnz},), got {tuple(values.shape)}")
    return RoutedSpMM.apply(A, values, B, backward)
