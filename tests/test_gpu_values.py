"""New values without routing again, and the autograd op's `values` argument.

RoutedSparse.set_values (trspmm_plan_set_values on the plans of A and A^T,
trspmm_bwd_plan_set_values on the fused-backward plan) rewrites only the
16-bit values, each slot from the CSR edge the router recorded for it.

Pinned by:
  * the router decides routes from the pattern alone, so a plan refreshed to
    values v1 must give the product of a plan routed from v1 (same routes, same
    conversion of every value): compared within 1e-6 relative (fp32 atomics of
    split windows may reorder sums), for every route mode, both operand types,
    both residual engines, both record orders and both tile kernels; a value
    landing in the wrong slot moves the product by O(1);
  * the refreshed product against its FP64 closed form on the rounded operands
    (1e-5 relative; exact and tuned modes keep every nonzero);
  * A^T's plan refreshed through perm, and the fused backward plan, against
    plans built from the new values;
  * autograd: output and both gradients against FP64 closed forms when `values`
    is not A's own tensor, after an in-place update (an optimiser step), and
    when a second forward with other values runs before the first backward.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from graphs_small import CASES
from trspmm import _lib
from trspmm import autograd as tag
from trspmm import graphs as g
from trspmm._lib import lib, ptr
from trspmm.core import Plan, route
from trspmm.sparse import RoutedSparse

pytestmark = pytest.mark.gpu


def _routed(csr, vals, dtype, mode="exact", **kw):
    return RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.as_tensor(vals), csr.n,
                        m=csr.m, mode=mode, dtype=dtype, **kw)


def _q(x, dtype):
    return torch.as_tensor(x).to(dtype).double().numpy()


def _new_values(csr, seed):
    return np.random.default_rng(seed).uniform(-2, 2, csr.nnz).astype(np.float32)


def _ref_product(csr, vals, B, dtype):
    Aq = sp.csr_matrix((_q(vals, dtype), csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
    return np.asarray(Aq @ B.detach().cpu().double().numpy())[: csr.m]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _close(x, y, rtol=1e-6):
    scale = max(float(y.abs().max()), 1e-30)
    return float((x - y).abs().max()) <= rtol * scale


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("mode", ["compat", "exact", "tuned"])
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_refreshed_plan_equals_plan_routed_from_new_values(cuda, case, mode, dtype):
    csr = CASES[case]()
    print(f"[values] {case} {mode} {dtype} nnz={csr.nnz}", flush=True)
    v1 = _new_values(csr, 1)
    A = _routed(csr, csr.val, dtype, mode)
    B = torch.randn(csr.n, 64, generator=torch.Generator().manual_seed(2)).to(dtype).to(cuda)
    C0 = A @ B
    A.set_values(torch.from_numpy(v1).to(cuda))
    C1 = A @ B
    F = _routed(csr, v1, dtype, mode)
    assert _close(C1, F @ B)
    if mode != "compat" and csr.nnz:
        assert _rel(C1.double().cpu().numpy(), _ref_product(csr, v1, B, dtype)) < 1e-5
    A.set_values(torch.from_numpy(csr.val).to(cuda))        # and back
    assert _close(A @ B, C0)


@pytest.mark.parametrize("engine", ["valu", "mfma"])
@pytest.mark.parametrize("res_order", [0, 1])
@pytest.mark.parametrize("tile_kernel", [0, 1])
def test_refresh_every_engine_and_layout(cuda, engine, res_order, tile_kernel):
    csr = g.rmat_gcn(12, seed=6)
    v1 = _new_values(csr, 3)
    opts = dict(res_order=res_order, tile_kernel=tile_kernel)
    A = _routed(csr, csr.val, torch.float16, "tuned", engine=engine, plan_opts=opts)
    F = _routed(csr, v1, torch.float16, "tuned", engine=engine, plan_opts=opts)
    for K in (64, 256):
        B = torch.randn(csr.n, K, generator=torch.Generator().manual_seed(K)).half().to(cuda)
        A.set_values(torch.from_numpy(v1).to(cuda))
        C = A @ B
        assert _close(C, F @ B)
        assert _rel(C.double().cpu().numpy(), _ref_product(csr, v1, B, torch.float16)) < 1e-5


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_refresh_reaches_transpose_and_backward_plans(cuda, dtype):
    csr = CASES["hub"]()
    v1 = _new_values(csr, 5)
    K = 64
    A = _routed(csr, csr.val, dtype)
    gen = torch.Generator().manual_seed(7)
    dY = torch.randn(csr.m, K, generator=gen).to(dtype).to(cuda)
    B = torch.randn(csr.n, K, generator=gen).to(dtype).to(cuda)
    A.transpose()
    A._backward_plan()                     # both built before the new values arrive
    A.set_values(torch.from_numpy(v1).to(cuda))
    F = _routed(csr, v1, dtype)
    assert _close(A.transpose() @ dY, F.transpose() @ dY)
    dB, dv = A.spmm_t_sddmm(dY, B)
    dB_f, dv_f = F.spmm_t_sddmm(dY, B)
    assert _close(dB, dB_f)
    assert torch.equal(dv, dv_f)           # dvals does not depend on the values


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_autograd_follows_values(cuda, dtype):
    csr = g.rmat_gcn(11, seed=8)
    K = 64
    A = _routed(csr, csr.val, dtype)
    gen = torch.Generator().manual_seed(11)
    B = torch.randn(csr.n, K, generator=gen).to(dtype).to(cuda).requires_grad_()
    dY = torch.randn(csr.m, K, generator=gen)
    dYq = np.zeros((csr.m_pad, K))
    dYq[: csr.m] = _q(dY, dtype)
    Bq = B.detach().cpu().double().numpy()
    rows = np.repeat(np.arange(csr.m_pad), np.diff(csr.rowptr))
    dv_ref = np.einsum("ek,ek->e", dYq[rows], Bq[csr.col])      # independent of the values
    tol_b = 4e-3 if dtype == torch.float16 else 8e-3

    def check(C, vals_np):
        assert _rel(C.detach().double().cpu().numpy(), _ref_product(csr, vals_np, B, dtype)) < 1e-5
        Aq = sp.csr_matrix((_q(vals_np, dtype), csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
        dB_ref = np.asarray(Aq.T @ dYq)
        assert _rel(B.grad.double().cpu().numpy(), dB_ref) < tol_b

    vals = torch.from_numpy(_new_values(csr, 9)).to(cuda).requires_grad_()
    C = tag.spmm(A, B, vals)
    C.backward(dY.to(cuda))
    check(C, vals.detach().cpu().numpy())
    assert _rel(vals.grad.double().cpu().numpy(), dv_ref) < 1e-5

    with torch.no_grad():                   # an optimiser step on the values
        vals.mul_(-0.5).add_(0.25)
    B.grad = None
    vals.grad = None
    C = tag.spmm(A, B, vals)
    C.backward(dY.to(cuda))
    check(C, vals.detach().cpu().numpy())
    assert _rel(vals.grad.double().cpu().numpy(), dv_ref) < 1e-5

    # a second forward with other values before the first backward
    other = torch.from_numpy(_new_values(csr, 10)).to(cuda).requires_grad_()
    B.grad = None
    C1 = tag.spmm(A, B, vals)
    C2 = tag.spmm(A, B, other)
    assert _rel(C2.detach().double().cpu().numpy(), _ref_product(csr, other.detach().cpu().numpy(), B, dtype)) < 1e-5
    C1.backward(dY.to(cuda))
    check(C1, vals.detach().cpu().numpy())


def test_set_values_needs_edge_maps(cuda):
    csr = g.er_gcn(512, 8.0, seed=1)
    meta = route(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n)
    plan = Plan.from_meta(meta, csr.m, csr.n)
    v = torch.zeros(csr.nnz, device=cuda)
    with pytest.raises(ValueError, match="edge maps"):
        plan.set_values(v)
    rc = lib().trspmm_plan_set_values(plan.handle, ptr(v), csr.nnz, None, csr.nnz, _lib.current_stream())
    assert rc == _lib.EUNSUPPORTED
    meta_e = route(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                   edges=True)
    plan_e = Plan.from_meta(meta_e, csr.m, csr.n)
    rc = lib().trspmm_plan_set_values(plan_e.handle, ptr(v), csr.nnz, None, csr.nnz + 1, _lib.current_stream())
    assert rc == _lib.EINVAL


def test_set_values_index_outside_vals_gives_zero(cuda):
    """index entries outside [0, n_vals) read 0: equal to routing with those values zeroed."""
    csr = CASES["dense"]()
    rng = np.random.default_rng(4)
    v = rng.uniform(-1, 1, csr.nnz).astype(np.float32)
    idx = rng.permutation(csr.nnz).astype(np.int32)
    idx[::7] = csr.nnz + 3                                    # out of range
    meta = route(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                 edges=True)
    plan = Plan.from_meta(meta, csr.m, csr.n)
    plan.set_values(torch.from_numpy(v).to(cuda), torch.from_numpy(idx).to(cuda))
    want = np.where(idx < csr.nnz, v[np.minimum(idx, csr.nnz - 1)], 0).astype(np.float32)
    F = _routed(csr, want, torch.float16)
    B = torch.randn(csr.n, 32, generator=torch.Generator().manual_seed(1)).half().to(cuda)
    assert _close(plan.matmul(B), F @ B)


def test_bwd_plan_rejects_bad_columns_and_perms(cuda):
    n = 64
    t_rowptr = torch.arange(n + 1, dtype=torch.int32, device=cuda)
    t_val = torch.ones(n, device=cuda)
    good_col = torch.arange(n, dtype=torch.int32, device=cuda)
    good_perm = torch.arange(n, dtype=torch.int32, device=cuda).flip(0).contiguous()

    def create(col, perm):
        h = ctypes.c_void_p()
        rc = lib().trspmm_bwd_plan_create(ptr(t_rowptr), ptr(col), ptr(t_val), ptr(perm), n, n, n, _lib.F16,
                                          _lib.current_stream(), ctypes.byref(h))
        if rc == 0:
            lib().trspmm_bwd_plan_free(h)
        return rc, lib().trspmm_last_error()

    assert create(good_col, good_perm)[0] == 0
    bad_col = good_col.clone()
    bad_col[5] = n
    rc, msg = create(bad_col, good_perm)
    assert rc == _lib.EINVAL and b"column" in msg
    dup = good_perm.clone()
    dup[3] = dup[4]
    rc, msg = create(good_col, dup)
    assert rc == _lib.EINVAL and b"permutation" in msg
    neg = good_perm.clone()
    neg[0] = -1
    assert create(good_col, neg)[0] == _lib.EINVAL


@pytest.mark.parametrize("K", [8, 64, 256])
def test_sddmm_out_of_range_columns_give_zero(cuda, K):
    rows, ncols = 40, 50
    rng = np.random.default_rng(K)
    deg = rng.integers(0, 9, rows)
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    col = rng.integers(-3, ncols + 4, rowptr[-1]).astype(np.int32)
    X = torch.randn(rows, K, generator=torch.Generator().manual_seed(1)).half()
    Y = torch.randn(ncols, K, generator=torch.Generator().manual_seed(2)).half()
    out = torch.full((int(rowptr[-1]),), 7.0, device=cuda)
    Xd, Yd = X.to(cuda), Y.to(cuda)
    rp, cd = torch.from_numpy(rowptr).to(cuda), torch.from_numpy(col).to(cuda)
    rc = lib().trspmm_sddmm(ptr(rp), rows, ptr(cd), int(rowptr[-1]), ncols, ptr(Xd), K, ptr(Yd), K, K, _lib.F16,
                            ptr(out), _lib.current_stream())
    assert rc == 0
    r = np.repeat(np.arange(rows), deg)
    ok = (col >= 0) & (col < ncols)
    Xn, Yn = X.double().numpy(), Y.double().numpy()
    want = np.zeros(col.size)
    want[ok] = np.einsum("ek,ek->e", Xn[r[ok]], Yn[col[ok]])
    got = out.double().cpu().numpy()
    assert np.all(got[~ok] == 0)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-5)
