"""Row reordering on the device (TRSPMM_ROUTE_REORDER) against the oracle.

  * the router's keep-or-not decision, its row order and, for the routing it
    keeps, the 29 arrays bit-exact against oracle.route_reordered (the oracle
    routes the permuted CSR itself);
  * products of reordered plans against the FP64 CSR product of the input
    (every row back at its own index), fp16 and bf16, K = 64 / 128 / 256,
    with windows split into many work items (fp32 atomics through the row
    map, rows of split windows zeroed through it) and without;
  * new values through the translated edge maps (product against FP64 with
    the new values), and autograd gradients against FP64 closed forms;
  * the options that cannot run a reordered routing are refused, and the
    default stays off where it cannot run.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from graphs_small import CASES
from trspmm import _lib
from trspmm import autograd as tag
from trspmm import graphs as g
from trspmm._lib import TrspmmError
from trspmm.core import META_NAMES, Plan, route
from trspmm.sparse import RoutedSparse

pytestmark = pytest.mark.gpu

GRAPHS = dict(CASES, rmat14=lambda: g.rmat_gcn(14, seed=1), rmat12=lambda: g.rmat_gcn(12, seed=1))


def _q(x, dtype):
    return torch.as_tensor(x).to(dtype).double().numpy()


def _ref(csr, vals, B, dtype):
    A = sp.csr_matrix((_q(vals, dtype), csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
    return np.asarray(A @ B.detach().cpu().double().numpy())[: csr.m]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("case", sorted(GRAPHS))
@pytest.mark.parametrize("mode", ["exact", "tuned"])
def test_reordered_routing_bit_exact(cuda, case, mode):
    csr = GRAPHS[case]()
    print(f"[row-order] {case} {mode} nnz={csr.nnz}", flush=True)
    kw = dict(exact=True, sptc_min_nnz=8 if mode == "tuned" else 0)
    ref, order = oracle.route_reordered(csr.rowptr, csr.col, csr.val, **kw)
    meta = route(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                 mode=mode, reorder=True, edges=True)
    assert meta.reordered == (order is not None)
    if order is not None:
        assert np.array_equal(meta.row_order("cpu").numpy(), order)
    got = meta.tensors("cpu")
    bad = [n for n, t in zip(META_NAMES, got) if not (t.numpy().shape == ref[n].shape and np.array_equal(t.numpy(), ref[n]))]
    assert not bad, f"arrays differ: {bad}"


def test_reordering_is_taken_and_refused_where_expected(cuda):
    """The R-MAT graphs keep the reordered routing (a fifth or more fewer
    gathered rows); ER, the dense block and the hub window keep their order
    (the oracle's counts decide; checked above); compat mode and narrow windows
    refuse the flag."""
    for case, want in (("rmat14", True), ("rmat10", True), ("er512", False), ("dense", False), ("hub", False)):
        csr = GRAPHS[case]()
        meta = route(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     mode="tuned", reorder=True)
        assert meta.reordered == want, case
    csr = GRAPHS["er512"]()
    args = (torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n)
    with pytest.raises(TrspmmError):
        route(*args, mode="compat", reorder=True)
    with pytest.raises(TrspmmError):
        route(*args, mode="exact", window=8, reorder=True)


@pytest.mark.parametrize("case", sorted(GRAPHS))
@pytest.mark.parametrize("K", [64, 128, 256])
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_reordered_products(cuda, case, K, dtype):
    csr = GRAPHS[case]()
    print(f"[row-order] product {case} K={K} {dtype}", flush=True)
    A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode="tuned", dtype=dtype, plan_opts={"k_hint": K})
    assert A.reorder
    B = torch.randn(csr.n, K, generator=torch.Generator().manual_seed(K)).to(dtype).to(cuda)
    C = A @ B
    assert _rel(C.double().cpu().numpy(), _ref(csr, csr.val, B, dtype)) < 1e-5, A.stats


@pytest.mark.parametrize("case", ["rmat14", "rmat12", "rmat10"])
def test_reordered_split_windows(cuda, case):
    """Small caps split most windows into several work items: their rows are
    zeroed and accumulated atomically, both through the row map."""
    csr = GRAPHS[case]()
    meta_check = route(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                       mode="tuned", reorder=True)
    assert meta_check.reordered
    A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode="tuned", plan_opts={"tile_cap": 1, "tc_cap": 1, "res_cap": 32})
    B = torch.randn(csr.n, 128, generator=torch.Generator().manual_seed(3)).half().to(cuda)
    C = torch.full((csr.m, 128), float("nan"), device=cuda)
    A.matmul(B, C)
    got = C.double().cpu().numpy()
    assert np.isfinite(got).all()
    assert _rel(got, _ref(csr, csr.val, B, torch.float16)) < 1e-5


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_reordered_plan_takes_new_values(cuda, dtype):
    csr = GRAPHS["rmat14"]()
    A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode="tuned", dtype=dtype)
    assert A.meta.reordered
    B = torch.randn(csr.n, 64, generator=torch.Generator().manual_seed(5)).to(dtype).to(cuda)
    v1 = np.random.default_rng(9).uniform(-2, 2, csr.nnz).astype(np.float32)
    A.set_values(torch.from_numpy(v1).to(cuda))
    assert _rel((A @ B).double().cpu().numpy(), _ref(csr, v1, B, dtype)) < 1e-5


def test_reordered_autograd(cuda):
    csr = GRAPHS["rmat10"]()
    A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode="tuned")
    assert A.meta.reordered
    K = 32
    vals = A.values.clone().requires_grad_()
    B = torch.randn(csr.n, K, generator=torch.Generator().manual_seed(1)).half().to(cuda).requires_grad_()
    dY = torch.randn(csr.m, K, generator=torch.Generator().manual_seed(2)).to(cuda)
    C = tag.spmm(A, B, vals)
    C.backward(dY)
    # FP64 closed forms on the rounded operands
    vq = _q(csr.val, torch.float16)
    A64 = sp.csr_matrix((vq, csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))[: csr.m]
    B64 = B.detach().cpu().double().numpy()
    dY64 = dY.half().cpu().double().numpy()
    assert _rel(C.detach().double().cpu().numpy(), np.asarray(A64 @ B64)) < 1e-5
    assert _rel(B.grad.double().cpu().numpy(), np.asarray(A64.T @ dY64)) < 1e-2
    rows = np.repeat(np.arange(csr.m_pad), np.diff(csr.rowptr))
    keep = rows < csr.m
    dv = np.zeros(csr.nnz)
    dv[keep] = np.einsum("ij,ij->i", dY64[rows[keep]], B64[csr.col[keep]])
    assert _rel(vals.grad.double().cpu().numpy(), dv) < 1e-5


def test_reordered_routing_needs_the_deep_kernel(cuda):
    csr = GRAPHS["rmat10"]()
    meta = route(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                 mode="tuned", reorder=True)
    assert meta.reordered
    for kw in ({"tile_kernel": 1}, {"item_order": 2}, {"engine": "valu"}):
        with pytest.raises(TrspmmError) as e:
            Plan.from_meta(meta, csr.m, csr.n, **kw)
        assert "reordered" in str(e.value)
    # the default leaves the reordering off where the options cannot run it
    A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode="tuned", engine="valu")
    assert not A.reorder and not A.meta.reordered
    B = torch.randn(csr.n, 64, generator=torch.Generator().manual_seed(4)).half().to(cuda)
    assert _rel((A @ B).double().cpu().numpy(), _ref(csr, csr.val, B, torch.float16)) < 1e-5
    # exact routing stays in the input order unless asked
    E = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode="exact")
    assert not E.meta.reordered


def test_reordered_routing_reports_bad_input_and_recovers(cuda):
    """Both row orders are routed at the same time (the input order on a second
    host thread): a column out of range is reported as an error, not a hang,
    and the next routing works."""
    csr = GRAPHS["rmat10"]()
    col = csr.col.copy()
    col[len(col) // 2] = csr.n + 5
    with pytest.raises(TrspmmError):
        route(torch.from_numpy(csr.rowptr), torch.from_numpy(col), torch.from_numpy(csr.val), csr.n,
              mode="tuned", reorder=True)
    meta = route(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                 mode="tuned", reorder=True)
    assert meta.reordered
