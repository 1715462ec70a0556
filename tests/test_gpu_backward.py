"""Backward pass of the routed SpMM: A^T on the device, SDDMM, autograd.

Pinned by mathematics (the reference has no backward):
  * the device transpose equals scipy's CSR transpose, bit for bit, and
    perm maps every A^T entry to the A edge it came from;
  * SDDMM and dB = A^T·dY equal their FP64 closed forms on the operands
    rounded to the operand type (fp32 accumulation: 1e-5 relative);
  * the gradients of L = sum(dY * C) agree with torch.sparse's FP64
    gradients (CPU) within 1e-2 relative, and with FP64 central finite
    differences (eps = 1e-3, 64 value probes and 64 B-entry probes): each
    probe within the rounding bound of its fp16/bf16 operands, all probes
    together within 1e-2 relative.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from graphs_small import CASES
from trspmm import graphs as g
from trspmm import autograd as tag
from trspmm.sparse import RoutedSparse

pytestmark = pytest.mark.gpu


def _routed(csr, dtype=torch.float16, mode="exact"):
    return RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                        m=csr.m, mode=mode, dtype=dtype)


def _q(x, dtype):
    """Values as the kernels consume them (rounded to the operand type), in FP64."""
    return torch.as_tensor(x).to(dtype).double().numpy()


def _rect(m=700, n=1100, deg=9, seed=21):
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(m), deg)
    cols = rng.integers(0, n, rows.size)
    return g.csr_from_pairs(rows, cols, rng.uniform(-1, 1, rows.size), m, n)


def _star(n=1500, seed=23):
    """Column 0 in every row (A^T row 0 has n entries: split into atomic chunks),
    column 1 in every other row, plus a sparse random rest."""
    rng = np.random.default_rng(seed)
    rows = [np.arange(n), np.arange(0, n, 2), np.repeat(np.arange(n), 3)]
    cols = [np.zeros(n, np.int64), np.ones((n + 1) // 2, np.int64), rng.integers(2, n, 3 * n)]
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    return g.csr_from_pairs(rows, cols, rng.uniform(-1, 1, rows.size), n, n)


MATS = dict(CASES, rect=_rect, star=_star, star800=lambda: _star(800, seed=5))   # star800: a whole long row


@pytest.mark.parametrize("case", sorted(MATS))
def test_csr_transpose_matches_scipy(cuda, case):
    csr = MATS[case]()
    print(f"[transpose] {case} nnz={csr.nnz}", flush=True)
    A = _routed(csr)
    At = A.transpose()
    ref = sp.csr_matrix((csr.val, csr.col, csr.rowptr), shape=(csr.m_pad, csr.n)).T.tocsr()
    ref.sort_indices()
    assert np.array_equal(At.rowptr.cpu().numpy(), ref.indptr.astype(np.int32))
    assert np.array_equal(At.col.cpu().numpy(), ref.indices.astype(np.int32))
    assert np.array_equal(At.values.cpu().numpy(), ref.data.astype(np.float32))
    perm = A._t_perm.cpu().numpy()
    assert np.array_equal(np.sort(perm), np.arange(csr.nnz))
    rows_of_edge = np.repeat(np.arange(csr.m_pad), np.diff(csr.rowptr))
    assert np.array_equal(csr.col[perm], np.repeat(np.arange(csr.n), np.diff(ref.indptr)))
    assert np.array_equal(rows_of_edge[perm], ref.indices)
    assert At.m == csr.n and At.ncols == csr.m


def test_csr_transpose_empty(cuda):
    csr = g.csr_from_pairs(np.zeros(0, np.int64), np.zeros(0, np.int64), None, 32, 48)
    A = _routed(csr)
    At = A.transpose()
    assert At.nnz == 0 and np.array_equal(At.rowptr.cpu().numpy(), np.zeros(49, np.int32))


def test_csr_transpose_rejects_bad_column(cuda):
    from trspmm import TrspmmError
    from trspmm._lib import check, current_stream, lib, ptr
    rp = torch.tensor([0, 2], dtype=torch.int32, device=cuda)
    ci = torch.tensor([1, 9], dtype=torch.int32, device=cuda)
    out = [torch.empty(n, dtype=torch.int32, device=cuda) for n in (9, 2, 2)]
    with pytest.raises(TrspmmError, match="out of range"):
        check(lib().trspmm_csr_transpose(ptr(rp), ptr(ci), None, 1, 2, 8, ptr(out[0]), ptr(out[1]), None,
                                         ptr(out[2]), current_stream()), "transpose")


def _sddmm_ref(csr, X, Y):
    rows = np.repeat(np.arange(csr.m_pad), np.diff(csr.rowptr))
    return np.einsum("ef,ef->e", X[rows], Y[csr.col])


@pytest.mark.parametrize("K", [8, 24, 64, 128, 256, 264])
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_sddmm_closed_form(cuda, K, dtype):
    csr = CASES["hub"]()
    print(f"[sddmm] K={K} {dtype} nnz={csr.nnz}", flush=True)
    A = _routed(csr, dtype)
    gen = torch.Generator().manual_seed(K)
    X = torch.randn(csr.m, K, generator=gen).to(dtype)
    Y = torch.randn(csr.n, K, generator=gen).to(dtype)
    got = A.sddmm(X.to(cuda), Y.to(cuda)).cpu().numpy().astype(np.float64)
    Xp = np.zeros((csr.m_pad, K))
    Xp[: csr.m] = X.double().numpy()
    ref = _sddmm_ref(csr, Xp, Y.double().numpy())
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    worst = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-5 and worst < 1e-4, (err, worst)


@pytest.mark.parametrize("case", sorted(MATS))
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
@pytest.mark.parametrize("path", ["fused", "routed"])
def test_backward_closed_form(cuda, case, dtype, path):
    csr = MATS[case]()
    print(f"[backward] {case} {dtype} {path} nnz={csr.nnz}", flush=True)
    K = 64
    A = _routed(csr, dtype)
    gen = torch.Generator().manual_seed(3)
    B = torch.randn(csr.n, K, generator=gen).to(dtype).to(cuda).requires_grad_()
    vals = A.values.clone().requires_grad_()
    dY = torch.randn(csr.m, K, generator=gen)
    C = tag.spmm(A, B, vals, backward=path)
    C.backward(dY.to(cuda))
    dYq = _q(dY, dtype)
    Bq = B.detach().cpu().double().numpy()
    Aq = sp.csr_matrix((_q(torch.from_numpy(csr.val), dtype), csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
    dYp = np.zeros((csr.m_pad, K))
    dYp[: csr.m] = dYq
    dB_ref = np.asarray(Aq.T @ dYp)
    dv_ref = _sddmm_ref(csr, dYp, Bq)
    dB = B.grad.double().cpu().numpy()
    dv = vals.grad.double().cpu().numpy()
    assert B.grad.dtype == dtype and vals.grad.dtype == torch.float32
    # dB is rounded to the operand type when handed back (B.grad has B's dtype)
    tol_b = 4e-3 if dtype == torch.float16 else 8e-3
    assert np.linalg.norm(dB - dB_ref) / max(np.linalg.norm(dB_ref), 1e-30) < tol_b
    if csr.nnz:
        assert np.linalg.norm(dv - dv_ref) / max(np.linalg.norm(dv_ref), 1e-30) < 1e-5


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_backward_vs_torch_sparse(cuda, dtype):
    """Gradients against torch.sparse's autograd in FP64 on the CPU."""
    csr = _rect()
    K = 64
    A = _routed(csr, dtype)
    gen = torch.Generator().manual_seed(9)
    B0 = torch.randn(csr.n, K, generator=gen)
    dY = torch.randn(csr.m, K, generator=gen)
    B = B0.to(dtype).to(cuda).requires_grad_()
    vals = A.values.clone().requires_grad_()
    tag.spmm(A, B, vals).backward(dY.to(cuda))

    rows = np.repeat(np.arange(csr.m_pad), np.diff(csr.rowptr))
    v64 = torch.from_numpy(csr.val.astype(np.float64)).requires_grad_()
    B64 = B0.double().requires_grad_()
    S = torch.sparse_coo_tensor(torch.from_numpy(np.stack([rows, csr.col]).astype(np.int64)), v64,
                                (csr.m_pad, csr.n))
    C64 = torch.sparse.mm(S, B64)[: csr.m]
    (C64 * dY.double()).sum().backward()
    dv_ref = v64.grad.numpy()
    dB_ref = B64.grad.numpy()
    dv = vals.grad.double().cpu().numpy()
    dB = B.grad.double().cpu().numpy()
    assert np.linalg.norm(dv - dv_ref) / np.linalg.norm(dv_ref) < 1e-2
    assert np.linalg.norm(dB - dB_ref) / np.linalg.norm(dB_ref) < 1e-2


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_backward_finite_differences(cuda, dtype):
    """FP64 central differences of L(vals, B) = sum(dY * (A(vals) B)), eps = 1e-3,
    at 64 random values and 64 random entries of B."""
    csr = CASES["rmat10"]()
    K = 32
    A = _routed(csr, dtype)
    gen = torch.Generator().manual_seed(11)
    B0 = torch.randn(csr.n, K, generator=gen)
    dY = torch.randn(csr.m, K, generator=gen)
    B = B0.to(dtype).to(cuda).requires_grad_()
    vals = A.values.clone().requires_grad_()
    tag.spmm(A, B, vals).backward(dY.to(cuda))
    dv = vals.grad.double().cpu().numpy()
    dB = B.grad.double().cpu().numpy()

    B64 = B0.double().numpy()
    dY64 = np.zeros((csr.m_pad, K))
    dY64[: csr.m] = dY.double().numpy()
    v64 = csr.val.astype(np.float64)

    def loss(v, Bm):
        Am = sp.csr_matrix((v, csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
        return float(np.sum(dY64 * (Am @ Bm)))

    # each probe: |grad - fd| within the rounding bound of the operands the
    # kernels consume (x, y rounded to unit roundoff u: |x'y' - xy| <= (2u + u^2)|xy|,
    # plus u|grad| when the gradient itself is handed back in the operand type);
    # over all probes: relative error <= 1e-2
    u = 2.0 ** -11 if dtype == torch.float16 else 2.0 ** -8
    absA = sp.csr_matrix((np.abs(v64), csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
    rows = np.repeat(np.arange(csr.m_pad), np.diff(csr.rowptr))
    eps = 1e-3
    rng = np.random.default_rng(12)
    got, want = [], []
    for e in rng.choice(csr.nnz, 64, replace=False):
        vp, vm = v64.copy(), v64.copy()
        vp[e] += eps
        vm[e] -= eps
        fd = (loss(vp, B64) - loss(vm, B64)) / (2 * eps)
        bound = 3 * u * float(np.abs(dY64[rows[e]]) @ np.abs(B64[csr.col[e]])) + 1e-6
        assert abs(dv[e] - fd) <= bound, (e, dv[e], fd, bound)
        got.append(dv[e])
        want.append(fd)
    absdY = np.abs(dY64)
    for _ in range(64):
        i, f = int(rng.integers(csr.n)), int(rng.integers(K))
        Bp, Bm_ = B64.copy(), B64.copy()
        Bp[i, f] += eps
        Bm_[i, f] -= eps
        fd = (loss(v64, Bp) - loss(v64, Bm_)) / (2 * eps)
        bound = 3 * u * float((absA[:, i].T @ absdY[:, f]).sum()) + u * abs(fd) + 1e-6
        assert abs(dB[i, f] - fd) <= bound, (i, f, dB[i, f], fd, bound)
        got.append(dB[i, f])
        want.append(fd)
    got, want = np.array(got), np.array(want)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-2


@pytest.mark.parametrize("K", [8, 24, 64, 128, 200, 256, 264, 512])
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
@pytest.mark.parametrize("case", ["star", "hub", "ragged", "rect"])
def test_fused_backward_closed_form(cuda, K, dtype, case):
    """One pass over A^T: dB = A^T dY (fp32) and dvals = SDDMM(dY, B) against
    their FP64 closed forms on the operands as the kernel consumes them."""
    csr = MATS[case]()
    print(f"[fused] {case} K={K} {dtype} nnz={csr.nnz}", flush=True)
    A = _routed(csr, dtype)
    gen = torch.Generator().manual_seed(K + 1)
    B = torch.randn(csr.n, K, generator=gen).to(dtype)
    dY = torch.randn(csr.m, K, generator=gen).to(dtype)
    dB, dv = A.spmm_t_sddmm(dY.to(cuda), B.to(cuda))
    Aq = sp.csr_matrix((_q(torch.from_numpy(csr.val), dtype), csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
    dYp = np.zeros((csr.m_pad, K))
    dYp[: csr.m] = dY.double().numpy()
    dB_ref = np.asarray(Aq.T @ dYp)
    dv_ref = _sddmm_ref(csr, dYp, B.double().numpy())
    dB = dB.double().cpu().numpy()
    dv = dv.double().cpu().numpy()
    assert np.linalg.norm(dB - dB_ref) / np.linalg.norm(dB_ref) < 1e-5
    assert np.abs(dB - dB_ref).max() / np.abs(dB_ref).max() < 1e-4
    assert np.linalg.norm(dv - dv_ref) / np.linalg.norm(dv_ref) < 1e-5
    assert np.abs(dv - dv_ref).max() / np.abs(dv_ref).max() < 1e-4


def test_fused_backward_repeatable(cuda):
    """Two calls on one plan (shared scratch) give the same result."""
    csr = _star()
    A = _routed(csr)
    gen = torch.Generator().manual_seed(5)
    B = torch.randn(csr.n, 64, generator=gen).half().to(cuda)
    dY = torch.randn(csr.m, 64, generator=gen).half().to(cuda)
    dB1, dv1 = A.spmm_t_sddmm(dY, B)
    dB1, dv1 = dB1.clone(), dv1.clone()
    dB2, dv2 = A.spmm_t_sddmm(dY, B)
    assert torch.allclose(dB1, dB2, rtol=1e-6, atol=1e-6) and torch.equal(dv1, dv2)


def test_transpose_is_cached_and_routed(cuda):
    csr = CASES["dense"]()
    A = _routed(csr)
    At = A.transpose()
    assert A.transpose() is At
    st = At.plan.stats
    assert st["tc_subblocks"] + st["sptc_tiles"] + st["residual_nnz"] > 0


@pytest.mark.parametrize("case", ["rmat10", "dense", "star", "er512", "ragged"])
@pytest.mark.parametrize("K", [64, 256, 512])
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_fused_backward_tc_subblocks_and_reordered_rows(cuda, case, K, dtype):
    """The fused backward routes A^T into TC sub-blocks (columns two rows of a
    window share) and residual records, with its rows reordered where that
    gathers fewer dY rows: every nonzero in one of the two routes, and dB and
    dvals against their FP64 closed forms (the value gradient of a TC value
    read back through its rank among the sub-block's occupancy bits)."""
    csr = (lambda: g.rmat_gcn(10, seed=1)) if case == "rmat10" else MATS[case]
    csr = csr()
    A = _routed(csr, dtype)
    st = A.backward_stats
    print(f"[fused-tc] {case} K={K} {dtype} {st}", flush=True)
    assert st["routed"] and st["tc_subblocks"] > 0
    assert st["tc_nnz"] + st["residual_nnz"] == csr.nnz
    if case in ("rmat10", "ragged"):           # a tenth or more fewer gathered dY rows (oracle counts)
        assert st["reordered"]
    gen = torch.Generator().manual_seed(K + 7)
    B = torch.randn(csr.n, K, generator=gen).to(dtype)
    dY = torch.randn(csr.m, K, generator=gen).to(dtype)
    dB, dv = A.spmm_t_sddmm(dY.to(cuda), B.to(cuda))
    Aq = sp.csr_matrix((_q(torch.from_numpy(csr.val), dtype), csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
    dYp = np.zeros((csr.m_pad, K))
    dYp[: csr.m] = dY.double().numpy()
    dB_ref = np.asarray(Aq.T @ dYp)
    dv_ref = _sddmm_ref(csr, dYp, B.double().numpy())
    dB = dB.double().cpu().numpy()
    dv = dv.double().cpu().numpy()
    assert np.linalg.norm(dB - dB_ref) / np.linalg.norm(dB_ref) < 1e-5
    assert np.linalg.norm(dv - dv_ref) / np.linalg.norm(dv_ref) < 1e-5
    assert np.abs(dv - dv_ref).max() / np.abs(dv_ref).max() < 1e-4
