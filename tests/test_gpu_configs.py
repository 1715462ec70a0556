"""Parity at the full sizes of the five benchmark workloads.

  1. Erdos-Renyi n = 4096, degree 16, seed 0, GCN values, K = 64: the 29
     metadata arrays bit-exact against the oracle router; the compat product
     against the FP64 product over the oracle's metadata (the reference's
     routed result), the exact and tuned products against the FP64 CSR
     product; 1e-5 relative Frobenius (fp32 accumulation of fp16 operands).
  2. R-MAT scale 18 (2^18 nodes, 8.39M nonzeros), seed 1, K = 128: compat
     metadata bit-exact against the oracle router; compat, exact and tuned
     products as in 1.
  3. Three-route mix, 2^20 rows, K = 256: exact-mode metadata bit-exact
     against the oracle router; every route's nonzeros counted (together all
     of A's); each route launched alone against the FP64 product of that
     route's part of the metadata, and the full product against the FP64 CSR
     product, on 8192 seeded rows.
  4. Backward on the graph of 2, K = 256, fp16 and bf16: FP64 central
     differences (eps = 1e-3) at 64 values and 64 entries of B, each probe
     within the rounding bound of its operands and all together within 1e-2;
     and both gradients against torch.sparse's FP64 autograd (CPU) for a dY
     that is nonzero on 4096 seeded rows only (torch.sparse's backward of the
     whole graph does not fit the test's time), within 1e-2.
  5. 1024 Barabasi-Albert graphs (m = 3, 20..5000 nodes), K = 128: the batched
     launch, eager and hipGraph-replayed, against each graph's FP64 product
     (1e-5), and the replay within 1e-6 of the eager launch (split windows add
     atomically, in launch-dependent order).
The FP64 references run on the CPU (scipy); the oracle router is the C
restatement under oracle/.
"""
import functools

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from trspmm import _lib
from trspmm import autograd as tag
from trspmm import graphs as g
from trspmm.batch import RoutedBatch
from trspmm.core import META_NAMES, route
from trspmm.sparse import RoutedSparse

pytestmark = pytest.mark.gpu


@functools.lru_cache(maxsize=None)
def _graph(name):
    return {"er4096": lambda: g.er_gcn(4096, 16.0, seed=0), "rmat18": lambda: g.rmat_gcn(18, seed=1),
            "three": lambda: g.three_route(1 << 20, seed=2)}[name]()


def _meta_equal(csr, mode):
    meta = route(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                 mode=mode)
    got = meta.tensors("cpu")
    ref = oracle.route(csr.rowptr, csr.col, csr.val, exact=(mode != "compat"))
    bad = [n for n, gt in zip(META_NAMES, got)
           if not (gt.numpy().dtype == ref[n].dtype and gt.numpy().shape == ref[n].shape
                   and np.array_equal(gt.numpy(), ref[n]))]
    assert not bad, f"arrays differ: {bad}"
    return ref


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _routed(csr, mode, dtype=torch.float16, **kw):
    return RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                        m=csr.m, mode=mode, dtype=dtype, **kw)


def _csr64(csr, rows=None, dtype=torch.float16):
    vq = torch.from_numpy(csr.val).to(dtype).double().numpy()
    A = sp.csr_matrix((vq, csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
    return A if rows is None else A[rows]


@pytest.mark.parametrize("name,K", [("er4096", 64), ("rmat18", 128)])
def test_forward_configs_metadata_and_products(cuda, name, K):
    csr = _graph(name)
    print(f"[config] {name} nnz={csr.nnz} K={K}", flush=True)
    ref_meta = _meta_equal(csr, "compat")
    B16 = g.dense_features(csr.n, K, seed=1000 if name == "er4096" else 1001)
    B = torch.from_numpy(B16).to(cuda)
    B64 = B16.astype(np.float64)
    C = _routed(csr, "compat").matmul(B).double().cpu().numpy()
    assert _rel(C, oracle.spmm64(ref_meta, B64, csr.m)) < 1e-5
    ref = np.asarray(_csr64(csr) @ B64)[: csr.m]
    for mode in ("exact", "tuned"):
        print(f"[config] {name} {mode}", flush=True)
        C = _routed(csr, mode, plan_opts={"k_hint": K}).matmul(B).double().cpu().numpy()
        assert _rel(C, ref) < 1e-5, mode


def _route_only(meta, which):
    """The metadata with the other routes' part lists emptied (decode reads
    n_tparts, n_cparts, n_csparts and n_groups from these arrays' lengths)."""
    m = dict(meta.arrays) if hasattr(meta, "arrays") else dict(meta)
    empty = np.zeros(0, np.int32)
    if which != "tc":
        m["t_window_row"] = empty
    if which != "residual":
        m["c_row"] = empty
        m["c_row_short"] = empty
    if which != "sptc":
        m["s_window"] = empty
    return m


def test_three_route_mix_per_route(cuda):
    csr = _graph("three")
    K = 256
    print(f"[config] three-route nnz={csr.nnz}", flush=True)
    ref_meta = _meta_equal(csr, "exact")
    rows = np.sort(np.random.default_rng(7).choice(csr.m, 8192, replace=False))
    B16 = g.dense_features(csr.n, K, seed=1002)
    B = torch.from_numpy(B16).to(cuda)
    B64 = B16.astype(np.float64)
    A = _routed(csr, "exact", plan_opts={"k_hint": K})
    idx = torch.from_numpy(rows).to(cuda)
    total = 0
    for which, mask in (("tc", _lib.ROUTE_TC), ("sptc", _lib.ROUTE_SPTC), ("residual", _lib.ROUTE_RESIDUAL)):
        r, c, v = oracle.decode(_route_only(ref_meta, which))
        total += r.size
        print(f"[config] route {which}: {r.size} nonzeros in {np.unique(r).size} rows", flush=True)
        assert r.size > 0, which
        vq = v.astype(np.float16).astype(np.float64)
        Ar = sp.csr_matrix((vq, (r, c)), shape=(csr.m_pad, csr.n))[rows]
        got = A.matmul(B, routes=mask)[idx].double().cpu().numpy()
        assert _rel(got, np.asarray(Ar @ B64)) < 1e-5, which
    assert total == csr.nnz                  # exact mode drops nothing
    got = A.matmul(B)[idx].double().cpu().numpy()
    assert _rel(got, np.asarray(_csr64(csr, rows) @ B64)) < 1e-5


def _row_loss_delta(csr, v64, B64, dY64, r, e, eps):
    """L(v + eps e_e) - L(v - eps e_e) for L = sum(dY * (A(v) B)): only row r of
    A(v) B changes, so only its term is evaluated."""
    a, b = csr.rowptr[r], csr.rowptr[r + 1]
    cols = csr.col[a:b]

    def row_term(v):
        return float(dY64[r] @ (v[a:b] @ B64[cols]))

    vp, vm = v64.copy(), v64.copy()
    vp[e] += eps
    vm[e] -= eps
    return row_term(vp) - row_term(vm)


def _entry_loss_delta(A64, rows_i, B64, dY64, i, f, eps):
    """L(B + eps E_if) - L(B - eps E_if): only column f of the rows of A that
    hold column i changes."""
    Ai = A64[rows_i]

    def col_term(b):
        return float(dY64[rows_i, f] @ (Ai @ b))

    bp, bm = B64[:, f].copy(), B64[:, f].copy()
    bp[i] += eps
    bm[i] -= eps
    return col_term(bp) - col_term(bm)


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_backward_config_finite_differences_and_torch_sparse(cuda, dtype):
    csr = _graph("rmat18")
    K = 256
    print(f"[config] backward rmat18 K={K} {dtype}", flush=True)
    A = _routed(csr, "exact", dtype=dtype)
    gen = torch.Generator().manual_seed(31)
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
    rows_of = np.repeat(np.arange(csr.m_pad), np.diff(csr.rowptr))
    A64 = sp.csr_matrix((v64, csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
    colA = A64.tocsc()
    colA.sort_indices()
    u = 2.0 ** -11 if dtype == torch.float16 else 2.0 ** -8
    eps = 1e-3
    rng = np.random.default_rng(32)
    got, want = [], []
    for e in rng.choice(csr.nnz, 64, replace=False):
        r = int(rows_of[e])
        fd = _row_loss_delta(csr, v64, B64, dY64, r, e, eps) / (2 * eps)
        bound = 3 * u * float(np.abs(dY64[r]) @ np.abs(B64[csr.col[e]])) + 1e-6
        assert abs(dv[e] - fd) <= bound, (e, dv[e], fd, bound)
        got.append(dv[e])
        want.append(fd)
    for _ in range(64):
        i, f = int(rng.integers(csr.n)), int(rng.integers(K))
        rows_i = colA.indices[colA.indptr[i]: colA.indptr[i + 1]]
        absv = np.abs(colA.data[colA.indptr[i]: colA.indptr[i + 1]])      # |A[r, i]| for r in rows_i
        fd = _entry_loss_delta(A64, rows_i, B64, dY64, i, f, eps) / (2 * eps)
        bound = 3 * u * float(absv @ np.abs(dY64[rows_i, f])) + u * abs(fd) + 1e-6
        assert abs(dB[i, f] - fd) <= bound, (i, f, dB[i, f], fd, bound)
        got.append(dB[i, f])
        want.append(fd)
    got, want = np.array(got), np.array(want)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-2

    # torch.sparse FP64 autograd over the rows where dY is nonzero
    sample = np.sort(rng.choice(csr.m, 4096, replace=False))
    dYs = torch.zeros(csr.m, K)
    dYs[sample] = dY[sample]
    B.grad = None
    vals.grad = None
    tag.spmm(A, B, vals).backward(dYs.to(cuda))
    dv = vals.grad.double().cpu().numpy()
    dB = B.grad.double().cpu().numpy()
    sub = sp.csr_matrix((np.arange(csr.nnz, dtype=np.float64), csr.col, csr.rowptr),
                        shape=(csr.m_pad, csr.n))[sample]
    edge = sub.data.astype(np.int64)                                  # A's edge of each sampled entry
    sub_rows = np.repeat(np.arange(sample.size), np.diff(sub.indptr))
    v_s = torch.from_numpy(v64[edge]).requires_grad_()
    B_s = torch.from_numpy(B64).requires_grad_()
    S = torch.sparse_coo_tensor(torch.from_numpy(np.stack([sub_rows, sub.indices]).astype(np.int64)), v_s,
                                (sample.size, csr.n))
    (torch.sparse.mm(S, B_s) * dYs[sample].double()).sum().backward()
    assert _rel(dv[edge], v_s.grad.numpy()) < 1e-2
    outside = np.ones(csr.nnz, bool)
    outside[edge] = False
    assert np.all(dv[outside] == 0)
    assert _rel(dB, B_s.grad.numpy()) < 1e-2


def test_batched_1024_graphs(cuda):
    gs = g.ba_batch(1024)
    K = 128
    print(f"[config] batched graphs={len(gs)} nnz={sum(c.nnz for c in gs)}", flush=True)
    batch = RoutedBatch(gs)
    Bs = [g.dense_features(c.n, K, seed=7000 + i) for i, c in enumerate(gs)]
    Bcat = batch.stack_operands([torch.from_numpy(b).to(cuda) for b in Bs])
    eager = batch.matmul(Bcat)
    C = torch.full_like(eager, float("nan"))
    graph = batch.capture(Bcat, C)
    graph.replay()
    torch.cuda.synchronize()
    # windows over the caps (the hubs of the larger graphs) are split into work
    # items that add atomically, in an order that may change between launches
    diff = (C - eager).abs()
    print(f"[config] replay vs eager: {int((diff.amax(1) > 0).sum())} rows differ, max {float(diff.max()):.2e}",
          flush=True)
    assert float(diff.max()) <= 1e-6 * float(eager.abs().max())
    outs = batch.outputs(C)
    worst = 0.0
    for i, c in enumerate(gs):
        ref = oracle.csr_spmm64(c.rowptr, c.col, torch.from_numpy(c.val).half().double().numpy(),
                                Bs[i].astype(np.float64), c.m)
        err = _rel(outs[i].double().cpu().numpy(), ref)
        worst = max(worst, err)
        assert err < 1e-5, (i, c.n, err)
    print(f"[config] batched worst per-graph error {worst:.2e}", flush=True)
