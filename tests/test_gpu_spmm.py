"""Routed multiply (fused MFMA / SMFMAC / residual kernel) vs the oracle.

The oracle decodes the same routed metadata and multiplies in FP64 with the
values rounded to the operand type, so the only difference left is fp32
accumulation order: tolerance 1e-5 relative Frobenius, 1e-4 elementwise
(relative to the column scale)."""
import numpy as np
import pytest
import torch

import oracle
from graphs_small import CASES
from trspmm import ROUTE_ALL, ROUTE_RESIDUAL, ROUTE_SPTC, ROUTE_TC
from trspmm.sparse import RoutedSparse

pytestmark = pytest.mark.gpu

TOL_FRO = 1e-5


def _ref(csr, B, mask, exact, kw=None, dtype=np.float16):
    meta = oracle.route(csr.rowptr, csr.col, csr.val, exact=exact, **(kw or {}))
    arr = dict(meta.arrays)
    if not (mask & ROUTE_TC):
        for k in ("t_row_offset",):
            arr[k] = np.zeros(1, np.int32)
        arr["t_window_row"] = np.zeros(0, np.int32)
    if not (mask & ROUTE_SPTC):
        arr["s_window"] = np.zeros(0, np.int32)
        arr["s_pos"] = np.zeros((0, 16, 2), np.int8)
    if not (mask & ROUTE_RESIDUAL):
        for k in ("c_row", "c_row_short"):
            arr[k] = np.zeros(0, np.int32)
    rows, cols, vals = oracle.decode(arr, meta.window)
    if dtype == np.float16:
        vals = vals.astype(np.float16).astype(np.float64)
    else:
        vals = torch.from_numpy(vals).to(torch.bfloat16).double().numpy()
    keep = rows < csr.m
    return oracle.coo_spmm64(rows[keep], cols[keep], vals[keep], B.astype(np.float64), csr.m)


def _run(csr, K, mask=ROUTE_ALL, mode="exact", dtype=torch.float16, seed=1000):
    print(f"[spmm] nnz={csr.nnz} m={csr.m} K={K} mask={mask} mode={mode} dtype={dtype}", flush=True)
    A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode=mode, dtype=dtype)
    rng = np.random.default_rng(seed)
    Bf = rng.standard_normal((csr.n, K)).astype(np.float32)
    Bt = torch.from_numpy(Bf).to(dtype)
    C = A.matmul(Bt.cuda(), routes=mask)
    torch.cuda.synchronize()
    Bq = Bt.float().numpy()
    kw = {"sptc_min_nnz": 8} if mode == "tuned" else {}
    ref = _ref(csr, Bq, mask, mode != "compat", kw, dtype=np.float16 if dtype == torch.float16 else "bf16")
    got = C.cpu().numpy()
    den = max(np.linalg.norm(ref), 1e-30)
    err = np.linalg.norm(got - ref) / den
    scale = np.abs(ref).max() + 1e-30
    worst = np.abs(got - ref).max() / scale
    return err, worst, A


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("K", [64, 128])
def test_spmm_all_routes(cuda, case, K):
    err, worst, A = _run(CASES[case](), K)
    assert err < TOL_FRO and worst < 1e-4, (err, worst, A.stats)


@pytest.mark.parametrize("case", sorted(CASES))
def test_spmm_tuned_routing(cuda, case):
    err, worst, A = _run(CASES[case](), 128, mode="tuned")
    assert err < TOL_FRO and worst < 1e-4, (err, worst, A.stats)


@pytest.mark.parametrize("case", ["hub", "rmat10", "ragged", "dense"])
@pytest.mark.parametrize("mode", ["exact", "tuned"])
def test_spmm_vector_alu_residual_engine(cuda, case, mode, monkeypatch):
    """The vector-ALU residual engine (separate CSR kernel) gives the same product."""
    monkeypatch.setenv("TRSPMM_RESIDUAL", "valu")
    err, worst, A = _run(CASES[case](), 128, mode=mode)
    assert err < TOL_FRO and worst < 1e-4, (err, worst, A.stats)


@pytest.mark.parametrize("mask", [ROUTE_TC, ROUTE_SPTC, ROUTE_RESIDUAL])
def test_spmm_single_route(cuda, mask):
    err, worst, A = _run(CASES["dense"](), 64, mask=mask)
    assert err < TOL_FRO and worst < 1e-4, (mask, err, worst, A.stats)


@pytest.mark.parametrize("K", [8, 24, 72, 256])
def test_spmm_feature_widths(cuda, K):
    err, worst, _ = _run(CASES["dense"](), K)
    assert err < TOL_FRO and worst < 1e-4


def test_spmm_bf16(cuda):
    err, worst, _ = _run(CASES["dense"](), 64, dtype=torch.bfloat16)
    assert err < TOL_FRO and worst < 1e-4


def test_spmm_compat_mode_drops(cuda):
    """compat metadata encodes A_kept (overflow of 2:4 rows dropped): the
    product matches the decode of that metadata."""
    err, worst, _ = _run(CASES["er512"](), 64, mode="compat")
    assert err < TOL_FRO and worst < 1e-4


def test_exact_mode_vs_fp64_csr(cuda):
    """exact mode represents every nonzero once: compare with the plain FP64
    CSR product (values rounded to fp16 as the multiply consumes them)."""
    from trspmm import graphs as g
    csr = g.er_gcn(1024, 16.0, seed=2)
    A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode="exact")
    B = torch.randn(csr.n, 64, generator=torch.Generator().manual_seed(0)).half()
    C = (A @ B.cuda()).cpu().numpy()
    ref = oracle.csr_spmm64(csr.rowptr, csr.col, csr.val.astype(np.float16).astype(np.float64),
                            B.float().numpy(), csr.m)
    err = np.linalg.norm(C - ref) / np.linalg.norm(ref)
    assert err < TOL_FRO, err


@pytest.mark.parametrize("window", [1, 5, 8, 12])
@pytest.mark.parametrize("case", ["dense", "ragged", "rmat10", "hub"])
@pytest.mark.parametrize("mode", ["compat", "exact"])
def test_narrow_windows(cuda, window, case, mode):
    """Windows narrower than 16 rows (the reference takes any window in (0, 16]):
    the metadata's nonzeros, TC bits included (TC rows t_window_row + r), run
    as residual records of 16-row windows; windows of 5 and 12 rows straddle
    the 16-row tiles.  Against the FP64 product of the oracle's decode of the
    same window's metadata.  The reference codes a TC part's rows from
    16 * window_id whatever the window (block_online.cpp:494, read so by
    mGCNkernel_online.cu:207-209), so the CSR product itself is expected only
    where no TC part was emitted."""
    csr = CASES[case]()
    print(f"[narrow] {case} window={window} mode={mode}", flush=True)
    K = 64
    A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode=mode, window=window)
    Bt = torch.from_numpy(np.random.default_rng(window).standard_normal((csr.n, K)).astype(np.float32)).half()
    got = A.matmul(Bt.cuda()).cpu().numpy()
    meta = oracle.route(csr.rowptr, csr.col, csr.val, window=window, exact=mode != "compat")
    ref = oracle.spmm64(meta, Bt.float().numpy(), csr.m, window)
    assert np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30) < TOL_FRO
    if mode == "exact" and meta["t_window_row"].size == 0:     # every nonzero kept where it is
        vq = torch.from_numpy(csr.val).half().double().numpy()
        full = oracle.csr_spmm64(csr.rowptr, csr.col, vq, Bt.double().numpy(), csr.m)
        assert np.linalg.norm(got - full) / max(np.linalg.norm(full), 1e-30) < TOL_FRO
