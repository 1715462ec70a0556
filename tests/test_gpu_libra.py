"""The legacy Libra kernel family (LibraSpMM / trspmm_libra_fp16_*) on the GPU.

Each host function of TR-source/SpMM/mGCNkernel.cu this build replaces (fp16
and tf32 members) runs on
legacy-format metadata written by tests/libra_format.py and is checked
against the FP64 product of what oracle.decode_libra reads from that metadata
(values as fp16): the full products (tcu_cuda, _seq, _hybrid, _hybrid_seq on
the matrix cores, cc_only on the vector ALU), the TC route alone (nomap) and
the residual lists alone (split, split_v1); fp32 accumulation, 1e-5 relative.
Also: dimN not a multiple of 8 (padded operand copies), rows beyond mOri
dropped, the returned mean ms, and errors raised for malformed offsets.
"""
import numpy as np
import pytest
import torch

import LibraSpMM as L
import oracle
from graphs_small import CASES
from libra_format import encode
from trspmm import graphs as g

pytestmark = pytest.mark.gpu

GRAPHS = {"rmat10": lambda: g.rmat_gcn(10, seed=1), "er512": CASES["er512"], "dense": CASES["dense"],
          "ragged": CASES["ragged"], "hub": CASES["hub"]}
FULL = ["spmm_forward_fp16_tcu_cuda_kernel", "spmm_forward_fp16_tcu_cuda_kernel_seq",
        "spmm_forward_fp16_tcu_cuda_hybrid_kernel", "spmm_forward_fp16_tcu_cuda_hybrid_kernel_seq",
        "spmm_forward_fp16_tcu_cuda_kernel_cc_only"]
NAMES = ["t_row_offset", "t_blockNew_offset", "t_value", "t_column", "t_window_row", "t_atomic", "t_binary",
         "c_row_offset", "c_row", "c_atomic", "c_column", "c_value",
         "c_row_offset_short", "c_row_short", "c_atomic_short", "c_column_short", "c_value_short"]


def _ref(meta, B, mOri, which="all", tf32=False):
    m = dict(meta)
    if which == "tc":
        for k in ("c_row", "c_row_short"):
            m[k] = torch.zeros(0, dtype=torch.int32)
        m["c_row_offset"] = m["c_row_offset_short"] = torch.zeros(1, dtype=torch.int32)
    if which == "res":
        m["t_window_row"] = torch.zeros(0, dtype=torch.int32)
        m["t_row_offset"] = torch.zeros(1, dtype=torch.int32)
    gt = lambda k: m[k].numpy()            # noqa: E731
    r, c, v = oracle.decode_libra(gt("t_row_offset"), gt("t_blockNew_offset"), m["t_value"].float().numpy(),
                                  gt("t_column"), gt("t_window_row"), gt("t_binary"), gt("c_row_offset"), gt("c_row"),
                                  gt("c_column"), m["c_value"].float().numpy(), gt("c_row_offset_short"),
                                  gt("c_row_short"), gt("c_column_short"), m["c_value_short"].float().numpy(),
                                  tf32=tf32)
    keep = r < mOri
    return oracle.coo_spmm64(r[keep], c[keep], v[keep], B.double().numpy(), mOri)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _setup(case, dimN, density=4):
    csr = GRAPHS[case]()
    meta, counts = encode(csr, density=density)
    B = torch.randn(csr.n, dimN, generator=torch.Generator().manual_seed(dimN)).half()
    return csr, meta, counts, B


@pytest.mark.parametrize("fn", FULL)
@pytest.mark.parametrize("case", sorted(GRAPHS))
@pytest.mark.parametrize("dimN", [64, 128, 40])
def test_full_products(cuda, fn, case, dimN):
    csr, meta, counts, B = _setup(case, dimN)
    mOri = csr.m
    print(f"[libra] {fn} {case} dimN={dimN} {counts}", flush=True)
    out = torch.full((mOri, dimN), float("nan"), device=cuda)
    ms = getattr(L, fn)(*[meta[n].to(cuda) for n in NAMES], B.to(cuda), out, counts["parts_t"], counts["parts_c"], 32,
                        counts["parts_c_short"], dimN, mOri, 3)
    assert ms > 0
    got = out.double().cpu().numpy()
    assert np.isfinite(got).all()
    assert _rel(got, _ref(meta, B, mOri)) < 1e-5


@pytest.mark.parametrize("case", ["dense", "rmat10", "hub"])
def test_nomap_is_the_tc_route(cuda, case):
    csr, meta, counts, B = _setup(case, 64, density=2)
    assert counts["parts_t"] > 0
    out = torch.empty((csr.m, 64), device=cuda)
    L.spmm_forward_fp16_tcu_cuda_nomap_kernel(*[meta[n].to(cuda) for n in NAMES[:7]], B.to(cuda), out,
                                              counts["parts_t"], 64, csr.m, 2)
    assert _rel(out.double().cpu().numpy(), _ref(meta, B, csr.m, "tc")) < 1e-5


@pytest.mark.parametrize("fn", ["spmm_forward_fp16_cuda_kernel_split", "spmm_forward_fp16_cuda_kernel_split_v1"])
@pytest.mark.parametrize("case", ["er512", "rmat10", "ragged"])
def test_split_is_the_residual_route(cuda, fn, case):
    csr, meta, counts, B = _setup(case, 128)
    out = torch.empty((csr.m, 128), device=cuda)
    getattr(L, fn)(*[meta[n].to(cuda) for n in NAMES[7:]], B.to(cuda), out, 32, 128, csr.m, 2, counts["parts_c"],
                   counts["parts_c_short"], True)
    assert _rel(out.double().cpu().numpy(), _ref(meta, B, csr.m, "res")) < 1e-5


def test_rows_beyond_mOri_are_dropped(cuda):
    csr, meta, counts, B = _setup("rmat10", 64)
    mOri = csr.m - 100
    out = torch.empty((mOri, 64), device=cuda)
    L.spmm_forward_fp16_tcu_cuda_kernel(*[meta[n].to(cuda) for n in NAMES], B.to(cuda), out, counts["parts_t"],
                                        counts["parts_c"], 32, counts["parts_c_short"], 64, mOri, 1)
    assert _rel(out.double().cpu().numpy(), _ref(meta, B, mOri)) < 1e-5


def test_malformed_offsets_raise(cuda):
    csr, meta, counts, B = _setup("er512", 64)
    bad = dict(meta)
    bad["c_row_offset"] = meta["c_row_offset"].clone()
    bad["c_row_offset"][1] = -5
    out = torch.empty((csr.m, 64), device=cuda)
    with pytest.raises(RuntimeError, match="non-decreasing|start at 0"):
        L.spmm_forward_fp16_tcu_cuda_kernel(*[bad[n].to(cuda) for n in NAMES], B.to(cuda), out, counts["parts_t"],
                                            counts["parts_c"], 32, counts["parts_c_short"], 64, csr.m, 1)
    with pytest.raises(RuntimeError, match="output_matrix"):
        L.spmm_forward_fp16_tcu_cuda_kernel(*[meta[n].to(cuda) for n in NAMES], B.to(cuda), out[:10],
                                            counts["parts_t"], counts["parts_c"], 32, counts["parts_c_short"], 64,
                                            csr.m, 1)


TF32 = ["spmm_forward_tf32_tcu_cuda_kernel", "spmm_forward_tf32_tcu_cuda_kernel_seq"]


@pytest.mark.parametrize("fn", TF32)
@pytest.mark.parametrize("case", sorted(GRAPHS))
@pytest.mark.parametrize("dimN", [64, 40])
def test_tf32_products(cuda, fn, case, dimN):
    """fp32 values and operand, 8 x 4 TC blocks: three bf16 products summed, against
    the FP64 product of the fp32 metadata (finer than tf32: 3e-5 relative)."""
    csr = GRAPHS[case]()
    meta, counts = encode(csr, density=2, tf32=True)
    B = torch.randn(csr.n, dimN, generator=torch.Generator().manual_seed(dimN + 5))
    out = torch.full((csr.m, dimN), float("nan"), device=cuda)
    print(f"[libra-tf32] {fn} {case} dimN={dimN} {counts}", flush=True)
    ms = getattr(L, fn)(*[meta[n].to(cuda) for n in NAMES], B.to(cuda), out, counts["parts_t"], counts["parts_c"], 32,
                        counts["parts_c_short"], dimN, csr.m, 2)
    assert ms > 0
    got = out.double().cpu().numpy()
    assert np.isfinite(got).all()
    assert _rel(got, _ref(meta, B, csr.m, tf32=True)) < 3e-5


@pytest.mark.parametrize("case", ["er512", "rmat10"])
def test_tf32_split_is_the_residual_route(cuda, case):
    csr = GRAPHS[case]()
    meta, counts = encode(csr, density=2, tf32=True)
    B = torch.randn(csr.n, 64, generator=torch.Generator().manual_seed(9))
    out = torch.empty((csr.m, 64), device=cuda)
    L.spmm_forward_tf32_cuda_kernel_split(*[meta[n].to(cuda) for n in NAMES[7:]], B.to(cuda), out, 32, 64, csr.m, 2,
                                          counts["parts_c"], counts["parts_c_short"], False)
    assert _rel(out.double().cpu().numpy(), _ref(meta, B, csr.m, "res", tf32=True)) < 3e-5
