"""The drop-in modules, called the way the reference's own callers call them.

Libra5BlockOnline.block_sptc_2to4_online must return the reference router's
29 arrays bit for bit (the golden fixtures were written by the reference's own
block_online.cpp) plus a float64 duration.  Libra6SpMMOnline's ops must return
the product that metadata encodes: [out f32[mOri, dimN], ms, ...], checked
against the oracle's FP64 product of the decoded metadata (values rounded to
fp16 as the reference's caller hands them over, mdataset2_online.py:146-153),
tolerance 1e-5 relative Frobenius (fp32 accumulation order).
"""
import ast
import glob
import os

import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "router_*.npz")))
TOL = 1e-5


def _fixture(path):
    z = np.load(path)
    kw = ast.literal_eval(str(z["kwargs"]))
    env = ast.literal_eval(str(z["env"]))
    arrays = {n: z["out_" + n] for n in oracle.NAMES}
    return z, kw, env, arrays


def _libra5(z, kw):
    import Libra5BlockOnline
    return Libra5BlockOnline.block_sptc_2to4_online(
        torch.from_numpy(z["rowptr"]), torch.from_numpy(z["col"]), torch.from_numpy(z["val"]),
        kw.get("partsize_t", 32), kw.get("partsize_c", 32), kw.get("shortsize", 3), kw.get("density", 8),
        kw.get("window", 16), 16, float(kw.get("tc_min_util", 0.0)), 0.5, 12)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[7:-4] for p in GOLDEN])
def test_libra5_matches_reference_router(cuda, path, monkeypatch):
    z, kw, env, arrays = _fixture(path)
    print(f"[libra5] {os.path.basename(path)}", flush=True)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    monkeypatch.delenv("TRSPMM_ROUTE_MODE", raising=False)
    out = _libra5(z, kw)
    assert len(out) == 30
    for name, got in zip(oracle.NAMES, out[:29]):
        want = arrays[name]
        assert got.device.type == "cpu"
        assert got.numpy().dtype == want.dtype, name
        assert tuple(got.shape) == want.shape, (name, tuple(got.shape), want.shape)
        assert np.array_equal(got.numpy(), want), name
    assert out[29].dtype == torch.float64 and out[29].dim() == 0 and float(out[29]) >= 0


def _info(z, kw):
    """The reference caller's preprocessing (mdataset2_online.py:100-153)."""
    t = _libra5(z, kw)
    names = oracle.NAMES
    d = dict(zip(names, t[:29]))
    for k in ("t_value", "c_value", "c_value_short"):
        d[k] = d[k].half()
    d["parts_t"] = d["t_atomic"].shape[0]
    d["parts_c"] = d["c_atomic"].shape[0]
    d["parts_c_short"] = d["c_atomic_short"].shape[0]
    return d


def _rhs(kOri, dimN, seed=1000):
    return torch.randn(kOri, dimN, generator=torch.Generator().manual_seed(seed)).half()


def _ref_product(arrays, x, mOri, window=16, routes="all"):
    arr = {k: np.asarray(v.float() if v.dtype == torch.float16 else v) if isinstance(v, torch.Tensor) else v
           for k, v in arrays.items() if k in oracle.NAMES}
    if routes == "sptc":
        arr["t_row_offset"] = np.zeros(1, np.int32)
        arr["t_window_row"] = np.zeros(0, np.int32)
        arr["c_row"] = np.zeros(0, np.int32)
        arr["c_row_short"] = np.zeros(0, np.int32)
    return oracle.spmm64(arr, x.float().numpy(), mOri, window)


PARALLEL = GOLDEN          # every fixture, the window-8 one included


@pytest.mark.parametrize("path", PARALLEL, ids=[os.path.basename(p)[7:-4] for p in PARALLEL])
@pytest.mark.parametrize("dimN", [64, 20])
def test_libra6_parallel_op(cuda, path, dimN, monkeypatch):
    import Libra6SpMMOnline
    z, kw, env, _ = _fixture(path)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    monkeypatch.delenv("TRSPMM_ROUTE_MODE", raising=False)
    print(f"[libra6] {os.path.basename(path)} dimN={dimN}", flush=True)
    d = _info(z, kw)
    mOri = z["rowptr"].size - 1
    kOri = int(z["col"].max()) + 1
    x = _rhs(kOri, dimN)
    window = kw.get("window", 16)
    res = Libra6SpMMOnline.forward_fp16_tcu_cuda_sptc_mma_parallel_online(
        d["t_row_offset"], d["t_block_offset"], d["t_column"], d["t_value"], d["t_window_row"], d["t_atomic"],
        d["t_binary"],
        d["c_row_offset"], d["c_row"], d["c_atomic"], d["c_column"], d["c_value"],
        d["c_row_offset_short"], d["c_row_short"], d["c_atomic_short"], d["c_column_short"], d["c_value_short"],
        d["s_column"], d["s_value"], d["s_pos"], d["s_offsets"], d["s_tile_window"], d["s_tile_column"],
        d["s_packed_a"], d["s_packed_meta"], d["s_window_tile_offset"], x,
        d["parts_t"], d["parts_c"], kw.get("partsize_c", 32), d["parts_c_short"], window, dimN, mOri, kOri, 3)
    assert len(res) == 6
    out = res[0]
    assert out.device.type == "cpu" and out.dtype == torch.float32 and tuple(out.shape) == (mOri, dimN)
    for ms in res[1:]:
        assert ms.dtype == torch.float32 and ms.dim() == 0 and float(ms) >= 0
    ref = _ref_product(d, x, mOri, window)
    err = np.linalg.norm(out.numpy() - ref) / max(np.linalg.norm(ref), 1e-30)
    assert err < TOL, err


@pytest.mark.parametrize("op", ["sptc", "sptc_mma", "sptc_mma_packed"])
@pytest.mark.parametrize("path", [p for p in PARALLEL if "er512" in p or "hub" in p or "w8" in p])
def test_libra6_sptc_ops(cuda, op, path):
    import Libra6SpMMOnline
    z, kw, env, _ = _fixture(path)
    print(f"[libra6 {op}] {os.path.basename(path)}", flush=True)
    d = _info(z, kw)
    mOri = z["rowptr"].size - 1
    kOri = int(z["col"].max()) + 1
    dimN = 32
    x = _rhs(kOri, dimN, seed=7)
    window = kw.get("window", 16)
    if op == "sptc":
        out, ms = Libra6SpMMOnline.forward_fp16_sptc_online(d["s_column"], d["s_value"], d["s_pos"], d["s_window"],
                                                             x, window, dimN, mOri, kOri, 2)
    elif op == "sptc_mma":
        out, ms = Libra6SpMMOnline.forward_fp16_sptc_mma_online(d["s_column"], d["s_value"], d["s_pos"],
                                                                 d["s_offsets"], x, window, dimN, mOri, kOri, 2)
    else:
        out, ms = Libra6SpMMOnline.forward_fp16_sptc_mma_packed_online(
            d["s_tile_column"], d["s_packed_a"], d["s_packed_meta"], d["s_window_tile_offset"], x, window, dimN,
            mOri, kOri, 2)
    assert tuple(out.shape) == (mOri, dimN) and out.dtype == torch.float32 and float(ms) >= 0
    ref = _ref_product(d, x, mOri, window, routes="sptc")
    err = np.linalg.norm(out.numpy() - ref) / max(np.linalg.norm(ref), 1e-30)
    assert err < TOL, err


def test_libra6_rejects_device_tensors(cuda):
    import Libra6SpMMOnline
    z, kw, _, _ = _fixture(GOLDEN[0])
    d = _info(z, kw)
    x = _rhs(int(z["col"].max()) + 1, 16).cuda()
    with pytest.raises(RuntimeError, match="CPU tensor"):
        Libra6SpMMOnline.forward_fp16_sptc_mma_packed_online(
            d["s_tile_column"], d["s_packed_a"], d["s_packed_meta"], d["s_window_tile_offset"], x, 16, 16,
            z["rowptr"].size - 1, int(z["col"].max()) + 1, 1)
