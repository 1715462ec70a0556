"""Argument contract of the drop-in modules (CPU: every case raises before any
device work), following the reference's checks: block_online.cpp:551-553 and
mGCN_online.cpp:802-932 (TORCH_CHECK -> RuntimeError)."""
import os

import numpy as np
import pytest
import torch

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "router_er512.npz")


def _args(**over):
    z = np.load(GOLDEN)
    d = {n: torch.from_numpy(z["out_" + n]) for n in oracle.NAMES}
    for k in ("t_value", "c_value", "c_value_short"):
        d[k] = d[k].half()
    kOri = int(z["col"].max()) + 1
    mOri = z["rowptr"].size - 1
    a = dict(d)
    a.update(rhs=torch.zeros(kOri * 32, dtype=torch.float16), parts_t=d["t_atomic"].numel(),
             parts_c=d["c_atomic"].numel(), partsize_c=32, parts_c_short=d["c_atomic_short"].numel(), window=16,
             dimN=32, mOri=mOri, kOri=kOri, epoches=1)
    a.update(over)
    return a


def _parallel(a):
    import Libra6SpMMOnline
    return Libra6SpMMOnline.forward_fp16_tcu_cuda_sptc_mma_parallel_online(
        a["t_row_offset"], a["t_block_offset"], a["t_column"], a["t_value"], a["t_window_row"], a["t_atomic"],
        a["t_binary"], a["c_row_offset"], a["c_row"], a["c_atomic"], a["c_column"], a["c_value"],
        a["c_row_offset_short"], a["c_row_short"], a["c_atomic_short"], a["c_column_short"], a["c_value_short"],
        a["s_column"], a["s_value"], a["s_pos"], a["s_offsets"], a["s_tile_window"], a["s_tile_column"],
        a["s_packed_a"], a["s_packed_meta"], a["s_window_tile_offset"], a["rhs"], a["parts_t"], a["parts_c"],
        a["partsize_c"], a["parts_c_short"], a["window"], a["dimN"], a["mOri"], a["kOri"], a["epoches"])


@pytest.mark.parametrize("over, msg", [
    ({"t_value": None}, "unexpected dtype"),
    ({"window": 8}, "s_value must have shape"),       # window 8 runs, but not on 16-row metadata
    ({"window": 17}, "0 < window <= 16"),
    ({"window": 0}, "window"),
    ({"epoches": 0}, "epoches must be positive"),
    ({"dimN": 0}, "dimN must be positive"),
    ({"rhs": torch.zeros(7, dtype=torch.float16)}, "kOri \\* dimN"),
    ({"parts_t": 5}, "parts_t"),
    ({"parts_c": -1}, "part count"),
    ({"s_offsets": None}, "s_offsets"),
    ({"s_packed_meta": None}, "s_packed_meta"),
    ({"s_pos": None}, "unexpected dtype"),
])
def test_parallel_op_rejects(over, msg):
    a = _args()
    if "t_value" in over and over["t_value"] is None:
        over = {"t_value": a["t_value"].float()}
    if "s_offsets" in over and over["s_offsets"] is None:
        so = a["s_offsets"].clone()
        so[-1] += 1
        over = {"s_offsets": so}
    if "s_packed_meta" in over and over["s_packed_meta"] is None:
        over = {"s_packed_meta": a["s_packed_meta"][:-1]}
    if "s_pos" in over and over["s_pos"] is None:
        over = {"s_pos": a["s_pos"].int()}
    a.update(over)
    with pytest.raises(RuntimeError, match=msg):
        _parallel(a)


def test_sptc_ops_reject():
    import Libra6SpMMOnline
    a = _args()
    with pytest.raises(RuntimeError, match="unexpected dtype"):
        Libra6SpMMOnline.forward_fp16_sptc_online(a["s_column"], a["s_value"].half(), a["s_pos"], a["s_window"],
                                                  a["rhs"], 16, 32, a["mOri"], a["kOri"], 1)
    so = a["s_offsets"].clone()
    so[0] = 1
    with pytest.raises(RuntimeError, match="start with 0"):
        Libra6SpMMOnline.forward_fp16_sptc_mma_online(a["s_column"], a["s_value"], a["s_pos"], so, a["rhs"], 16, 32,
                                                      a["mOri"], a["kOri"], 1)
    with pytest.raises(RuntimeError, match="unexpected dtype"):
        Libra6SpMMOnline.forward_fp16_sptc_mma_packed_online(a["s_tile_column"], a["s_packed_a"].float(),
                                                             a["s_packed_meta"], a["s_window_tile_offset"], a["rhs"],
                                                             16, 32, a["mOri"], a["kOri"], 1)


@pytest.mark.parametrize("change, msg", [
    ({"row": torch.int64}, "kInt32"),
    ({"val": torch.float64}, "kFloat32"),
    ({"window": 0}, "window > 0"),
    ({"window": 33}, "window <= 32"),
    ({"window": 24}, "window > 16"),
    ({"wide": 0}, "wide > 0"),
])
def test_router_rejects(change, msg):
    import Libra5BlockOnline
    z = np.load(GOLDEN)
    row = torch.from_numpy(z["rowptr"]).to(change.get("row", torch.int32))
    val = torch.from_numpy(z["val"]).to(change.get("val", torch.float32))
    with pytest.raises(RuntimeError, match=msg):
        Libra5BlockOnline.block_sptc_2to4_online(row, torch.from_numpy(z["col"]), val, 32, 32, 3, 8,
                                                 change.get("window", 16), change.get("wide", 16), 0.0, 0.5, 12)
