"""matching_utils_online.match_2to4_online: the device statistics op.

CPU part: the oracle (oracle/matching_online.py) is pinned to the reference's
own match_2to4 (block_matching_window_para_dr.cpp, through the golden
statistics tests/golden/matching_stats.json and, where oracle/_ref is built,
live).  The two heuristics agree wherever the online kernel's table order is
the column order and its 32 padding slots per group equal 2 x window rows:
whole 16-row windows with every column below 2048.  Renumbering column c as
c + 2048 k keeps every table slot, so the online statistics must not change
(while match_2to4's do); a hand-computed window fixes the order where columns
share a slot.

GPU part: the HIP kernel (trspmm_match_2to4_online) equals the oracle on
seeded matrices with and without slot collisions, equals the reference's
golden statistics on the eligible cases, and its error paths."""
import json
import os
import sys

import numpy as np
import pytest

from matching_cases import CASES
from trspmm import graphs as g

import matching_utils

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import matching_online as oracle  # noqa: E402

GOLDEN = json.load(open(os.path.join(HERE, "golden", "matching_stats.json")))
# the reference's match_2to4 keys for the online op's
AS_WINDOW = {"num_sptc_groups": "num_sptc_groups", "num_tc_blocks": "num_tc_blocks", "fine_fb": "fine_fallback",
             "coarse_fb_cols": "coarse_fallback", "fake_zeros": "fake_zeros", "num_16x16_blocks": "num_16x16_blocks",
             "block_padding_groups": "block_padding_groups", "total_rows": "total_rows",
             "total_cols": "total_cols", "total_nnz": "total_nnz"}


def _eligible(c):
    return (c.rowptr.size - 1) % 16 == 0 and int(c.col.max()) < 2048


ELIGIBLE = [r for r in GOLDEN if r["kwargs"].get("window_size", 16) == 16 and _eligible(CASES[r["case"]]())]


def _online_kwargs(kw):
    return {k: v for k, v in kw.items() if k in ("dense_threshold", "t_max")}


def _check_against_window(ref, got, kw):
    for ko, kr in AS_WINDOW.items():
        assert got[ko] == ref[kr], ko
    assert got["dense_to_tc_cols"] == ref["num_tc_blocks"] * kw.get("dense_threshold", 8)


def _renumbered(c, seed):
    """Column c -> c + 2048 * k_c (k_c in 0..7): every table slot stays."""
    k = np.random.default_rng(seed).integers(0, 8, size=int(c.col.max()) + 1)
    col = (c.col.astype(np.int64) + 2048 * k[c.col]).astype(np.int32)
    return c.rowptr, col, c.val


def _window(masks):
    r, cc = [], []
    for col, m in masks.items():
        for b in range(16):
            if m >> b & 1:
                r.append(b)
                cc.append(col)
    return g.csr_from_pairs(np.array(r), np.array(cc), None, 16, max(masks) + 1)


# ---------------------------------------------------------------- oracle (CPU)

def test_eligible_golden_records_exist():
    assert len(ELIGIBLE) >= 20


@pytest.mark.parametrize("rec", ELIGIBLE, ids=[f"{r['case']}-{i}" for i, r in enumerate(ELIGIBLE)])
def test_oracle_equals_reference_match_2to4(rec):
    c = CASES[rec["case"]]()
    got = oracle.match_2to4_online(c.rowptr, c.col, **_online_kwargs(rec["kwargs"]))
    _check_against_window(rec["stats"], got, rec["kwargs"])


@pytest.mark.parametrize("name", ["er512", "rmat10", "hub"])
def test_oracle_keeps_slots_under_renumbering(name):
    c = CASES[name]()
    base = oracle.match_2to4_online(c.rowptr, c.col)
    rp, col, val = _renumbered(c, 7)
    assert oracle.match_2to4_online(rp, col) == {**base, "total_cols": int(col.max()) + 1}
    # the column-ordered heuristic does see the renumbering: the check has teeth
    assert matching_utils.match_2to4(rp, col, val)["fine_fallback"] != base["fine_fb"] or \
        matching_utils.match_2to4(rp, col, val)["fake_zeros"] != base["fake_zeros"]


COLLIDING = {5: 0b0011, 6: 0b1010, 2052: 0b1110, 2053: 0b1101}


def test_oracle_hand_computed_collisions():
    """Columns 5, 6, 2052, 2053 (homes 5, 6, 4, 5).  Inserted by ascending
    column: 5 -> slot 5, 6 -> 6, 2052 -> 4, 2053 -> 5, 6 taken -> 7; table
    order 2052 (rows 1-3), 5 (rows 0-1), 6 (rows 1,3), 2053 (rows 0,2,3).
    With t_max 1: 2052 takes 5 (one common row; 6 and 2053 share two), 6
    takes 2053 (one).  The pairs (both 0010 / any 1111, both 1000 / any 1111)
    conflict: both go to the pool (coarse 4), which forms two padded groups
    of 5 nonzeros each: 2 groups, 2 x 27 zeros, fine 0.  (In column order 5
    would take 6 and 2052, 2053 would fall back: fine 2, coarse 2.)"""
    c = _window(COLLIDING)
    s = oracle.match_2to4_online(c.rowptr, c.col, t_max=1)
    assert (s["fine_fb"], s["coarse_fb_cols"], s["num_sptc_groups"], s["fake_zeros"]) == (0, 4, 2, 54)
    assert (s["dense_to_sptc_cols"], s["num_16x16_blocks"], s["block_padding_groups"]) == (4, 1, 2)


def test_oracle_table_order_without_collisions_is_slot_order():
    assert oracle.table_order({4100: 1, 3: 2, 2050: 4}) == [4, 2, 1]      # slots 2, 3, 4


@pytest.mark.parametrize("seed", range(3))
def test_oracle_equals_reference_live(seed):
    import glob
    ref_dir = os.path.join(os.path.dirname(HERE), "oracle", "_ref")
    if not glob.glob(os.path.join(ref_dir, "matching_utils_ref*.so")):
        pytest.skip("oracle/_ref not built (make -C oracle ref needs the reference sources)")
    sys.path.insert(0, ref_dir)
    import matching_utils_ref as R
    from test_matching import _quiet
    rng = np.random.default_rng(200 + seed)
    c = g.er_gcn(16 * int(rng.integers(5, 60)), float(rng.uniform(2, 40)), seed=seed)
    for kw in ({}, {"dense_threshold": 4, "t_max": 2}, {"t_max": 0}):
        ref = _quiet(R.match_2to4, c.rowptr, c.col, c.val, **kw)
        _check_against_window(ref, oracle.match_2to4_online(c.rowptr, c.col, **kw), kw)


# ---------------------------------------------------------------- device (GPU)

def _gpu():
    import matching_utils_online
    return matching_utils_online.match_2to4_online


@pytest.mark.gpu
@pytest.mark.parametrize("rec", ELIGIBLE, ids=[f"{r['case']}-{i}" for i, r in enumerate(ELIGIBLE)])
def test_gpu_equals_reference_golden(rec):
    c = CASES[rec["case"]]()
    got = _gpu()(c.rowptr, c.col, c.val, **_online_kwargs(rec["kwargs"]))
    _check_against_window(rec["stats"], got, rec["kwargs"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("kw", [{}, {"dense_threshold": 4, "t_max": 2}, {"t_max": 0}, {"dense_threshold": 16, "t_max": 16}])
def test_gpu_equals_oracle(name, kw):
    c = CASES[name]()
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)   # rmat12 has a window of 2086 distinct columns
        got = _gpu()(c.rowptr, c.col, c.val, **kw)
    assert got == oracle.match_2to4_online(c.rowptr, c.col, **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_gpu_equals_oracle_with_collisions(seed):
    """Columns spread over [0, 2^20): windows whose columns share table slots."""
    rng = np.random.default_rng(300 + seed)
    c = g.er_gcn(16 * int(rng.integers(4, 40)), float(rng.uniform(20, 120)), seed=seed)
    k = rng.integers(0, 512, size=int(c.col.max()) + 1)
    col = (c.col.astype(np.int64) + 2048 * k[c.col]).astype(np.int32)
    # and a few columns forced onto one slot per window
    col[::7] = (col[::7] // 2048) * 2048 + 17
    rp = c.rowptr
    order = [np.sort(np.unique(col[rp[r]:rp[r + 1]])) for r in range(rp.size - 1)]
    rp2 = np.concatenate([[0], np.cumsum([o.size for o in order])]).astype(np.int32)
    col2 = np.concatenate(order).astype(np.int32)
    for kw in ({}, {"t_max": 1}, {"dense_threshold": 3, "t_max": 4}):
        assert _gpu()(rp2, col2, np.ones(col2.size, np.float32), **kw) == oracle.match_2to4_online(rp2, col2, **kw)


@pytest.mark.gpu
def test_gpu_hand_computed_collisions():
    c = _window(COLLIDING)
    s = _gpu()(c.rowptr, c.col, c.val, t_max=1)
    assert (s["fine_fb"], s["coarse_fb_cols"], s["num_sptc_groups"], s["fake_zeros"]) == (0, 4, 2, 54)


@pytest.mark.gpu
def test_gpu_renumbering_keeps_statistics():
    c = CASES["rmat10"]()
    rp, col, val = _renumbered(c, 11)
    a = _gpu()(c.rowptr, c.col, c.val)
    b = _gpu()(rp, col, val)
    assert {**a, "total_cols": b["total_cols"]} == b


@pytest.mark.gpu
def test_gpu_device_tensors_and_repeat():
    import torch
    c = CASES["rmat12"]()
    rp = torch.from_numpy(c.rowptr).cuda()
    col = torch.from_numpy(c.col).cuda()
    with pytest.warns(RuntimeWarning):
        first = _gpu()(rp, col, None)
    for _ in range(3):
        with pytest.warns(RuntimeWarning):
            assert _gpu()(c.rowptr, c.col, c.val) == first


@pytest.mark.gpu
def test_gpu_rejects():
    f = _gpu()
    c = CASES["er512"]()
    with pytest.raises(RuntimeError, match="window_size"):
        f(c.rowptr, c.col, c.val, window_size=8)
    with pytest.raises(RuntimeError, match="dense_threshold"):
        f(c.rowptr, c.col, c.val, dense_threshold=0)
    bad = c.col.copy()
    bad[5] = -3
    with pytest.raises(RuntimeError, match="negative column"):
        f(c.rowptr, bad, c.val)
    with pytest.raises(RuntimeError, match="row_ptr"):
        f(c.rowptr[::-1].copy(), c.col, c.val)
    with pytest.raises(RuntimeError, match="invalid matrix"):
        f(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(3))
def test_gpu_window_overflow_keeps_smallest_columns(seed):
    """Windows with more distinct columns than the 2048-slot table: the 2048
    smallest column ids are kept (so the statistics equal those of the matrix
    with the larger ids removed from that window), and a warning says so."""
    rng = np.random.default_rng(400 + seed)
    rows, cols = [], []
    for r in range(32):                      # window 0 overflows, window 1 does not
        k = 220 if r < 16 else 40
        cs = np.sort(rng.choice(1 << 16, size=k, replace=False))
        rows += [r] * k
        cols += cs.tolist()
    w = g.csr_from_pairs(np.array(rows), np.array(cols), None, 32, 1 << 16)
    with pytest.warns(RuntimeWarning, match="2048 distinct"):
        got = _gpu()(w.rowptr, w.col, w.val)
    assert got == oracle.match_2to4_online(w.rowptr, w.col)
    keep = np.sort(np.unique(w.col[: w.rowptr[16]]))[:2048]
    mask = np.ones(w.col.size, bool)
    mask[: w.rowptr[16]] = np.isin(w.col[: w.rowptr[16]], keep)
    r = np.repeat(np.arange(32), np.diff(w.rowptr))[mask]
    cut = g.csr_from_pairs(r, w.col[mask], None, 32, 1 << 16)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        trimmed = _gpu()(cut.rowptr, cut.col, cut.val)
    assert {k: v for k, v in got.items() if k != "total_nnz" and k != "total_cols"} == \
        {k: v for k, v in trimmed.items() if k != "total_nnz" and k != "total_cols"}


def test_oracle_overflowing_window_keeps_smallest_columns():
    """A window with 16 x 200 distinct columns 0..3199 keeps 0..2047 (in slot
    order = column order): the same statistics as the window cut to them."""
    rows = np.repeat(np.arange(16), 200)
    cols = np.arange(3200)
    full = g.csr_from_pairs(rows, cols, None, 16, 3200)
    keep = cols < 2048
    cut = g.csr_from_pairs(rows[keep], cols[keep], None, 16, 2048)
    a = oracle.match_2to4_online(full.rowptr, full.col)
    b = oracle.match_2to4_online(cut.rowptr, cut.col)
    assert {k: v for k, v in a.items() if not k.startswith("total")} == \
        {k: v for k, v in b.items() if not k.startswith("total")}
    assert matching_utils.match_2to4(cut.rowptr, cut.col, cut.val)["fine_fallback"] == b["fine_fb"]
