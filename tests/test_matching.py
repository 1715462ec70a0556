"""matching_utils.match_2to4 (the setup_24matching statistics op) on the CPU.

Pinned by the reference itself: tests/golden/matching_stats.json holds the
statistics the reference's block_matching_window_para_dr.cpp returns for
seeded matrices and parameter sets (written by make_golden_matching.py from
oracle/_ref); where oracle/_ref is built the comparison also runs live on
further matrices.  Plus the reference test's own invariants
(test_24_matching.py:177-180, 331-332), a hand-computed window and the error
paths."""
import ctypes
import glob
import json
import os
import sys

import numpy as np
import pytest

from matching_cases import CASES, PARAMS
from trspmm import graphs as g

import matching_utils

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "matching_stats.json")))
REF_DIR = os.path.join(os.path.dirname(HERE), "oracle", "_ref")


def _same(ref, got):
    for k, v in ref.items():
        if isinstance(v, float):
            assert got[k] == pytest.approx(v, rel=1e-12, abs=0), k
        else:
            assert got[k] == v, k


@pytest.mark.parametrize("rec", GOLDEN, ids=[f"{r['case']}-{i}" for i, r in enumerate(GOLDEN)])
def test_matches_reference_golden(rec):
    c = CASES[rec["case"]]()
    got = matching_utils.match_2to4(c.rowptr, c.col, c.val, **rec["kwargs"])
    _same(rec["stats"], got)


def _ref_module():
    if not glob.glob(os.path.join(REF_DIR, "matching_utils_ref*.so")):
        pytest.skip("oracle/_ref not built (make -C oracle ref needs the reference sources)")
    if REF_DIR not in sys.path:
        sys.path.insert(0, REF_DIR)
    import matching_utils_ref
    return matching_utils_ref


def _quiet(f, *a, **k):
    libc = ctypes.CDLL(None)
    fd = os.dup(1)
    nul = os.open(os.devnull, os.O_WRONLY)
    os.dup2(nul, 1)
    try:
        return f(*a, **k)
    finally:
        libc.fflush(None)
        os.dup2(fd, 1)
        os.close(fd)
        os.close(nul)


@pytest.mark.parametrize("seed", range(6))
def test_matches_reference_live(seed):
    R = _ref_module()
    rng = np.random.default_rng(seed)
    n = int(rng.integers(50, 700))
    deg = float(rng.uniform(2, 30))
    c = g.er_gcn(n, deg, seed=seed) if seed % 2 else g.rmat_gcn(9 + seed % 3, seed=seed)
    for kw in PARAMS:
        _same(_quiet(R.match_2to4, c.rowptr, c.col, c.val, **kw), matching_utils.match_2to4(c.rowptr, c.col, c.val, **kw))


@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_invariants(name):
    c = CASES[name]()
    s = matching_utils.match_2to4(c.rowptr, c.col, c.val)
    assert s["total_nnz"] == c.nnz
    assert s["block_padding_groups"] == s["num_16x16_blocks"] * 4 - s["num_sptc_groups"]
    assert s["num_row_panels"] == (c.m_pad + 15) // 16
    assert s["total_cols"] == int(c.col.max()) + 1


def test_hand_computed_window():
    """16 rows; columns 0..3 touch rows {0}, {1}, {2}, {3}: pairs (0,1), (2,3)
    with no common row, grouped (no conflict) -> one 2:4 group with
    2*16 - 4 = 28 padding zeros; column 9 touches rows 0..15: no partner among
    later columns -> pool -> one padded group (2*16 - 16 = 16 zeros)."""
    rows = [0, 1, 2, 3] + list(range(16))
    cols = [0, 1, 2, 3] + [9] * 16
    c = g.csr_from_pairs(np.array(rows), np.array(cols), None, 16, 10)
    s = matching_utils.match_2to4(c.rowptr, c.col, c.val)
    assert s["num_sptc_groups"] == 2 and s["num_tc_blocks"] == 0
    assert s["fine_fallback"] == 1 and s["coarse_fallback"] == 0
    assert s["fake_zeros"] == 28 + 16
    assert s["num_16x16_blocks"] == 1 and s["block_padding_groups"] == 2


def test_thread_count_does_not_change_result():
    c = CASES["rmat12"]()
    a = matching_utils.match_2to4(c.rowptr, c.col, c.val, threads=1)
    b = matching_utils.match_2to4(c.rowptr, c.col, c.val, threads=7)
    assert a == b


@pytest.mark.parametrize("kw, msg", [({"window_size": 0}, "window_size"), ({"window_size": 65}, "window_size"),
                                     ({"dense_threshold": 0}, "dense_threshold")])
def test_rejects(kw, msg):
    c = CASES["er512"]()
    with pytest.raises(RuntimeError, match=msg):
        matching_utils.match_2to4(c.rowptr, c.col, c.val, **kw)
    with pytest.raises(RuntimeError, match="invalid matrix"):
        matching_utils.match_2to4(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
