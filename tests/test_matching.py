"""The setup_24matching statistics ops on the CPU: matching_utils.match_2to4,
matching_utils_hash.match_2to4_hash and matching_utils_bucket.match_2to4_bucket.

Pinned by the reference itself: tests/golden/matching{,_hash,_bucket}_stats.json
hold the statistics the reference's block_matching_window_para_dr.cpp,
block_matching_hash.cpp and block_matching_bucket.cpp return for seeded
matrices and parameter sets (written by make_golden_matching.py from
oracle/_ref); where oracle/_ref is built the comparison also runs live on
further matrices.  Plus the reference test's own invariants
(test_24_matching.py:177-180, 331-332), hand-computed windows for each
variant's distinguishing rule and the error paths."""
import ctypes
import glob
import json
import os
import sys

import numpy as np
import pytest

from matching_cases import CASES, PARAMS, WINDOWS
from trspmm import graphs as g

import matching_utils
import matching_utils_bucket
import matching_utils_hash

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "matching_stats.json")))
GOLDEN_HASH = json.load(open(os.path.join(HERE, "golden", "matching_hash_stats.json")))
GOLDEN_BUCKET = json.load(open(os.path.join(HERE, "golden", "matching_bucket_stats.json")))
VARIANTS = {"hash": (matching_utils_hash.match_2to4_hash, GOLDEN_HASH, "matching_utils_hash_ref", "match_2to4_hash"),
            "bucket": (matching_utils_bucket.match_2to4_bucket, GOLDEN_BUCKET, "matching_utils_bucket_ref",
                       "match_2to4_bucket")}
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


def _ref_module(name="matching_utils_ref"):
    if not glob.glob(os.path.join(REF_DIR, name + "*.so")):
        pytest.skip("oracle/_ref not built (make -C oracle ref needs the reference sources)")
    if REF_DIR not in sys.path:
        sys.path.insert(0, REF_DIR)
    import importlib
    return importlib.import_module(name)


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


# ---------------------------------------------------------------- hash / bucket variants

def _golden_ids(recs):
    return [f"{r['case']}-w{r['kwargs']['window_size']}" for r in recs]


@pytest.mark.parametrize("rec", GOLDEN_HASH, ids=_golden_ids(GOLDEN_HASH))
def test_hash_matches_reference_golden(rec):
    c = CASES[rec["case"]]()
    _same(rec["stats"], matching_utils_hash.match_2to4_hash(c.rowptr, c.col, c.val, **rec["kwargs"]))


@pytest.mark.parametrize("rec", GOLDEN_BUCKET, ids=_golden_ids(GOLDEN_BUCKET))
def test_bucket_matches_reference_golden(rec):
    c = CASES[rec["case"]]()
    _same(rec["stats"], matching_utils_bucket.match_2to4_bucket(c.rowptr, c.col, c.val, **rec["kwargs"]))


@pytest.mark.parametrize("variant", sorted(VARIANTS))
@pytest.mark.parametrize("seed", range(4))
def test_variant_matches_reference_live(variant, seed):
    ours, _, mod, fn = VARIANTS[variant]
    ref = getattr(_ref_module(mod), fn)
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(50, 700))
    c = g.er_gcn(n, float(rng.uniform(2, 30)), seed=seed) if seed % 2 else g.rmat_gcn(9 + seed % 3, seed=seed)
    for w in WINDOWS + [33]:
        _same(_quiet(ref, c.rowptr, c.col, c.val, window_size=w), ours(c.rowptr, c.col, c.val, window_size=w))


@pytest.mark.parametrize("variant", sorted(VARIANTS))
@pytest.mark.parametrize("name", sorted(CASES))
def test_variant_invariants(variant, name):
    c = CASES[name]()
    s = VARIANTS[variant][0](c.rowptr, c.col, c.val)
    assert s["total_nnz"] == c.nnz
    assert s["block_padding_groups"] == s["num_16x16_blocks"] * 4 - s["num_groups"]
    assert s["num_row_panels"] == (c.m_pad + 15) // 16
    # a group holds at most 4 columns of <= 16 rows: slots are nnz + padding, 32 per group
    # (the hash variant's complement pairs count no nonzeros, so there the slots are an upper bound)
    if variant == "bucket":
        assert s["fake_zeros"] + s["total_nnz"] >= 0 and s["fake_zeros"] <= 32 * s["num_groups"]


def _window(masks, rows=16):
    """CSR of one window: column -> 16-bit row mask."""
    r, cc = [], []
    for col, m in masks.items():
        for b in range(16):
            if m >> b & 1:
                r.append(b)
                cc.append(col)
    return g.csr_from_pairs(np.array(r), np.array(cc), None, rows, max(masks) + 1)


def test_hash_complement_pairs_count_no_nonzeros():
    """Column 0 on rows 0-7, column 1 on rows 8-15: exact complements, paired on
    the chain path, which leaves the pair's nonzero count at 0
    (block_matching_hash.cpp:150-156), so its lone group counts 2*16 - 0 = 32
    padding zeros; match_2to4 pairs the same columns, sends them to the pool
    and counts 32 - 16 = 16."""
    c = _window({0: 0x00FF, 1: 0xFF00})
    h = matching_utils_hash.match_2to4_hash(c.rowptr, c.col, c.val)
    assert (h["num_groups"], h["fine_fallback"], h["coarse_fallback"], h["fake_zeros"]) == (1, 0, 2, 32)
    assert matching_utils.match_2to4(c.rowptr, c.col, c.val)["fake_zeros"] == 16


def test_hash_chain_takes_the_last_column_first():
    """Column 0 (rows 0-6) has two exact complements, columns 1 and 50 (rows
    7-15); the chain lists the later one first, so 0 pairs with 50.  Columns
    2..49 pair off as complements (0x0F0F / 0xF0F0).  Column 1 then finds only
    taken columns in its 32-column reach and column 51 (row 0) nothing after
    it: two half pairs (fine fallback 2).  Had the chain started from column 1,
    column 50 would pair with 51 and nothing would fall back.  Groups: 25 pairs
    with no common rows group in twos (12), the 25th with column 1's half pair
    (32 - 9 zeros), column 51's half pair alone (32 - 1): 14 groups, 12*32 + 23
    + 31 = 438 zeros."""
    masks = {0: 0x007F, 1: 0xFF80, 50: 0xFF80, 51: 0x0001}
    masks.update({k: (0x0F0F if k % 2 == 0 else 0xF0F0) for k in range(2, 50)})
    c = _window(masks)
    h = matching_utils_hash.match_2to4_hash(c.rowptr, c.col, c.val)
    assert (h["fine_fallback"], h["coarse_fallback"], h["num_groups"], h["fake_zeros"]) == (2, 2, 14, 438)
    assert (h["num_16x16_blocks"], h["block_padding_groups"]) == (4, 2)


def test_bucket_counts_32_slots_per_group_in_a_short_window():
    """4 rows, columns 0 (row 0) and 1 (row 1): both top-heavy, no bottom or
    balanced partner, so two half pairs (fine 2) that group together; the
    bucket variant counts 32 - 2 = 30 zeros whatever the window height
    (block_matching_bucket.cpp:341-350), the hash variant pairs the two
    columns directly and counts 2*4 - 2 = 6 for the lone group."""
    rp = np.array([0, 1, 2, 2, 2], np.int32)
    ci = np.array([0, 1], np.int32)
    v = np.ones(2, np.float32)
    b = matching_utils_bucket.match_2to4_bucket(rp, ci, v)
    assert (b["num_groups"], b["fine_fallback"], b["coarse_fallback"], b["fake_zeros"]) == (1, 2, 0, 30)
    h = matching_utils_hash.match_2to4_hash(rp, ci, v)
    assert (h["num_groups"], h["fine_fallback"], h["coarse_fallback"], h["fake_zeros"]) == (1, 0, 2, 6)


def test_bucket_pairs_top_with_bottom_first():
    """Columns 0 (rows 0-1, top), 1 (rows 0-3 | 8, top), 2 (rows 8-9,
    bottom), 3 (rows 0 | 8, balanced).  Top x bottom first: column 0 takes 2
    (no common row); column 1 then finds no bottom column left and takes the
    balanced 3 (common rows 0 and 8).  Pair (0,2) touches rows {0,1,8,9},
    pair (1,3) has rows 0 and 8 twice: they conflict, so each becomes its own
    group: 2 groups, coarse 4, 32 - 4 and 32 - 7 zeros."""
    c = _window({0: 0x0003, 1: 0x010F, 2: 0x0300, 3: 0x0101})
    b = matching_utils_bucket.match_2to4_bucket(c.rowptr, c.col, c.val)
    assert (b["num_groups"], b["fine_fallback"], b["coarse_fallback"], b["fake_zeros"]) == (2, 0, 4, 28 + 25)


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_variant_thread_count_does_not_change_result(variant):
    c = CASES["rmat12"]()
    f = VARIANTS[variant][0]
    assert f(c.rowptr, c.col, c.val, threads=1) == f(c.rowptr, c.col, c.val, threads=5)


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_variant_rejects(variant):
    f = VARIANTS[variant][0]
    c = CASES["er512"]()
    for w in (0, 65):
        with pytest.raises(RuntimeError, match="window_size"):
            f(c.rowptr, c.col, c.val, window_size=w)
    with pytest.raises(RuntimeError, match="invalid matrix"):
        f(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
