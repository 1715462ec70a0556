"""Device router vs the oracle: all 29 metadata arrays bit-exact."""
import numpy as np
import pytest
import torch

import oracle
from graphs_small import CASES, VARIANTS
from trspmm.core import META_NAMES, route

pytestmark = pytest.mark.gpu


def _to_trspmm_kwargs(kw):
    m = {"partsize_t": "part_t", "partsize_c": "part_c", "shortsize": "short_c"}
    return {m.get(k, k): v for k, v in kw.items()}


def _compare(csr, kw, mode):
    print(f"[router] nnz={csr.nnz} rows={csr.m_pad} kw={kw} mode={mode}", flush=True)
    tuned = {"sptc_min_nnz": 8} if mode == "tuned" else {}
    meta = route(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                 mode=mode, **_to_trspmm_kwargs(kw), **tuned)
    got = meta.tensors("cpu")
    ref = oracle.route(csr.rowptr, csr.col, csr.val, exact=(mode != "compat"), **kw, **tuned)
    bad = [n for n, gt in zip(META_NAMES, got)
           if not (gt.numpy().dtype == ref[n].dtype and gt.numpy().shape == ref[n].shape
                   and np.array_equal(gt.numpy(), ref[n]))]
    assert not bad, f"arrays differ: {bad}"


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("mode", ["compat", "exact", "tuned"])
def test_router_bit_exact(cuda, case, mode):
    _compare(CASES[case](), {}, mode)


@pytest.mark.parametrize("vi", range(len(VARIANTS)))
@pytest.mark.parametrize("case", ["rmat10", "dense", "ragged"])
def test_router_variants(cuda, case, vi):
    _compare(CASES[case](), VARIANTS[vi], "compat")


@pytest.mark.parametrize("mode", ["compat", "exact"])
def test_router_more_nonzeros_than_one_scan_block(cuda, mode):
    """442 884 nonzeros: the per-nonzero prefix sums take the multi-pass scan
    (more than 2^18 entries)."""
    from trspmm import graphs as g
    _compare(g.rmat_gcn(14, seed=4), {}, mode)


@pytest.mark.parametrize("cols_per_row", [300, 5000])
def test_router_hub_windows_sorted_by_the_wave(cuda, cols_per_row):
    """Windows whose 2:4 units (up to ~40 000, more than LDS holds at 5000
    columns per row) are ordered by the wave-parallel std::sort emulation;
    the 16 rows of window 0 draw from 2^20 columns, the other windows are
    sparse."""
    from trspmm import graphs as g
    rng = np.random.default_rng(cols_per_row)
    rows, cols = [], []
    for r in range(16):
        c = rng.choice(1 << 20, size=cols_per_row, replace=False)
        rows.append(np.full(c.size, r))
        cols.append(c)
    rows.append(rng.integers(16, 64, 400))
    cols.append(rng.integers(0, 1 << 20, 400))
    csr = g.csr_from_pairs(np.concatenate(rows), np.concatenate(cols), None, 64, 1 << 20)
    for mode in ("compat", "exact"):
        _compare(csr, {}, mode)


def test_router_window8(cuda):
    csr = CASES["er512"]()
    _compare(csr, {"window": 8}, "compat")


def test_router_empty_rows(cuda):
    from trspmm import graphs as g
    csr = g.csr_from_pairs(np.array([0, 40]), np.array([3, 5]), np.array([1.0, 2.0]), 64, 64)
    _compare(csr, {}, "compat")


def test_router_rejects_unsorted(cuda):
    rp = torch.tensor([0, 2], dtype=torch.int32)
    col = torch.tensor([5, 3], dtype=torch.int32)
    with pytest.raises(RuntimeError, match="ascending"):
        route(rp, col, torch.ones(2), 8)


@pytest.mark.parametrize("cols, msg", [
    ([1, 9, 3], "out of"),                 # 9 >= ncols comes first
    ([5, 3, 9], "ascending"),              # the descent comes before the out-of-range column
    ([2, 4, 8], "out of"),
], ids=["range-first", "order-first", "range-only"])
def test_router_reports_the_first_bad_entry_of_a_row(cuda, cols, msg):
    """One long row (checked by a whole wave): the first offending entry
    decides the message, as a scan along the row would."""
    n = 200
    head = list(range(10, 10 + n))           # a valid prefix longer than a wave
    c = np.array(head + [x + 300 for x in cols], np.int32)
    c[n:] = np.array(cols, np.int32) + np.array([300, 300, 300])
    ncols = 300 + 8                           # columns >= 308 are out of range
    bad = c.copy()
    rp = torch.tensor([0, bad.size], dtype=torch.int32)
    with pytest.raises(RuntimeError, match=msg):
        route(rp, torch.from_numpy(bad), torch.ones(bad.size), ncols)


def test_router_rejects_bad_row_pointers(cuda):
    rp = torch.tensor([0, 3, 2], dtype=torch.int32)
    with pytest.raises(RuntimeError, match="row pointers"):
        route(rp, torch.tensor([1, 2], dtype=torch.int32), torch.ones(2), 8)


def test_router_block_cache_reuse_and_trim(cuda):
    """Router scratch and results come from the device block cache: results
    that stay alive keep their blocks while other routes reuse the idle ones,
    trimming frees the idle blocks, and routing after a trim is unchanged."""
    import ctypes

    from trspmm import _lib
    from trspmm import graphs as g
    a, b = CASES["rmat10"](), g.rmat_gcn(12, seed=9)
    metas = []
    for csr in (a, b, a, b):
        metas.append(route(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val),
                           csr.n, mode="exact"))
    for csr, meta in zip((a, b, a, b), metas):
        ref = oracle.route(csr.rowptr, csr.col, csr.val, exact=True)
        got = meta.tensors("cpu")
        assert all(np.array_equal(gt.numpy(), ref[n]) for n, gt in zip(META_NAMES, got))
    del metas
    freed = ctypes.c_int64(-1)
    assert _lib.lib().trspmm_cache_trim(ctypes.byref(freed)) == 0
    assert freed.value > 0                      # the scratch of four routes and the freed results
    assert _lib.lib().trspmm_cache_trim(ctypes.byref(freed)) == 0 and freed.value == 0
    _compare(b, {}, "exact")
