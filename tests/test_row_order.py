"""The oracle of the router's row reordering (TRSPMM_ROUTE_REORDER), on the CPU.

The reordering has no reference counterpart; its definition is restated in
oracle/__init__.py (reorder_rows, permute_rows, gathered_rows,
route_reordered).  Pinned here by:
  * the order against Python's own sort of the (first column, -nonzeros,
    row) tuples and against the invariants (a permutation; keys
    non-decreasing; empty rows last);
  * the permuted CSR against scipy's row indexing A[order], and each entry's
    input edge against the input's column and value;
  * gathered_rows against hand-counted routings: a dense 16 x 8 block is one
    TC sub-block (8 gathered rows), a lone nonzero is one residual entry, four
    columns under the TC density shared by rows two at a time form one 2:4
    tile (its 4 non-empty columns);
  * the keep-or-not decision against the two counts it compares (kept at
    a tenth fewer gathered rows or more).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from graphs_small import CASES


def _random_csr(seed, m=96, n=80, density=0.05):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, format="csr", random_state=rng, dtype=np.float32)
    # a few empty rows and a few heavy rows
    A = A.tolil()
    for r in rng.choice(m, 5, replace=False):
        A[r, :] = 0
    for r in rng.choice(m, 3, replace=False):
        A[r, rng.choice(n, 30, replace=False)] = 1.0
    A = A.tocsr()
    A.sort_indices()
    A.eliminate_zeros()
    return A


@pytest.mark.parametrize("seed", range(5))
def test_reorder_rows_is_the_sort_of_its_keys(seed):
    A = _random_csr(seed)
    order = oracle.reorder_rows(A.indptr, A.indices)
    deg = np.diff(A.indptr)
    big = 1 << 62
    keys = [((int(A.indices[A.indptr[r]]) if deg[r] else big), -int(deg[r]), r) for r in range(A.shape[0])]
    assert list(order) == sorted(range(A.shape[0]), key=lambda r: keys[r])
    assert sorted(order.tolist()) == list(range(A.shape[0]))
    k = [keys[r] for r in order]
    assert all(a <= b for a, b in zip(k, k[1:]))
    empty = deg[order] == 0
    assert not np.any(empty[:-1] & ~empty[1:])          # empty rows only at the end


@pytest.mark.parametrize("seed", range(3))
def test_permute_rows_matches_scipy(seed):
    A = _random_csr(seed)
    order = oracle.reorder_rows(A.indptr, A.indices)
    prow, pcol, pval, edge = oracle.permute_rows(A.indptr, A.indices, A.data, order)
    P = A[order]
    assert np.array_equal(prow, P.indptr) and np.array_equal(pcol, P.indices) and np.array_equal(pval, P.data)
    assert np.array_equal(A.indices[edge], pcol) and np.array_equal(A.data[edge], pval)
    assert sorted(edge.tolist()) == list(range(A.nnz))


def _csr(rows, cols, m, n):
    A = sp.csr_matrix((np.ones(len(rows), np.float32), (rows, cols)), shape=(m, n))
    A.sort_indices()
    return A


def test_gathered_rows_hand_counts():
    # a dense 16 x 8 block: one TC sub-block, 8 gathered rows
    r, c = np.meshgrid(np.arange(16), np.arange(8), indexing="ij")
    A = _csr(r.ravel(), c.ravel(), 16, 8)
    meta = oracle.route(A.indptr, A.indices, A.data, exact=True, sptc_min_nnz=8)
    assert oracle.gathered_rows(meta) == 8
    # one lone nonzero: a unit of one column with one nonzero stays below the
    # 2:4 unit minimum, so it is one residual entry
    A = _csr([0], [3], 16, 8)
    meta = oracle.route(A.indptr, A.indices, A.data, exact=True, sptc_min_nnz=8)
    assert oracle.gathered_rows(meta) == 1 and meta["c_column"].size + meta["c_column_short"].size == 1
    # columns 0..3 on rows 0..7, each row holding two of them alternately (4 rows
    # per column, under the TC density): one 2:4 group of 4 columns with 16 of its
    # 32 slots kept, one tile with 4 non-empty columns
    rows = np.repeat(np.arange(8), 2)
    cols = np.array([[0, 1] if i % 2 == 0 else [2, 3] for i in range(8)]).ravel()
    A = _csr(rows, cols, 16, 8)
    meta = oracle.route(A.indptr, A.indices, A.data, exact=True, sptc_min_nnz=8)
    assert meta["s_tile_column"].shape[0] == 1
    assert oracle.gathered_rows(meta) == 4


@pytest.mark.parametrize("case", sorted(CASES))
def test_route_reordered_keeps_the_cheaper_routing(case):
    csr = CASES[case]()
    kw = dict(exact=True, sptc_min_nnz=8)
    kept, order = oracle.route_reordered(csr.rowptr, csr.col, csr.val, **kw)
    natural = oracle.route(csr.rowptr, csr.col, csr.val, **kw)
    o = oracle.reorder_rows(csr.rowptr, csr.col)
    prow, pcol, pval, _ = oracle.permute_rows(csr.rowptr, csr.col, csr.val, o)
    other = oracle.route(prow, pcol, pval, **kw)
    g0, g1 = oracle.gathered_rows(natural), oracle.gathered_rows(other)
    if 10 * g1 <= 9 * g0:
        assert order is not None and np.array_equal(order, o)
        assert oracle.gathered_rows(kept) == g1
    else:
        assert order is None and oracle.gathered_rows(kept) == g0
    # every nonzero is still routed exactly once (exact mode)
    rows, cols, vals = oracle.decode(kept)
    assert rows.size == csr.nnz
