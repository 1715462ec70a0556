"""Block-diagonal concatenation of a batch of graphs (CPU): equals
scipy.sparse.block_diag with every graph's rows padded to a multiple of 16."""
import numpy as np
import scipy.sparse as sp

from trspmm import graphs as g
from trspmm.batch import concat_block_diagonal


def _graphs():
    out = [g.ba_gcn(n, 3, seed=i) for i, n in enumerate([20, 37, 64, 5, 130])]
    empty = g.csr_from_pairs(np.zeros(0, np.int64), np.zeros(0, np.int64), None, 9, 11)
    out.insert(2, empty)
    return out


def test_concat_matches_block_diag():
    gs = _graphs()
    rowptr, col, val, row_off, col_off, m_list, n_list = concat_block_diagonal(gs)
    assert np.all(row_off % 16 == 0)
    assert list(m_list) == [c.m for c in gs] and list(n_list) == [c.n for c in gs]
    blocks = []
    for c in gs:
        a = sp.csr_matrix((c.val, c.col, c.rowptr), shape=(c.m_pad, c.n))[: c.m]
        pad = (-c.m) % 16
        blocks.append(sp.vstack([a, sp.csr_matrix((pad, c.n))]) if pad else a)
    ref = sp.block_diag(blocks, format="csr")
    ref.sort_indices()
    got = sp.csr_matrix((val, col, rowptr), shape=(int(row_off[-1]), int(col_off[-1])))
    assert got.shape == ref.shape
    assert np.array_equal(got.indptr, ref.indptr) and np.array_equal(got.indices, ref.indices)
    assert np.array_equal(got.data, ref.data.astype(np.float32))


def test_concat_rejects_out_of_range_column():
    bad = g.csr_from_pairs(np.array([0]), np.array([3]), None, 4, 8)
    bad.n = 2
    try:
        concat_block_diagonal([bad])
    except ValueError as e:
        assert "outside" in str(e)
    else:
        raise AssertionError("expected ValueError")
