"""trspmm_match_2to4_online_noalloc: the allocation-free form of the device
statistics op.  Called through the C ABI, it must return the same counters as
trspmm_match_2to4_online (integer counters: bit-exact) on two calls in a row,
the second on a smaller matrix (its fixed scratch keeps the first call's
contents, so a counter it failed to clear would show), and reject a matrix
larger than its scratch before any device work."""
import ctypes

import numpy as np
import pytest
import torch

from trspmm import _lib
from trspmm import graphs as g

N_COUNTERS = 11
MAX_ROWS = ((1 << 21) - 64) * 16


def _counters(entry, csr, dev):
    rp = torch.from_numpy(csr.rowptr).to(dev)
    col = torch.from_numpy(csr.col).to(dev)
    out = torch.full((N_COUNTERS,), -1, dtype=torch.int64, device=dev)
    fn = getattr(_lib.lib(), entry)
    _lib.check(fn(_lib.ptr(rp), _lib.ptr(col), csr.rowptr.size - 1, csr.col.size, 8, 8, _lib.ptr(out),
                  _lib.current_stream()), entry)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.gpu
def test_noalloc_equals_allocating_entry_on_two_calls(cuda):
    big = g.rmat_gcn(14, seed=3)                 # hub windows with > 2048 columns included
    small = g.er_gcn(1000, 12.0, seed=4)
    print(f"[noalloc] rmat14 rows={big.m_pad} nnz={big.nnz}; er1000 rows={small.m_pad} nnz={small.nnz}", flush=True)
    ref_big = _counters("trspmm_match_2to4_online", big, cuda)
    ref_small = _counters("trspmm_match_2to4_online", small, cuda)
    got_big = _counters("trspmm_match_2to4_online_noalloc", big, cuda)
    got_small = _counters("trspmm_match_2to4_online_noalloc", small, cuda)
    assert np.array_equal(got_big, ref_big), (got_big, ref_big)
    assert np.array_equal(got_small, ref_small), (got_small, ref_small)
    assert ref_big[0] > 0 and ref_small[0] > 0


@pytest.mark.gpu
def test_noalloc_rejects_more_rows_than_its_scratch(cuda):
    rp = torch.zeros(2, dtype=torch.int32, device=cuda)
    out = torch.zeros(N_COUNTERS, dtype=torch.int64, device=cuda)
    L = _lib.lib()
    rc = L.trspmm_match_2to4_online_noalloc(_lib.ptr(rp), ctypes.c_void_p(0), MAX_ROWS + 16, 0, 8, 8, _lib.ptr(out),
                                            _lib.current_stream())
    assert rc == _lib.EINVAL
    assert b"at most" in L.trspmm_last_error()
