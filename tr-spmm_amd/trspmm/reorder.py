"""TCA row reordering (the reference's spmm/tr_csr_fp16/tr_tca_reorder.py:44-357).

    apply_tca_reorder(row_ptr, col_ind, num_nodes, thres=16, verbose=True)
        -> (new_row_ptr int32 [num_nodes + 1], new_col_ind int32 [nnz], reorder_id int32 [num_nodes])

Same arguments and returns as the reference: rows with similar column sets
are clustered (clusters of at most `thres` rows, then clusters of clusters of
at most 128) and laid out one cluster after the other; reorder_id[new] = old
row; columns are not renumbered.  Row r of the result holds the columns of
row reorder_id[r] in ascending order (the canonical CSR the router takes).
MinHash signatures run on the GPU (trspmm_minhash), the LSH, Jaccard scoring
and clustering on the host (trspmm_tca_order).  The reference hashes through
datasketch (not available here), so the permutation is not the reference's;
invariants and the clustering quality are what the tests check.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _lib
from ._lib import lib


def minhash_coefficients(num_perm: int = 128, seed: int = 2022) -> np.ndarray:
    """The (a_k, b_k) pairs of the signatures' hash family, uint64 [num_perm, 2]."""
    out = np.empty(2 * num_perm, np.uint64)
    _lib.check(lib().trspmm_minhash_coefficients(int(num_perm), int(seed), out.ctypes.data_as(ctypes.c_void_p)),
               "trspmm_minhash_coefficients")
    return out.reshape(num_perm, 2)


def minhash_signatures(row_ptr, col_ind, num_perm: int = 128, seed: int = 2022):
    """MinHash signatures of every row's column set on the current GPU:
    uint32-valued int64 tensor [rows, num_perm] (0xffffffff for an empty row)."""
    import torch
    rp = torch.as_tensor(np.ascontiguousarray(np.asarray(row_ptr), dtype=np.int32)).cuda()
    ci = torch.as_tensor(np.ascontiguousarray(np.asarray(col_ind), dtype=np.int32)).cuda()
    rows = rp.numel() - 1
    sig = torch.empty((max(rows, 0), num_perm), dtype=torch.int32, device=rp.device)
    _lib.check(lib().trspmm_minhash(_lib.ptr(rp), _lib.ptr(ci), rows, int(num_perm), int(seed), _lib.ptr(sig),
                                    _lib.current_stream()), "trspmm_minhash")
    return sig.cpu().numpy().view(np.uint32)


def tca_order(row_ptr, col_ind, num_nodes: int, thres: int = 16, *, cluster_thres: int = 128,
              lsh_threshold: float = 0.2, num_perm: int = 128, seed: int = 2022, signatures=None) -> np.ndarray:
    """order[new] = old row.  `signatures` (uint32 [num_nodes, num_perm]) skips the device pass."""
    rp = np.ascontiguousarray(np.asarray(row_ptr)[: num_nodes + 1], dtype=np.int32)
    nnz = int(rp[-1]) if rp.size else 0
    ci = np.ascontiguousarray(np.asarray(col_ind)[:nnz], dtype=np.int32)
    order = np.empty(num_nodes, np.int32)
    sig_ptr = None
    if signatures is not None:
        sig = np.ascontiguousarray(signatures, dtype=np.uint32)
        if sig.shape != (num_nodes, num_perm):
            raise ValueError(f"signatures must be [{num_nodes}, {num_perm}]")
        sig_ptr = sig.ctypes.data_as(ctypes.c_void_p)
    _lib.check(lib().trspmm_tca_order(rp.ctypes.data_as(ctypes.c_void_p), ci.ctypes.data_as(ctypes.c_void_p),
                                      num_nodes, nnz, sig_ptr, int(num_perm), int(thres), int(cluster_thres),
                                      float(lsh_threshold), int(seed), order.ctypes.data_as(ctypes.c_void_p)),
               "trspmm_tca_order")
    return order


def permute_rows(row_ptr, col_ind, order):
    """CSR whose row r is row order[r] of the input, columns ascending."""
    rp = np.asarray(row_ptr, np.int64)
    ci = np.asarray(col_ind)
    counts = np.diff(rp)[order]
    new_rp = np.zeros(order.size + 1, np.int64)
    np.cumsum(counts, out=new_rp[1:])
    new_ci = np.empty(int(new_rp[-1]), np.int32)
    for r, old in enumerate(order):
        new_ci[new_rp[r]:new_rp[r + 1]] = np.sort(ci[rp[old]:rp[old + 1]])
    return new_rp.astype(np.int32), new_ci


def apply_tca_reorder(row_ptr, col_ind, num_nodes, thres=16, verbose=True, *, cluster_thres=128,
                      lsh_threshold=0.2, num_perm=128, seed=2022, signatures=None):
    t0 = time.time()
    order = tca_order(row_ptr, col_ind, int(num_nodes), int(thres), cluster_thres=cluster_thres,
                      lsh_threshold=lsh_threshold, num_perm=num_perm, seed=seed, signatures=signatures)
    new_rp, new_ci = permute_rows(np.asarray(row_ptr)[: int(num_nodes) + 1], col_ind, order)
    if verbose:
        print(f"TCA: reordered {int(num_nodes)} rows in {time.time() - t0:.2f}s")
    return new_rp, new_ci, order
