"""Drop-in replacement for the reference's `matching_utils` extension
(setup_24matching, TR-source/Block/block_matching_window_para_dr.cpp:346-467).

    match_2to4(row_ptr, col_ind, values, window_size=16, dense_threshold=8, t_max=8) -> dict

Same signature, keyword defaults, dict keys and errors: the arrays are taken
as C-contiguous int32 / int32 / float32 (force-cast like pybind11's
`forcecast`), `values` is not read, and RuntimeError is raised for
window_size outside [1, 64] or an empty matrix (no rows, or no column
index).  Computed by libtrspmm (trspmm_match_2to4, host threads).
Differences: dense_threshold < 1 raises RuntimeError (the reference loops
forever); the statistics table is printed only with verbose=True (the
reference always prints it).
"""
from __future__ import annotations

import ctypes

import numpy as np

from trspmm import _lib
from trspmm._lib import lib

KEYS = ["total_rows", "total_cols", "total_nnz", "sparsity", "num_row_panels", "num_sptc_groups", "num_tc_blocks",
        "num_16x16_blocks", "block_padding_groups", "fine_fallback", "coarse_fallback", "fake_zeros"]


def _arrays(row_ptr, col_ind, values, window_size):
    """C-contiguous int32 row_ptr / col_ind (values force-cast and dropped) and
    the reference's argument checks."""
    rp = np.ascontiguousarray(np.asarray(row_ptr), dtype=np.int32)
    ci = np.ascontiguousarray(np.asarray(col_ind), dtype=np.int32)
    np.ascontiguousarray(np.asarray(values), dtype=np.float32)      # forcecast, as the reference binding does
    if not 1 <= int(window_size) <= 64:
        raise RuntimeError("window_size must be in [1, 64]")
    if rp.size - 1 <= 0 or ci.size == 0 or int(ci.max()) < 0:
        raise RuntimeError("invalid matrix dimensions")
    return rp, ci


def _stats(fn, rp, ci, *args, keys=KEYS, rename=None, verbose=False):
    """Calls a trspmm_match_2to4* entry point and returns its dict; `rename`
    maps struct fields to the op's own key names."""
    st = _lib.MatchStats()
    rc = fn(rp.ctypes.data_as(ctypes.c_void_p), ci.ctypes.data_as(ctypes.c_void_p), rp.size - 1, ci.size, *args,
            ctypes.byref(st))
    if rc != _lib.OK:
        raise RuntimeError(lib().trspmm_last_error().decode())
    rename = rename or {}
    out = {}
    for k in keys:
        out[k] = getattr(st, next((f for f, name in rename.items() if name == k), k))
    out["sparsity"] = float(out["sparsity"])
    if verbose:
        for k in keys:
            print(f"{k:22s} {out[k]}")
    return out


def match_2to4(row_ptr, col_ind, values, window_size=16, dense_threshold=8, t_max=8, *, threads=0,
               verbose=False):
    rp, ci = _arrays(row_ptr, col_ind, values, window_size)
    return _stats(lib().trspmm_match_2to4, rp, ci, int(window_size), int(dense_threshold), int(t_max), int(threads),
                  verbose=verbose)
