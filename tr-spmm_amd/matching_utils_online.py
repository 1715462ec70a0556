"""Drop-in replacement for the reference's `matching_utils_online` extension
(TR-source/Block/block_matching_online.cu:437-683).

    match_2to4_online(row_ptr, col_ind, values, window_size=16, dense_threshold=8, t_max=8) -> dict

Same signature, keyword defaults and dict keys (total_rows, total_cols,
total_nnz, num_sptc_groups, num_tc_blocks, dense_to_tc_cols,
dense_to_sptc_cols, fine_fb, coarse_fb_cols, num_16x16_blocks,
block_padding_groups, fake_zeros).  Computed on the GPU by libtrspmm
(trspmm_match_2to4_online, one workgroup per 16-row window); the arrays may
be host arrays (copied to the current device, as the reference does) or
device tensors.

Where columns of a window share a slot of the 2048-slot table (column mod
2048), the reference's column order depends on thread scheduling; here it is
the order sequential insertion by ascending column gives, one of the
reference's outcomes.  A 16-row window with more than 2048 distinct columns
overflows the table; the reference then drops the columns that find no slot
(which ones depends on scheduling), here the window keeps its 2048 smallest
column ids (again one of its outcomes) and a RuntimeWarning names how many
windows dropped columns.  Differences, each an error here where the
reference returns an undefined result or hangs: window_size other than 16
(its kernel always takes 16 rows), dense_threshold < 1, a negative column
index, a malformed row_ptr, and an empty matrix (RuntimeError "invalid
matrix dimensions", as the other matching ops).  An empty window counts zero (the reference reads counters it
never wrote).  The statistics are printed only with verbose=True.
"""
from __future__ import annotations

import ctypes
import warnings

import numpy as np
import torch

from trspmm import _lib
from trspmm._lib import lib

KEYS = ["num_sptc_groups", "num_tc_blocks", "dense_to_tc_cols", "dense_to_sptc_cols", "fine_fb", "coarse_fb_cols",
        "num_16x16_blocks", "block_padding_groups", "fake_zeros"]
COUNTERS = 11      # TRSPMM_ONLINE_COUNTERS


def _device_int32(a, device):
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.int32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a), dtype=np.int32)).to(device)


def match_2to4_online(row_ptr, col_ind, values, window_size=16, dense_threshold=8, t_max=8, *, device=None,
                      verbose=False):
    if int(window_size) != 16:
        raise RuntimeError("match_2to4_online: the kernel takes 16-row windows; window_size must be 16")
    if int(dense_threshold) < 1:
        raise RuntimeError("match_2to4_online: dense_threshold must be >= 1")
    device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    rp = _device_int32(row_ptr, device)
    ci = _device_int32(col_ind, device)
    if isinstance(values, np.ndarray) or not isinstance(values, torch.Tensor):
        np.ascontiguousarray(np.asarray(values), dtype=np.float32)      # forcecast, as the reference binding does
    rows, nnz = rp.numel() - 1, ci.numel()
    if rows <= 0 or nnz == 0:
        raise RuntimeError("invalid matrix dimensions")
    lo, hi, first, last, mono = torch.stack([ci.min(), ci.max(), rp[0], rp[-1],
                                             (rp[1:] >= rp[:-1]).all().to(torch.int32)]).tolist()
    if hi < 0:
        raise RuntimeError("invalid matrix dimensions")
    if lo < 0:
        raise RuntimeError("match_2to4_online: negative column index")
    if first != 0 or last != nnz or not mono:
        raise RuntimeError("match_2to4_online: row_ptr must rise from 0 to len(col_ind)")
    counters = torch.empty(COUNTERS, dtype=torch.int64, device=device)
    with torch.cuda.device(device):
        rc = lib().trspmm_match_2to4_online(ctypes.c_void_p(rp.data_ptr()), ctypes.c_void_p(ci.data_ptr()), rows,
                                            nnz, int(dense_threshold), int(t_max),
                                            ctypes.c_void_p(counters.data_ptr()), _lib.current_stream())
    if rc != _lib.OK:
        raise RuntimeError(lib().trspmm_last_error().decode())
    c = counters.cpu().tolist()
    if c[9]:
        warnings.warn(f"match_2to4_online: {c[9]} window(s) hold more than 2048 distinct columns; "
                      "each keeps its 2048 smallest column ids", RuntimeWarning, stacklevel=2)
    if c[10]:
        raise RuntimeError("match_2to4_online: negative column index")
    out = {"total_rows": rows, "total_cols": hi + 1, "total_nnz": nnz}
    out.update(zip(KEYS, c[:9]))
    if verbose:
        for k, v in out.items():
            print(f"{k:22s} {v}")
    return out
