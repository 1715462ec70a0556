"""Drop-in replacement for the reference's `matching_utils_bucket` extension
(TR-source/Block/block_matching_bucket.cpp:72-537).

    match_2to4_bucket(row_ptr, col_ind, values, window_size=16) -> dict

Same signature, keyword default, dict keys (`num_groups` where match_2to4
reports num_sptc_groups / num_tc_blocks) and errors (RuntimeError for
window_size outside [1, 64] or an empty matrix).  Computed by libtrspmm
(trspmm_match_2to4_bucket); windows run on host threads where the reference
runs one thread, with the same result.  Each group counts 32 - nnz padding
zeros whatever the window height, as in the reference
(block_matching_bucket.cpp:341-350).  The statistics table is printed only
with verbose=True.
"""
from __future__ import annotations

from matching_utils import _arrays, _stats
from trspmm._lib import lib

KEYS = ["total_rows", "total_cols", "total_nnz", "sparsity", "num_row_panels", "num_groups", "num_16x16_blocks",
        "block_padding_groups", "fine_fallback", "coarse_fallback", "fake_zeros"]


def match_2to4_bucket(row_ptr, col_ind, values, window_size=16, *, threads=0, verbose=False):
    rp, ci = _arrays(row_ptr, col_ind, values, window_size)
    return _stats(lib().trspmm_match_2to4_bucket, rp, ci, int(window_size), int(threads), keys=KEYS,
                  rename={"num_sptc_groups": "num_groups"}, verbose=verbose)
