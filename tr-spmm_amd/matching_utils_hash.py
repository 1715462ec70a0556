"""Drop-in replacement for the reference's `matching_utils_hash` extension
(TR-source/Block/block_matching_hash.cpp:104-478).

    match_2to4_hash(row_ptr, col_ind, values, window_size=16) -> dict

Same signature, keyword default, dict keys (`num_groups` where match_2to4
reports num_sptc_groups / num_tc_blocks) and errors (RuntimeError for
window_size outside [1, 64] or an empty matrix).  Computed by libtrspmm
(trspmm_match_2to4_hash, host threads).  The reference's quirk is kept: a
pair made on the exact-complement path counts 0 nonzeros in fake_zeros
(block_matching_hash.cpp:150-156 leave the pair's count unset).  The
statistics table is printed only with verbose=True.
"""
from __future__ import annotations

from matching_utils import _arrays, _stats
from trspmm._lib import lib

KEYS = ["total_rows", "total_cols", "total_nnz", "sparsity", "num_row_panels", "num_groups", "num_16x16_blocks",
        "block_padding_groups", "fine_fallback", "coarse_fallback", "fake_zeros"]


def match_2to4_hash(row_ptr, col_ind, values, window_size=16, *, threads=0, verbose=False):
    rp, ci = _arrays(row_ptr, col_ind, values, window_size)
    return _stats(lib().trspmm_match_2to4_hash, rp, ci, int(window_size), int(threads), keys=KEYS,
                  rename={"num_sptc_groups": "num_groups"}, verbose=verbose)
