"""TEST INFRASTRUCTURE ONLY — CPU restatement of the reference's device
statistics op matching_utils_online.match_2to4_online
(TR-source/Block/block_matching_online.cu: kernel :85-431, totals :554-649).

Pure Python, for the small matrices the tests use; nothing in the product
imports it.  One 16-row window at a time:
  table_order    the window's columns in the order of the kernel's 2048-slot
                 table (slot = column mod 2048, linear probing, :121-139;
                 compacted in slot order, :143-160).  Where columns share a
                 slot the reference's order depends on thread scheduling; the
                 restatement inserts them one by one in ascending column
                 order, one of its possible outcomes (which, in a window
                 with more than 2048 distinct columns, keeps the 2048
                 smallest: the reference drops the columns that find no
                 slot);
  window_stats   1:2 pairs over the next 32 columns (first cost 0, else the
                 first of least cost <= t_max, :172-203), 2:4 groups over the
                 next 1024 pairs (:211-268), the dense pool (:270-302), 32 -
                 nnz padding zeros per 2:4 group.
Pinned by tests/test_matching_online.py: equal to the reference's own
match_2to4 (block_matching_window_para_dr.cpp, compiled from its source
into oracle/_ref) on every matrix whose windows are whole and whose columns
are below 2048 (there the table order is the column order and 32 = 2 x 16),
invariant under renumbering column c as c + 2048 k (the slot stays; the
reference's match_2to4 result changes), and a hand-computed window with
slot collisions.
"""
from __future__ import annotations

import numpy as np

ROWS = 16
SLOTS = 2048
PAIR_SEARCH = 32
GROUP_SEARCH = 1024


def _popc(x: int) -> int:
    return bin(x).count("1")


def table_order(col_masks: dict) -> list:
    """Row masks of a window's columns in table order (column -> mask in)."""
    table = [-1] * SLOTS
    for c in sorted(col_masks)[:SLOTS]:       # once the table is full, later columns find no slot
        s = c % SLOTS
        while table[s] != -1:
            s = (s + 1) % SLOTS
        table[s] = c
    return [col_masks[c] for c in table if c != -1]


def window_stats(masks: list, dense_threshold: int, t_max: int) -> dict:
    n = len(masks)
    st = dict.fromkeys(["num_sptc_groups", "num_tc_blocks", "dense_to_tc_cols", "dense_to_sptc_cols", "fine_fb",
                        "coarse_fb_cols", "num_16x16_blocks", "block_padding_groups", "fake_zeros"], 0)
    if n == 0:
        return st
    matched = [False] * n
    pairs, pool = [], []
    for i in range(n):
        if matched[i]:
            continue
        best, best_cost = -1, None
        for j in range(i + 1, min(n, i + 1 + PAIR_SEARCH)):
            if matched[j]:
                continue
            cost = _popc(masks[i] & masks[j])
            if cost <= t_max and (best_cost is None or cost < best_cost):
                best, best_cost = j, cost
        matched[i] = True
        if best >= 0:
            matched[best] = True
            pairs.append((i, best, masks[i] & masks[best], masks[i] | masks[best]))
        else:
            pool.append(i)
            st["fine_fb"] += 1
    used = [False] * len(pairs)
    groups = 0
    for i, (a1, a2, g1, g2) in enumerate(pairs):
        if used[i]:
            continue
        used[i] = True
        mate = -1
        for j in range(i + 1, min(len(pairs), i + 1 + GROUP_SEARCH)):
            if not used[j] and ((g1 & pairs[j][3]) | (g2 & pairs[j][2])) == 0:
                mate = j
                break
        if mate < 0:
            pool += [a1, a2]
            st["coarse_fb_cols"] += 2
            continue
        used[mate] = True
        groups += 1
        b1, b2 = pairs[mate][:2]
        st["fake_zeros"] += 32 - sum(_popc(masks[k]) for k in (a1, a2, b1, b2))
    tc = len(pool) // dense_threshold
    processed = tc * dense_threshold
    rest = pool[processed:]
    rgroups = (len(rest) + 1) // 2
    st["fake_zeros"] += 32 * rgroups - sum(_popc(masks[k]) for k in rest)
    st["num_tc_blocks"] = tc
    st["dense_to_tc_cols"] = processed
    st["dense_to_sptc_cols"] = len(rest)
    st["num_sptc_groups"] = groups + rgroups
    st["num_16x16_blocks"] = (st["num_sptc_groups"] + 3) // 4
    st["block_padding_groups"] = st["num_16x16_blocks"] * 4 - st["num_sptc_groups"]
    return st


def match_2to4_online(row_ptr, col_ind, dense_threshold: int = 8, t_max: int = 8) -> dict:
    rp = np.asarray(row_ptr, dtype=np.int64)
    ci = np.asarray(col_ind, dtype=np.int64)
    rows = rp.size - 1
    tot = None
    for r0 in range(0, rows, ROWS):
        col_masks: dict = {}
        for r in range(r0, min(rows, r0 + ROWS)):
            bit = 1 << (r - r0)
            for c in ci[rp[r]: rp[r + 1]].tolist():
                col_masks[c] = col_masks.get(c, 0) | bit
        st = window_stats(table_order(col_masks), dense_threshold, t_max)
        tot = st if tot is None else {k: tot[k] + v for k, v in st.items()}
    out = {"total_rows": rows, "total_cols": int(ci.max()) + 1, "total_nnz": int(ci.size)}
    out.update(tot)
    return out
