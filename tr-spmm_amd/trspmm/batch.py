"""Many graphs of uneven size in one routed launch.

    batch = RoutedBatch(graphs)                 # list of (rowptr, col, val, ncols[, m]) or graphs.CSR
    Bcat = batch.stack_operands([B_0, ..., B_{n-1}])   # or build Bcat directly: rows [col_off[i], +ncols_i)
    C = batch.matmul(Bcat)                      # fp32, rows [row_off[i], row_off[i] + m_i) belong to graph i
    g = batch.capture(Bcat, C)                  # hipGraph of the launch; g.replay() re-runs it

The graphs are concatenated block-diagonally: graph i's rows start at
row_off[i], a multiple of 16, so no 16-row window spans two graphs, and its
columns are shifted by col_off[i] into the concatenated operand.  One router
call and one plan cover the whole batch; the plan's work-item list is the
concatenated tile queue of every graph, so one launch (at most three
kernels on one stream) runs all of them and the grid stays full however
small the individual graphs are.  Rows added for padding are empty.

The rows keep this order by default (reorder=False), so each graph's rows sit
in windows of their own and its output equals its own routed product bit for
bit; reorder=True lets the router reorder the concatenated rows (see
RoutedSparse), which can mix graphs in a window.
"""
from __future__ import annotations

import numpy as np
import torch

from .sparse import RoutedSparse

WINDOW = 16


def _as_triplet(g):
    if hasattr(g, "rowptr"):
        return (np.asarray(g.rowptr), np.asarray(g.col), np.asarray(g.val), int(g.n), int(g.m))
    rowptr, col, val, ncols = g[:4]
    rowptr = np.asarray(rowptr)
    m = int(g[4]) if len(g) > 4 else rowptr.size - 1
    return rowptr, np.asarray(col), np.asarray(val), int(ncols), m


def concat_block_diagonal(graphs):
    """(rowptr, col, val, row_off, col_off, m_list, n_list) of the block-diagonal
    matrix; row_off[i] is a multiple of 16."""
    parts = [_as_triplet(g) for g in graphs]
    m_list = np.array([p[4] for p in parts], np.int64)
    n_list = np.array([p[3] for p in parts], np.int64)
    rows_list = np.array([p[0].size - 1 for p in parts], np.int64)
    if np.any(rows_list < m_list):
        raise ValueError("a graph's CSR has fewer rows than its m")
    for rp, _, _, _, m in parts:
        if rp.size - 1 > m and rp[m] != rp[-1]:
            raise ValueError("a graph holds nonzeros in rows beyond its m")
    padded = (m_list + WINDOW - 1) // WINDOW * WINDOW
    row_off = np.concatenate([[0], np.cumsum(padded)])
    col_off = np.concatenate([[0], np.cumsum(n_list)])
    nnz_list = np.array([p[0][-1] - p[0][0] for p in parts], np.int64)
    nnz_off = np.concatenate([[0], np.cumsum(nnz_list)])
    total_rows = int(row_off[-1])
    counts = np.zeros(total_rows, np.int64)
    cols = np.empty(int(nnz_off[-1]), np.int64)
    vals = np.empty(int(nnz_off[-1]), np.float32)
    for i, (rp, col, val, ncols, m) in enumerate(parts):
        rp = rp.astype(np.int64)
        counts[row_off[i]: row_off[i] + m] = np.diff(rp[: m + 1])
        a, b = int(rp[0]), int(rp[m])
        cols[nnz_off[i]: nnz_off[i + 1]] = col[a:b].astype(np.int64) + col_off[i]
        vals[nnz_off[i]: nnz_off[i + 1]] = val[a:b]
        if np.any(col[a:b] >= ncols) or np.any(col[a:b] < 0):
            raise ValueError(f"graph {i}: column index outside [0, {ncols})")
    rowptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    if cols.size and cols.max() >= 2 ** 31:
        raise ValueError("the batch has more than 2^31 operand rows")
    return rowptr, cols.astype(np.int32), vals, row_off, col_off, m_list, n_list


class RoutedBatch:
    def __init__(self, graphs, *, mode: str = "exact", dtype: torch.dtype = torch.float16, engine=None, slab=None,
                 plan_opts: dict | None = None, reorder: bool = False, **router_kwargs):
        rowptr, col, val, self.row_off, self.col_off, self.m_list, self.n_list = concat_block_diagonal(graphs)
        self.count = len(self.m_list)
        self.dtype = dtype
        self.A = RoutedSparse(torch.from_numpy(rowptr), torch.from_numpy(col), torch.from_numpy(val),
                              int(self.col_off[-1]), int(self.row_off[-1]), mode=mode, dtype=dtype, engine=engine,
                              slab=slab, plan_opts=plan_opts, reorder=reorder, **router_kwargs)

    @property
    def rows(self) -> int:
        return int(self.row_off[-1])

    @property
    def ncols(self) -> int:
        return int(self.col_off[-1])

    @property
    def nnz(self) -> int:
        return self.A.nnz

    def stack_operands(self, Bs) -> torch.Tensor:
        if len(Bs) != self.count:
            raise ValueError(f"expected {self.count} operands, got {len(Bs)}")
        for i, B in enumerate(Bs):
            if B.shape[0] != self.n_list[i]:
                raise ValueError(f"operand {i} must have {self.n_list[i]} rows, got {B.shape[0]}")
        return torch.cat(list(Bs), 0).contiguous()

    def matmul(self, Bcat: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        return self.A.matmul(Bcat, out)

    __matmul__ = matmul

    def outputs(self, C: torch.Tensor):
        """Per-graph views [m_i, K] of the concatenated result."""
        return [C[int(self.row_off[i]): int(self.row_off[i]) + int(self.m_list[i])] for i in range(self.count)]

    def capture(self, Bcat: torch.Tensor, C: torch.Tensor, warmup: int = 2) -> torch.cuda.CUDAGraph:
        """hipGraph of one batched launch writing C; replay() re-runs it (same
        buffers).  The launch records kernels only (no memset nodes)."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self.A.matmul(Bcat, C)
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.A.matmul(Bcat, C)
        return g
