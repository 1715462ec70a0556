"""Seeded synthetic sparse matrices for the routed SpMM workloads.

Every generator returns a canonical CSR (rows ascending, column indices
strictly ascending within a row, no duplicates) as numpy int32/float32 arrays,
the form the reference's caller builds with scipy `coo_matrix(...).tocsr()`
(spmm/tr_csr_fp16/mdataset2_online.py:68-70).  Row counts are padded to a
multiple of the window height the same way (mdataset2_online.py:62-63).

Workloads (sizes as used by bench.py and the tests):
  er_gcn       Erdos-Renyi n=4096, mean degree 16, symmetric + self loops,
               GCN values 1/sqrt(d_i d_j); B 4096x64.
  rmat_gcn     R-MAT scale 18 (a,b,c,d)=(.57,.19,.19,.05), self edges dropped,
               symmetrised, deduplicated, self loops added (~2^23 nnz), GCN values.
  three_route  2^20 rows: 64-node dense communities (p=0.6) on the first 20 %,
               2:4 banded quads on the next 30 %, uniform degree-8 tail.
  ba_batch     1024 Barabasi-Albert graphs (m=3), n_i log-uniform in [20, 5000].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class CSR:
    rowptr: np.ndarray       # int32 [m_pad + 1]
    col: np.ndarray          # int32 [nnz]
    val: np.ndarray          # float32 [nnz]
    m: int                   # logical rows (before padding)
    n: int                   # columns (rows of B)

    @property
    def nnz(self) -> int:
        return int(self.col.size)

    @property
    def m_pad(self) -> int:
        return int(self.rowptr.size - 1)


def _pad_rows(m: int, window: int) -> int:
    return m if m % window == 0 else m + window - m % window


def csr_from_pairs(rows: np.ndarray, cols: np.ndarray, vals: np.ndarray | None, m: int, n: int,
                   window: int = 16) -> CSR:
    """Canonical CSR from (row, col) pairs; duplicate pairs collapse to one
    entry (their values are summed, as scipy's tocsr does)."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    if vals is None:
        vals = np.ones(rows.size, np.float64)
    key = rows * n + cols
    order = np.argsort(key, kind="stable")
    key = key[order]
    v = np.asarray(vals, np.float64)[order]
    uniq, start = np.unique(key, return_index=True)
    vsum = np.add.reduceat(v, start) if uniq.size else np.zeros(0)
    r = (uniq // n).astype(np.int64)
    c = (uniq % n).astype(np.int32)
    m_pad = _pad_rows(m, window)
    counts = np.bincount(r, minlength=m_pad)
    rowptr = np.zeros(m_pad + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    if rowptr[-1] > np.iinfo(np.int32).max:
        raise ValueError("nnz exceeds int32 range")
    return CSR(rowptr.astype(np.int32), c, vsum.astype(np.float32), m, n)


def gcn_values(csr: CSR) -> CSR:
    """a_ij = 1/sqrt(d_i d_j) with d = row nonzero count (self loop included)."""
    deg = np.diff(csr.rowptr).astype(np.float64)
    deg_n = np.ones(max(csr.n, csr.m_pad), np.float64)
    deg_n[: deg.size] = np.maximum(deg, 1.0)
    rows = np.repeat(np.arange(csr.m_pad), np.diff(csr.rowptr))
    val = 1.0 / np.sqrt(deg_n[rows] * deg_n[csr.col])
    return CSR(csr.rowptr, csr.col, val.astype(np.float32), csr.m, csr.n)


def _symmetric_with_loops(src, dst, n, window):
    keep = src != dst
    src, dst = src[keep], dst[keep]
    loops = np.arange(n, dtype=np.int64)
    rows = np.concatenate([src, dst, loops])
    cols = np.concatenate([dst, src, loops])
    return csr_from_pairs(rows, cols, None, n, n, window)


def er_gcn(n: int = 4096, mean_degree: float = 16.0, seed: int = 0, window: int = 16) -> CSR:
    """G(n, p) with p = mean_degree/(n-1), undirected, plus self loops."""
    rng = np.random.default_rng(seed)
    p = mean_degree / (n - 1)
    iu, ju = np.triu_indices(n, k=1)
    hit = rng.random(iu.size) < p
    return gcn_values(_symmetric_with_loops(iu[hit].astype(np.int64), ju[hit].astype(np.int64), n, window))


# Raw R-MAT edge count that gives ~2^23 nonzeros after symmetrising,
# deduplication and self loops at scale 18 with seed 1 (measured: 8,393,760).
RMAT_S18_EDGES = 4_500_000


def rmat_pairs(scale: int, edges: int, abcd=(0.57, 0.19, 0.19, 0.05), seed: int = 1):
    rng = np.random.default_rng(seed)
    a, b, c, _ = abcd
    src = np.zeros(edges, np.int64)
    dst = np.zeros(edges, np.int64)
    for level in range(scale):
        r = rng.random(edges)
        sbit = r >= a + b
        dbit = ((r >= a) & (r < a + b)) | (r >= a + b + c)
        src |= sbit.astype(np.int64) << level
        dst |= dbit.astype(np.int64) << level
    return src, dst


def rmat_gcn(scale: int = 18, edges: int | None = None, seed: int = 1, window: int = 16) -> CSR:
    n = 1 << scale
    if edges is None:
        edges = RMAT_S18_EDGES if scale == 18 else 16 * n
    src, dst = rmat_pairs(scale, edges, seed=seed)
    return gcn_values(_symmetric_with_loops(src, dst, n, window))


def three_route(n: int = 1 << 20, seed: int = 2, window: int = 16, community: int = 64,
                p_in: float = 0.6, frac_comm: float = 0.2, frac_band: float = 0.3,
                tail_degree: int = 8) -> CSR:
    """Rows [0, 20%): 64-node communities, each (i, j) inside a community with
    probability p_in.  Rows [20%, 50%): quads of 4 rows share 4 groups of 4
    consecutive columns near the diagonal and every row holds exactly 2 of the
    4 columns of each group (2:4 pattern, 8 nonzeros per row, ~2 hits per
    column per 16-row window).  Remaining rows: `tail_degree` distinct uniform
    columns.  Values U(-1, 1)."""
    rng = np.random.default_rng(seed)
    n_comm = (int(n * frac_comm) // community) * community
    n_band_end = n_comm + ((int(n * frac_band)) // 16) * 16
    rows_l, cols_l = [], []
    # communities
    nb = n_comm // community
    for b0 in range(0, nb, 256):                      # chunks keep memory modest
        bs = np.arange(b0, min(nb, b0 + 256))
        hit = rng.random((bs.size, community, community)) < p_in
        bi, ii, jj = np.nonzero(hit)
        rows_l.append(bs[bi] * community + ii)
        cols_l.append(bs[bi] * community + jj)
    # banded 2:4 quads: window w, quad q = (row % 16) // 4 owns columns
    # [16 w + 16 q, 16 w + 16 q + 16), i.e. 4 groups of 4
    band_rows = np.arange(n_comm, n_band_end, dtype=np.int64)
    if band_rows.size:
        base = (band_rows // 16) * 16 + ((band_rows % 16) // 4) * 16
        pairs = np.array([[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]], np.int64)
        sel = pairs[rng.integers(0, 6, size=(band_rows.size, 4))]      # [rows, 4 groups, 2]
        grp = np.arange(4, dtype=np.int64)[None, :, None] * 4
        rows_l.append(np.repeat(band_rows, 8))
        cols_l.append(((base[:, None, None] + grp + sel) % n).reshape(-1))
    # uniform tail: tail_degree uniform draws per row (a repeated draw, about
    # 3e-5 of the rows at n = 2^20, collapses to one entry)
    tail_rows = np.arange(n_band_end, n, dtype=np.int64)
    if tail_rows.size:
        rows_l.append(np.repeat(tail_rows, tail_degree))
        cols_l.append(rng.integers(0, n, size=tail_rows.size * tail_degree))
    rows = np.concatenate(rows_l)
    cols = np.concatenate(cols_l)
    pattern = csr_from_pairs(rows, cols, None, n, n, window)
    vals = rng.uniform(-1.0, 1.0, size=pattern.nnz).astype(np.float32)
    return CSR(pattern.rowptr, pattern.col, vals, n, n)


def barabasi_albert_pairs(n: int, m: int, rng: np.random.Generator):
    """Undirected BA edges: node v >= m attaches to m distinct earlier nodes
    drawn proportionally to degree (repeated-endpoint list); seed graph is a
    star on nodes 0..m."""
    if n <= m:
        n = m + 1
    src = np.empty(m * (n - m), np.int64)
    dst = np.empty(m * (n - m), np.int64)
    rep = np.empty(2 * m * (n - m) + 2 * m, np.int64)
    # star: node m connected to 0..m-1
    src[:m] = m
    dst[:m] = np.arange(m)
    rep[: 2 * m] = np.concatenate([np.arange(m), np.full(m, m)])
    nrep = 2 * m
    e = m
    for v in range(m + 1, n):
        chosen: list[int] = []
        while len(chosen) < m:
            t = int(rep[rng.integers(0, nrep)])
            if t not in chosen:
                chosen.append(t)
        for t in chosen:
            src[e] = v
            dst[e] = t
            e += 1
        rep[nrep: nrep + m] = chosen
        rep[nrep + m: nrep + 2 * m] = v
        nrep += 2 * m
    return src[:e], dst[:e]


def ba_sizes(count: int = 1024, lo: int = 20, hi: int = 5000, seed: int = 5000) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(np.int64)


def ba_gcn(n: int, m: int = 3, seed: int = 0, window: int = 16) -> CSR:
    rng = np.random.default_rng(seed)
    src, dst = barabasi_albert_pairs(n, m, rng)
    return gcn_values(_symmetric_with_loops(src, dst, n, window))


def ba_batch(count: int = 1024, m: int = 3, seed: int = 5000, window: int = 16):
    """List of `count` BA graphs; graph i is seeded with i."""
    return [ba_gcn(int(n), m, seed=i, window=window) for i, n in enumerate(ba_sizes(count, seed=seed))]


def dense_features(n: int, k: int, seed: int, dtype=np.float16) -> np.ndarray:
    """B ~ N(0, 1), n x k, cast to `dtype` (fp16 by default)."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, k), dtype=np.float32).astype(dtype)


# Workload table used by bench.py and the tests: name -> (builder, K, B seed)
WORKLOADS = {
    "er4096_k64": (lambda: er_gcn(4096, 16.0, seed=0), 64, 1000),
    "rmat18_k128": (lambda: rmat_gcn(18, seed=1), 128, 1001),
    "three_route_k256": (lambda: three_route(1 << 20, seed=2), 256, 1002),
    "rmat18_k256": (lambda: rmat_gcn(18, seed=1), 256, 1001),
}
