"""Small seeded matrices that exercise every routing branch (shared by tests)."""
import numpy as np

from trspmm import graphs as g


def banded(n=512, seed=5):
    return g.three_route(n=n, seed=seed)


def hub_window(seed=7, n=2048, hub_cols=1500):
    """16 rows that touch many columns (a window with > 16 units and heavy
    ties: exercises the introsort partition path) plus a sparse rest."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for r in range(16):
        c = rng.choice(n, size=hub_cols // 4, replace=False)
        rows.append(np.full(c.size, r))
        cols.append(c)
    for r in range(16, n):
        c = rng.choice(n, size=4, replace=False)
        rows.append(np.full(4, r))
        cols.append(c)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = rng.standard_normal(rows.size)
    return g.csr_from_pairs(rows, cols, vals, n, n)


def dense_block(seed=3, n=256):
    """Dense 16-row windows (TC route) with a 2:4 tail and residual leftovers."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for r in range(n):
        w = r // 16
        base = (w * 16) % n
        dense = base + np.nonzero(rng.random(24) < 0.7)[0]
        extra = rng.choice(n, size=3, replace=False)
        c = np.unique(np.concatenate([dense % n, extra]))
        rows.append(np.full(c.size, r))
        cols.append(c)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    return g.csr_from_pairs(rows, cols, rng.uniform(-1, 1, rows.size), n, n)


def ragged(seed=11, n=300, m=37):
    """Row count not a multiple of 16 (padded), empty rows, a long row (> 32
    residual entries possible with unit_min_nnz raised)."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for r in range(m):
        if r % 5 == 3:
            continue                                    # empty row
        k = 40 if r == 7 else int(rng.integers(1, 12))
        c = rng.choice(n, size=k, replace=False)
        rows.append(np.full(k, r))
        cols.append(c)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    return g.csr_from_pairs(rows, cols, rng.standard_normal(rows.size), m, n)


CASES = {
    "er512": lambda: g.er_gcn(512, 16.0, seed=0),
    "rmat10": lambda: g.rmat_gcn(10, seed=4),
    "banded": banded,
    "hub": hub_window,
    "dense": dense_block,
    "ragged": ragged,
}

# router argument variants (kwargs for oracle.route / trspmm.route)
VARIANTS = [
    {},
    {"density": 4},
    {"density": 3, "tc_min_util": 0.3},
    {"skip_tc": True},
    {"unit_min_nnz": 5, "partsize_c": 4, "shortsize": 2},
    {"partsize_t": 2, "density": 2},
]
