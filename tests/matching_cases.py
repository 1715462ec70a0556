"""Seeded matrices and parameter sets for the matching-statistics tests."""
from graphs_small import CASES as _SMALL
from trspmm import graphs as g

CASES = dict(_SMALL)
CASES["rmat12"] = lambda: g.rmat_gcn(12, seed=3)
CASES["er4096"] = lambda: g.er_gcn(4096, 16.0, seed=0)

PARAMS = [
    {},
    {"window_size": 8},
    {"window_size": 20},
    {"dense_threshold": 4, "t_max": 2},
    {"t_max": 0},
    {"dense_threshold": 16, "t_max": 16},
]

# window heights for the variants that take only window_size
WINDOWS = [16, 8, 20, 1, 64]
