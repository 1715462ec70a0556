"""Writes the router golden fixtures (tests/golden/router_*.npz).

Each fixture holds a seeded input CSR and the 29 arrays the reference router
returns for it, produced by the reference's own block_online.cpp compiled into
oracle/_ref (`make -C oracle ref`, needs the reference sources mounted).  The
environment knobs the reference reads during the call (LIBRA_SKIP_TC,
LIBRA_SPTC_UNIT_MIN_NNZ) are set around the call for the variants that use
them.  Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tr-spmm_amd"), os.path.join(ROOT, "tests"), ROOT]

import oracle  # noqa: E402
import oracle.ref as ref  # noqa: E402
from graphs_small import CASES  # noqa: E402

# (fixture name, case, router kwargs, environment)
FIXTURES = [
    ("er512", "er512", {}, {}),
    ("er512_d4", "er512", {"density": 4}, {}),
    ("rmat10", "rmat10", {}, {}),
    ("rmat10_pt2", "rmat10", {"partsize_t": 2, "density": 2}, {}),
    ("dense", "dense", {}, {}),
    ("dense_util", "dense", {"density": 3, "tc_min_util": 0.3}, {}),
    ("dense_skiptc", "dense", {}, {"LIBRA_SKIP_TC": "1"}),
    ("dense_unit5", "dense", {"partsize_c": 4, "shortsize": 2}, {"LIBRA_SPTC_UNIT_MIN_NNZ": "5"}),
    ("hub", "hub", {}, {}),
    ("ragged", "ragged", {}, {}),
    ("ragged_w8", "ragged", {"window": 8}, {}),
]


def main():
    if not ref.available():
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` first")
    for name, case, kw, env in FIXTURES:
        csr = CASES[case]()
        saved = {k: os.environ.get(k) for k in env}
        os.environ.update(env)
        try:
            arrays, _ = ref.route(csr.rowptr, csr.col, csr.val, **kw)
        finally:
            for k, v in saved.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
        out = {"rowptr": csr.rowptr, "col": csr.col, "val": csr.val,
               "kwargs": np.array(repr(kw)), "env": np.array(repr(env))}
        for n, a in zip(oracle.NAMES, arrays):
            out["out_" + n] = a
        np.savez_compressed(os.path.join(HERE, f"router_{name}.npz"), **out)
        print(name, "nnz", csr.nnz, "groups", arrays[17].shape[0], "tc subs", arrays[2].size // 8)


if __name__ == "__main__":
    main()
