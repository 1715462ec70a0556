"""Writes tests/golden/matching_stats.json: statistics of the reference's own
matching_utils.match_2to4 (TR-source/Block/block_matching_window_para_dr.cpp,
compiled into oracle/_ref by `make -C oracle ref`) for seeded matrices and
parameter sets.  Run from the repository root:
    python tests/golden/make_golden_matching.py
"""
import ctypes
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tr-spmm_amd"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle", "_ref")]

from matching_cases import CASES, PARAMS  # noqa: E402


def quiet(f, *a, **k):
    """Calls f with C-level stdout discarded (the reference prints a table)."""
    libc = ctypes.CDLL(None)
    fd = os.dup(1)
    nul = os.open(os.devnull, os.O_WRONLY)
    os.dup2(nul, 1)
    try:
        return f(*a, **k)
    finally:
        libc.fflush(None)
        os.dup2(fd, 1)
        os.close(fd)
        os.close(nul)


def main():
    import matching_utils_ref as R  # noqa: E402
    out = []
    for name, mk in CASES.items():
        c = mk()
        for kw in PARAMS:
            out.append({"case": name, "kwargs": kw, "stats": quiet(R.match_2to4, c.rowptr, c.col, c.val, **kw)})
    with open(os.path.join(HERE, "matching_stats.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(len(out), "records")


if __name__ == "__main__":
    main()
