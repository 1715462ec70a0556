"""Writes the statistics of the reference's own setup_24matching ops for
seeded matrices and parameter sets (the ops compiled into oracle/_ref by
`make -C oracle ref`):
    matching_stats.json         matching_utils.match_2to4
                                (TR-source/Block/block_matching_window_para_dr.cpp)
    matching_hash_stats.json    matching_utils_hash.match_2to4_hash
                                (TR-source/Block/block_matching_hash.cpp)
    matching_bucket_stats.json  matching_utils_bucket.match_2to4_bucket
                                (TR-source/Block/block_matching_bucket.cpp)
Run from the repository root:
    python tests/golden/make_golden_matching.py
"""
import ctypes
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tr-spmm_amd"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle", "_ref")]

from matching_cases import CASES, PARAMS, WINDOWS  # noqa: E402


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


def write(fname, rows):
    with open(os.path.join(HERE, fname), "w") as f:
        json.dump(rows, f, indent=0)
    print(fname, len(rows), "records")


def main():
    import matching_utils_ref as R  # noqa: E402
    import matching_utils_hash_ref as RH  # noqa: E402
    import matching_utils_bucket_ref as RB  # noqa: E402
    out, hashed, bucketed = [], [], []
    for name, mk in CASES.items():
        c = mk()
        for kw in PARAMS:
            out.append({"case": name, "kwargs": kw, "stats": quiet(R.match_2to4, c.rowptr, c.col, c.val, **kw)})
        for w in WINDOWS:
            kw = {"window_size": w}
            hashed.append({"case": name, "kwargs": kw,
                           "stats": quiet(RH.match_2to4_hash, c.rowptr, c.col, c.val, **kw)})
            bucketed.append({"case": name, "kwargs": kw,
                             "stats": quiet(RB.match_2to4_bucket, c.rowptr, c.col, c.val, **kw)})
    write("matching_stats.json", out)
    write("matching_hash_stats.json", hashed)
    write("matching_bucket_stats.json", bucketed)


if __name__ == "__main__":
    main()
