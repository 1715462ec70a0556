"""The legacy Libra format's oracle (oracle.decode_libra) on the CPU.

decode_libra restates how TR-source/SpMM/mGCNkernel.cu's kernels read the
metadata.  Pinned by:
  * the fragment arithmetic it relies on: m16n8k8, lane l reads mask bits
    2 l + j as B-fragment row k = 2 (l % 4) + j, column n = l / 4, i.e. bit
    8 n + k; m16n8k4 (tf32), bit l as row k = l % 4, column n = l / 4, i.e.
    bit 4 n + k (all 32 lanes);
  * hand-built blocks (three bits, one of them the sign bit of the mask) of
    both widths (8 columns / int64, the tf32 members' 4 columns / int32)
    decoded to their rows, columns and values in bit order;
  * round trips through the test encoder (tests/libra_format.py), which is
    written from the format description independently: the decoded triplets
    equal the CSR's (values rounded to fp16), on graphs with and without TC
    blocks, rows of every residual kind and a row count that is not a
    multiple of 8.
"""
import numpy as np
import pytest
import torch

import oracle
from graphs_small import CASES
from libra_format import decode, encode
from trspmm import graphs as g


def test_fragment_bit_layout():
    for lane in range(32):
        for j in range(2):
            k, n = 2 * (lane % 4) + j, lane // 4
            assert 2 * lane + j == 8 * n + k


def test_fragment_bit_layout_tf32():
    """m16n8k4 (tf32): lane l holds B-fragment row k = l % 4 of column n = l / 4,
    read as mask bit l (mGCNkernel.cu:2937-2949), i.e. bit 4 n + k."""
    for lane in range(32):
        k, n = lane % 4, lane // 4
        assert lane == 4 * n + k


def test_hand_block_tf32():
    bits = (1 << 2) | (1 << 13) | (1 << 31)            # (row 0, col 2), (row 3, col 1), (row 7, col 3)
    cols = np.array([11, 12, 13, 14], np.int32)
    r, c, v = oracle.decode_libra([0, 1], [0, 3], np.array([1.5, 2.5, -3.5]), cols, [1], np.array([bits - (1 << 32)]),
                                  [0], [], [], [], [0], [], [], [], tf32=True)
    assert r.tolist() == [8, 11, 15] and c.tolist() == [13, 12, 14] and v.tolist() == [1.5, 2.5, -3.5]


def test_hand_block():
    bits = (1 << 1) | (1 << 24) | (1 << 63)
    cols = np.array([5, 9, -1, -1, -1, -1, -1, 42], np.int32)
    r, c, v = oracle.decode_libra([0, 1], [0, 3], np.array([0.5, -1.0, 2.0]), cols, [3], np.array([bits - (1 << 64)]),
                                  [0], [], [], [], [0], [], [], [])
    assert r.tolist() == [24 + 0, 24 + 3, 24 + 7]
    assert c.tolist() == [9, 5, 42]
    assert v.tolist() == [0.5, -1.0, 2.0]


GRAPHS = {"rmat10": lambda: g.rmat_gcn(10, seed=1), "er512": CASES["er512"], "dense": CASES["dense"],
          "ragged": CASES["ragged"], "hub": CASES["hub"]}


@pytest.mark.parametrize("case", sorted(GRAPHS))
@pytest.mark.parametrize("density", [2, 4, 9])
@pytest.mark.parametrize("tf32", [False, True])
def test_round_trip(case, density, tf32):
    csr = GRAPHS[case]()
    meta, counts = encode(csr, density=density, tf32=tf32)
    r, c, v = decode(meta)
    rows = np.repeat(np.arange(csr.rowptr.size - 1), np.diff(csr.rowptr))
    vals = torch.from_numpy(csr.val).float() if tf32 else torch.from_numpy(csr.val).half().float()
    want = sorted(zip(rows.tolist(), csr.col.tolist(), vals.tolist()))
    got = sorted(zip(r.tolist(), c.tolist(), v.tolist()))
    assert got == want
    assert counts["parts_t"] == meta["t_window_row"].numel()
    if case == "dense" and density <= 4:
        assert counts["parts_t"] > 0
