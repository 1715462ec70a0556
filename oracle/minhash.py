"""TEST INFRASTRUCTURE ONLY — MinHash signatures restated in numpy.

h_k(c) = ((a_k * x(c) + b_k) mod (2^61 - 1)) mod 2^32, x(c) the murmur3 32-bit
finaliser of the column index; a row's signature is the minimum over its
columns (0xffffffff for an empty row).  The coefficients come from
trspmm.reorder.minhash_coefficients (an input, like a seed); the check is the
hash evaluation and the minimum, which the device computes in its own
arithmetic (64-bit halves, folding 2^61 = 1).  Python integers here, so the
restatement has no overflow of its own.
"""
from __future__ import annotations

import numpy as np

P61 = (1 << 61) - 1


def mix32(x: int) -> int:
    x &= 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x85EBCA6B) & 0xFFFFFFFF
    x ^= x >> 13
    x = (x * 0xC2B2AE35) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def signatures(row_ptr, col_ind, coeffs) -> np.ndarray:
    rp = np.asarray(row_ptr, np.int64)
    ci = np.asarray(col_ind, np.int64)
    ab = [(int(a), int(b)) for a, b in np.asarray(coeffs, np.uint64)]
    out = np.full((rp.size - 1, len(ab)), 0xFFFFFFFF, np.uint32)
    for r in range(rp.size - 1):
        xs = [mix32(int(c)) for c in ci[rp[r]:rp[r + 1]]]
        if not xs:
            continue
        for k, (a, b) in enumerate(ab):
            out[r, k] = min(((a * x + b) % P61) & 0xFFFFFFFF for x in xs)
    return out
