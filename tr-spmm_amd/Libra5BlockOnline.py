"""Drop-in replacement for the reference's `Libra5BlockOnline` extension.

    block_sptc_2to4_online(row1, column1, values1, partSize_t, partSize_c, shortSize,
                           density, window, wide, tc_min_util, sptc_min_util, sptc_threshold)
        -> list of 30 tensors

Same signature, argument meaning, return order and error behaviour as
TR-source/Block/block_online.cpp:545-553, :843-876: 29 routed-metadata CPU
tensors followed by the routing wall time (float64 scalar, seconds).  The
routing itself runs on the GPU (trspmm_route) and is bit-exact with the
reference router.  The reference reads LIBRA_SKIP_TC and
LIBRA_SPTC_UNIT_MIN_NNZ from the environment during the call; so does this.

Differences, all raising RuntimeError where the reference would not:
  * window must be <= 16 (the reference accepts up to 32 but shifts 32/64-bit
    words by >= their width there, block_online.cpp:600, :482-486);
  * each row's column indices must be strictly ascending (the canonical CSR
    its caller builds with scipy, mdataset2_online.py:68-70);
  * a GPU is required.
`sptc_min_util` and `sptc_threshold` are accepted and unused, as in the
reference.  Keyword-only `mode` selects "compat" (default: the reference's
output, 3rd/4th hits of a 2:4 group row dropped) or "exact" (those hits moved
to the residual route); the environment variable TRSPMM_ROUTE_MODE sets the
default.
"""
from __future__ import annotations

import os

import torch

from trspmm import core


def block_sptc_2to4_online(row1, column1, values1, partSize_t, partSize_c, shortSize, density, window, wide,
                           tc_min_util, sptc_min_util, sptc_threshold, *, mode=None):
    if row1.dtype != torch.int32 or column1.dtype != torch.int32:
        raise RuntimeError("Expected row1.dtype() == torch::kInt32 && column1.dtype() == torch::kInt32")
    if values1.dtype != torch.float32:
        raise RuntimeError("Expected values1.dtype() == torch::kFloat32")
    if not (0 < window <= 32 and 0 < wide <= 32):
        raise RuntimeError("Expected window > 0 && window <= 32 && wide > 0 && wide <= 32")
    if window > 16:
        raise RuntimeError("window > 16 is not supported (the reference's behaviour there is undefined)")
    mode = mode or os.environ.get("TRSPMM_ROUTE_MODE", "compat")
    ncols = int(column1.max().item()) + 1 if column1.numel() else 1
    meta = core.route(row1, column1, values1, max(ncols, 1), part_t=int(partSize_t), part_c=int(partSize_c),
                      short_c=int(shortSize), density=int(density), window=int(window), wide=int(wide),
                      tc_min_util=float(tc_min_util), mode=mode)
    out = meta.tensors("cpu")
    return out + [torch.tensor(meta.seconds, dtype=torch.float64)]
