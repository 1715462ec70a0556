"""TEST INFRASTRUCTURE ONLY — the reference router, built from its own source.

`make -C oracle ref` compiles TR-source/Block/block_online.cpp (unchanged, from
/root/reference) into oracle/_ref/Libra5BlockOnline_ref*.so.  This module loads
that build and calls `block_sptc_2to4_online` with stdout silenced (the
reference prints per window).  It is used to pin the oracle and to write the
golden fixtures under tests/golden/; nothing in the product imports it.
"""
from __future__ import annotations

import contextlib
import glob
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REF_DIR = os.path.join(_HERE, "_ref")
_MOD = None


def available() -> bool:
    return bool(glob.glob(os.path.join(_REF_DIR, "Libra5BlockOnline_ref*.so")))


def module():
    global _MOD
    if _MOD is None:
        if not available():
            raise ImportError("oracle/_ref not built (make -C oracle ref needs /root/reference)")
        import torch  # noqa: F401  (the extension links libtorch)
        if _REF_DIR not in sys.path:
            sys.path.insert(0, _REF_DIR)
        import Libra5BlockOnline_ref  # type: ignore
        _MOD = Libra5BlockOnline_ref
    return _MOD


@contextlib.contextmanager
def _quiet_stdout():
    sys.stdout.flush()
    fd = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    try:
        os.dup2(devnull, 1)
        yield
    finally:
        os.dup2(fd, 1)
        os.close(fd)
        os.close(devnull)


def route(rowptr, col, val, partsize_t=32, partsize_c=32, shortsize=3, density=8, window=16, wide=16,
          tc_min_util=0.0, sptc_min_util=0.5, sptc_threshold=12):
    """Returns (list of 29 numpy arrays, duration seconds)."""
    import torch
    m = module()
    rp = torch.from_numpy(np.ascontiguousarray(rowptr, np.int32))
    ci = torch.from_numpy(np.ascontiguousarray(col, np.int32))
    vv = torch.from_numpy(np.ascontiguousarray(val, np.float32))
    with _quiet_stdout():
        out = m.block_sptc_2to4_online(rp, ci, vv, partsize_t, partsize_c, shortsize, density, window, wide,
                                       float(tc_min_util), float(sptc_min_util), int(sptc_threshold))
    arrays = [t.numpy().copy() for t in out[:29]]
    return arrays, float(out[29].item())
