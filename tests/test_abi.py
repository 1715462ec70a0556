"""The C ABI: library loads, exports every declared symbol, host-side logic
and argument errors behave (no GPU needed: nothing here launches a kernel)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from trspmm import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "trspmm.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(trspmm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_what_the_binding_lists():
    assert _declared() == sorted(_lib.EXPORTS)


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(L, s)]
    assert not missing, missing


def test_version():
    assert _lib.lib().trspmm_version() >= 1


def _sort(keys):
    keys = np.ascontiguousarray(keys, np.int32)
    out = np.empty(keys.size, np.int32)
    rc = _lib.lib().trspmm_host_sort_order(keys.ctypes.data_as(ctypes.c_void_p), keys.size,
                                           out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0
    return out


def _organ_pipe(n):
    h = np.arange(n // 2)
    return np.concatenate([h, h[::-1]])


@pytest.mark.parametrize("pattern", ["sorted", "random", "few", "equal", "organ", "saw", "rev"])
def test_order_emulation_equals_std_sort(pattern):
    rng = np.random.default_rng(len(pattern) * 7919 + ord(pattern[0]))
    for n in [0, 1, 2, 15, 16, 17, 33, 100, 257, 1000, 4099, 20000]:
        k = {"sorted": lambda: np.sort(rng.integers(1, 33, n))[::-1],
             "random": lambda: rng.integers(0, 256, n),
             "few": lambda: rng.integers(1, 4, n),
             "equal": lambda: np.full(n, 9),
             "organ": lambda: _organ_pipe(n) % 256 if n > 1 else np.zeros(n),
             "saw": lambda: np.arange(n) % 17,
             "rev": lambda: np.arange(n)[::-1] % 256}[pattern]()
        k = np.asarray(k, np.int32)
        assert np.array_equal(_sort(k), oracle.sort_order(k)), (pattern, n)


def test_errors_without_device_work():
    L = _lib.lib()
    out = ctypes.c_void_p()
    rc = L.trspmm_route(None, None, None, 0, 0, 1, None, None, ctypes.byref(out))
    assert rc == _lib.EINVAL and b"cfg" in L.trspmm_last_error()
    rc = L.trspmm_plan_create(None, 1, 1, 0, None, ctypes.byref(out))
    assert rc == _lib.EINVAL and b"view" in L.trspmm_last_error()
    rc = L.trspmm_spmm(None, None, 8, 8, None, 8, 7, None)
    assert rc == _lib.EINVAL and b"plan" in L.trspmm_last_error()
    cfg = _lib.RouteCfg(32, 32, 3, 8, 20, 16, 0.0, 0, 2, 0, 0, 0)
    rp = np.zeros(2, np.int32)
    rc = L.trspmm_route(rp.ctypes.data_as(ctypes.c_void_p), None, None, 1, 0, 1, ctypes.byref(cfg), None,
                        ctypes.byref(out))
    assert rc == _lib.EINVAL and b"window" in L.trspmm_last_error()


def test_check_raises_with_message():
    L = _lib.lib()
    rc = L.trspmm_spmm(None, None, 8, 8, None, 8, 7, None)
    with pytest.raises(_lib.TrspmmError, match="plan"):
        _lib.check(rc, "trspmm_spmm")


def test_cache_trim_without_blocks():
    freed = ctypes.c_int64(-1)
    assert _lib.lib().trspmm_cache_trim(ctypes.byref(freed)) == 0
    assert freed.value == 0
    assert _lib.lib().trspmm_cache_trim(None) == 0
