"""ctypes binding of libtrspmm.so (C ABI declared in include/trspmm.h).

The library is required: there is no CPU fallback.  Loading fails loudly with
an ImportError naming the missing file when it has not been built
(`make -C tr-spmm_amd` or `python -c "import __graft_entry__ as g; g.build()"`).
"""
from __future__ import annotations

import ctypes
import os

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libtrspmm.so")

OK, EINVAL, EHIP, ENOMEM, EUNSUPPORTED = 0, 1, 2, 3, 4
F16, BF16 = 0, 1
ROUTE_TC, ROUTE_SPTC, ROUTE_RESIDUAL, ROUTE_ALL = 1, 2, 4, 7
MODE_COMPAT, MODE_EXACT, MODE_TUNED = 0, 1, 2
ROUTE_EDGES = 1   # trspmm_route_cfg.flags: record the CSR edge behind every value slot
ROUTE_REORDER = 2  # trspmm_route_cfg.flags: also route the rows in (first column, degree) order; keep the cheaper

# every symbol include/trspmm.h declares (checked by tests/test_abi.py)
EXPORTS = [
    "trspmm_last_error", "trspmm_version", "trspmm_device_arch",
    "trspmm_route", "trspmm_meta_count", "trspmm_meta_data", "trspmm_meta_seconds", "trspmm_meta_free",
    "trspmm_plan_create", "trspmm_plan_from_meta", "trspmm_plan_create_packed", "trspmm_plan_free",
    "trspmm_plan_stats", "trspmm_spmm", "trspmm_spmm_timed", "trspmm_memcpy", "trspmm_host_sort_order",
    "trspmm_csr_transpose", "trspmm_sddmm", "trspmm_bwd_plan_create", "trspmm_bwd_plan_free",
    "trspmm_spmm_t_sddmm", "trspmm_plan_create_ex", "trspmm_plan_from_meta_ex", "trspmm_match_2to4",
    "trspmm_match_2to4_hash", "trspmm_match_2to4_bucket", "trspmm_match_2to4_online", "trspmm_match_2to4_online_noalloc",
    "trspmm_minhash", "trspmm_minhash_coefficients", "trspmm_tca_order", "trspmm_plan_set_values",
    "trspmm_bwd_plan_set_values", "trspmm_cache_trim", "trspmm_meta_row_order", "trspmm_bwd_plan_stats",
    "trspmm_libra_fp16_tcu_cuda", "trspmm_libra_fp16_tcu_cuda_seq", "trspmm_libra_fp16_tcu_cuda_hybrid",
    "trspmm_libra_fp16_tcu_cuda_hybrid_seq", "trspmm_libra_fp16_tcu_cuda_cc_only", "trspmm_libra_fp16_tcu_nomap",
    "trspmm_libra_fp16_cuda_split", "trspmm_libra_fp16_cuda_split_v1", "trspmm_libra_tf32_tcu_cuda",
    "trspmm_libra_tf32_tcu_cuda_seq", "trspmm_libra_tf32_cuda_split",
]

ENGINE_DEFAULT, ENGINE_VALU, ENGINE_MFMA = 0, 1, 2


class RouteCfg(ctypes.Structure):
    _fields_ = [
        ("part_t", ctypes.c_int32), ("part_c", ctypes.c_int32), ("short_c", ctypes.c_int32),
        ("density", ctypes.c_int32), ("window", ctypes.c_int32), ("wide", ctypes.c_int32),
        ("tc_min_util", ctypes.c_float), ("skip_tc", ctypes.c_int32), ("unit_min_nnz", ctypes.c_int32),
        ("mode", ctypes.c_int32), ("sptc_min_nnz", ctypes.c_int32), ("flags", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64


class MetaView(ctypes.Structure):
    _fields_ = [
        ("t_row_offset", _P), ("n_tparts", _I64),
        ("t_block_offset", _P),
        ("t_column", _P), ("n_tsub", _I64),
        ("t_value", _P), ("n_tval", _I64),
        ("t_window_row", _P),
        ("t_binary", _P),
        ("c_row_offset", _P), ("n_cparts", _I64),
        ("c_row", _P),
        ("c_column", _P), ("n_cval", _I64),
        ("c_value", _P),
        ("cs_row_offset", _P), ("n_csparts", _I64),
        ("cs_row", _P),
        ("cs_column", _P), ("n_csval", _I64),
        ("cs_value", _P),
        ("s_column", _P), ("n_groups", _I64),
        ("s_value", _P),
        ("s_pos", _P),
        ("s_offsets", _P), ("n_windows", _I64),
        ("t_value_is_f32", ctypes.c_int32), ("c_value_is_f32", ctypes.c_int32),
        ("window", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class PlanOpts(ctypes.Structure):
    _fields_ = [("residual_engine", ctypes.c_int32), ("slab", ctypes.c_int32), ("tile_cap", ctypes.c_int32),
                ("tc_cap", ctypes.c_int32), ("res_cap", ctypes.c_int32), ("item_order", ctypes.c_int32),
                ("k_hint", ctypes.c_int32), ("res_order", ctypes.c_int32),
                ("tile_kernel", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class MatchStats(ctypes.Structure):
    _fields_ = [("total_rows", _I64), ("total_cols", _I64), ("total_nnz", _I64), ("sparsity", ctypes.c_double),
                ("num_row_panels", _I64), ("num_sptc_groups", _I64), ("num_tc_blocks", _I64),
                ("num_16x16_blocks", _I64), ("block_padding_groups", _I64), ("fine_fallback", _I64),
                ("coarse_fallback", _I64), ("fake_zeros", _I64)]


class TrspmmError(RuntimeError):
    """A libtrspmm call returned an error status."""


_LIB = None


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libtrspmm.so is not built ({LIB_PATH}); run `make -C tr-spmm_amd`")
    L = ctypes.CDLL(LIB_PATH)
    L.trspmm_last_error.restype = ctypes.c_char_p
    L.trspmm_version.restype = ctypes.c_int
    L.trspmm_cache_trim.argtypes = [ctypes.POINTER(_I64)]
    L.trspmm_device_arch.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.trspmm_route.argtypes = [_P, _P, _P, _I64, _I64, _I64, ctypes.POINTER(RouteCfg), _P, ctypes.POINTER(_P)]
    L.trspmm_meta_count.restype = _I64
    L.trspmm_meta_count.argtypes = [_P, ctypes.c_int]
    L.trspmm_meta_data.restype = _P
    L.trspmm_meta_data.argtypes = [_P, ctypes.c_int]
    L.trspmm_meta_seconds.restype = ctypes.c_double
    L.trspmm_meta_seconds.argtypes = [_P]
    L.trspmm_meta_free.argtypes = [_P]
    L.trspmm_meta_row_order.restype = _P
    L.trspmm_meta_row_order.argtypes = [_P]
    L.trspmm_plan_create.argtypes = [ctypes.POINTER(MetaView), _I64, _I64, ctypes.c_int, _P, ctypes.POINTER(_P)]
    L.trspmm_plan_from_meta.argtypes = [_P, _I64, _I64, ctypes.c_int, _P, ctypes.POINTER(_P)]
    L.trspmm_plan_create_packed.argtypes = [_P, _P, _P, _P, _I64, _I64, ctypes.c_int32, _I64, _I64, _P,
                                            ctypes.POINTER(_P)]
    L.trspmm_plan_create_ex.argtypes = [ctypes.POINTER(MetaView), _I64, _I64, ctypes.c_int, ctypes.POINTER(PlanOpts),
                                        _P, ctypes.POINTER(_P)]
    L.trspmm_plan_from_meta_ex.argtypes = [_P, _I64, _I64, ctypes.c_int, ctypes.POINTER(PlanOpts), _P,
                                           ctypes.POINTER(_P)]
    L.trspmm_plan_free.argtypes = [_P]
    L.trspmm_plan_stats.argtypes = [_P, ctypes.POINTER(_I64)]
    L.trspmm_spmm.argtypes = [_P, _P, _I64, _I64, _P, _I64, ctypes.c_uint32, _P]
    L.trspmm_spmm_timed.argtypes = [_P, _P, _I64, _I64, _P, _I64, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, _P,
                                    ctypes.POINTER(ctypes.c_float)]
    L.trspmm_memcpy.argtypes = [_P, _P, _I64, _P]
    L.trspmm_host_sort_order.argtypes = [_P, ctypes.c_int32, _P]
    L.trspmm_csr_transpose.argtypes = [_P, _P, _P, _I64, _I64, _I64, _P, _P, _P, _P, _P]
    L.trspmm_sddmm.argtypes = [_P, _I64, _P, _I64, _I64, _P, _I64, _P, _I64, _I64, ctypes.c_int, _P, _P]
    L.trspmm_bwd_plan_create.argtypes = [_P, _P, _P, _P, _I64, _I64, _I64, ctypes.c_int, _P, ctypes.POINTER(_P)]
    L.trspmm_bwd_plan_free.argtypes = [_P]
    L.trspmm_bwd_plan_set_values.argtypes = [_P, _P, _I64, _P]
    L.trspmm_bwd_plan_stats.argtypes = [_P, ctypes.POINTER(_I64)]
    L.trspmm_plan_set_values.argtypes = [_P, _P, _I64, _P, _I64, _P]
    L.trspmm_spmm_t_sddmm.argtypes = [_P, _P, _I64, _P, _I64, _I64, _P, _I64, _P, _P]
    L.trspmm_match_2to4.argtypes = [_P, _P, _I64, _I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(MatchStats)]
    for fn in (L.trspmm_match_2to4_hash, L.trspmm_match_2to4_bucket):
        fn.argtypes = [_P, _P, _I64, _I64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(MatchStats)]
    L.trspmm_match_2to4_online.argtypes = [_P, _P, _I64, _I64, ctypes.c_int, ctypes.c_int, _P, _P]
    L.trspmm_match_2to4_online_noalloc.argtypes = [_P, _P, _I64, _I64, ctypes.c_int, ctypes.c_int, _P, _P]
    L.trspmm_minhash.argtypes = [_P, _P, _I64, ctypes.c_int, ctypes.c_uint64, _P, _P]
    L.trspmm_minhash_coefficients.argtypes = [ctypes.c_int, ctypes.c_uint64, _P]
    L.trspmm_tca_order.argtypes = [_P, _P, _I64, _I64, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_uint64, _P]
    _LIB = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != OK:
        msg = lib().trspmm_last_error()
        msg = msg.decode() if msg else "unknown error"
        raise TrspmmError(f"{what + ': ' if what else ''}{msg} (status {rc})")


def current_stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> ctypes.c_void_p:
    """Device (or host) address of a tensor; None for an empty tensor."""
    if t is None or t.numel() == 0:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(t.data_ptr())
