"""The legacy Libra kernel family (TR-source/SpMM/mGCNkernel.cu) on MI355X.

The reference's host functions that multiply with the original Libra
metadata in fp16 take device pointers, write the product into a caller's
output buffer and return the mean milliseconds per launch; they are not bound
to Python in the reference.  The functions here keep their names, argument
order and return value, take device (or CPU) tensors for the pointers, and
run on libtrspmm's C entry points (`trspmm_libra_fp16_*`, include/trspmm.h):

  spmm_forward_fp16_tcu_cuda_kernel(...)            (:1357-1566)
  spmm_forward_fp16_tcu_cuda_kernel_seq(...)        (:1778-1935)
  spmm_forward_fp16_tcu_cuda_hybrid_kernel(...)     (:1936-2100)
  spmm_forward_fp16_tcu_cuda_hybrid_kernel_seq(...) (:2181-2338)
  spmm_forward_fp16_tcu_cuda_kernel_cc_only(...)    (:1567-1777)
     (t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_atomic, t_binary,
      c_row_offset, c_row, c_atomic, c_column, c_value,
      c_row_offset_short, c_row_short, c_atomic_short, c_column_short, c_value_short,
      rhs_matrix, output_matrix, parts_t, parts_c, partsize_c, parts_c_short, dimN, mOri, epoches) -> ms
  spmm_forward_fp16_tcu_cuda_nomap_kernel(t_row_offset, t_blockNew_offset, t_value, t_column,
      t_window_row, t_atomic, t_binary, rhs_matrix, output_matrix, parts_t, dimN, mOri, epoches) -> ms
                                                    (:2101-2180)
  spmm_forward_fp16_cuda_kernel_split(c_row_offset, c_row, c_atomic, c_column, c_value,
      c_row_offset_short, c_row_short, c_atomic_short, c_column_short, c_value_short,
      rhs_matrix, output_matrix, partsize, dimN, mOri, epoches, parts, parts_short, swizzle) -> ms
                                                    (:2651-2766; _v1 :2767-2884)

Format (as the reference's kernels read it): TC blocks of 8-row windows
(output rows 8 t_window_row[p] + n; block b = 8 columns t_column[8 b + k]
with -1 for an empty column, and the mask t_binary[b] whose bit 8 n + k is
row n, column k; values at t_blockNew_offset[b] in bit order), residual parts
as in the online format (c_row[p] gets c_value[e] B[c_column[e]] over the
part's entries).  output_matrix is [mOri, dimN] fp32 on the device; rhs_matrix
holds (largest column + 1) rows of dimN fp16 values.  The products
accumulate in fp32 (the reference's kernels accumulate in fp16); the seq and
hybrid variants, which schedule the same kernels differently in the
reference, run as one launch here; cc_only uses no matrix instructions.
Errors raise RuntimeError (the reference prints CUDA errors and goes on).

The tf32 members take fp32 values and rhs_matrix, and TC blocks of 8 rows x 4
columns (t_column[4 b + k], int32 masks whose bit 4 n + k is row n, column k):

  spmm_forward_tf32_tcu_cuda_kernel(...), spmm_forward_tf32_tcu_cuda_kernel_seq(...)   (:3513-3837)
  spmm_forward_tf32_cuda_kernel_split(...)                                            (:3388-3512)

They run as three bf16 products (hi/lo split of values and operand) summed in
fp32, finer than tf32, because gfx950 has no tf32 matrix instruction.
"""
from __future__ import annotations

import ctypes

import torch

from trspmm._lib import lib

_P = ctypes.c_void_p
_I = ctypes.c_int
_FULL = [_P] * 19 + [_I] * 7
_SIGS = {
    "trspmm_libra_fp16_tcu_cuda": _FULL,
    "trspmm_libra_fp16_tcu_cuda_seq": _FULL,
    "trspmm_libra_fp16_tcu_cuda_hybrid": _FULL,
    "trspmm_libra_fp16_tcu_cuda_hybrid_seq": _FULL,
    "trspmm_libra_fp16_tcu_cuda_cc_only": _FULL,
    "trspmm_libra_fp16_tcu_nomap": [_P] * 9 + [_I] * 4,
    "trspmm_libra_fp16_cuda_split": [_P] * 12 + [_I] * 7,
    "trspmm_libra_fp16_cuda_split_v1": [_P] * 12 + [_I] * 7,
    "trspmm_libra_tf32_tcu_cuda": _FULL,
    "trspmm_libra_tf32_tcu_cuda_seq": _FULL,
    "trspmm_libra_tf32_cuda_split": [_P] * 12 + [_I] * 7,
}


def _fn(name):
    f = getattr(lib(), name)
    f.restype = ctypes.c_float
    f.argtypes = _SIGS[name]
    return f


def _dev(t, dtype, name):
    """A contiguous device tensor of the exact dtype (CPU tensors are uploaded)."""
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a tensor")
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_cuda:
        t = t.to("cuda")
    return t.contiguous()


def _p(t):
    return ctypes.c_void_p(t.data_ptr() if t.numel() else 0)


def _out(output_matrix, mOri, dimN):
    if not (isinstance(output_matrix, torch.Tensor) and output_matrix.is_cuda and output_matrix.dtype == torch.float32
            and output_matrix.is_contiguous() and output_matrix.numel() >= int(mOri) * int(dimN)):
        raise RuntimeError("output_matrix must be a contiguous float32 device tensor of mOri * dimN elements")
    return output_matrix


def _call(name, args, keep):
    ms = _fn(name)(*args)
    torch.cuda.synchronize()
    del keep
    if ms < 0:
        msg = lib().trspmm_last_error()
        raise RuntimeError(f"{name}: {msg.decode() if msg else 'failed'}")
    return float(ms)


_TC = [("t_row_offset", torch.int32), ("t_blockNew_offset", torch.int32), ("t_value", torch.float16),
       ("t_column", torch.int32), ("t_window_row", torch.int32), ("t_atomic", torch.int32), ("t_binary", torch.int64)]
_RES = [("c_row_offset", torch.int32), ("c_row", torch.int32), ("c_atomic", torch.int32), ("c_column", torch.int32),
        ("c_value", torch.float16)]
_RES_S = [(n + "_short", dt) for n, dt in _RES]


_TC32 = [(n, torch.float32 if n == "t_value" else (torch.int32 if n == "t_binary" else dt)) for n, dt in _TC]
_RES32 = [(n, torch.float32 if n.startswith("c_value") else dt) for n, dt in _RES]
_RES32_S = [(n, torch.float32 if n.startswith("c_value") else dt) for n, dt in _RES_S]


def _full(name, tensors, rhs_matrix, output_matrix, parts_t, parts_c, partsize_c, parts_c_short, dimN, mOri, epoches,
          tf32=False):
    spec = _TC32 + _RES32 + _RES32_S if tf32 else _TC + _RES + _RES_S
    ts = [_dev(t, dt, n) for t, (n, dt) in zip(tensors, spec)]
    rhs = _dev(rhs_matrix, torch.float32 if tf32 else torch.float16, "rhs_matrix")
    out = _out(output_matrix, mOri, dimN)
    args = [_p(t) for t in ts] + [_p(rhs), _p(out)] + [int(parts_t), int(parts_c), int(partsize_c),
                                                       int(parts_c_short), int(dimN), int(mOri), int(epoches)]
    return _call(name, args, (ts, rhs))


def spmm_forward_fp16_tcu_cuda_kernel(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_atomic,
                                      t_binary, c_row_offset, c_row, c_atomic, c_column, c_value,
                                      c_row_offset_short, c_row_short, c_atomic_short, c_column_short, c_value_short,
                                      rhs_matrix, output_matrix, parts_t, parts_c, partsize_c, parts_c_short, dimN,
                                      mOri, epoches):
    return _full("trspmm_libra_fp16_tcu_cuda", locals_list(locals()), rhs_matrix, output_matrix, parts_t, parts_c,
                 partsize_c, parts_c_short, dimN, mOri, epoches)


def spmm_forward_fp16_tcu_cuda_kernel_seq(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_atomic,
                                          t_binary, c_row_offset, c_row, c_atomic, c_column, c_value,
                                          c_row_offset_short, c_row_short, c_atomic_short, c_column_short,
                                          c_value_short, rhs_matrix, output_matrix, parts_t, parts_c, partsize_c,
                                          parts_c_short, dimN, mOri, epoches):
    return _full("trspmm_libra_fp16_tcu_cuda_seq", locals_list(locals()), rhs_matrix, output_matrix, parts_t, parts_c,
                 partsize_c, parts_c_short, dimN, mOri, epoches)


def spmm_forward_fp16_tcu_cuda_hybrid_kernel(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row,
                                             t_atomic, t_binary, c_row_offset, c_row, c_atomic, c_column, c_value,
                                             c_row_offset_short, c_row_short, c_atomic_short, c_column_short,
                                             c_value_short, rhs_matrix, output_matrix, parts_t, parts_c, partsize_c,
                                             parts_c_short, dimN, mOri, epoches):
    return _full("trspmm_libra_fp16_tcu_cuda_hybrid", locals_list(locals()), rhs_matrix, output_matrix, parts_t,
                 parts_c, partsize_c, parts_c_short, dimN, mOri, epoches)


def spmm_forward_fp16_tcu_cuda_hybrid_kernel_seq(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row,
                                                 t_atomic, t_binary, c_row_offset, c_row, c_atomic, c_column, c_value,
                                                 c_row_offset_short, c_row_short, c_atomic_short, c_column_short,
                                                 c_value_short, rhs_matrix, output_matrix, parts_t, parts_c,
                                                 partsize_c, parts_c_short, dimN, mOri, epoches):
    return _full("trspmm_libra_fp16_tcu_cuda_hybrid_seq", locals_list(locals()), rhs_matrix, output_matrix, parts_t,
                 parts_c, partsize_c, parts_c_short, dimN, mOri, epoches)


def spmm_forward_fp16_tcu_cuda_kernel_cc_only(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row,
                                              t_atomic, t_binary, c_row_offset, c_row, c_atomic, c_column, c_value,
                                              c_row_offset_short, c_row_short, c_atomic_short, c_column_short,
                                              c_value_short, rhs_matrix, output_matrix, parts_t, parts_c, partsize_c,
                                              parts_c_short, dimN, mOri, epoches):
    return _full("trspmm_libra_fp16_tcu_cuda_cc_only", locals_list(locals()), rhs_matrix, output_matrix, parts_t,
                 parts_c, partsize_c, parts_c_short, dimN, mOri, epoches)


def spmm_forward_fp16_tcu_cuda_nomap_kernel(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row,
                                            t_atomic, t_binary, rhs_matrix, output_matrix, parts_t, dimN, mOri,
                                            epoches):
    ts = [_dev(t, dt, n) for t, (n, dt) in zip((t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row,
                                                 t_atomic, t_binary), _TC)]
    rhs = _dev(rhs_matrix, torch.float16, "rhs_matrix")
    out = _out(output_matrix, mOri, dimN)
    args = [_p(t) for t in ts] + [_p(rhs), _p(out), int(parts_t), int(dimN), int(mOri), int(epoches)]
    return _call("trspmm_libra_fp16_tcu_nomap", args, (ts, rhs))


def _split(name, tensors, rhs_matrix, output_matrix, partsize, dimN, mOri, epoches, parts, parts_short, swizzle,
           tf32=False):
    ts = [_dev(t, dt, n) for t, (n, dt) in zip(tensors, _RES32 + _RES32_S if tf32 else _RES + _RES_S)]
    rhs = _dev(rhs_matrix, torch.float32 if tf32 else torch.float16, "rhs_matrix")
    out = _out(output_matrix, mOri, dimN)
    args = [_p(t) for t in ts] + [_p(rhs), _p(out), int(partsize), int(dimN), int(mOri), int(epoches), int(parts),
                                  int(parts_short), int(bool(swizzle))]
    return _call(name, args, (ts, rhs))


def spmm_forward_fp16_cuda_kernel_split(c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short,
                                        c_row_short, c_atomic_short, c_column_short, c_value_short, rhs_matrix,
                                        output_matrix, partsize, dimN, mOri, epoches, parts, parts_short, swizzle):
    return _split("trspmm_libra_fp16_cuda_split",
                  (c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short, c_row_short, c_atomic_short,
                   c_column_short, c_value_short), rhs_matrix, output_matrix, partsize, dimN, mOri, epoches, parts,
                  parts_short, swizzle)


def spmm_forward_fp16_cuda_kernel_split_v1(c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short,
                                           c_row_short, c_atomic_short, c_column_short, c_value_short, rhs_matrix,
                                           output_matrix, partsize, dimN, mOri, epoches, parts, parts_short, swizzle):
    return _split("trspmm_libra_fp16_cuda_split_v1",
                  (c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short, c_row_short, c_atomic_short,
                   c_column_short, c_value_short), rhs_matrix, output_matrix, partsize, dimN, mOri, epoches, parts,
                  parts_short, swizzle)


def spmm_forward_tf32_tcu_cuda_kernel(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_atomic,
                                      t_binary, c_row_offset, c_row, c_atomic, c_column, c_value,
                                      c_row_offset_short, c_row_short, c_atomic_short, c_column_short, c_value_short,
                                      rhs_matrix, output_matrix, parts_t, parts_c, partsize_c, parts_c_short, dimN,
                                      mOri, epoches):
    return _full("trspmm_libra_tf32_tcu_cuda", locals_list(locals()), rhs_matrix, output_matrix, parts_t, parts_c,
                 partsize_c, parts_c_short, dimN, mOri, epoches, tf32=True)


def spmm_forward_tf32_tcu_cuda_kernel_seq(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_atomic,
                                          t_binary, c_row_offset, c_row, c_atomic, c_column, c_value,
                                          c_row_offset_short, c_row_short, c_atomic_short, c_column_short,
                                          c_value_short, rhs_matrix, output_matrix, parts_t, parts_c, partsize_c,
                                          parts_c_short, dimN, mOri, epoches):
    return _full("trspmm_libra_tf32_tcu_cuda_seq", locals_list(locals()), rhs_matrix, output_matrix, parts_t, parts_c,
                 partsize_c, parts_c_short, dimN, mOri, epoches, tf32=True)


def spmm_forward_tf32_cuda_kernel_split(c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short,
                                        c_row_short, c_atomic_short, c_column_short, c_value_short, rhs_matrix,
                                        output_matrix, partsize, dimN, mOri, epoches, parts, parts_short, swizzle):
    return _split("trspmm_libra_tf32_cuda_split",
                  (c_row_offset, c_row, c_atomic, c_column, c_value, c_row_offset_short, c_row_short, c_atomic_short,
                   c_column_short, c_value_short), rhs_matrix, output_matrix, partsize, dimN, mOri, epoches, parts,
                  parts_short, swizzle, tf32=True)


def locals_list(d):
    """The 17 metadata tensors of a full-signature call, in argument order."""
    return [d[n] for n, _ in _TC + _RES + _RES_S]
