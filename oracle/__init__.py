"""TEST INFRASTRUCTURE ONLY — the oracle.

CPU restatement of the reference router (TR-source/Block/block_online.cpp:545-871)
plus FP64 products over the routed metadata.  Imported only by tests/, the
smoke check in __graft_entry__.py and bench.py's CPU-baseline leg, always as
the checker; the product package under tr-spmm_amd/ never imports it.

Functions
---------
route(...)        -> RoutedMeta   (29 arrays in the reference's return order)
decode(meta)      -> (rows, cols, vals) COO triplets the metadata encodes
spmm64(meta, B)   -> FP64 C = A_routed @ B
csr_spmm64(...)   -> FP64 C = A @ B for a plain CSR (torch.sparse / scipy)
sort_order(keys)  -> libstdc++ std::sort order of iota(n) by key desc
f32_to_f16_bits   -> IEEE RNE binary16 bits (the router's packing conversion)
reorder_rows(...) -> row order of the reordered routing (first column, degree desc, row)
permute_rows(...) -> the CSR with its rows in that order, and each entry's input edge
gathered_rows(m)  -> B rows one multiply over a routing gathers (the reorder decision)
route_reordered   -> the routing the device router keeps under TRSPMM_ROUTE_REORDER
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NAMES = [
    "t_row_offset", "t_block_offset", "t_column", "t_value", "t_window_row", "t_atomic", "t_binary",
    "c_row_offset", "c_row", "c_atomic", "c_column", "c_value",
    "c_row_offset_short", "c_row_short", "c_atomic_short", "c_column_short", "c_value_short",
    "s_column", "s_mask", "s_value", "s_pos", "s_window", "s_offsets",
    "s_tile_window", "s_tile_group", "s_tile_column", "s_packed_a", "s_packed_meta",
    "s_window_tile_offset",
]
DTYPES = [
    np.int32, np.int32, np.int32, np.float32, np.int32, np.int32, np.int64,
    np.int32, np.int32, np.int32, np.int32, np.float32,
    np.int32, np.int32, np.int32, np.int32, np.float32,
    np.int32, np.uint8, np.float32, np.int8, np.int32, np.int32,
    np.int32, np.int32, np.int32, np.int32, np.int32,
    np.int32,
]


def _shape(name: str, n: int, window: int):
    return {
        "s_column": (-1, 4), "s_mask": (-1, window), "s_value": (-1, window, 2), "s_pos": (-1, window, 2),
        "s_tile_column": (-1, 16), "s_packed_a": (-1, 8, 4, 2), "s_packed_meta": (-1, 8),
    }.get(name, (n,))


def lib():
    """Loads oracle/liboracle.so, building it with make if it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    L.oracle_route.restype = P
    L.oracle_route.argtypes = [P, P, P, ctypes.c_int64] + [ctypes.c_int] * 6 + \
        [ctypes.c_float] + [ctypes.c_int] * 5
    L.oracle_error.restype = ctypes.c_char_p
    L.oracle_error.argtypes = [P]
    L.oracle_get.restype = ctypes.c_int
    L.oracle_get.argtypes = [P, ctypes.c_int, ctypes.POINTER(P), ctypes.POINTER(ctypes.c_int64)]
    L.oracle_duration.restype = ctypes.c_double
    L.oracle_duration.argtypes = [P]
    L.oracle_free.argtypes = [P]
    L.oracle_decode.restype = ctypes.c_int64
    L.oracle_decode.argtypes = [P, ctypes.c_int64, P, P, P, P, P,
                                P, P, ctypes.c_int64, P, P,
                                P, P, ctypes.c_int64, P, P,
                                P, P, P, P, ctypes.c_int64,
                                ctypes.c_int, P, P, P, ctypes.c_int64]
    L.oracle_sort_order.argtypes = [P, ctypes.c_int32, P]
    L.oracle_f32_to_f16.argtypes = [P, P, ctypes.c_int64]
    _LIB = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p) if a.size else None


@dataclass
class RoutedMeta:
    arrays: dict = field(default_factory=dict)
    window: int = 16
    duration: float = 0.0

    def __getitem__(self, k):
        return self.arrays[k]

    def as_list(self):
        return [self.arrays[n] for n in NAMES]


def env_knobs():
    """LIBRA_SKIP_TC / LIBRA_SPTC_UNIT_MIN_NNZ parsed as block_online.cpp:118-128 does."""
    v = os.environ.get("LIBRA_SKIP_TC")
    skip = v in ("1", "true", "TRUE") if v is not None else False
    u = os.environ.get("LIBRA_SPTC_UNIT_MIN_NNZ")
    unit = int(u) if u is not None and u.strip().lstrip("-").isdigit() else (2 if u is None else 0)
    return skip, unit


def route(rowptr, col, val, partsize_t=32, partsize_c=32, shortsize=3, density=8, window=16, wide=16,
          tc_min_util=0.0, skip_tc=None, unit_min_nnz=None, exact=False, faithful=False,
          sptc_min_nnz=0) -> RoutedMeta:
    """sptc_min_nnz > 0 (tuned routing, needs exact=True): 2:4 groups with fewer
    kept slots leave the 2:4 route; their nonzeros join the residual route."""
    L = lib()
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float32)
    es, eu = env_knobs()
    skip_tc = es if skip_tc is None else bool(skip_tc)
    unit_min_nnz = eu if unit_min_nnz is None else int(unit_min_nnz)
    h = L.oracle_route(_ptr(rowptr), _ptr(col), _ptr(val), rowptr.size - 1,
                       partsize_t, partsize_c, shortsize, density, window, wide, float(tc_min_util),
                       int(skip_tc), unit_min_nnz, int(bool(exact)), int(bool(faithful)), int(sptc_min_nnz))
    try:
        err = L.oracle_error(h)
        if err:
            raise RuntimeError(err.decode())
        out = RoutedMeta(window=window, duration=L.oracle_duration(h))
        for i, (name, dt) in enumerate(zip(NAMES, DTYPES)):
            p = ctypes.c_void_p()
            n = ctypes.c_int64()
            L.oracle_get(h, i, ctypes.byref(p), ctypes.byref(n))
            if n.value:
                buf = (ctypes.c_char * (n.value * np.dtype(dt).itemsize)).from_address(p.value)
                a = np.frombuffer(bytes(buf), dtype=dt).copy()
            else:
                a = np.zeros(0, dtype=dt)
            out.arrays[name] = a.reshape(_shape(name, n.value, window))
        return out
    finally:
        L.oracle_free(h)


def decode(meta, window=None):
    """COO triplets encoded by the routed metadata (dict or RoutedMeta).

    Values are taken as stored (float32 arrays; pass f16-rounded copies to see
    what a half-precision multiply sees)."""
    L = lib()
    m = meta.arrays if isinstance(meta, RoutedMeta) else meta
    window = window or (meta.window if isinstance(meta, RoutedMeta) else 16)
    a = {k: np.ascontiguousarray(m[k]) for k in NAMES}
    f = {k: np.ascontiguousarray(a[k], dtype=np.float32) for k in ("t_value", "c_value", "c_value_short", "s_value")}
    args = [_ptr(a["t_row_offset"]), a["t_window_row"].size, _ptr(a["t_block_offset"]), _ptr(a["t_column"]),
            _ptr(f["t_value"]), _ptr(a["t_window_row"]), _ptr(a["t_binary"]),
            _ptr(a["c_row_offset"]), _ptr(a["c_row"]), a["c_row"].size, _ptr(a["c_column"]), _ptr(f["c_value"]),
            _ptr(a["c_row_offset_short"]), _ptr(a["c_row_short"]), a["c_row_short"].size,
            _ptr(a["c_column_short"]), _ptr(f["c_value_short"]),
            _ptr(a["s_column"]), _ptr(f["s_value"]), _ptr(a["s_pos"]), _ptr(a["s_window"]), a["s_window"].size,
            window]
    n = L.oracle_decode(*args, None, None, None, 0)
    rows = np.empty(n, np.int32)
    cols = np.empty(n, np.int32)
    vals = np.empty(n, np.float32)
    L.oracle_decode(*args, _ptr(rows), _ptr(cols), _ptr(vals), n)
    return rows, cols, vals


def coo_spmm64(rows, cols, vals, B, m):
    """FP64 C[m, K] = sum over triplets of vals * B[cols]  (scipy CSR product)."""
    import scipy.sparse as sp
    A = sp.csr_matrix((np.asarray(vals, np.float64), (rows, cols)), shape=(m, B.shape[0]))
    return np.asarray(A @ np.asarray(B, np.float64))


def spmm64(meta, B, m, window=None, half_values=True):
    """FP64 product of the routed matrix with B; values rounded to fp16 first
    when `half_values` (what the half-precision multiply consumes)."""
    rows, cols, vals = decode(meta, window)
    if half_values:
        vals = vals.astype(np.float16).astype(np.float32)
    keep = rows < m
    return coo_spmm64(rows[keep], cols[keep], vals[keep], B, m)


def csr_spmm64(rowptr, col, val, B, m=None):
    import scipy.sparse as sp
    m = m if m is not None else len(rowptr) - 1
    A = sp.csr_matrix((np.asarray(val, np.float64), np.asarray(col), np.asarray(rowptr)),
                      shape=(len(rowptr) - 1, B.shape[0]))
    return np.asarray(A @ np.asarray(B, np.float64))[:m]


def sort_order(keys):
    L = lib()
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    out = np.empty(keys.size, np.int32)
    L.oracle_sort_order(_ptr(keys), keys.size, _ptr(out))
    return out


def f32_to_f16_bits(x):
    L = lib()
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.size, np.uint16)
    L.oracle_f32_to_f16(_ptr(x), _ptr(out), x.size)
    return out


# ---------------------------------------------------------------- row reordering
# No reference counterpart: the device router's TRSPMM_ROUTE_REORDER option
# (tr-spmm_amd/csrc/router.hip, route_device).  Restated here from its
# definition: rows by (first column, nonzeros descending, row index), empty
# rows last; the reordered routing is kept when it gathers at least a tenth
# fewer B rows.


def reorder_rows(rowptr, col):
    """order[i] = input row of routed row i."""
    rowptr = np.asarray(rowptr, np.int64)
    col = np.asarray(col, np.int64)
    rows = rowptr.size - 1
    deg = np.diff(rowptr)
    first = np.full(rows, np.iinfo(np.int64).max, np.int64)
    nz = deg > 0
    first[nz] = col[rowptr[:-1][nz]]
    return np.lexsort((np.arange(rows), -deg, first)).astype(np.int32)


def permute_rows(rowptr, col, val, order):
    """(rowptr', col', val', edge'): the CSR with routed row i = input row
    order[i]; edge'[k] = the input edge of entry k."""
    rowptr = np.asarray(rowptr, np.int64)
    deg = np.diff(rowptr)[order]
    prow = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    edge = np.concatenate([np.arange(rowptr[r], rowptr[r + 1]) for r in order]).astype(np.int32) \
        if len(order) else np.zeros(0, np.int32)
    return prow, np.asarray(col)[edge].astype(np.int32), np.asarray(val)[edge].astype(np.float32), edge


def gathered_rows(meta):
    """TC sub-block columns + 2:4 tile columns + residual entries of a routing."""
    a = meta.arrays if isinstance(meta, RoutedMeta) else meta
    return int((a["t_column"] >= 0).sum() + (a["s_tile_column"] >= 0).sum()
               + a["c_column"].size + a["c_column_short"].size)


def route_reordered(rowptr, col, val, **kw):
    """(meta, order): the routing the device router keeps with row reordering
    (order None when the input order is kept), routed by this oracle."""
    natural = route(rowptr, col, val, **kw)
    if len(rowptr) - 1 < 2 or len(col) == 0:
        return natural, None
    order = reorder_rows(rowptr, col)
    prow, pcol, pval, _ = permute_rows(rowptr, col, val, order)
    other = route(prow, pcol, pval, **kw)
    if 10 * gathered_rows(other) <= 9 * gathered_rows(natural):
        return other, order
    return natural, None


# ---------------------------------------------------------------- legacy Libra format
def decode_libra(t_row_offset, t_blockNew_offset, t_value, t_column, t_window_row, t_binary,
                 c_row_offset, c_row, c_column, c_value, cs_row_offset, cs_row, cs_column, cs_value, tf32=False):
    """COO triplets the legacy Libra metadata encodes, as the reference's
    kernels read it (TR-source/SpMM/mGCNkernel.cu):
      * TC (spmm_forward_fp16_csr_v2_kernel_tcu, :60-150): part p, blocks
        t_row_offset[p] .. [p+1]; block b's mask bit 8 n + k (lane l holds bits
        2 l, 2 l + 1, the m16n8k8 B-fragment rows k = 2 (l % 4) + j of column
        n = l / 4) is row 8 t_window_row[p] + n (:132), column t_column[8 b + k]
        (:62, :113-116), value t_value[t_blockNew_offset[b] + popcount of the
        lower bits] (:71-95);
      * residual (spmm_forward_fp16_csr_v2_kernel_cuda{,_short}, :977-1043,
        :1120-1171): part p adds c_value[e] B[c_column[e]] to row c_row[p].
    tf32=True reads the tf32 members' blocks (spmm_forward_tf32_csr_v2_kernel_tcu,
    :2886-3030): 4 columns t_column[4 b + k] and a 32-bit mask whose bit
    l = 4 n + k is lane l's m16n8k4 B-fragment element (row k = l % 4 of
    column n = l / 4, :2937-2949), i.e. row n, column k.
    Empty TC columns (-1) carry no nonzero."""
    rows, cols, vals = [], [], []
    t_row_offset = np.asarray(t_row_offset)
    for p in range(len(t_window_row)):
        base = 8 * int(t_window_row[p])
        for b in range(int(t_row_offset[p]), int(t_row_offset[p + 1])):
            width = 4 if tf32 else 8
            bits = int(np.asarray(t_binary)[b]) & ((1 << (8 * width)) - 1)
            rank = 0
            for bit in range(8 * width):
                if (bits >> bit) & 1:
                    n, k = bit // width, bit % width
                    c = int(t_column[width * b + k])
                    if c >= 0:
                        rows.append(base + n)
                        cols.append(c)
                        vals.append(float(t_value[int(t_blockNew_offset[b]) + rank]))
                    rank += 1
    for off, rr, cc, vv in ((c_row_offset, c_row, c_column, c_value), (cs_row_offset, cs_row, cs_column, cs_value)):
        for p in range(len(rr)):
            for e in range(int(off[p]), int(off[p + 1])):
                rows.append(int(rr[p]))
                cols.append(int(cc[e]))
                vals.append(float(vv[e]))
    return np.asarray(rows, np.int64), np.asarray(cols, np.int64), np.asarray(vals, np.float64)
