"""Device-side router, plans and the routed multiply (thin wrappers over the C ABI)."""
from __future__ import annotations

import ctypes
import os

import torch

from . import _lib
from ._lib import check, lib, ptr

# names and dtypes of the 29 routed-metadata arrays, in the reference's
# return order (TR-source/Block/block_online.cpp:844-868)
META_NAMES = [
    "t_row_offset", "t_block_offset", "t_column", "t_value", "t_window_row", "t_atomic", "t_binary",
    "c_row_offset", "c_row", "c_atomic", "c_column", "c_value",
    "c_row_offset_short", "c_row_short", "c_atomic_short", "c_column_short", "c_value_short",
    "s_column", "s_mask", "s_value", "s_pos", "s_window", "s_offsets",
    "s_tile_window", "s_tile_group", "s_tile_column", "s_packed_a", "s_packed_meta",
    "s_window_tile_offset",
]
META_DTYPES = [
    torch.int32, torch.int32, torch.int32, torch.float32, torch.int32, torch.int32, torch.int64,
    torch.int32, torch.int32, torch.int32, torch.int32, torch.float32,
    torch.int32, torch.int32, torch.int32, torch.int32, torch.float32,
    torch.int32, torch.uint8, torch.float32, torch.int8, torch.int32, torch.int32,
    torch.int32, torch.int32, torch.int32, torch.int32, torch.int32,
    torch.int32,
]


def meta_shape(name: str, n: int, window: int):
    return {
        "s_column": (-1, 4), "s_mask": (-1, window), "s_value": (-1, window, 2), "s_pos": (-1, window, 2),
        "s_tile_column": (-1, 16), "s_packed_a": (-1, 8, 4, 2), "s_packed_meta": (-1, 8),
    }.get(name, (n,))


def _require_cuda():
    if not torch.cuda.is_available():
        raise RuntimeError("trspmm needs a ROCm GPU (gfx950); no device is visible")


def _atoi(s: str) -> int:
    """C atoi: optional whitespace and sign, then leading digits (0 if none)."""
    s = s.lstrip()
    sign, i = 1, 0
    if i < len(s) and s[i] in "+-":
        sign = -1 if s[i] == "-" else 1
        i += 1
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    return sign * int(s[i:j]) if j > i else 0


def env_router_knobs():
    """LIBRA_SKIP_TC and LIBRA_SPTC_UNIT_MIN_NNZ, parsed as the reference does
    (TR-source/Block/block_online.cpp:118-128, :564-567)."""
    v = os.environ.get("LIBRA_SKIP_TC")
    skip = v in ("1", "true", "TRUE") if v is not None else False
    u = os.environ.get("LIBRA_SPTC_UNIT_MIN_NNZ")
    unit = 2 if u is None else _atoi(u)
    return skip, unit


MODES = {"compat": _lib.MODE_COMPAT, "exact": _lib.MODE_EXACT, "tuned": _lib.MODE_TUNED}


class RoutedMeta:
    """The router's 29 arrays, resident on the GPU (owned by the library)."""

    def __init__(self, handle, window: int, nnz: int = 0, edges: bool = False, rows: int = 0):
        self._h = handle
        self.window = window
        self.nnz = nnz              # nonzeros of the routed CSR
        self.edges = edges          # routed with edge maps (plans accept new values)
        self.rows = rows            # rows of the routed CSR
        self.seconds = lib().trspmm_meta_seconds(handle)
        self._views = None

    @property
    def reordered(self) -> bool:
        """The router kept the reordered routing (route(..., reorder=True))."""
        return bool(lib().trspmm_meta_row_order(self._h))

    def row_order(self, device="cpu"):
        """Input row of each routed row (int32 [rows]) when the routing was
        reordered, else None."""
        src = lib().trspmm_meta_row_order(self._h)
        if not src:
            return None
        t = torch.empty(self.rows, dtype=torch.int32, device=device)
        if self.rows:
            check(lib().trspmm_memcpy(ptr(t), ctypes.c_void_p(src), 4 * self.rows, _lib.current_stream()),
                  "trspmm_memcpy")
        return t

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                lib().trspmm_meta_free(h)
            except Exception:
                pass

    @property
    def handle(self):
        return self._h

    def count(self, i: int) -> int:
        return int(lib().trspmm_meta_count(self._h, i))

    def tensors(self, device="cpu"):
        """Copies the 29 arrays into fresh tensors (reference shapes) on `device`."""
        out = []
        stream = _lib.current_stream()
        for i, (name, dt) in enumerate(zip(META_NAMES, META_DTYPES)):
            n = self.count(i)
            t = torch.empty(n, dtype=dt, device=device)
            if n:
                src = lib().trspmm_meta_data(self._h, i)
                check(lib().trspmm_memcpy(ptr(t), ctypes.c_void_p(src), n * t.element_size(), stream),
                      "trspmm_memcpy")
            out.append(t.reshape(meta_shape(name, n, self.window)))
        return out


TUNED_SPTC_MIN_NNZ = 8   # default 2:4 occupancy floor of the tuned router (kept slots of 32)


def route(rowptr: torch.Tensor, col: torch.Tensor, val: torch.Tensor, ncols: int, *, part_t=32, part_c=32,
          short_c=3, density=8, window=16, wide=16, tc_min_util=0.0, skip_tc=None, unit_min_nnz=None,
          mode="exact", sptc_min_nnz=None, edges=False, reorder=False) -> RoutedMeta:
    """Routes a canonical CSR on the GPU (see trspmm_route).

    mode "compat": the reference's metadata bit for bit; "exact": the same t_*/s_*
    with the reference's dropped 2:4 overflow moved to the residual route;
    "tuned": exact plus 2:4 groups with fewer than `sptc_min_nnz` kept slots
    moved to the residual route.  edges: also record the CSR edge behind every
    value slot, so that plans built from the result accept new values
    (Plan.set_values) without routing again.  reorder (exact / tuned, window 16):
    also route the rows in order of (first column, nonzeros descending, row)
    and keep that routing when it gathers at least a tenth fewer B rows per
    multiply; plans built
    from it write every row to its input index (RoutedMeta.row_order)."""
    _require_cuda()
    es, eu = env_router_knobs()
    mode_code = MODES[mode] if isinstance(mode, str) else int(mode)
    if sptc_min_nnz is None:
        sptc_min_nnz = TUNED_SPTC_MIN_NNZ if mode_code == _lib.MODE_TUNED else 0
    cfg = _lib.RouteCfg(part_t, part_c, short_c, density, window, wide, float(tc_min_util),
                        int(es if skip_tc is None else bool(skip_tc)), int(eu if unit_min_nnz is None else unit_min_nnz),
                        mode_code, int(sptc_min_nnz),
                        (_lib.ROUTE_EDGES if edges else 0) | (_lib.ROUTE_REORDER if reorder else 0))
    dev = torch.device("cuda", torch.cuda.current_device())
    rp = rowptr.to(dev, torch.int32).contiguous()
    ci = col.to(dev, torch.int32).contiguous()
    vv = val.to(dev, torch.float32).contiguous()
    h = ctypes.c_void_p()
    check(lib().trspmm_route(ptr(rp), ptr(ci), ptr(vv), rp.numel() - 1, ci.numel(), int(ncols), ctypes.byref(cfg),
                             _lib.current_stream(), ctypes.byref(h)), "trspmm_route")
    return RoutedMeta(h, window, ci.numel(), bool(edges), rp.numel() - 1)


class Plan:
    """Device plan of a routed matrix (gfx950 lane layouts); owns its memory."""

    def __init__(self, handle, m_out: int, ncols: int, dtype: torch.dtype, keep=()):
        self._h = handle
        self.m_out = m_out
        self.ncols = ncols
        self.dtype = dtype
        self._keep = keep          # device tensors the plan was built from (kept alive by the caller's choice)
        c = (ctypes.c_int64 * 6)()
        check(lib().trspmm_plan_stats(handle, c), "trspmm_plan_stats")
        self.stats = dict(zip(["windows", "tc_subblocks", "tc_nnz", "sptc_tiles", "residual_items",
                               "residual_nnz"], list(c)))

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                lib().trspmm_plan_free(h)
            except Exception:
                pass

    @property
    def handle(self):
        return self._h

    @staticmethod
    def from_meta(meta: RoutedMeta, m_out: int, ncols: int, dtype=torch.float16, engine: str | None = None,
                  slab: int | None = None, tile_cap: int = 0, tc_cap: int = 0, res_cap: int = 0,
                  item_order: int = 0, k_hint: int = 0, res_order: int = 0, tile_kernel: int = 0) -> "Plan":
        """engine: None (default), "valu" or "mfma" residual engine; slab: None, 64 or 128
        features per pass of the tile kernel; *_cap: split a window into work items of at
        most this many 2:4 tiles / TC sub-blocks / residual entries (0 = default);
        item_order: 0 = one workgroup per item, items heaviest first, round robin over the
        XCDs; 1 = window order, contiguous item ranges per XCD; 2 = persistent workgroups taking
        the heaviest-first items from the plan's queue counters (launches of such a plan must
        not overlap on concurrent streams); 3 = as 0 in window order; k_hint: the feature width
        K the plan will mostly run at (0 = not known): a residual cap left at 0 is then sized
        for it, 4096 * 128 / max(K, 64) entries (any K still runs); res_order: 0 = the
        matrix-core residual engine's entries of a window by column (the default), 1 = in
        row order; tile_kernel: 0 = the deep-pipelined tile kernel (default), 1 = the earlier
        one-sub-step-ahead kernel (persistent plans always run it)."""
        opts = _lib.PlanOpts()
        opts.residual_engine = {None: _lib.ENGINE_DEFAULT, "valu": _lib.ENGINE_VALU, "mfma": _lib.ENGINE_MFMA}[engine]
        opts.slab = int(slab or 0)
        opts.tile_cap, opts.tc_cap, opts.res_cap, opts.item_order = int(tile_cap), int(tc_cap), int(res_cap), int(item_order)
        opts.k_hint = int(k_hint)
        opts.res_order = int(res_order)
        opts.tile_kernel = int(tile_kernel)
        h = ctypes.c_void_p()
        check(lib().trspmm_plan_from_meta_ex(meta.handle, int(m_out), int(ncols), dtype_code(dtype), ctypes.byref(opts),
                                             _lib.current_stream(), ctypes.byref(h)), "trspmm_plan_from_meta_ex")
        plan = Plan(h, m_out, ncols, dtype)
        plan.src_nnz = meta.nnz if meta.edges else None
        return plan

    def set_values(self, vals: torch.Tensor, index: torch.Tensor | None = None):
        """New values for a plan built from a meta routed with edges=True: the
        routed CSR's edge e takes vals[index[e]] (index=None: vals[e]); fp32 device
        tensors, index int32 of the routed CSR's nnz entries.  Stream-ordered."""
        nnz = getattr(self, "src_nnz", None)
        if nnz is None:
            raise ValueError("this plan has no edge maps: route with edges=True and build it with Plan.from_meta")
        if not vals.is_cuda or vals.dtype != torch.float32 or vals.dim() != 1 or not vals.is_contiguous():
            raise ValueError("vals must be a contiguous 1-D float32 device tensor")
        if index is not None and (not index.is_cuda or index.dtype != torch.int32 or index.shape != (nnz,)
                                  or not index.is_contiguous()):
            raise ValueError(f"index must be a contiguous int32 device tensor of shape ({nnz},)")
        if index is None and vals.numel() != nnz:
            raise ValueError(f"vals must have {nnz} entries (the routed CSR's nonzeros)")
        check(lib().trspmm_plan_set_values(self._h, ptr(vals), vals.numel(), ptr(index) if index is not None else None,
                                           nnz, _lib.current_stream()), "trspmm_plan_set_values")

    def matmul(self, B: torch.Tensor, out: torch.Tensor | None = None, route_mask: int = _lib.ROUTE_ALL):
        check_operand(B, self.ncols, self.dtype)
        K = B.shape[1]
        if out is None:
            out = torch.empty((self.m_out, K), device=B.device, dtype=torch.float32)
        check(lib().trspmm_spmm(self._h, ptr(B), B.stride(0), K, ptr(out), out.stride(0), route_mask,
                                _lib.current_stream()), "trspmm_spmm")
        return out

    def timed(self, B, out, route_mask=_lib.ROUTE_ALL, warmup=10, iters=10) -> float:
        ms = ctypes.c_float()
        check(lib().trspmm_spmm_timed(self._h, ptr(B), B.stride(0), B.shape[1], ptr(out), out.stride(0), route_mask,
                                      warmup, iters, _lib.current_stream(), ctypes.byref(ms)), "trspmm_spmm_timed")
        return float(ms.value)


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float16:
        return _lib.F16
    if dtype == torch.bfloat16:
        return _lib.BF16
    raise ValueError(f"unsupported operand dtype {dtype} (float16 or bfloat16)")


def check_operand(B: torch.Tensor, ncols: int, dtype: torch.dtype):
    if not B.is_cuda:
        raise ValueError("B must be a device tensor")
    if B.dtype != dtype:
        raise ValueError(f"B dtype {B.dtype} does not match the plan dtype {dtype}")
    if B.dim() != 2 or B.shape[0] != ncols:
        raise ValueError(f"B must be [{ncols}, K], got {tuple(B.shape)}")
    if B.stride(1) != 1 or B.shape[1] % 8 or B.stride(0) % 8 or B.data_ptr() % 16:
        raise ValueError("B must be row-major with K and the row stride multiples of 8, 16-byte aligned")
