"""Drop-in replacement for the reference's `Libra6SpMMOnline` extension.

Four ops with the signatures, argument contract, return tuples and timing
semantics of TR-source/SpMM/mGCN_online.cpp (PYBIND11_MODULE :1142-1152):

  forward_fp16_tcu_cuda_sptc_mma_parallel_online(26 metadata tensors, rhs_matrix,
        parts_t, parts_c, partsize_c, parts_c_short, window, dimN, mOri, kOri, epoches)
        -> [out f32[mOri, dimN], ms, ms_tc, ms_cuda_long, ms_cuda_short, ms_sptc]   (:760-1138)
  forward_fp16_sptc_online(s_column, s_value, s_pos, s_window, rhs, window, dimN, mOri, kOri, epoches)
        -> [out, ms]                                                                 (:202-400)
  forward_fp16_sptc_mma_online(s_column, s_value, s_pos, s_offsets, rhs, ...) -> [out, ms]   (:404-652)
  forward_fp16_sptc_mma_packed_online(s_tile_column, s_packed_a, s_packed_meta,
        s_window_tile_offset, rhs, ...) -> [out, ms]                                 (:654-757)

As in the reference: inputs are CPU tensors (a device tensor raises), are made
contiguous, must have the exact dtypes, and the output is a CPU fp32 tensor of
mOri rows; each call uploads, runs 10 untimed warm-up launches, times
`epoches` launches (ms = mean per launch) and returns the result of one final
launch.  The multiply runs as one fused routed launch (libtrspmm, gfx950):
every output row is written once with fp32 accumulation.  ms_tc / ms_sptc /
ms_cuda_* time the same launch restricted to one route; the residual route
serves the long and short part lists together, so ms_cuda_long and
ms_cuda_short report the same residual-route time.

Any window in (0, 16] runs, as in the reference.  At window 16 the routes
keep their tiles (16-row MFMA tiles); a narrower window's nonzeros (TC bits,
kept 2:4 slots, residual lists, packed tiles) run as residual records of
16-row windows, each at the row the reference's kernels give it (2:4 and
packed rows window_index * window + r, TC rows t_window_row + r).

Stricter than the reference (RuntimeError instead of undefined behaviour):
part counts must match the tensors they describe, offsets must be monotone
and in range, TC parts and residual parts must be in window / row order (as
the router emits them).
"""
from __future__ import annotations

import ctypes

import torch

from trspmm import _lib
from trspmm._lib import check, lib, ptr
from trspmm.core import Plan

_WARMUP = 10


def _cpu(t, name):
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a tensor")
    if t.is_cuda:
        raise RuntimeError(f"{name} must be a CPU tensor")
    return t.contiguous()


def _dtype(t, dt, name):
    if t.dtype != dt:
        raise RuntimeError(f"{name} has an unexpected dtype")


def _common_scalars(window, dimN, mOri, kOri, epoches, rhs):
    if not (0 < window <= 16):
        raise RuntimeError("parallel mma path requires 0 < window <= 16")
    for v, n in ((dimN, "dimN"), (mOri, "mOri"), (kOri, "kOri"), (epoches, "epoches")):
        if int(v) <= 0:
            raise RuntimeError(f"{n} must be positive")
    if rhs.numel() != int(kOri) * int(dimN):
        raise RuntimeError("rhs_matrix must have kOri * dimN elements")


def _nondecreasing(t, name):
    if t.numel() > 1 and bool((t[1:] < t[:-1]).any()):
        raise RuntimeError(f"{name} must be non-decreasing")


def _offsets(t, total, name):
    if t.numel() < 1 or int(t[0]) != 0 or int(t[-1]) != total:
        raise RuntimeError(f"{name} must start at 0 and end at {total}")
    _nondecreasing(t, name)


def _dev(t):
    return t.to("cuda", non_blocking=False)


def _operand(rhs, kOri, dimN):
    """B on the device, padded to a multiple of 8 features (16-byte rows)."""
    B = rhs.reshape(int(kOri), int(dimN))
    k8 = (int(dimN) + 7) // 8 * 8
    if k8 != int(dimN):
        B = torch.nn.functional.pad(B, (0, k8 - int(dimN)))
    return _dev(B.contiguous()), k8


def _run(plan: Plan, B, k8, mOri, dimN, epoches, masks):
    out = torch.empty((int(mOri), k8), device=B.device, dtype=torch.float32)
    times = [plan.timed(B, out, m, _WARMUP, int(epoches)) for m in masks]
    plan.matmul(B, out)                       # the returned product comes from one final launch
    torch.cuda.synchronize()
    return out[:, : int(dimN)].cpu().contiguous(), times


def _plan_from_view(view_tensors, n_windows, mOri, kOri, keep, dimN, window=16):
    """Windows narrower than 16 rows run as residual records of 16-row windows
    (the plan builder decodes the metadata's rows; see trspmm_plan_create)."""
    v = _lib.MetaView()
    for k, val in view_tensors.items():
        setattr(v, k, val)
    v.n_windows = n_windows
    v.window = int(window)
    opts = _lib.PlanOpts()
    opts.k_hint = int(dimN)                   # the op runs at one K: size the work items for it
    h = ctypes.c_void_p()
    check(lib().trspmm_plan_create_ex(ctypes.byref(v), int(mOri), int(kOri), _lib.F16, ctypes.byref(opts),
                                      _lib.current_stream(), ctypes.byref(h)), "trspmm_plan_create_ex")
    return Plan(h, int(mOri), int(kOri), torch.float16, keep)


def forward_fp16_tcu_cuda_sptc_mma_parallel_online(
        t_row_offset, t_blockNew_offset, t_column, t_value, t_window_row, t_atomic, t_binary,
        c_row_offsets, c_row, c_atomic, c_column, c_value,
        c_row_offsets_short, c_row_short, c_atomic_short, c_column_short, c_value_short,
        s_column, s_value, s_pos, s_offsets, s_tile_window, s_tile_column, s_packed_a, s_packed_meta,
        s_window_tile_offset, rhs_matrix,
        parts_t, parts_c, partsize_c, parts_c_short, window, dimN, mOri, kOri, epoches):
    names = ["t_row_offset", "t_blockNew_offset", "t_column", "t_value", "t_window_row", "t_atomic", "t_binary",
             "c_row_offsets", "c_row", "c_atomic", "c_column", "c_value",
             "c_row_offsets_short", "c_row_short", "c_atomic_short", "c_column_short", "c_value_short",
             "s_column", "s_value", "s_pos", "s_offsets", "s_tile_window", "s_tile_column", "s_packed_a",
             "s_packed_meta", "s_window_tile_offset", "rhs_matrix"]
    vals = [t_row_offset, t_blockNew_offset, t_column, t_value, t_window_row, t_atomic, t_binary,
            c_row_offsets, c_row, c_atomic, c_column, c_value,
            c_row_offsets_short, c_row_short, c_atomic_short, c_column_short, c_value_short,
            s_column, s_value, s_pos, s_offsets, s_tile_window, s_tile_column, s_packed_a, s_packed_meta,
            s_window_tile_offset, rhs_matrix]
    T = {n: _cpu(v, n) for n, v in zip(names, vals)}                 # mGCN_online.cpp:802-857
    i32, f16 = torch.int32, torch.float16
    for n, dt in [("t_row_offset", i32), ("t_blockNew_offset", i32), ("t_column", i32), ("t_value", f16),
                  ("t_window_row", i32), ("t_atomic", i32), ("t_binary", torch.int64),
                  ("c_row_offsets", i32), ("c_row", i32), ("c_atomic", i32), ("c_column", i32), ("c_value", f16),
                  ("c_row_offsets_short", i32), ("c_row_short", i32), ("c_atomic_short", i32),
                  ("c_column_short", i32), ("c_value_short", f16), ("s_column", i32), ("s_value", torch.float32),
                  ("s_pos", torch.int8), ("s_offsets", i32), ("s_tile_window", i32), ("s_tile_column", i32),
                  ("s_packed_a", i32), ("s_packed_meta", i32), ("rhs_matrix", f16)]:       # :859-883
        _dtype(T[n], dt, n)
    _common_scalars(window, dimN, mOri, kOri, epoches, T["rhs_matrix"])                   # :885-890
    sc, sv, sp, so = T["s_column"], T["s_value"], T["s_pos"], T["s_offsets"]
    if not (sc.dim() == 2 and sc.size(1) == 4):
        raise RuntimeError("s_column must have shape [groups, 4]")
    if not (sv.dim() == 3 and sv.size(1) == window and sv.size(2) == 2):
        raise RuntimeError("s_value must have shape [groups, window, 2]")
    if not (sp.dim() == 3 and sp.size(1) == window and sp.size(2) == 2):
        raise RuntimeError("s_pos must have shape [groups, window, 2]")
    if not (so.dim() == 1 and so.size(0) >= 1):
        raise RuntimeError("s_offsets must be 1-D with at least one item")
    if T["s_tile_window"].dim() != 1:
        raise RuntimeError("s_tile_window must be 1-D")
    if not (T["s_tile_column"].dim() == 2 and T["s_tile_column"].size(1) == 16):
        raise RuntimeError("s_tile_column must have shape [tiles, 16]")
    pa = T["s_packed_a"]
    if not (pa.dim() == 4 and tuple(pa.shape[1:]) == (8, 4, 2)):
        raise RuntimeError("s_packed_a must have shape [tiles, 8, 4, 2]")
    if not (T["s_packed_meta"].dim() == 2 and T["s_packed_meta"].size(1) == 8):
        raise RuntimeError("s_packed_meta must have shape [tiles, 8]")
    G = sc.size(0)
    if sv.size(0) != G:
        raise RuntimeError("s_value group count mismatch")
    if sp.size(0) != G:
        raise RuntimeError("s_pos group count mismatch")
    W = so.size(0) - 1
    if int(so[0]) != 0:
        raise RuntimeError("s_offsets must start with 0")
    if int(so[W]) != G:
        raise RuntimeError("s_offsets end must equal group count")
    _nondecreasing(so, "s_offsets")
    ntiles = int(((so[1:] - so[:-1] + 3) // 4).sum()) if W > 0 else 0
    nt = T["s_tile_window"].size(0)
    if T["s_tile_column"].size(0) != nt:
        raise RuntimeError("s_tile_column tile count mismatch")
    if pa.size(0) != nt:
        raise RuntimeError("s_packed_a tile count mismatch")
    if T["s_packed_meta"].size(0) != nt:
        raise RuntimeError("s_packed_meta tile count mismatch")
    if nt != ntiles:
        raise RuntimeError("packed tile count must match offsets-derived tile count")
    # stricter consistency checks (device safety)
    twr, tro, tbo = T["t_window_row"], T["t_row_offset"], T["t_blockNew_offset"]
    if int(parts_t) != twr.numel() or tro.numel() != twr.numel() + 1:
        raise RuntimeError("parts_t must equal len(t_window_row) and len(t_row_offset) - 1")
    nsub = T["t_column"].numel() // 8
    if T["t_column"].numel() != 8 * nsub or T["t_binary"].numel() != 2 * nsub:
        raise RuntimeError("t_column / t_binary must hold 8 / 2 entries per sub-block")
    _offsets(tro, nsub, "t_row_offset")
    _offsets(tbo if tbo.numel() else torch.zeros(1, dtype=i32), T["t_value"].numel(), "t_blockNew_offset")
    if tbo.numel() != nsub + 1:
        raise RuntimeError("t_blockNew_offset must have one entry per sub-block plus one")
    if twr.numel() and (bool((twr % 16 != 0).any()) or bool((twr < 0).any())):
        raise RuntimeError("t_window_row must hold 16 * window indices")
    _nondecreasing(twr, "t_window_row")
    for pre, cnt in (("c_row", parts_c), ("c_row_short", parts_c_short)):
        off = T["c_row_offsets" if pre == "c_row" else "c_row_offsets_short"]
        rows = T[pre]
        colt = T["c_column" if pre == "c_row" else "c_column_short"]
        valt = T["c_value" if pre == "c_row" else "c_value_short"]
        if int(cnt) != rows.numel() or off.numel() != rows.numel() + 1:
            raise RuntimeError(f"part count must equal len({pre}) and its offsets length - 1")
        _offsets(off, colt.numel(), pre + " offsets")
        if valt.numel() != colt.numel():
            raise RuntimeError(f"{pre}: column and value counts differ")
        _nondecreasing(rows, pre)
        if rows.numel() and (int(rows.min()) < 0 or int(rows.max()) >= int(window) * W):
            raise RuntimeError(f"{pre} rows must lie inside the windows")
    if int(window) * W < int(mOri):
        raise RuntimeError("the windows of s_offsets must cover mOri rows")

    D = {n: _dev(T[n]) for n in names if n != "rhs_matrix"}
    view = {
        "t_row_offset": ptr(D["t_row_offset"]), "n_tparts": twr.numel(),
        "t_block_offset": ptr(D["t_blockNew_offset"]), "t_column": ptr(D["t_column"]), "n_tsub": nsub,
        "t_value": ptr(D["t_value"]), "n_tval": T["t_value"].numel(), "t_window_row": ptr(D["t_window_row"]),
        "t_binary": ptr(D["t_binary"]),
        "c_row_offset": ptr(D["c_row_offsets"]), "n_cparts": T["c_row"].numel(), "c_row": ptr(D["c_row"]),
        "c_column": ptr(D["c_column"]), "n_cval": T["c_column"].numel(), "c_value": ptr(D["c_value"]),
        "cs_row_offset": ptr(D["c_row_offsets_short"]), "n_csparts": T["c_row_short"].numel(),
        "cs_row": ptr(D["c_row_short"]), "cs_column": ptr(D["c_column_short"]),
        "n_csval": T["c_column_short"].numel(), "cs_value": ptr(D["c_value_short"]),
        "s_column": ptr(D["s_column"]), "n_groups": G, "s_value": ptr(D["s_value"]), "s_pos": ptr(D["s_pos"]),
        "s_offsets": ptr(D["s_offsets"]), "t_value_is_f32": 0, "c_value_is_f32": 0,
    }
    plan = _plan_from_view(view, W, mOri, kOri, keep=tuple(D.values()), dimN=dimN, window=window)
    B, k8 = _operand(T["rhs_matrix"], kOri, dimN)
    out, (ms, ms_tc, ms_res, ms_sptc) = _run(plan, B, k8, mOri, dimN, epoches,
                                             [_lib.ROUTE_ALL, _lib.ROUTE_TC, _lib.ROUTE_RESIDUAL, _lib.ROUTE_SPTC])
    f = lambda x: torch.tensor(x, dtype=torch.float32)   # noqa: E731
    return [out, f(ms), f(ms_tc), f(ms_res), f(ms_res), f(ms_sptc)]


def _sptc_only(sc, sv, sp, so, rhs, window, dimN, mOri, kOri, epoches):
    G = sc.size(0)
    W = so.numel() - 1
    if int(window) * W < int(mOri):
        raise RuntimeError("the windows of s_offsets must cover mOri rows")
    z = torch.zeros(1, dtype=torch.int32, device="cuda")
    Dsc, Dsv, Dsp, Dso = _dev(sc), _dev(sv), _dev(sp), _dev(so)
    view = {"t_row_offset": ptr(z), "t_block_offset": ptr(z), "c_row_offset": ptr(z), "cs_row_offset": ptr(z),
            "s_column": ptr(Dsc), "n_groups": G, "s_value": ptr(Dsv), "s_pos": ptr(Dsp), "s_offsets": ptr(Dso)}
    plan = _plan_from_view(view, W, mOri, kOri, keep=(z, Dsc, Dsv, Dsp, Dso), dimN=dimN, window=window)
    B, k8 = _operand(rhs, kOri, dimN)
    out, (ms,) = _run(plan, B, k8, mOri, dimN, epoches, [_lib.ROUTE_ALL])
    return [out, torch.tensor(ms, dtype=torch.float32)]


def _check_groups(sc, sv, sp, window):
    if not (sc.dim() == 2 and sc.size(1) == 4):
        raise RuntimeError("s_column must have shape [groups, 4]")
    if not (sv.dim() == 3 and sv.size(1) == window and sv.size(2) == 2):
        raise RuntimeError("s_value must have shape [groups, window, 2]")
    if not (sp.dim() == 3 and sp.size(1) == window and sp.size(2) == 2):
        raise RuntimeError("s_pos must have shape [groups, window, 2]")
    G = sc.size(0)
    if sv.size(0) != G:
        raise RuntimeError("s_value group count mismatch")
    if sp.size(0) != G:
        raise RuntimeError("s_pos group count mismatch")


def forward_fp16_sptc_online(s_column, s_value, s_pos, s_window, rhs_matrix, window, dimN, mOri, kOri, epoches):
    sc, sv, sp, sw, rhs = (_cpu(s_column, "s_column"), _cpu(s_value, "s_value"), _cpu(s_pos, "s_pos"),
                           _cpu(s_window, "s_window"), _cpu(rhs_matrix, "rhs_matrix"))
    _dtype(sc, torch.int32, "s_column"), _dtype(sv, torch.float32, "s_value"), _dtype(sp, torch.int8, "s_pos")
    _dtype(sw, torch.int32, "s_window"), _dtype(rhs, torch.float16, "rhs_matrix")
    if window <= 0:
        raise RuntimeError("window must be positive")
    _common_scalars(window, dimN, mOri, kOri, epoches, rhs)
    _check_groups(sc, sv, sp, window)
    if sw.dim() != 1 or sw.size(0) != sc.size(0):
        raise RuntimeError("s_column group count mismatch")
    # groups may come in any window order here (each is an independent product term)
    W = max((int(mOri) + int(window) - 1) // int(window), int(sw.max()) + 1 if sw.numel() else 0)
    if sw.numel() and int(sw.min()) < 0:
        raise RuntimeError("s_window entries must be non-negative")
    order = torch.argsort(sw, stable=True)
    sc, sv, sp, sw = sc[order].contiguous(), sv[order].contiguous(), sp[order].contiguous(), sw[order]
    so = torch.zeros(W + 1, dtype=torch.int32)
    so[1:] = torch.cumsum(torch.bincount(sw.long(), minlength=W)[:W], 0).to(torch.int32)
    return _sptc_only(sc, sv, sp, so, rhs, window, dimN, mOri, kOri, epoches)


def forward_fp16_sptc_mma_online(s_column, s_value, s_pos, s_offsets, rhs_matrix, window, dimN, mOri, kOri,
                                 epoches):
    sc, sv, sp, so, rhs = (_cpu(s_column, "s_column"), _cpu(s_value, "s_value"), _cpu(s_pos, "s_pos"),
                           _cpu(s_offsets, "s_offsets"), _cpu(rhs_matrix, "rhs_matrix"))
    _dtype(sc, torch.int32, "s_column"), _dtype(sv, torch.float32, "s_value"), _dtype(sp, torch.int8, "s_pos")
    _dtype(so, torch.int32, "s_offsets"), _dtype(rhs, torch.float16, "rhs_matrix")
    if not (0 < window <= 16):
        raise RuntimeError("mma path requires 0 < window <= 16")
    _common_scalars(window, dimN, mOri, kOri, epoches, rhs)
    _check_groups(sc, sv, sp, window)
    if not (so.dim() == 1 and so.size(0) >= 1):
        raise RuntimeError("s_offsets must be 1-D with at least one item")
    if int(so[0]) != 0:
        raise RuntimeError("s_offsets must start with 0")
    if int(so[-1]) != sc.size(0):
        raise RuntimeError("s_offsets end must equal group count")
    _nondecreasing(so, "s_offsets")
    return _sptc_only(sc, sv, sp, so, rhs, window, dimN, mOri, kOri, epoches)


def forward_fp16_sptc_mma_packed_online(s_tile_column, s_packed_a, s_packed_meta, s_window_tile_offset,
                                        rhs_matrix, window, dimN, mOri, kOri, epoches):
    tc, pa, pm, wo, rhs = (_cpu(s_tile_column, "s_tile_column"), _cpu(s_packed_a, "s_packed_a"),
                           _cpu(s_packed_meta, "s_packed_meta"), _cpu(s_window_tile_offset, "s_window_tile_offset"),
                           _cpu(rhs_matrix, "rhs_matrix"))
    for t, n in ((tc, "s_tile_column"), (pa, "s_packed_a"), (pm, "s_packed_meta"), (wo, "s_window_tile_offset")):
        _dtype(t, torch.int32, n)
    _dtype(rhs, torch.float16, "rhs_matrix")
    if not (0 < window <= 16):
        raise RuntimeError("packed mma path requires 0 < window <= 16")
    _common_scalars(window, dimN, mOri, kOri, epoches, rhs)
    if not (tc.dim() == 2 and tc.size(1) == 16):
        raise RuntimeError("s_tile_column must have shape [tiles, 16]")
    if not (pa.dim() == 4 and tuple(pa.shape[1:]) == (8, 4, 2)):
        raise RuntimeError("s_packed_a must have shape [tiles, 8, 4, 2]")
    if not (pm.dim() == 2 and pm.size(1) == 8):
        raise RuntimeError("s_packed_meta must have shape [tiles, 8]")
    if not (wo.dim() == 1 and wo.size(0) >= 2):
        raise RuntimeError("s_window_tile_offset must be 1-D with at least 2 elements")
    T = tc.size(0)
    if pa.size(0) != T:
        raise RuntimeError("s_packed_a tile count mismatch")
    if pm.size(0) != T:
        raise RuntimeError("s_packed_meta tile count mismatch")
    _offsets(wo, T, "s_window_tile_offset")
    W = wo.numel() - 1
    if int(window) * W < int(mOri):
        raise RuntimeError("the windows of s_window_tile_offset must cover mOri rows")
    Dtc, Dpa, Dpm, Dwo = _dev(tc), _dev(pa), _dev(pm), _dev(wo)
    h = ctypes.c_void_p()
    check(lib().trspmm_plan_create_packed(ptr(Dtc), ptr(Dpa), ptr(Dpm), ptr(Dwo), T, W, int(window), int(mOri),
                                          int(kOri),
                                          _lib.current_stream(), ctypes.byref(h)), "trspmm_plan_create_packed")
    plan = Plan(h, int(mOri), int(kOri), torch.float16, keep=(Dtc, Dpa, Dpm, Dwo))
    B, k8 = _operand(rhs, kOri, dimN)
    out, (ms,) = _run(plan, B, k8, mOri, dimN, epoches, [_lib.ROUTE_ALL])
    return [out, torch.tensor(ms, dtype=torch.float32)]
