"""Routed SpMM benchmark (one MI355X).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rmat18_k128] [--mode tuned]
                    [--detail FILE]

A step is one routed multiply C = A·B over the whole matrix: R-MAT scale 18
(262 144 rows, 8 393 760 nnz), K = 128, fp16 operands, fp32 accumulation and
output, with A's routed plan and B already resident in HBM.  Preprocessing
(device router + plan) runs once before the timed region.  The default
routing is the tuned router (2:4 groups with fewer than 8 of 32 slots filled
go to the residual route, which runs as one-hot MFMA fragments in the same
kernel); config.exact_ms times the same product on the reference's tiling.

Prints ONE JSON line (at most about 2 KB) on stdout:
  value        GFLOP/s = 2·nnz·K / ms_per_step (HIP events on the launch stream)
  roofline     HBM-bound: algorithmic bytes per multiply
               nnz·(2+4) + 4·(M+1) + N·K·2 + M·K·4   (fp16 values/B, int32 cols, fp32 C)
               over ms_per_step, against the 8 TB/s MI355X peak; traffic = measured
               HBM bytes per multiply (rocprofv3 FETCH_SIZE + WRITE_SIZE, recorded in
               profiles/) when recorded for this workload, else null; copy_gbs = a
               plain device copy measured in the same run
  config       workload, sizes, FP64 error, router time, exact-routing time and the
               multiply timed after flushing the L2 and the 256 MB MALL (ms_cold)
  cpu_baseline the reference's CPU path on the host: torch.sparse.mm (fp32) over the
               whole workload with the host threads of one GPU's share, and the
               oracle's faithful restatement of the reference router, single-threaded
               like the reference, on the leading (hub) rows
  fwd_bwd      (extras) the same graph at K = 256, fp16 and bf16: tuned forward vs
               tuned forward + fused backward (dB = A^T·dY and dvals = SDDMM(dY, B)),
               the ratio on the reference's tiling beside it, errors vs FP64
  batched      (extras) 1024 Barabasi-Albert graphs (m=3, n_i log-uniform in
               [20, 5000]), K=128: one autotuned launch replayed from a captured
               hipGraph vs 1024 sequential per-graph launches, error vs FP64
  configs      (extras) the ER n=4096 K=64 and three-route 2^20 K=256 products:
               time, roofline fraction and error vs FP64
Full tables (autotuner candidates, per-route figures) go to stderr and, with
--detail FILE, into FILE as JSON.  --no-extras times the headline only.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "tr-spmm_amd"), ROOT]

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    # the first ~60 multiplies after routing run about 5 % slower than the rest
    # (measured: 0.215 ms for 10 warm-up + 50 timed steps, 0.205 ms for every
    # later set of 50); the default warm-up reaches steady state, a shorter one
    # given on the command line is honoured as given
    p.add_argument("--warmup", type=int, default=200)
    p.add_argument("--workload", default="rmat18_k128")
    p.add_argument("--mode", default="tuned", choices=["exact", "compat", "tuned"],
                   help="routing of the measured multiply: tuned (the default: 2:4 groups with fewer than 8 of "
                        "32 slots filled go to the residual route), exact (the reference's tiles, every nonzero "
                        "kept) or compat (the reference's metadata bit for bit)")
    p.add_argument("--dtype", default="fp16", choices=["fp16", "bf16"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-check", action="store_true")
    p.add_argument("--no-extras", action="store_true",
                   help="skip the exact-routing comparison, fwd+bwd, the batched graphs and the other configs")
    p.add_argument("--detail", default=None, help="write the run's full tables to this JSON file")
    p.add_argument("--batched", action="store_true", help="measure only the batched workload")
    p.add_argument("--per-route", action="store_true",
                   help="per-route rows, nnz and times of --workload (exact and tuned routing)")
    p.add_argument("--backward", action="store_true",
                   help="measure only forward vs forward+backward on --workload (fp16 and bf16)")
    p.add_argument("--matching", action="store_true",
                   help="time the four setup_24matching statistics ops on --workload's matrix")
    p.add_argument("--library", action="store_true",
                   help="time the GPU sparse library's fp32 SpMM (torch.sparse.mm) on --workload's matrix")
    p.add_argument("--copy-calibration", action="store_true",
                   help="also run a plain 256 MiB device copy (the reference kernel for scaling "
                        "rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    return p.parse_args()


def algorithmic_bytes(nnz, m, n, k, vb=2, cb=4):
    return nnz * (vb + 4) + 4 * (m + 1) + n * k * vb + m * k * cb


def cpu_baseline(csr, K, B16, budget_s=20.0):
    """The reference's CPU path on the host cores: its router as preprocessing
    (the oracle's faithful restatement of block_online.cpp, single-threaded like
    the reference, printf-free) on a bounded leading row sample, and
    torch.sparse.mm(A_csr fp32, B fp32) over the whole workload with every host
    thread of this GPU's share (median of 5 after one warm-up).  `value` is the
    multiply's GFLOP/s."""
    import torch

    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    torch.set_num_threads(threads)
    A = torch.sparse_csr_tensor(torch.from_numpy(csr.rowptr.astype(np.int64)), torch.from_numpy(csr.col.astype(np.int64)),
                                torch.from_numpy(csr.val.astype(np.float32)), size=(csr.m_pad, csr.n))
    Bf = torch.from_numpy(B16.astype(np.float32))
    torch.sparse.mm(A, Bf)
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        torch.sparse.mm(A, Bf)
        times.append(time.perf_counter() - t0)
    t_mm = float(np.median(times))
    # router on a leading row sample (grown x4 while it takes under budget_s / 8;
    # R-MAT's leading rows are its hubs, where the reference's matching is quadratic)
    rows = 16 * 16
    while True:
        rows = min(rows, csr.m_pad)
        rp = csr.rowptr[: rows + 1]
        nnz = int(rp[-1])
        t0 = time.perf_counter()
        oracle.route(rp, csr.col[:nnz], csr.val[:nnz], faithful=True)
        dt = time.perf_counter() - t0
        if dt * 4 > budget_s / 2 or rows >= csr.m_pad:
            break
        rows *= 4
    return {"value": round(2.0 * csr.nnz * K / t_mm / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"torch.sparse.mm fp32, whole workload, {threads} threads, median {t_mm * 1e3:.1f} ms; "
                      f"router (faithful port, 1 thread) on rows [0,{rows}) ({nnz} nnz): {dt:.2f} s, "
                      f"{nnz / dt / 1e6:.3f} M nnz/s"}


def traffic_record(workload, mode, dtype, reordered):
    """Measured HBM bytes per routed multiply (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    scaled by a plain copy's readings), recorded for this workload, routing mode,
    row order and operand type; None when no such measurement is recorded."""
    path = os.path.join(ROOT, "profiles", f"spmm_{workload}_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            rec = json.load(f)
        if (rec.get("mode") == mode and rec.get("dtype", "fp16") == dtype
                and bool(rec.get("rows_reordered", False)) == bool(reordered)):
            return rec.get("traffic_bytes")
    return None


def time_steps(fn, steps, warmup):
    """Mean ms per call of fn() over `steps` calls after `warmup`, HIP events on
    the current stream (the stream the library launches on)."""
    import torch
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def time_cold(fn, iters, flush_bytes=512 << 20):
    """Mean ms of fn() when it starts with the L2s and the 256 MB MALL holding
    none of its data: before each call a 512 MiB buffer is read and written
    (outside the events), so only fn() itself is timed."""
    import torch
    stream = torch.cuda.current_stream()
    flush = torch.zeros(flush_bytes // 4, dtype=torch.float32, device="cuda")
    ev = []
    for _ in range(iters):
        flush.add_(1.0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        ev.append((e0, e1))
    torch.cuda.synchronize()
    ms = sum(a.elapsed_time(b) for a, b in ev) / iters
    del flush
    return ms


def copy_bandwidth(nbytes=1 << 30, iters=10):
    """Achievable HBM bandwidth of a plain device copy (GB/s, read + write bytes)."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    ms = time_steps(lambda: dst.copy_(src), iters, 2)
    del src, dst
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def measure_fwd_bwd(csr, K, dtype_name, steps, warmup, check=True, full=False):
    """Forward vs forward + backward on one graph at width K: the tuned
    router's forward (the headline routing) with the fused backward (dB =
    A^T dY and dvals = SDDMM(dY, B) in one pass over A^T's rows), and the same
    on the reference's tiling (exact routing).  Preprocessing (plans) outside
    the timing.  full=True also times the routed backward (dB on A^T's routed
    plan, then the SDDMM kernel) and each backward kernel alone."""
    import torch

    from trspmm import graphs
    from trspmm.sparse import RoutedSparse
    dtype = torch.float16 if dtype_name == "fp16" else torch.bfloat16
    dev = torch.device("cuda", 0)
    B = torch.from_numpy(graphs.dense_features(csr.n, K, seed=1001).astype(np.float32)).to(dev).to(dtype)
    dY = torch.from_numpy(graphs.dense_features(csr.m, K, seed=1003).astype(np.float32)).to(dev).to(dtype)
    C = torch.empty((csr.m, K), device=dev, dtype=torch.float32)
    dB = torch.empty((csr.n, K), device=dev, dtype=torch.float32)
    dv = torch.empty(csr.nnz, device=dev, dtype=torch.float32)
    out = {"K": K, "dtype": dtype_name}
    for mode in ("tuned", "exact"):
        A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                         m=csr.m, mode=mode, dtype=dtype, plan_opts={"k_hint": K})
        A.spmm_t_sddmm(dY, B, dB, dv)                   # builds the fused plan (preprocessing)
        torch.cuda.synchronize()
        ms_f = time_steps(lambda: A.matmul(B, C), steps, warmup)
        ms_fb = time_steps(lambda: (A.matmul(B, C), A.spmm_t_sddmm(dY, B, dB, dv)), steps, warmup)
        out[mode] = {"ms_fwd": round(ms_f, 5), "ms_fwd_bwd": round(ms_fb, 5), "ratio": round(ms_fb / ms_f, 3)}
        if mode == "tuned":
            out["ms_bwd_fused"] = round(time_steps(lambda: A.spmm_t_sddmm(dY, B, dB, dv), steps, warmup), 5)
            if full:
                At = A.transpose()
                torch.cuda.synchronize()
                out["ms_fwd_bwd_routed"] = round(time_steps(
                    lambda: (A.matmul(B, C), At.matmul(dY, dB), A.sddmm(dY, B, dv)), steps, warmup), 5)
                out["ms_dB_routed"] = round(time_steps(lambda: At.matmul(dY, dB), steps, warmup), 5)
                out["ms_sddmm"] = round(time_steps(lambda: A.sddmm(dY, B, dv), steps, warmup), 5)
                del At
            if check:
                A.spmm_t_sddmm(dY, B, dB, dv)        # the fused results are the ones checked below
                torch.cuda.synchronize()
                import scipy.sparse as sp
                Af = sp.csr_matrix((csr.val.astype(np.float64), csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
                dY64 = np.zeros((csr.m_pad, K))
                dY64[: csr.m] = dY.float().cpu().numpy()
                B64 = B.float().cpu().numpy().astype(np.float64)
                ref = np.asarray(Af.T @ dY64)
                got = dB.cpu().numpy().astype(np.float64)
                out["err_dB"] = float(f"{np.linalg.norm(got - ref) / np.linalg.norm(ref):.3e}")
                rng = np.random.default_rng(0)
                sample = rng.choice(csr.nnz, min(csr.nnz, 50000), replace=False)
                rows = np.searchsorted(csr.rowptr, sample, side="right") - 1
                ref_v = np.einsum("ef,ef->e", dY64[rows], B64[csr.col[sample]])
                got_v = dv.cpu().numpy()[sample].astype(np.float64)
                out["err_dvals"] = float(f"{np.linalg.norm(got_v - ref_v) / np.linalg.norm(ref_v):.3e}")
        del A
        torch.cuda.empty_cache()
    return out


def route_rows(meta, rows):
    """Rows holding at least one nonzero of each route, from the routed metadata
    (device tensors): TC sub-block occupancy bits, 2:4 per-row slots, residual parts."""
    import torch
    t = dict(zip(__import__("trspmm").META_NAMES, meta.tensors("cuda")))
    W = t["s_offsets"].numel() - 1
    dev = t["s_offsets"].device
    out = {}
    # TC: sub-block s of part p covers window t_window_row[p] / 16; bit (r%8)*8+q of word r/8
    tc = torch.zeros(W, 16, dtype=torch.int32, device=dev)
    nsub = t["t_binary"].numel() // 2
    if nsub:
        part_of = torch.repeat_interleave(torch.arange(t["t_window_row"].numel(), device=dev),
                                          torch.diff(t["t_row_offset"]).long())
        win = (t["t_window_row"][part_of] // 16).long()
        words = t["t_binary"].view(nsub, 2)
        r = torch.arange(16, device=dev)
        occ = ((words[:, (r // 8)] >> ((r % 8) * 8).long()) & 0xff) != 0          # [S, 16]
        tc.index_put_((win[:, None].expand(-1, 16).reshape(-1), r.repeat(nsub)), occ.reshape(-1).int(),
                      accumulate=True)
    sp = torch.zeros(W, 16, dtype=torch.int32, device=dev)
    G = t["s_column"].shape[0]
    if G:
        hit = (t["s_pos"].view(G, 16, 2) >= 0).any(-1)
        sw = t["s_window"].long()
        r = torch.arange(16, device=dev)
        sp.index_put_((sw[:, None].expand(-1, 16).reshape(-1), r.repeat(G)), hit.reshape(-1).int(), accumulate=True)
    res = torch.zeros(W * 16, dtype=torch.bool, device=dev)
    for rows_t, off_t in ((t["c_row"], t["c_row_offset"]), (t["c_row_short"], t["c_row_offset_short"])):
        if rows_t.numel():
            nz = torch.diff(off_t) > 0
            res[rows_t[nz].long()] = True
    m = rows
    out["tc"] = int((tc.view(-1)[:m] > 0).sum())
    out["sptc"] = int((sp.view(-1)[:m] > 0).sum())
    out["residual"] = int(res[:m].sum())
    return out


def measure_routes(csr, K, mode, steps, warmup):
    """Per-route view of one routing (the three-route workload): per-route nnz, rows and launch times."""
    import torch

    from trspmm import ROUTE_RESIDUAL, ROUTE_SPTC, ROUTE_TC, graphs
    from trspmm.sparse import RoutedSparse
    dev = torch.device("cuda", 0)
    B = torch.from_numpy(graphs.dense_features(csr.n, K, seed=1002)).to(dev)
    A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode=mode, plan_opts={"k_hint": K})
    C = torch.empty((csr.m, K), device=dev, dtype=torch.float32)
    st = A.plan.stats
    ms = {name: time_steps(lambda mk=mk: A.matmul(B, C, routes=mk), steps, warmup)
          for name, mk in (("all", 7), ("tc", ROUTE_TC), ("sptc", ROUTE_SPTC), ("residual", ROUTE_RESIDUAL))}
    nnz = {"tc": st["tc_nnz"], "residual": st["residual_nnz"]}
    nnz["sptc"] = csr.nnz - nnz["tc"] - nnz["residual"]
    out = {"mode": mode, "ms": {k: round(v, 5) for k, v in ms.items()}, "nnz": nnz,
           "rows_with_route": route_rows(A.meta, csr.m), "tc_subblocks": st["tc_subblocks"],
           "sptc_tiles": st["sptc_tiles"], "router_s": round(A.meta.seconds, 4),
           "gflops": round(2.0 * csr.nnz * K / (ms["all"] * 1e-3) / 1e9, 1),
           "roofline_frac": round(algorithmic_bytes(csr.nnz, csr.m, csr.n, K) / (ms["all"] * 1e-3) / 1e9
                                  / HBM_PEAK_GBS, 4)}
    for name in ("tc", "sptc", "residual"):
        if ms[name] > 0 and nnz[name]:
            out.setdefault("gflops_per_route", {})[name] = round(2.0 * nnz[name] * K / (ms[name] * 1e-3) / 1e9, 1)
    return out


def measure_batched(steps, warmup, check=True, count=1024, K=128):
    """`count` Barabasi-Albert graphs (m = 3) in one launch vs one launch per graph."""
    import torch

    from trspmm import graphs
    from trspmm.batch import RoutedBatch, concat_block_diagonal
    from trspmm.sparse import RoutedSparse
    from trspmm.tune import autotune
    dev = torch.device("cuda", 0)
    t0 = time.time()
    gs = graphs.ba_batch(count)
    gen_s = time.time() - t0
    rowptr, col, val, row_off, col_off, m_list, n_list = concat_block_diagonal(gs)
    ncols, rows, nnz = int(col_off[-1]), int(row_off[-1]), int(col.size)
    Bcat = torch.from_numpy(graphs.dense_features(ncols, K, seed=5001)).to(dev)
    t0 = time.time()
    tune = autotune(torch.from_numpy(rowptr), torch.from_numpy(col), torch.from_numpy(val), ncols, rows, Bcat,
                    iters=10, warmup=3)
    tune_s = time.time() - t0
    batch = RoutedBatch(gs, engine=tune.best["engine"], slab=tune.best["slab"], plan_opts=tune.plan_opts(),
                        **tune.router_kwargs())
    C = torch.empty((rows, K), device=dev, dtype=torch.float32)
    ms_eager = time_steps(lambda: batch.matmul(Bcat, C), steps, warmup)
    graph = batch.capture(Bcat, C)
    ms_graph = time_steps(graph.replay, steps, warmup)
    # plain: one routed matrix per graph (default routing and kernels), launched one after another
    t0 = time.time()
    plains = [RoutedSparse(torch.from_numpy(c.rowptr), torch.from_numpy(c.col), torch.from_numpy(c.val), c.n, m=c.m)
              for c in gs]
    plain_prep_s = time.time() - t0
    Cs = torch.empty((rows, K), device=dev, dtype=torch.float32)
    Bv = [Bcat[int(col_off[i]): int(col_off[i]) + int(n_list[i])] for i in range(count)]
    Cv = [Cs[int(row_off[i]): int(row_off[i]) + int(m_list[i])] for i in range(count)]

    def sequential():
        for A, b, c in zip(plains, Bv, Cv):
            A.matmul(b, c)

    ms_seq = time_steps(sequential, max(3, steps // 10), 1)
    flops = 2.0 * nnz * K
    byts = algorithmic_bytes(nnz, rows, ncols, K)
    out = {"workload": f"{count} BA graphs m=3, n_i in [20, 5000], K={K}", "rows_padded": rows, "nnz": nnz,
           "ms_batched_graph": round(ms_graph, 5), "ms_batched_eager": round(ms_eager, 5),
           "ms_sequential_plain": round(ms_seq, 4), "speedup_vs_sequential": round(ms_seq / ms_graph, 2),
           "gflops_batched": round(flops / (ms_graph * 1e-3) / 1e9, 1),
           "roofline_frac_batched": round(byts / (ms_graph * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "autotuned": tune.best, "autotune_table_ms": [[c, round(ms, 5)] for c, ms in tune.table],
           "routes": batch.A.plan.stats, "gen_s": round(gen_s, 1), "autotune_s": round(tune_s, 1),
           "plain_prep_s": round(plain_prep_s, 1)}
    if check:
        import scipy.sparse as sp
        A64 = sp.csr_matrix((val.astype(np.float64), col, rowptr), shape=(rows, ncols))
        ref = np.asarray(A64 @ Bcat.float().cpu().numpy().astype(np.float64))
        graph.replay()
        torch.cuda.synchronize()
        out["rel_err_batched_vs_fp64"] = float(np.linalg.norm(C.cpu().numpy() - ref) / np.linalg.norm(ref))
        sequential()
        torch.cuda.synchronize()
        out["rel_err_sequential_vs_fp64"] = float(np.linalg.norm(Cs.cpu().numpy() - ref) / np.linalg.norm(ref))
    return out


def gpu_library_baseline(csr, B16, steps=10):
    """The GPU's own sparse library on the same product: torch.sparse.mm of
    the CSR matrix (hipSPARSE SpMM, its default algorithm), in fp32 because
    hipSPARSE SpMM has no fp16 CSR path on this image.  A reported reference
    point beside the routed kernel, not the measured value."""
    import torch
    dev = torch.device("cuda", 0)
    A = torch.sparse_csr_tensor(torch.from_numpy(csr.rowptr).to(dev), torch.from_numpy(csr.col).to(dev),
                                torch.from_numpy(csr.val).to(dev), size=(csr.m_pad, csr.n))
    B = torch.from_numpy(B16).to(dev).float()
    for _ in range(2):
        C = A @ B
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        C = A @ B
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    del A, B, C
    torch.cuda.empty_cache()
    return {"what": "torch.sparse.mm (hipSPARSE SpMM, default algorithm), fp32 CSR x fp32 B (no fp16 CSR SpMM "
                    "there), same matrix, one GPU", "ms": round(ms, 4),
            "gflops": round(2.0 * csr.nnz * B16.shape[1] / (ms * 1e-3) / 1e9, 1)}


def measure_matching(csr, workload, steps, warmup):
    """Wall time of each matching statistics op on the workload's CSR: the
    host ops over their thread pool (best of `steps`, min 3), the device op
    with HIP events around the kernels alone, inputs already on the device."""
    import ctypes
    import warnings

    import torch

    import matching_utils
    import matching_utils_bucket
    import matching_utils_hash
    import matching_utils_online
    from trspmm import _lib

    rp, col, val = csr.rowptr, csr.col, csr.val
    out = {"workload": workload, "rows": int(rp.size - 1), "nnz": int(col.size), "host_threads": 16, "ms": {}}
    reps = max(3, min(steps, 10))
    for name, fn in (("match_2to4", matching_utils.match_2to4), ("match_2to4_hash", matching_utils_hash.match_2to4_hash),
                     ("match_2to4_bucket", matching_utils_bucket.match_2to4_bucket)):
        best = float("inf")
        for _ in range(reps):
            t = time.perf_counter()
            fn(rp, col, val, threads=16)
            best = min(best, time.perf_counter() - t)
        out["ms"][name] = round(best * 1e3, 3)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)     # hub windows with > 2048 columns: counted below
        stats = matching_utils_online.match_2to4_online(rp, col, val)
    drp = torch.from_numpy(rp).cuda()
    dcol = torch.from_numpy(col).cuda()
    cnt = torch.empty(11, dtype=torch.int64, device="cuda")
    L = _lib.lib()

    def call():
        _lib.check(L.trspmm_match_2to4_online(ctypes.c_void_p(drp.data_ptr()), ctypes.c_void_p(dcol.data_ptr()),
                                              rp.size - 1, col.size, 8, 8, ctypes.c_void_p(cnt.data_ptr()),
                                              _lib.current_stream()), "trspmm_match_2to4_online")
    for _ in range(max(1, warmup)):
        call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        call()
    e1.record()
    torch.cuda.synchronize()
    out["ms"]["match_2to4_online"] = round(e0.elapsed_time(e1) / reps, 3)
    c = cnt.cpu().tolist()
    out["online_num_sptc_groups"] = stats["num_sptc_groups"]
    out["online_windows_over_2048_columns"] = c[9]
    assert c[0] == stats["num_sptc_groups"], "repeated online launches disagree"
    return out


def measure_config(workload, steps, warmup, mode="tuned", check=True, sample_rows=8192):
    """One product of the workload table (trspmm.graphs.WORKLOADS): routed
    multiply time, roofline fraction and relative Frobenius error against the
    FP64 CSR product (all rows, or a seeded sample of `sample_rows` rows for
    matrices of more than 2^18 rows)."""
    import torch

    from trspmm import graphs
    from trspmm.sparse import RoutedSparse
    builder, K, bseed = graphs.WORKLOADS[workload]
    csr = builder()
    B16 = graphs.dense_features(csr.n, K, seed=bseed)
    dev = torch.device("cuda", 0)
    Bt = torch.from_numpy(B16).to(dev)
    A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode=mode, plan_opts={"k_hint": K})
    C = torch.empty((csr.m, K), device=dev, dtype=torch.float32)
    ms = time_steps(lambda: A.matmul(Bt, C), steps, warmup)
    byts = algorithmic_bytes(csr.nnz, csr.m, csr.n, K)
    out = {"rows": csr.m, "nnz": csr.nnz, "K": K, "mode": mode, "ms": round(ms, 5),
           "gflops": round(2.0 * csr.nnz * K / (ms * 1e-3) / 1e9, 1),
           "frac": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "router_s": round(A.meta.seconds, 4)}
    if check:
        import scipy.sparse as sp
        rows = np.arange(csr.m)
        if csr.m > (1 << 18):
            rows = np.sort(np.random.default_rng(7).choice(csr.m, sample_rows, replace=False))
        Af = sp.csr_matrix((csr.val.astype(np.float64), csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))[rows]
        ref = np.asarray(Af @ B16.astype(np.float64))
        got = C[torch.from_numpy(rows).to(dev)].cpu().numpy().astype(np.float64)
        out["err"] = float(f"{np.linalg.norm(got - ref) / np.linalg.norm(ref):.3e}")
        out["err_rows"] = "all" if rows.size == csr.m else f"{rows.size} sampled"
    del A, C, Bt
    torch.cuda.empty_cache()
    return out


def emit_detail(detail, path):
    """The run's full tables: stderr always, `path` when given."""
    print("[bench] detail " + json.dumps(detail), file=sys.stderr, flush=True)
    if path:
        d = os.path.dirname(os.path.abspath(path))
        os.makedirs(d, exist_ok=True)
        with open(path, "w") as f:
            json.dump(detail, f, indent=1)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(one GPU's share of a sharded run is what this benchmark measures)"}))
        sys.exit(2)
    import torch

    from trspmm import graphs
    from trspmm.sparse import RoutedSparse

    if a.batched:
        rec = measure_batched(a.steps, a.warmup, not a.no_check)
        emit_detail({"batched": rec}, a.detail)
        rec.pop("autotune_table_ms", None)
        print(json.dumps(rec), flush=True)
        return
    builder, K, bseed = graphs.WORKLOADS[a.workload]
    t0 = time.time()
    csr = builder()
    gen_s = time.time() - t0
    if a.per_route:
        n = csr.m
        b1 = (int(n * 0.2) // 64) * 64
        b2 = b1 + (int(n * 0.3) // 16) * 16
        regions = {"communities": [0, b1], "banded": [b1, b2], "tail": [b2, n]} if a.workload.startswith(
            "three_route") else {}
        for mode in (["exact", "tuned"] if a.mode != "compat" else ["compat"]):
            rec = measure_routes(csr, K, mode, a.steps, a.warmup)
            rec.update(workload=a.workload, rows=csr.m, nnz_total=csr.nnz, K=K, regions_rows=regions)
            print(json.dumps(rec), flush=True)
        return
    if a.matching:
        print(json.dumps(measure_matching(csr, a.workload, a.steps, a.warmup)), flush=True)
        return
    if a.backward:
        for dt in ("fp16", "bf16"):
            print(json.dumps(measure_fwd_bwd(csr, K, dt, a.steps, a.warmup, not a.no_check, full=True)), flush=True)
        return
    if a.library:
        print(json.dumps(gpu_library_baseline(csr, graphs.dense_features(csr.n, K, seed=bseed))), flush=True)
        return
    dtype = torch.float16 if a.dtype == "fp16" else torch.bfloat16
    B16 = graphs.dense_features(csr.n, K, seed=bseed)                    # fp16 N(0,1)
    dev = torch.device("cuda", 0)
    Bt = torch.from_numpy(B16).to(dev)
    if dtype == torch.bfloat16:
        Bt = torch.from_numpy(B16.astype(np.float32)).to(dev).to(torch.bfloat16)
    print(f"[bench] workload={a.workload} m={csr.m} nnz={csr.nnz} K={K} gen={gen_s:.1f}s", file=sys.stderr,
          flush=True)
    t0 = time.time()
    # the plan's work items are sized for the K it runs at (plan option k_hint)
    A = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val), csr.n,
                     m=csr.m, mode=a.mode, dtype=dtype, plan_opts={"k_hint": K})
    torch.cuda.synchronize()
    prep_s = time.time() - t0
    print(f"[bench] routed on device in {A.meta.seconds:.3f}s (plan total {prep_s:.2f}s) stats={A.plan.stats}",
          file=sys.stderr, flush=True)
    # the same routing again: the first one of a process also pays the block
    # cache's first allocations and the router's stream creation
    from trspmm.core import route
    dev_csr = [torch.from_numpy(x).to(dev) for x in (csr.rowptr, csr.col, csr.val)]
    router_s_repeat = route(*dev_csr, csr.n, mode=a.mode, edges=True, reorder=A.reorder).seconds
    del dev_csr
    C = torch.empty((csr.m, K), device=dev, dtype=torch.float32)
    stream = torch.cuda.current_stream()
    # measured before the timed steps: the multiply after a MALL flush, and the
    # device's plain copy bandwidth (the first few dozen multiplies after routing
    # run ~5 % slower than later ones, and these runs cover that phase)
    ms_cold = time_cold(lambda: A.matmul(Bt, C), max(5, min(a.steps, 20)))
    copy_gbs = copy_bandwidth()
    for _ in range(a.warmup):
        A.matmul(Bt, C)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_wall = time.perf_counter()
    e0.record(stream)
    for _ in range(a.steps):
        A.matmul(Bt, C)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    ms = e0.elapsed_time(e1) / a.steps
    print(f"[bench] {a.steps} steps: {ms:.5f} ms per multiply", file=sys.stderr, flush=True)
    flops = 2.0 * csr.nnz * K
    gflops = flops / (ms * 1e-3) / 1e9
    byts = algorithmic_bytes(csr.nnz, csr.m, csr.n, K)
    achieved = byts / (ms * 1e-3) / 1e9

    if a.copy_calibration:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)                     # 256 MiB read + 256 MiB written per launch
        torch.cuda.synchronize()
        del src, dst

    err = None
    if not a.no_check:
        import scipy.sparse as sp
        Af = sp.csr_matrix((csr.val.astype(np.float64), csr.col, csr.rowptr), shape=(csr.m_pad, csr.n))
        ref = np.asarray(Af @ B16.astype(np.float64))[: csr.m]
        got = C.cpu().numpy().astype(np.float64)
        err = float(f"{np.linalg.norm(got - ref) / np.linalg.norm(ref):.3e}")
    detail = {"routes": A.plan.stats, "plan_opts": A.plan_opts}
    exact_ms = None
    if a.mode != "exact" and not a.no_extras:
        # the same product on the reference's tiling (exact routing), timed the same way
        A_x = RoutedSparse(torch.from_numpy(csr.rowptr), torch.from_numpy(csr.col), torch.from_numpy(csr.val),
                           csr.n, m=csr.m, mode="exact", dtype=dtype, plan_opts={"k_hint": K})
        exact_ms = round(time_steps(lambda: A_x.matmul(Bt, C), a.steps, a.warmup), 5)
        detail["exact_routes"] = A_x.plan.stats
        del A_x
    line = {
        "metric": "SpMM GFLOP/s (2·nnz·K) + %HBM roofline",
        "value": round(gflops, 2), "unit": "GFLOP/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": a.dtype, "data": "synthetic (seeded generators, trspmm.graphs)",
        "config": {"workload": a.workload, "rows": csr.m, "nnz": csr.nnz, "K": K, "mode": a.mode,
                   "rows_reordered": A.meta.reordered,
                   "rel_err_fp64": err, "router_s": round(A.meta.seconds, 4),
                   "router_s_repeat": round(router_s_repeat, 4), "exact_ms": exact_ms,
                   "ms_cold": round(ms_cold, 5), "wall_s_timed": round(wall, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_record(a.workload, a.mode, a.dtype, A.meta.reordered),
                     "algorithmic_bytes": int(byts),
                     "kernel": "k_routes", "copy_gbs": round(copy_gbs, 1)},
    }
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(csr, K, B16)
    del A, C, Bt
    torch.cuda.empty_cache()
    if not a.no_extras and a.workload == "rmat18_k128":
        print("[bench] fwd+bwd at K=256", file=sys.stderr, flush=True)
        fb = {dt: measure_fwd_bwd(csr, 256, dt, a.steps, a.warmup, not a.no_check) for dt in ("fp16", "bf16")}
        detail["fwd_bwd"] = fb
        line["fwd_bwd"] = {dt: {"ms_fwd": r["tuned"]["ms_fwd"], "ms_fwd_bwd": r["tuned"]["ms_fwd_bwd"],
                                "ratio": r["tuned"]["ratio"], "ratio_exact": r["exact"]["ratio"],
                                "err_dB": r.get("err_dB"), "err_dvals": r.get("err_dvals")}
                           for dt, r in fb.items()}
        del csr
        torch.cuda.empty_cache()
        print("[bench] batched graphs", file=sys.stderr, flush=True)
        bt = measure_batched(a.steps, a.warmup, not a.no_check)
        detail["batched"] = bt
        line["batched"] = {"ms_graph": bt["ms_batched_graph"], "ms_seq": bt["ms_sequential_plain"],
                           "speedup": bt["speedup_vs_sequential"], "frac": bt["roofline_frac_batched"],
                           "err": bt.get("rel_err_batched_vs_fp64")}
        torch.cuda.empty_cache()
        print("[bench] configs 1 and 3", file=sys.stderr, flush=True)
        cfgs = {w: measure_config(w, a.steps, a.warmup, check=not a.no_check)
                for w in ("er4096_k64", "three_route_k256")}
        detail["configs"] = cfgs
        line["configs"] = {w: {k: r[k] for k in ("ms", "frac", "err") if k in r} for w, r in cfgs.items()}
    emit_detail(detail, a.detail)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
