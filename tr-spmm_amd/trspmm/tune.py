"""Autotuner: picks routing thresholds and kernel variants for one matrix by
measuring them on the device.

The reference fixes its choices at build time (TC width 8, density 8, one
2:4 kernel).  Here the candidates are
  router    mode exact / tuned with a 2:4 occupancy floor (sptc_min_nnz),
            TC density threshold (columns with >= density rows of a window
            become dense 16x8 sub-blocks), row reordering (kept by the router
            only when it gathers fewer B rows; its plans run only on the deep
            tile kernel with the matrix-core residual engine, and the other
            kernel candidates are skipped for it)
  kernels   residual engine (vector ALU / matrix cores), tile-kernel
            features per pass (64 / 128), work-item cap (a window with more
            2:4 tiles / TC sub-blocks is split into chunks of this many), work
            item order (round robin over the XCDs, or contiguous ranges per
            XCD, which keeps a small graph's operand rows in one L2)
The window height stays 16: the MFMA/SMFMAC tiles are 16 rows tall.  Each
router configuration is routed once and every kernel variant gets its own
plan; each candidate is timed with HIP events over `iters` launches after a
warm-up.  The result names the fastest candidate and keeps the whole table.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import torch

from .core import Plan, route
from .sparse import reorder_supported

ROUTER_CANDIDATES = [
    {"mode": "exact", "density": 8},
    {"mode": "tuned", "density": 8, "sptc_min_nnz": 4},
    {"mode": "tuned", "density": 8, "sptc_min_nnz": 8},
    {"mode": "tuned", "density": 8, "sptc_min_nnz": 16},
    {"mode": "exact", "density": 6},
    {"mode": "tuned", "density": 8, "sptc_min_nnz": 8, "reorder": True},
]
KERNEL_CANDIDATES = [{"engine": e, "slab": s, "tile_cap": c, "item_order": o}
                     for e, s, c, o in itertools.product(["valu", "mfma"], [64, 128], [32, 64, 128], [0, 2, 3])]


@dataclass
class TuneResult:
    best: dict
    ms: float
    table: list = field(default_factory=list)     # (candidate, ms)
    k_hint: int = 0                                # the K the candidates were timed at

    def router_kwargs(self) -> dict:
        return {k: v for k, v in self.best.items() if k not in ("engine", "slab", "tile_cap", "item_order")}

    def plan_opts(self) -> dict:
        c = self.best.get("tile_cap", 0)
        return {"tile_cap": c, "tc_cap": c, "item_order": self.best.get("item_order", 0), "k_hint": self.k_hint}


def autotune(rowptr, col, val, ncols: int, m: int, B: torch.Tensor, *, dtype=torch.float16,
             router_candidates=None, kernel_candidates=None, warmup: int = 3, iters: int = 10) -> TuneResult:
    """Times every (router, kernel) candidate on C = A·B and returns the fastest."""
    router_candidates = ROUTER_CANDIDATES if router_candidates is None else router_candidates
    kernel_candidates = KERNEL_CANDIDATES if kernel_candidates is None else kernel_candidates
    C = torch.empty((m, B.shape[1]), device=B.device, dtype=torch.float32)
    table = []
    for rc in router_candidates:
        kw = dict(rc)
        mode = kw.pop("mode")
        meta = route(rowptr, col, val, ncols, mode=mode, **kw)
        for kc in kernel_candidates:
            cap = kc.get("tile_cap", 0)
            if meta.reordered and not reorder_supported(mode, meta.window, kc["engine"], kc):
                continue
            plan = Plan.from_meta(meta, m, ncols, dtype, engine=kc["engine"], slab=kc["slab"], tile_cap=cap, tc_cap=cap,
                                  item_order=kc.get("item_order", 0), k_hint=B.shape[1])
            ms = plan.timed(B, C, warmup=warmup, iters=iters)
            table.append(({**rc, **kc}, ms))
            del plan
        del meta
    best, ms = min(table, key=lambda t: t[1])
    return TuneResult(best, ms, table, k_hint=int(B.shape[1]))
