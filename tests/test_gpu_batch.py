"""Batched routed SpMM (one launch over many graphs), its hipGraph capture and
the autotuner.

The batch is block-diagonal with every graph padded to whole 16-row windows
and its columns shifted by a constant, which keeps the router's tie order, so
each graph is routed exactly as it is alone: the batched result must equal
the per-graph launches bit for bit, and both must be within 1e-5 of the FP64
product (values rounded to the operand type)."""
import numpy as np
import pytest
import torch

import oracle
from graphs_small import CASES
from trspmm import graphs as g
from trspmm.batch import RoutedBatch
from trspmm.sparse import RoutedSparse
from trspmm.tune import autotune

pytestmark = pytest.mark.gpu


def _batch_graphs():
    sizes = g.ba_sizes(24, lo=20, hi=600, seed=77)
    gs = [g.ba_gcn(int(n), 3, seed=i) for i, n in enumerate(sizes)]
    gs.insert(5, CASES["ragged"]())
    gs.insert(9, CASES["dense"]())
    gs.append(g.csr_from_pairs(np.zeros(0, np.int64), np.zeros(0, np.int64), None, 7, 13))
    return gs


@pytest.mark.parametrize("mode", ["exact", "tuned"])
@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_batched_equals_per_graph(cuda, mode, dtype):
    gs = _batch_graphs()
    print(f"[batch] graphs={len(gs)} mode={mode} {dtype}", flush=True)
    K = 64
    batch = RoutedBatch(gs, mode=mode, dtype=dtype)
    gen = torch.Generator().manual_seed(1)
    Bs = [torch.randn(c.n, K, generator=gen).to(dtype) for c in gs]
    Bcat = batch.stack_operands([b.to(cuda) for b in Bs])
    C = batch.matmul(Bcat)
    outs = batch.outputs(C)
    for i, c in enumerate(gs):
        # each graph routed on its own in its input order (the batch keeps the order)
        A = RoutedSparse(torch.from_numpy(c.rowptr), torch.from_numpy(c.col), torch.from_numpy(c.val), c.n, m=c.m,
                         mode=mode, dtype=dtype, reorder=False)
        Ci = A.matmul(Bs[i].to(cuda))
        assert torch.equal(outs[i], Ci), i
        vq = torch.from_numpy(c.val).to(dtype).double().numpy()
        ref = oracle.csr_spmm64(c.rowptr, c.col, vq, Bs[i].double().numpy(), c.m)
        if c.nnz:
            err = np.linalg.norm(outs[i].cpu().numpy() - ref) / max(np.linalg.norm(ref), 1e-30)
            assert err < 1e-5, (i, err)
        else:
            assert torch.count_nonzero(outs[i]) == 0


@pytest.mark.parametrize("item_order", [0, 2])
def test_batched_graph_replay(cuda, item_order):
    """Captured launch replayed with new operand contents and eager work in
    between; every replay equals the eager result.  item_order 2 (persistent
    workgroups on the plan's queue counters) checks that each launch leaves
    the queue clean for the next replay."""
    gs = _batch_graphs()
    batch = RoutedBatch(gs, plan_opts={"item_order": item_order})
    K = 128
    Bcat = torch.empty(batch.ncols, K, device=cuda, dtype=torch.float16)
    C = torch.empty(batch.rows, K, device=cuda, dtype=torch.float32)
    Bcat.normal_()
    graph = batch.capture(Bcat, C)
    for rep in range(4):
        print(f"[replay] {rep}", flush=True)
        Bcat.copy_(torch.randn(batch.ncols, K, device=cuda, generator=None).half())
        C.fill_(float("nan"))
        eager = batch.matmul(Bcat)                      # eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(C, eager), rep


def test_autotune_picks_fastest_and_is_correct(cuda):
    c = CASES["rmat10"]()
    K = 64
    B = torch.randn(c.n, K, generator=torch.Generator().manual_seed(2)).half().to(cuda)
    res = autotune(torch.from_numpy(c.rowptr), torch.from_numpy(c.col), torch.from_numpy(c.val), c.n, c.m, B,
                   iters=3, warmup=1)
    # 5 router configurations x 36 kernel variants, and the reordered routing
    # (rmat10 keeps it) with the 12 variants that run it (matrix-core residual
    # engine, item orders 0 and 3)
    from trspmm.sparse import reorder_supported
    from trspmm.tune import KERNEL_CANDIDATES
    runnable = [kc for kc in KERNEL_CANDIDATES if reorder_supported("tuned", 16, kc["engine"], kc)]
    assert len(runnable) == 12
    assert len(res.table) == 5 * 36 + 12
    assert res.ms == min(ms for _, ms in res.table)
    A = RoutedSparse(torch.from_numpy(c.rowptr), torch.from_numpy(c.col), torch.from_numpy(c.val), c.n, m=c.m,
                     engine=res.best["engine"], slab=res.best["slab"], plan_opts=res.plan_opts(),
                     **res.router_kwargs())
    got = A.matmul(B).cpu().numpy()
    vq = torch.from_numpy(c.val).half().double().numpy()
    ref = oracle.csr_spmm64(c.rowptr, c.col, vq, B.cpu().double().numpy(), c.m)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5


@pytest.mark.parametrize("engine", ["valu", "mfma"])
@pytest.mark.parametrize("slab", [64, 128])
def test_plan_options(cuda, engine, slab):
    c = CASES["hub"]()
    B = torch.randn(c.n, 128, generator=torch.Generator().manual_seed(3)).half().to(cuda)
    A = RoutedSparse(torch.from_numpy(c.rowptr), torch.from_numpy(c.col), torch.from_numpy(c.val), c.n, m=c.m,
                     engine=engine, slab=slab)
    got = A.matmul(B).cpu().numpy()
    vq = torch.from_numpy(c.val).half().double().numpy()
    ref = oracle.csr_spmm64(c.rowptr, c.col, vq, B.cpu().double().numpy(), c.m)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5


@pytest.mark.parametrize("item_order", [0, 1, 2, 3])
@pytest.mark.parametrize("cap", [1, 3, 64])
@pytest.mark.parametrize("engine", ["valu", "mfma"])
def test_work_item_options(cuda, item_order, cap, engine):
    """Small caps split every window into chunks that add atomically into
    zeroed rows; every item order; the product is unchanged, and stays so on
    a second launch of the same plan (the persistent order's queue counters
    must have been reset by the first)."""
    c = CASES["hub"]()
    B = torch.randn(c.n, 128, generator=torch.Generator().manual_seed(4)).half().to(cuda)
    A = RoutedSparse(torch.from_numpy(c.rowptr), torch.from_numpy(c.col), torch.from_numpy(c.val), c.n, m=c.m,
                     engine=engine, plan_opts={"tile_cap": cap, "tc_cap": cap, "res_cap": 16 * cap,
                                               "item_order": item_order})
    vq = torch.from_numpy(c.val).half().double().numpy()
    ref = oracle.csr_spmm64(c.rowptr, c.col, vq, B.cpu().double().numpy(), c.m)
    for rep in range(2):
        got = A.matmul(B).cpu().numpy()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5, rep


@pytest.mark.parametrize("k_hint", [64, 128, 256, 1024])
@pytest.mark.parametrize("K", [64, 256])
def test_k_hint_caps(cuda, k_hint, K):
    """Residual caps sized by k_hint (down to 512 entries per item at k_hint
    1024, so hub windows split into many atomic chunks) at a K other than the
    hint: the product is unchanged."""
    c = CASES["hub"]()
    B = torch.randn(c.n, K, generator=torch.Generator().manual_seed(6)).half().to(cuda)
    A = RoutedSparse(torch.from_numpy(c.rowptr), torch.from_numpy(c.col), torch.from_numpy(c.val), c.n, m=c.m,
                     mode="tuned", plan_opts={"k_hint": k_hint})
    vq = torch.from_numpy(c.val).half().double().numpy()
    ref = oracle.csr_spmm64(c.rowptr, c.col, vq, B.cpu().double().numpy(), c.m)
    got = A.matmul(B).cpu().numpy()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5


def test_persistent_queue_many_items_per_workgroup(cuda):
    """More work items than the chip holds workgroups (5000 windows, 80000
    rows), so each persistent workgroup takes many items from the queue; five
    launches in a row on one plan each equal the one-workgroup-per-item
    launch bit for bit (6 random columns per row: no window reaches a split
    cap, so every row has one writer and no atomics reorder sums)."""
    rng = np.random.default_rng(21)
    n = 80000
    c = g.csr_from_pairs(np.repeat(np.arange(n), 6), rng.integers(0, n, 6 * n), rng.standard_normal(6 * n), n, n)
    B = torch.randn(c.n, 64, generator=torch.Generator().manual_seed(5)).half().to(cuda)
    args = (torch.from_numpy(c.rowptr), torch.from_numpy(c.col), torch.from_numpy(c.val), c.n)
    # the persistent queue runs the one-sub-step-ahead tile kernel: compare it with
    # that kernel's one-workgroup-per-item launch (the deep kernel sums in another order)
    A0 = RoutedSparse(*args, m=c.m, plan_opts={"item_order": 0, "tile_kernel": 1})
    A2 = RoutedSparse(*args, m=c.m, plan_opts={"item_order": 2})
    A3 = RoutedSparse(*args, m=c.m, plan_opts={"item_order": 3, "tile_kernel": 1})
    want = A3.matmul(B)                                # window order: heaviest-first sorting changes nothing
    assert torch.equal(A0.matmul(B), want)
    for rep in range(5):
        print(f"[persistent] launch {rep}", flush=True)
        assert torch.equal(A2.matmul(B), want), rep
    # the deep kernel (the default) is likewise independent of the item order
    D0 = RoutedSparse(*args, m=c.m, plan_opts={"item_order": 0})
    D3 = RoutedSparse(*args, m=c.m, plan_opts={"item_order": 3})
    assert torch.equal(D0.matmul(B), D3.matmul(B))
