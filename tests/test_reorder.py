"""TCA row reordering (trspmm.reorder; the reference's tr_tca_reorder.py).

The reference's permutation comes from datasketch's SHA-1 MinHash, which is
not available here, so the reordering is checked by what must hold for any
correct TCA pass:
a valid permutation, every row carried over intact with ascending columns,
nnz preserved; and by its purpose, on a matrix with planted clusters whose
rows were shuffled: after reordering, 16-row windows hold far fewer
distinct columns.  The MinHash oracle (oracle/minhash.py) is pinned by the
murmur3 finaliser's published value and by exact MinHash properties
(min over a union, equal sets, the Jaccard estimate).  GPU: the device
signatures equal the oracle's bit for bit, and the dataset drop-in routes
and multiplies a TCA-reordered graph."""
import os
import sys

import numpy as np
import pytest

from trspmm import graphs as g
from trspmm import reorder

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import minhash  # noqa: E402


def _planted(groups=24, size=16, pool=24, per_row=14, ncols=4000, seed=0):
    """groups x size rows; a group's rows draw per_row columns from its own pool; rows shuffled."""
    rng = np.random.default_rng(seed)
    pools = [rng.choice(ncols, size=pool, replace=False) for _ in range(groups)]
    rows = [np.sort(rng.choice(pools[gi], size=per_row, replace=False)) for gi in range(groups) for _ in range(size)]
    perm = rng.permutation(len(rows))
    rows = [rows[p] for p in perm]
    rp = np.zeros(len(rows) + 1, np.int32)
    rp[1:] = np.cumsum([r.size for r in rows])
    return rp, np.concatenate(rows).astype(np.int32), ncols


def _distinct_per_window(rp, ci, w=16):
    return np.mean([np.unique(ci[rp[r]:rp[min(r + w, rp.size - 1)]]).size for r in range(0, rp.size - 1, w)])


def _check_permuted(rp, ci, new_rp, new_ci, order):
    n = rp.size - 1
    assert sorted(order.tolist()) == list(range(n))
    assert new_rp[-1] == rp[-1] and new_ci.size == ci.size
    for r in range(n):
        got = new_ci[new_rp[r]:new_rp[r + 1]]
        assert np.all(np.diff(got) > 0)
        assert np.array_equal(got, np.sort(ci[rp[order[r]]:rp[order[r] + 1]]))


# ---------------------------------------------------------------- oracle pinning (CPU)

def test_oracle_mix32_is_murmur3_fmix32():
    assert minhash.mix32(0) == 0
    assert minhash.mix32(1) == 0x514E28B7          # the published fmix32(1)


def test_oracle_minhash_properties():
    co = reorder.minhash_coefficients(64, seed=5)
    a, b = [3, 17, 99, 1000], [17, 99, 4000, 7]
    rp = np.array([0, 4, 8, 16, 16], np.int32)
    ci = np.array(a + sorted(b) + sorted(set(a) | set(b)), np.int32)
    s = minhash.signatures(rp, ci, co)
    assert np.array_equal(s[2], np.minimum(s[0], s[1]))          # union = elementwise min
    assert np.all(s[3] == 0xFFFFFFFF)                             # empty row
    s2 = minhash.signatures(np.array([0, 4], np.int32), np.array(a[::-1], np.int32), co)
    assert np.array_equal(s2[0], s[0])                            # a set, not a sequence


def test_oracle_minhash_estimates_jaccard():
    rng = np.random.default_rng(1)
    co = reorder.minhash_coefficients(256, seed=11)
    base = rng.choice(100000, size=60, replace=False)
    other = np.concatenate([base[:30], rng.choice(np.arange(100000, 200000), size=30, replace=False)])
    rp = np.array([0, 60, 120], np.int32)
    s = minhash.signatures(rp, np.concatenate([np.sort(base), np.sort(other)]).astype(np.int32), co)
    jac = 30 / 90
    est = np.mean(s[0] == s[1])
    assert abs(est - jac) < 0.1, (est, jac)


# ---------------------------------------------------------------- clustering (host, oracle signatures)

def test_tca_order_invariants_and_quality_host():
    rp, ci, _ = _planted()
    sig = minhash.signatures(rp, ci, reorder.minhash_coefficients(128, 2022))
    new_rp, new_ci, order = reorder.apply_tca_reorder(rp, ci, rp.size - 1, verbose=False, signatures=sig)
    _check_permuted(rp, ci, new_rp, new_ci, order)
    before, after = _distinct_per_window(rp, ci), _distinct_per_window(new_rp, new_ci)
    assert after <= 0.5 * before, (before, after)
    assert after <= 30, after                       # close to one 24-column pool per window


def test_tca_order_handles_empty_rows_and_tiny_inputs():
    rp = np.array([0, 0, 2, 2, 3], np.int32)
    ci = np.array([1, 5, 5], np.int32)
    sig = minhash.signatures(rp, ci, reorder.minhash_coefficients(128, 2022))
    new_rp, new_ci, order = reorder.apply_tca_reorder(rp, ci, 4, verbose=False, signatures=sig)
    _check_permuted(rp, ci, new_rp, new_ci, order)
    sig1 = minhash.signatures(np.array([0, 1], np.int32), np.array([3], np.int32),
                              reorder.minhash_coefficients(128, 2022))
    assert reorder.tca_order(np.array([0, 1], np.int32), np.array([3], np.int32), 1, signatures=sig1).tolist() == [0]


def test_tca_order_rejects():
    from trspmm._lib import TrspmmError
    rp, ci = np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32)
    with pytest.raises(TrspmmError, match="non-decreasing"):
        reorder.tca_order(rp, ci, 2, signatures=np.zeros((2, 128), np.uint32))
    with pytest.raises(TrspmmError, match="lsh_threshold"):
        reorder.tca_order(np.array([0, 1], np.int32), np.array([0], np.int32), 1, lsh_threshold=1.5,
                          signatures=np.zeros((1, 128), np.uint32))


# ---------------------------------------------------------------- device

@pytest.mark.gpu
def test_gpu_minhash_equals_oracle(cuda):
    rng = np.random.default_rng(3)
    c = g.er_gcn(300, 12.0, seed=3)
    rp = c.rowptr[:301].copy()
    rp[5:] = rp[5:] - (rp[5] - rp[4])                 # make row 4 empty
    ci = np.delete(c.col, np.arange(c.rowptr[4], c.rowptr[5]))[: rp[-1]]
    ci = (ci.astype(np.int64) * 7919 + rng.integers(0, 2, ci.size) * (1 << 30)).astype(np.int32) & 0x7FFFFFFF
    for perm, seed in ((128, 2022), (70, 9)):
        got = reorder.minhash_signatures(rp, ci, perm, seed)
        want = minhash.signatures(rp, ci, reorder.minhash_coefficients(perm, seed))
        assert np.array_equal(got, want)


@pytest.mark.gpu
def test_gpu_apply_tca_reorder_end_to_end(cuda):
    rp, ci, _ = _planted(groups=40, seed=4)
    new_rp, new_ci, order = reorder.apply_tca_reorder(rp, ci, rp.size - 1, verbose=False)
    _check_permuted(rp, ci, new_rp, new_ci, order)
    assert _distinct_per_window(new_rp, new_ci) <= 0.5 * _distinct_per_window(rp, ci)


@pytest.mark.gpu
def test_gpu_dataset_with_tca_reorder(tmp_path, cuda, monkeypatch):
    """LIBRA_TCA_REORDER=1 in the dataset drop-in: the reordered graph routes
    and multiplies; the product equals the oracle's FP64 product of its
    routed metadata, and more nonzeros land on the dense TC route than
    without the reordering."""
    import torch

    import oracle
    import Libra6SpMMOnline
    from trspmm import datasets
    from tr_csr_fp16.mdataset2_online import _NAMES, GCN_dataset_tcu_sptc_cuda_online
    monkeypatch.delenv("TRSPMM_ROUTE_MODE", raising=False)
    rp, ci, ncols = _planted(groups=32, per_row=14, pool=16, ncols=600, seed=6)
    src = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    datasets.save_graph_npz(tmp_path / "pl.npz", src, ci, rp.size - 1, ncols)
    infos = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("LIBRA_TCA_REORDER", flag)
        torch.manual_seed(0)
        infos[flag] = GCN_dataset_tcu_sptc_cuda_online("pl", 32, 8, 32, 32, 3, str(tmp_path), 16, 16)
    info = infos["1"]
    res = Libra6SpMMOnline.forward_fp16_tcu_cuda_sptc_mma_parallel_online(
        info.t_rowNew_offsetTensor, info.t_blockTensor, info.t_columnTensor, info.t_valueTensor,
        info.t_window_rowTensor, info.t_atomicTensor, info.t_binaryTensor,
        info.c_row_offsetTensor, info.c_rowTensor, info.c_atomicTensor, info.c_colTensor, info.c_valueTensor,
        info.c_row_offsetTensor_short, info.c_rowTensor_short, info.c_atomicTensor_short, info.c_colTensor_short,
        info.c_valueTensor_short,
        info.s_columnTensor, info.s_valueTensor, info.s_posTensor, info.s_offsetTensor, info.s_tile_windowTensor,
        info.s_tile_columnTensor, info.s_packed_aTensor, info.s_packed_metaTensor, info.s_window_tile_offsetTensor,
        info.x, info.parts_t, info.parts_c, info.partsize_c, info.parts_c_short, info.window, info.x.size(1),
        info.num_nodes_ori, info.num_nodes_dst, 2)
    arr = {n: (getattr(info, a).float().numpy() if getattr(info, a).dtype == torch.float16 else getattr(info, a).numpy())
           for n, a in zip(oracle.NAMES, _NAMES[:29])}
    ref = oracle.spmm64(arr, info.x.float().numpy(), info.num_nodes_ori, 16)
    assert np.linalg.norm(res[0].numpy() - ref) / np.linalg.norm(ref) < 1e-5
    assert infos["1"].t_valueTensor.numel() > 2 * infos["0"].t_valueTensor.numel()
