"""Graph input formats and loaders (trspmm.datasets, the tr_csr_fp16 drop-in).

The reference holds no dataset files, so the loaders are checked on files
written here in each format the reference reads: the graph npz
(num_nodes_src / num_nodes_dst / src_li / dst_li, tr_paths.py:22-30,
mdataset2_online.py:57-70), scipy's npz, the src_li/dst_li[/data]/shape npz
and MatrixMarket files (test_24_matching.py:88-146), against matrices written
out by hand.  The GPU test runs the reference driver's flow
(test_tr_csr_online.py:199-214, 83-128) on a loaded graph through the
drop-in dataset class and Libra6SpMMOnline, against the oracle's FP64
product of the routed metadata."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from trspmm import datasets


def _graph(tmp_path, name="g", window=16):
    # 37 source nodes, 50 destination nodes, one duplicate edge (3, 7) and a self loop
    src = np.array([0, 0, 3, 3, 3, 36, 20, 20, 11])
    dst = np.array([1, 49, 7, 7, 8, 0, 20, 21, 30])
    datasets.save_graph_npz(tmp_path / f"{name}.npz", src, dst, 37, 50)
    return src, dst


def test_graph_npz_round_trip_and_csr(tmp_path):
    _graph(tmp_path)
    g = datasets.load_graph_npz(str(tmp_path), "g")
    assert sorted(g.keys()) == sorted(datasets.GRAPH_KEYS)
    csr = datasets.graph_csr(g, window=16)
    assert (csr.m, csr.n, csr.m_pad) == (37, 50, 48)
    dense = np.zeros((48, 50), np.float32)
    for r, c in [(0, 1), (0, 49), (3, 7), (3, 7), (3, 8), (36, 0), (20, 20), (20, 21), (11, 30)]:
        dense[r, c] += 1.0
    got = sp.csr_matrix((csr.val, csr.col, csr.rowptr), shape=(48, 50)).toarray()
    assert np.array_equal(got, dense)                      # duplicate edge summed to 2
    for r in range(48):
        cols = csr.col[csr.rowptr[r]:csr.rowptr[r + 1]]
        assert np.all(np.diff(cols) > 0)                   # canonical: strictly ascending
    assert datasets.graph_csr(g, window=8).m_pad == 40


def test_relative_data_path_uses_tr_data_root(tmp_path, monkeypatch):
    (tmp_path / "sub").mkdir()
    _graph(tmp_path / "sub", "h")
    monkeypatch.setenv("TR_DATA_ROOT", str(tmp_path))
    assert datasets.dataset_path("sub", "h") == (tmp_path / "sub" / "h.npz").resolve()
    assert int(datasets.load_graph_npz("sub", "h")["num_nodes_dst"]) == 50
    with pytest.raises(FileNotFoundError, match="nothere"):
        datasets.load_graph_npz("sub", "nothere")


def test_edges_outside_the_shape_raise(tmp_path):
    # rows are checked against the padded count (48), as the reference's coo_matrix shape
    datasets.save_graph_npz(tmp_path / "bad.npz", [0, 60], [1, 2], 37, 50)
    with pytest.raises(ValueError):
        datasets.graph_csr(datasets.load_graph_npz(str(tmp_path), "bad"))
    datasets.save_graph_npz(tmp_path / "bad2.npz", [0, 1], [1, 50], 37, 50)
    with pytest.raises(ValueError):
        datasets.graph_csr(datasets.load_graph_npz(str(tmp_path), "bad2"))


MTX_PATTERN = """%%MatrixMarket matrix coordinate pattern general
3 4 3
1 1
2 4
3 2
"""
MTX_SYMMETRIC = """%%MatrixMarket matrix coordinate real symmetric
3 3 3
1 1 2.5
3 1 -1.0
3 2 4.0
"""
MTX_INTEGER = """%%MatrixMarket matrix coordinate integer general
2 2 2
1 2 7
2 1 -3
"""
MTX_ARRAY = """%%MatrixMarket matrix array real general
2 2
1.0
0.0
0.0
5.0
"""


@pytest.mark.parametrize("text, want", [
    (MTX_PATTERN, [[1, 0, 0, 0], [0, 0, 0, 1], [0, 1, 0, 0]]),
    (MTX_SYMMETRIC, [[2.5, 0, -1.0], [0, 0, 4.0], [-1.0, 4.0, 0]]),
    (MTX_INTEGER, [[0, 7], [-3, 0]]),
    (MTX_ARRAY, [[1.0, 0.0], [0.0, 5.0]]),           # column-major array: a11, a21, a12, a22
], ids=["pattern", "symmetric", "integer", "array"])
def test_load_matrix_mtx(tmp_path, text, want):
    p = tmp_path / "m.mtx"
    p.write_text(text)
    mat = datasets.load_matrix(str(p))
    assert sp.isspmatrix_csr(mat) and mat.dtype == np.float32 and mat.has_sorted_indices
    assert np.array_equal(mat.toarray(), np.array(want, np.float32))


def test_load_matrix_npz_formats(tmp_path):
    want = np.array([[0, 2, 0], [1, 0, 3]], np.float32)
    sp.save_npz(tmp_path / "s.npz", sp.csr_matrix(want))
    assert np.array_equal(datasets.load_matrix(str(tmp_path / "s.npz")).toarray(), want)
    np.savez(tmp_path / "d.npz", src_li=np.array([0, 1, 1]), dst_li=np.array([1, 0, 2]),
             data=np.array([2.0, 1.0, 3.0]), shape=np.array([2, 3]))
    assert np.array_equal(datasets.load_matrix(str(tmp_path / "d.npz")).toarray(), want)
    np.savez(tmp_path / "u.npz", src_li=np.array([0, 1, 1]), dst_li=np.array([1, 0, 2]), shape=np.array([2, 3]))
    assert np.array_equal(datasets.load_matrix(str(tmp_path / "u.npz")).toarray(), (want > 0).astype(np.float32))
    with pytest.raises(ValueError, match="unsupported"):
        datasets.load_matrix(str(tmp_path / "x.csv"))


def test_matrix_arrays_feed_the_matching_ops(tmp_path):
    """The reference test's flow: load_matrix -> matrix_to_csr_arrays ->
    match_2to4, with its nnz check (test_24_matching.py:177-180)."""
    import matching_utils
    rng = np.random.default_rng(5)
    m = sp.random(200, 300, density=0.05, random_state=rng, format="coo", dtype=np.float32)
    sp.save_npz(tmp_path / "r.npz", m.tocsr())
    mat = datasets.load_matrix(str(tmp_path / "r.npz"))
    rp, ci, v, rows, cols = datasets.matrix_to_csr_arrays(mat)
    assert (rp.dtype, ci.dtype, v.dtype, rows, cols) == (np.int32, np.int32, np.float32, 200, 300)
    s = matching_utils.match_2to4(rp, ci, v)
    assert s["total_nnz"] == mat.nnz and s["total_rows"] == 200
    csr = datasets.matrix_csr(mat)
    assert csr.m_pad == 208 and csr.rowptr[-1] == mat.nnz and (csr.m, csr.n) == (200, 300)


def test_tr_paths_helpers(tmp_path, monkeypatch):
    from tr_csr_fp16 import tr_paths
    monkeypatch.setenv("TR_RESULTS_DIR", str(tmp_path / "res"))
    assert tr_paths.result_path("a", "b.csv") == (tmp_path / "res" / "a" / "b.csv").resolve()
    assert (tmp_path / "res" / "a").is_dir()
    monkeypatch.setenv("TR_DATA_CSV", str(tmp_path / "list.csv"))
    assert tr_paths.index_csv() == (tmp_path / "list.csv").resolve()


@pytest.mark.gpu
@pytest.mark.parametrize("dimN", [64, 32])
def test_gpu_reference_driver_flow_on_a_loaded_graph(tmp_path, cuda, dimN, monkeypatch):
    """GCN_dataset_tcu_sptc_cuda_online from a graph npz, then
    Libra6SpMMOnline's parallel op with its attributes, as the reference's
    driver calls it; checked against the oracle's FP64 product of the routed
    metadata (values rounded to fp16 as the dataset hands them over)."""
    import oracle
    import Libra6SpMMOnline
    from tr_csr_fp16.mdataset2_online import GCN_dataset_tcu_sptc_cuda_online
    monkeypatch.delenv("TRSPMM_ROUTE_MODE", raising=False)
    monkeypatch.delenv("LIBRA_TCA_REORDER", raising=False)
    rng = np.random.default_rng(9)
    n_src, n_dst = 300, 280
    src = np.concatenate([rng.integers(0, n_src, 3000), np.repeat(np.arange(16), 40)])
    dst = np.concatenate([rng.integers(0, n_dst, 3000), np.tile(np.arange(40), 16)])
    datasets.save_graph_npz(tmp_path / "gr.npz", src, dst, n_src, n_dst)
    torch.manual_seed(0)
    info = GCN_dataset_tcu_sptc_cuda_online("gr", dimN, 8, 32, 32, 3, str(tmp_path), 16, 16)
    assert info.num_nodes == 304 and info.x.shape == (n_dst, dimN) and info.x.dtype == torch.float16
    assert info.t_valueTensor.dtype == torch.float16 and info.num_edges == info.column_index.numel()
    res = Libra6SpMMOnline.forward_fp16_tcu_cuda_sptc_mma_parallel_online(
        info.t_rowNew_offsetTensor, info.t_blockTensor, info.t_columnTensor, info.t_valueTensor,
        info.t_window_rowTensor, info.t_atomicTensor, info.t_binaryTensor,
        info.c_row_offsetTensor, info.c_rowTensor, info.c_atomicTensor, info.c_colTensor, info.c_valueTensor,
        info.c_row_offsetTensor_short, info.c_rowTensor_short, info.c_atomicTensor_short, info.c_colTensor_short,
        info.c_valueTensor_short,
        info.s_columnTensor, info.s_valueTensor, info.s_posTensor, info.s_offsetTensor, info.s_tile_windowTensor,
        info.s_tile_columnTensor, info.s_packed_aTensor, info.s_packed_metaTensor, info.s_window_tile_offsetTensor,
        info.x, info.parts_t, info.parts_c, info.partsize_c, info.parts_c_short, info.window, info.x.size(1),
        info.num_nodes_ori, info.num_nodes_dst, 3)
    out = res[0]
    assert tuple(out.shape) == (n_src, dimN)
    from tr_csr_fp16.mdataset2_online import _NAMES          # attribute names, in the router's return order
    arr = {}
    for n, attr in zip(oracle.NAMES, _NAMES[:29]):
        t = getattr(info, attr)
        arr[n] = t.float().numpy() if t.dtype == torch.float16 else t.numpy()
    ref = oracle.spmm64(arr, info.x.float().numpy(), n_src, 16)
    err = np.linalg.norm(out.numpy() - ref) / np.linalg.norm(ref)
    assert err < 1e-5, err
