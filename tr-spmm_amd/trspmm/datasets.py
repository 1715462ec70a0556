"""Graph input formats and loaders of the reference's callers.

  dataset_path, load_graph_npz   the graph npz of spmm/tr_csr_fp16/tr_paths.py:7-30:
                                 <data_path>/<data>.npz, a relative data_path
                                 resolved under $TR_DATA_ROOT (default
                                 $TR_ROOT/dgl_dataset); keys num_nodes_src,
                                 num_nodes_dst, src_li, dst_li
  graph_csr                      the caller's CSR of such a graph
                                 (mdataset2_online.py:57-70): rows padded to a
                                 multiple of the window, every edge value 1,
                                 duplicate edges summed (scipy coo -> csr)
  save_graph_npz                 writes that format
  load_matrix                    TR-source/test_24_matching.py:88-146: .npz in
                                 scipy's format or the src_li/dst_li[/data]/shape
                                 format, .mtx (MatrixMarket: pattern entries
                                 become 1, dense arrays are converted), as a
                                 float32 scipy CSR with sorted indices
  matrix_to_csr_arrays           test_24_matching.py:149-163

Loaded matrices come back as trspmm.graphs.CSR where the routed multiply
takes them (canonical CSR: indices sorted, no duplicates).
"""
from __future__ import annotations

import os
from pathlib import Path

import numpy as np
import scipy.sparse as sp

from .graphs import CSR, _pad_rows

GRAPH_KEYS = ("num_nodes_src", "num_nodes_dst", "src_li", "dst_li")


def project_root() -> Path:
    """$TR_ROOT, else the directory above this package (tr_paths.py:7-8)."""
    return Path(os.environ.get("TR_ROOT", Path(__file__).resolve().parents[2])).resolve()


def data_root() -> Path:
    """$TR_DATA_ROOT, else <project_root>/dgl_dataset (tr_paths.py:11-12)."""
    return Path(os.environ.get("TR_DATA_ROOT", project_root() / "dgl_dataset")).resolve()


def dataset_path(data_path: str, data: str) -> Path:
    """<data_path>/<data>.npz; a relative data_path is taken under data_root() (tr_paths.py:15-19)."""
    base = Path(data_path)
    if not base.is_absolute():
        base = data_root() / data_path
    return (base / f"{data}.npz").resolve()


def load_graph_npz(data_path: str, data: str):
    """The graph npz (tr_paths.py:22-30); FileNotFoundError naming the path
    when it is missing.  Loaded without pickles."""
    path = dataset_path(data_path, data)
    if not path.exists():
        raise FileNotFoundError(f"Cannot find tr dataset '{data}' at {path}. Set TR_DATA_ROOT or pass data_path.")
    return np.load(path, allow_pickle=False)


def save_graph_npz(path, src, dst, num_nodes_src: int, num_nodes_dst: int) -> None:
    """Writes the graph npz format load_graph_npz reads."""
    np.savez(path, num_nodes_src=np.int64(num_nodes_src), num_nodes_dst=np.int64(num_nodes_dst),
             src_li=np.asarray(src, np.int64), dst_li=np.asarray(dst, np.int64))


def graph_csr(graph, window: int = 16) -> CSR:
    """The caller's CSR of a graph npz (mdataset2_online.py:57-70): rows
    num_nodes_src padded up to a multiple of `window`, columns num_nodes_dst,
    value 1 per edge, duplicate edges summed.  Edges outside the shape raise
    ValueError, as scipy does for the reference."""
    for k in GRAPH_KEYS:
        if k not in graph:
            raise KeyError(f"graph npz has no '{k}' (keys: {sorted(graph.keys())})")
    m = int(graph["num_nodes_src"])
    n = int(graph["num_nodes_dst"])
    src = np.asarray(graph["src_li"], np.int64)
    dst = np.asarray(graph["dst_li"], np.int64)
    if src.shape != dst.shape:
        raise ValueError("src_li and dst_li differ in length")
    m_pad = _pad_rows(m, window)
    adj = sp.coo_matrix((np.ones(src.size, np.float32), (src, dst)), shape=(m_pad, n)).tocsr()
    adj.sum_duplicates()
    adj.sort_indices()
    if adj.nnz > np.iinfo(np.int32).max:
        raise ValueError("nnz exceeds int32 range")
    return CSR(adj.indptr.astype(np.int32), adj.indices.astype(np.int32), adj.data.astype(np.float32), m, n)


def _load_npz_matrix(path):
    raw = np.load(path, allow_pickle=False)
    keys = list(raw.keys())
    if "format" in keys or "indices" in keys:
        return sp.load_npz(path)
    src, dst = raw["src_li"], raw["dst_li"]
    val = raw["data"].astype(np.float32) if "data" in keys else np.ones(len(src), np.float32)
    shape = tuple(int(x) for x in raw["shape"].tolist())
    return sp.coo_matrix((val, (src, dst)), shape=shape).tocsr()


def load_matrix(path) -> sp.csr_matrix:
    """float32 scipy CSR with sorted indices from .npz or .mtx
    (test_24_matching.py:106-146)."""
    from scipy.io import mminfo, mmread
    ext = os.path.splitext(str(path))[1].lower()
    if ext == ".npz":
        mat = _load_npz_matrix(path)
    elif ext == ".mtx":
        field = mminfo(path)[4]
        mat = mmread(path)
        if field == "pattern" and hasattr(mat, "data"):
            mat.data = np.ones(mat.nnz, dtype=np.float32)
        if isinstance(mat, np.ndarray):
            mat = sp.csr_matrix(mat)
    else:
        raise ValueError(f"unsupported matrix format: {ext}")
    if not sp.isspmatrix_csr(mat):
        mat = mat.tocsr()
    mat.sort_indices()
    if mat.dtype != np.float32:
        mat = mat.astype(np.float32)
    return mat


def matrix_to_csr_arrays(mat):
    """(row_ptr int32, col_ind int32, values float32, rows, cols) (test_24_matching.py:149-163)."""
    rows, cols = mat.shape
    return mat.indptr.astype(np.int32), mat.indices.astype(np.int32), mat.data.astype(np.float32), rows, cols


def matrix_csr(mat, window: int = 16) -> CSR:
    """A scipy matrix as the routed multiply's canonical CSR, rows padded to a
    multiple of `window` (duplicates summed, indices sorted)."""
    mat = sp.csr_matrix(mat, dtype=np.float32)
    mat.sum_duplicates()
    mat.sort_indices()
    m, n = mat.shape
    m_pad = _pad_rows(m, window)
    rowptr = np.concatenate([mat.indptr, np.full(m_pad - m, mat.indptr[-1], mat.indptr.dtype)])
    return CSR(rowptr.astype(np.int32), mat.indices.astype(np.int32), mat.data.astype(np.float32), m, n)
