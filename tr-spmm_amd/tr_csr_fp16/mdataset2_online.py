"""Drop-in for spmm/tr_csr_fp16/mdataset2_online.py: the online dataset the
reference's driver (test_tr_csr_online.py) routes and multiplies.

    GCN_dataset_tcu_sptc_cuda_online(data, dimN, density, partsize_t, partsize_c, shortsize,
                                     data_path, window, wide, tc_min_util=0.0,
                                     sptc_min_util=0.50, sptc_threshold=12)

Same constructor and attributes (mdataset2_online.py:29-169): the graph npz
is loaded (trspmm.datasets.load_graph_npz), turned into the padded CSR with
unit edge values (graph_csr), optionally row-reordered (LIBRA_TCA_REORDER=1,
trspmm.reorder.apply_tca_reorder), routed by this repository's
Libra5BlockOnline with random normal values (`degrees`, as the reference),
and the 29 metadata tensors, part counts and fp16 value tensors are set; `x`
is a random normal fp16 feature matrix [num_nodes_dst, dimN].  The reference
also calls Rabbit.reorder when that module is importable but keeps the CSR
it built before (mdataset2_online.py:90); Rabbit is not part of this build,
so edge_index stays as loaded.
"""
from __future__ import annotations

import os

import numpy as np
import torch

import Libra5BlockOnline
from trspmm.datasets import graph_csr, load_graph_npz

_NAMES = ["t_rowNew_offsetTensor", "t_blockTensor", "t_columnTensor", "t_valueTensor", "t_window_rowTensor",
          "t_atomicTensor", "t_binaryTensor", "c_row_offsetTensor", "c_rowTensor", "c_atomicTensor", "c_colTensor",
          "c_valueTensor", "c_row_offsetTensor_short", "c_rowTensor_short", "c_atomicTensor_short",
          "c_colTensor_short", "c_valueTensor_short", "s_columnTensor", "s_maskTensor", "s_valueTensor",
          "s_posTensor", "s_windowTensor", "s_offsetTensor", "s_tile_windowTensor", "s_tile_groupTensor",
          "s_tile_columnTensor", "s_packed_aTensor", "s_packed_metaTensor", "s_window_tile_offsetTensor",
          "duration"]


class GCN_dataset_tcu_sptc_cuda_online(torch.nn.Module):
    def __init__(self, data, dimN, density, partsize_t, partsize_c, shortsize, data_path, window, wide,
                 tc_min_util=0.0, sptc_min_util=0.50, sptc_threshold=12):
        super().__init__()
        self.graph = load_graph_npz(data_path, data)
        self.num_features = dimN
        self.window = window
        self.init_edges(density, partsize_t, partsize_c, shortsize, window, wide, tc_min_util, sptc_min_util,
                        sptc_threshold)
        self.init_embedding()

    def init_edges(self, density, partsize_t, partsize_c, shortsize, window, wide, tc_min_util, sptc_min_util,
                   sptc_threshold):
        self.num_nodes_ori = int(self.graph["num_nodes_src"])
        self.num_nodes_dst = int(self.graph["num_nodes_dst"])
        csr = graph_csr(self.graph, window)
        self.num_nodes = csr.m_pad
        self.edge_index = np.stack([np.asarray(self.graph["src_li"]), np.asarray(self.graph["dst_li"])])
        rowptr, col = csr.rowptr, csr.col
        if os.environ.get("LIBRA_TCA_REORDER", "0") == "1":
            from trspmm.reorder import apply_tca_reorder
            rp, ci, _ = apply_tca_reorder(rowptr[: self.num_nodes_ori + 1], col, self.num_nodes_ori, thres=16,
                                          verbose=True)
            rowptr = np.concatenate([rp, np.full(self.num_nodes - self.num_nodes_ori, rp[-1], rp.dtype)])
            col = ci
            src = np.repeat(np.arange(self.num_nodes_ori, dtype=np.int64), np.diff(rp))
            self.edge_index = np.stack([src, ci.astype(np.int64)])
        else:
            print("TCA reorder disabled (LIBRA_TCA_REORDER=0)")
        self.column_index = torch.IntTensor(np.ascontiguousarray(col, dtype=np.int32))
        self.row_pointers = torch.IntTensor(np.ascontiguousarray(rowptr, dtype=np.int32))
        self.num_edges = self.column_index.numel()
        self.degrees = torch.randn(self.num_edges)

        tensors = Libra5BlockOnline.block_sptc_2to4_online(
            self.row_pointers, self.column_index, self.degrees, partsize_t, partsize_c, shortsize, density, window,
            wide, float(tc_min_util), float(sptc_min_util), int(sptc_threshold))
        for name, t in zip(_NAMES, tensors):
            setattr(self, name, t)

        self.parts_t = self.t_atomicTensor.shape[0]
        self.parts_c = self.c_atomicTensor.shape[0]
        self.parts_c_short = self.c_atomicTensor_short.shape[0]
        self.partsize_c = partsize_c

        self.t_valueTensor = self.t_valueTensor.half()
        self.c_valueTensor = self.c_valueTensor.half()
        self.c_valueTensor_short = self.c_valueTensor_short.half()

        tcu_nnz = self.t_valueTensor.numel()
        sptc_nnz = int((self.s_posTensor >= 0).sum().item())
        cuda_nnz = self.c_valueTensor.numel() + self.c_valueTensor_short.numel()
        print(f"tcu: {tcu_nnz};   sptc: {sptc_nnz};   cuda: {cuda_nnz}")
        print(f"parts_t: {self.parts_t};   parts_c: {self.parts_c};   parts_c_short: {self.parts_c_short};   "
              f"sptc_groups: {self.s_columnTensor.shape[0]};   sptc_tiles: {self.s_tile_windowTensor.shape[0]}")

    def init_embedding(self):
        self.x = torch.randn(self.num_nodes_dst, self.num_features).half()
