"""Drop-in for the reference's caller package spmm/tr_csr_fp16 (dataset
paths, the graph npz loader and the online dataset class), built on this
repository's Libra5BlockOnline and trspmm.datasets."""
