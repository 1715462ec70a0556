"""Drop-in for spmm/tr_csr_fp16/tr_paths.py: dataset and result paths
($TR_ROOT, $TR_DATA_ROOT, $TR_DATA_CSV, $TR_RESULTS_DIR) and the graph npz
loader."""
from __future__ import annotations

import os
from pathlib import Path

from trspmm.datasets import data_root, dataset_path, load_graph_npz, project_root

__all__ = ["project_root", "data_root", "dataset_path", "load_graph_npz", "index_csv", "results_root",
           "result_path"]


def index_csv(default_name: str = "data_filter.csv") -> Path:
    """$TR_DATA_CSV, else the first of <root>/<default_name>, data_filter.csv,
    data.csv that exists (else the first), as tr_paths.py:33-43."""
    explicit = os.environ.get("TR_DATA_CSV")
    if explicit:
        return Path(explicit).resolve()
    root = project_root()
    candidates = [root / default_name, root / "data_filter.csv", root / "data.csv"]
    for path in candidates:
        if path.exists():
            return path.resolve()
    return candidates[0].resolve()


def results_root() -> Path:
    """$TR_RESULTS_DIR, else <root>/res; created if missing (tr_paths.py:46-49)."""
    root = Path(os.environ.get("TR_RESULTS_DIR", project_root() / "res")).resolve()
    root.mkdir(parents=True, exist_ok=True)
    return root


def result_path(*parts: str) -> Path:
    path = results_root().joinpath(*parts)
    path.parent.mkdir(parents=True, exist_ok=True)
    return path
