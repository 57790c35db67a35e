"""On-disk formats around the path.

* Graph npz (``src_li``, ``dst_li``, ``num_nodes``): read as the reference's
  loader does (scripts/DTCSpMM/dataset.py:21-44): an all-ones COO matrix
  converted to CSR, so duplicate edges are summed into one entry and column
  indices come out sorted per row.  ``load_graph_npz`` returns that CSR
  pattern as int32 plus the summed counts.
* Reorder npz: ``<name>.reorder_id.npz`` holds ``reorder_id``, the list of
  original node ids in their new order (reordering/TCA_reorder.py:317); the
  reordered graph relabels both ends of every edge by its position in that
  list (:318-326).  ``apply_reorder`` does the same relabelling.
* ME-TCF cache: ``save_metcf`` / ``load_metcf`` store the arrays of a
  ``DTCMatrix`` with ``torch.save`` and read them back with
  ``torch.load(weights_only=True)`` (tensors only, nothing executed).
"""
from __future__ import annotations

import numpy as np

from .graphs import CSR

METCF_FIELDS = ["rowptr", "col", "row_window_offset", "tcblock_rowid", "tcblocktile_id",
                "tcblock_offset", "sparse_atox", "tc_edge", "block_partition", "edge_to_column",
                "edge_to_row"]


def load_graph_npz(path: str) -> CSR:
    """CSR of an npz edge list with duplicates summed (values = counts)."""
    import scipy.sparse as sp
    if not path.endswith(".npz"):
        raise ValueError("graph file must be a .npz file")
    z = np.load(path)  # allow_pickle stays False
    src, dst = np.asarray(z["src_li"]), np.asarray(z["dst_li"])
    n = int(z["num_nodes"])
    coo = sp.coo_matrix((np.ones(src.shape[0], np.float32), (src, dst)), shape=(n, n))
    csr = coo.tocsr()
    csr.sort_indices()
    return CSR(csr.indptr.astype(np.int32), csr.indices.astype(np.int32),
               csr.data.astype(np.float32), n, n)


def save_graph_npz(path: str, src, dst, num_nodes: int) -> None:
    np.savez(path, src_li=np.asarray(src), dst_li=np.asarray(dst), num_nodes=int(num_nodes))


def load_reorder_id(path: str) -> np.ndarray:
    return np.asarray(np.load(path)["reorder_id"], np.int64)


def save_reorder_id(path: str, reorder_id) -> None:
    np.savez(path, reorder_id=np.asarray(reorder_id, np.int64))


def save_reordered_graph(path: str, src, dst, num_nodes: int, reorder_id) -> None:
    """``<dataset>.reorder.npz``: the edge list with both ends relabelled
    (TCA_reorder.py:318-326)."""
    s, d = apply_reorder(src, dst, reorder_id)
    save_graph_npz(path, s, d, num_nodes)


def apply_reorder(src, dst, reorder_id):
    """Relabel edge endpoints: node reorder_id[i] becomes node i."""
    reorder_id = np.asarray(reorder_id, np.int64)
    pos = np.empty(reorder_id.shape[0], np.int64)
    pos[reorder_id] = np.arange(reorder_id.shape[0])
    return pos[np.asarray(src, np.int64)], pos[np.asarray(dst, np.int64)]


def save_metcf(path: str, A) -> None:
    import torch
    blob = {k: getattr(A, k).cpu() for k in METCF_FIELDS}
    blob["meta"] = torch.tensor([A.num_rows, A.num_cols, A.nnz, A.block_count], dtype=torch.int64)
    if A.values is not None:
        blob["values"] = A.values.cpu()
    torch.save(blob, path)


def load_metcf(path: str, device="cuda"):
    """A ``DTCMatrix`` rebuilt from a cache file without re-running the build."""
    import torch

    from .matrix import DTCMatrix
    blob = torch.load(path, map_location="cpu", weights_only=True)
    A = DTCMatrix.__new__(DTCMatrix)
    A.device = torch.device(device)
    for k in METCF_FIELDS:
        setattr(A, k, blob[k].to(A.device))
    A.num_rows, A.num_cols, A.nnz, A.block_count = (int(x) for x in blob["meta"].tolist())
    A.num_windows = A.row_window_offset.numel() - 1
    A.values = blob["values"].to(A.device) if "values" in blob else None
    A.values_tc = None
    if A.values is not None:
        from . import load_extension
        A.values_tc = load_extension().gather(A.values, A.tc_edge)
    A._reset_caches()
    return A
