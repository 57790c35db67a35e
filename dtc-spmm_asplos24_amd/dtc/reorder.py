"""TCA row reordering on the GPU (reordering/TCA_reorder.py).

``tca_reorder`` returns ``reorder_id``: the original row ids in their new
order, as the reference writes to ``<dataset>.reorder_id.npz``
(TCA_reorder.py:317).  ``relabel`` applies it to both ends of every edge, as
the reference does before writing ``<dataset>.reorder.npz`` (:306-326).

Stages (all in libdtc): MinHash signatures, LSH candidate pairs and exact
Jaccard similarity on the GPU; the greedy pair-merge clustering (most similar
pair first, clusters frozen at ``thres`` rows) on the host; the same again
over cluster unions with ``cluster_thres``; rows emitted cluster by cluster.
Defaults follow the reference: thres 16 (one row window), cluster_thres 128,
LSH threshold 0.2 at 128 permutations, i.e. 28 bands of 2 entries.
"""
from __future__ import annotations

import numpy as np

from . import load_extension
from .graphs import CSR


def canonical_rows(rowptr, col, num_rows, values=None):
    """Rows with sorted, duplicate-free columns (duplicates summed)."""
    import scipy.sparse as sp
    rowptr = np.asarray(rowptr, np.int64)
    col = np.asarray(col, np.int64)
    num_cols = int(col.max()) + 1 if col.size else 1
    data = np.ones(col.shape[0], np.float64) if values is None else np.asarray(values, np.float64)
    m = sp.csr_matrix((data, col, rowptr), shape=(num_rows, num_cols))
    m.sum_duplicates()
    m.sort_indices()
    return m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data.astype(np.float32)


def is_canonical(rowptr, col) -> bool:
    """Columns strictly increasing inside every row (vectorised check)."""
    col = np.asarray(col)
    if col.shape[0] < 2:
        return True
    inc = np.diff(col.astype(np.int64)) > 0
    cross = np.asarray(rowptr, np.int64)[1:-1] - 1  # diff index that enters a new row
    cross = cross[(cross >= 0) & (cross < inc.shape[0])]
    inc[cross] = True
    return bool(inc.all())


def tca_reorder(rowptr, col, num_rows=None, *, thres=16, cluster_thres=128, bands=28,
                rows_per_band=2, max_neighbors=8, seed=2022, max_pairs=0, device="cuda"):
    """(reorder_id int64 [num_rows], stats dict)."""
    import torch
    ext = load_extension()
    rowptr = np.asarray(rowptr)
    num_rows = rowptr.shape[0] - 1 if num_rows is None else int(num_rows)
    if is_canonical(rowptr, col):
        rp, cc = np.asarray(rowptr, np.int32), np.asarray(col, np.int32)
    else:
        rp, cc, _ = canonical_rows(rowptr, col, num_rows)
    order, stats = ext.tca_reorder(torch.from_numpy(rp).to(device), torch.from_numpy(cc).to(device),
                                   thres, cluster_thres, bands, rows_per_band, max_neighbors, seed,
                                   max_pairs)
    return order.numpy().astype(np.int64), stats


def relabel(a: CSR, reorder_id) -> CSR:
    """The symmetric relabelling: node reorder_id[i] becomes node i, for
    rows and columns alike (square matrices); values follow their edges."""
    reorder_id = np.asarray(reorder_id, np.int64)
    M = a.num_rows
    if a.num_cols != M or reorder_id.shape[0] != M:
        raise ValueError("relabel needs a square matrix and one id per row")
    pos = np.empty(M, np.int64)
    pos[reorder_id] = np.arange(M)
    if not np.array_equal(np.sort(reorder_id), np.arange(M)):
        raise ValueError("reorder_id is not a permutation")
    deg = np.diff(a.rowptr.astype(np.int64))
    rows = pos[np.repeat(np.arange(M), deg)]
    cols = pos[a.col.astype(np.int64)]
    o = np.lexsort((cols, rows))
    rowptr = np.zeros(M + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=M), out=rowptr[1:])
    vals = None if a.values is None else a.values[o]
    return CSR(rowptr.astype(np.int32), cols[o].astype(np.int32), vals, M, M)
