"""Seeded sparse-matrix generators producing int32 CSR.

These replace the reference's dataset loader (scripts/DTCSpMM/dataset.py:21-44,
which reads an npz edge list and lets scipy build CSR with duplicates summed
and columns sorted per row).  Every generator here returns a ``CSR`` whose
rows hold sorted, duplicate-free int32 column indices, the same shape of
input ``preprocess_gpu`` receives from that loader.

Generators:
  uniform_rows     fixed number of distinct uniform columns per row
  rmat             R-MAT power-law graph, symmetrised, no self loops
  sbm_hubs         stochastic block model with Zipf community sizes + hub rows
  banded           band matrix with random half-width fill
  block_diagonal   dense-ish diagonal blocks plus uniform noise
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class CSR:
    rowptr: np.ndarray  # int32 [M+1]
    col: np.ndarray  # int32 [nnz], sorted and unique within each row
    values: np.ndarray | None  # float32 [nnz] or None (all ones)
    num_rows: int
    num_cols: int

    @property
    def nnz(self) -> int:
        return int(self.col.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.rowptr).astype(np.int64)

    def to_scipy(self):
        import scipy.sparse as sp
        vals = np.ones(self.nnz, np.float32) if self.values is None else self.values
        return sp.csr_matrix((vals, self.col, self.rowptr), shape=(self.num_rows, self.num_cols))


def _from_keys(keys: np.ndarray, num_rows: int, num_cols: int) -> CSR:
    """CSR from unique int64 keys row * num_cols + col (already sorted)."""
    rows = (keys // num_cols).astype(np.int64)
    col = (keys % num_cols).astype(np.int32)
    counts = np.bincount(rows, minlength=num_rows)
    rowptr = np.zeros(num_rows + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    if rowptr[-1] > np.iinfo(np.int32).max:
        raise ValueError("nnz exceeds int32")
    return CSR(rowptr.astype(np.int32), col, None, num_rows, num_cols)


def _dedup_pairs(rows: np.ndarray, cols: np.ndarray, num_rows: int, num_cols: int) -> CSR:
    keys = rows.astype(np.int64) * np.int64(num_cols) + cols.astype(np.int64)
    keys = np.unique(keys)
    return _from_keys(keys, num_rows, num_cols)


def uniform_rows(num_rows: int = 4096, num_cols: int = 4096, per_row: int = 16, seed: int = 0,
                 values: str | None = "uniform") -> CSR:
    """Each row gets ``per_row`` distinct columns drawn uniformly (PCG64(seed)),
    sorted; values U[-1, 1] float32 drawn from the same stream afterwards."""
    rng = np.random.Generator(np.random.PCG64(seed))
    per_row = min(per_row, num_cols)
    col = np.empty(num_rows * per_row, np.int32)
    for r in range(num_rows):
        col[r * per_row:(r + 1) * per_row] = np.sort(rng.choice(num_cols, per_row, replace=False))
    rowptr = (np.arange(num_rows + 1, dtype=np.int64) * per_row).astype(np.int32)
    vals = None
    if values == "uniform":
        vals = rng.uniform(-1.0, 1.0, col.shape[0]).astype(np.float32)
    return CSR(rowptr, col, vals, num_rows, num_cols)


def uniform_random(num_rows: int, num_cols: int, avg_per_row: float, seed: int = 0) -> CSR:
    """Vectorised uniform pattern: about ``avg_per_row`` random columns per row
    (rows drawn uniformly, duplicates removed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = int(round(num_rows * avg_per_row))
    rows = rng.integers(0, num_rows, m)
    cols = rng.integers(0, num_cols, m)
    return _dedup_pairs(rows, cols, num_rows, num_cols)


def rmat(scale: int = 20, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 1, symmetric: bool = True, values: str | None = "gcn") -> CSR:
    """R-MAT graph on 2**scale nodes with edge_factor * 2**scale drawn edges.

    Each edge descends ``scale`` levels choosing quadrant (0,0), (0,1), (1,0),
    (1,1) with probabilities a, b, c, d = 1 - a - b - c.  The edge list is
    symmetrised, self loops and duplicates removed.  ``values="gcn"`` gives
    1 / sqrt(deg_i * deg_j) (float32), the symmetric normalisation.
    """
    n = 1 << scale
    m = edge_factor * n
    rng = np.random.Generator(np.random.PCG64(seed))
    src = np.zeros(m, np.int64)
    dst = np.zeros(m, np.int64)
    ab, abc = a + b, a + b + c
    chunk = 1 << 22
    for start in range(0, m, chunk):
        stop = min(m, start + chunk)
        s = np.zeros(stop - start, np.int64)
        d = np.zeros(stop - start, np.int64)
        for _ in range(scale):
            r = rng.random(stop - start)
            sbit = r >= ab
            dbit = ((r >= a) & (r < ab)) | (r >= abc)
            s = (s << 1) | sbit
            d = (d << 1) | dbit
        src[start:stop] = s
        dst[start:stop] = d
    if symmetric:
        src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
    keep = src != dst
    csr = _dedup_pairs(src[keep], dst[keep], n, n)
    if values == "gcn":
        csr.values = gcn_values(csr)
    return csr


def gcn_values(csr: CSR) -> np.ndarray:
    """1 / sqrt(deg_i * deg_j) per edge (row degree i, column degree j)."""
    deg_r = csr.degrees().astype(np.float64)
    deg_c = np.bincount(csr.col, minlength=csr.num_cols).astype(np.float64)
    rows = np.repeat(np.arange(csr.num_rows), np.diff(csr.rowptr))
    v = 1.0 / np.sqrt(np.maximum(deg_r[rows], 1.0) * np.maximum(deg_c[csr.col], 1.0))
    return v.astype(np.float32)


def sbm_hubs(num_rows: int = 232965, communities: int = 50, zipf_s: float = 1.1,
             avg_per_row: float = 490.0, cross_frac: float = 0.05, hubs: int = 200,
             hub_min: int = 20000, hub_max: int = 50000, seed: int = 3,
             values: str | None = "uniform01") -> CSR:
    """Stochastic block model with hub rows.

    Communities take contiguous node ids with sizes proportional to
    k**-zipf_s (k = 1..communities; every community holds at least one row,
    the first takes the rounding remainder).  A pair (i, j) with j in i's
    community is an entry with probability p_in(c) = avg_per_row *
    (1 - cross_frac) / size_c (capped below 1), and any pair with probability
    p_out = avg_per_row * cross_frac / num_rows, independently -- so a row
    holds about avg_per_row entries, a fraction cross_frac of them drawn
    over all columns.  Each pair set is sampled by drawing Poisson(lam)
    columns with replacement, lam = -size * ln(1 - p), and dropping repeats:
    a column's draw count is then Poisson(lam / size) and it is present
    with probability 1 - exp(-lam / size) = p, independently of the others.
    ``hubs`` distinct rows get U[hub_min, hub_max] further distinct uniform
    columns.  Values U[0, 1] float32."""
    rng = np.random.Generator(np.random.PCG64(seed))
    bounds = sbm_community_starts(num_rows, communities, zipf_s)
    sizes, starts = np.diff(bounds), bounds[:-1]
    comm_of = np.repeat(np.arange(communities), sizes)
    p_in = np.minimum(avg_per_row * (1.0 - cross_frac) / sizes, 0.999)
    lam_in = -sizes * np.log1p(-p_in)
    p_out = min(avg_per_row * cross_frac / num_rows, 0.999)
    lam_out = -num_rows * np.log1p(-p_out)
    n_in = rng.poisson(lam_in[comm_of])
    n_out = rng.poisson(lam_out, num_rows)
    rows_in = np.repeat(np.arange(num_rows), n_in)
    c_in = comm_of[rows_in]
    cols_in = starts[c_in] + (rng.random(rows_in.shape[0]) * sizes[c_in]).astype(np.int64)
    rows_out = np.repeat(np.arange(num_rows), n_out)
    cols_out = rng.integers(0, num_rows, rows_out.shape[0])
    hub_rows = rng.choice(num_rows, min(hubs, num_rows), replace=False)
    hub_n = np.minimum(rng.integers(hub_min, hub_max + 1, hub_rows.shape[0]), num_rows)
    rows_h = np.repeat(hub_rows, hub_n)
    cols_h = np.concatenate([rng.choice(num_rows, k, replace=False) for k in hub_n]
                            or [np.zeros(0, np.int64)])
    rows = np.concatenate([rows_in, rows_out, rows_h])
    cols = np.concatenate([cols_in, cols_out, cols_h])
    csr = _dedup_pairs(rows, cols, num_rows, num_rows)
    if values == "uniform01":
        csr.values = rng.random(csr.nnz).astype(np.float32)
    return csr


def sbm_community_starts(num_rows: int = 232965, communities: int = 50,
                         zipf_s: float = 1.1) -> np.ndarray:
    """First row of every community of ``sbm_hubs`` (and num_rows at the end)."""
    w = 1.0 / np.arange(1, communities + 1) ** zipf_s
    sizes = np.maximum(1, np.floor(w / w.sum() * num_rows)).astype(np.int64)
    sizes[0] += num_rows - sizes.sum()
    return np.concatenate([[0], np.cumsum(sizes)])


def banded(num_rows: int, half_width: int, fill: float = 0.5, seed: int = 0) -> CSR:
    """Band matrix: each entry within |i - j| <= half_width kept with
    probability ``fill``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    offs = np.arange(-half_width, half_width + 1, dtype=np.int64)
    rows = np.repeat(np.arange(num_rows, dtype=np.int64), offs.shape[0])
    cols = rows + np.tile(offs, num_rows)
    keep = (cols >= 0) & (cols < num_rows) & (rng.random(rows.shape[0]) < fill)
    return _dedup_pairs(rows[keep], cols[keep], num_rows, num_rows)


def block_diagonal(num_rows: int, block: int, fill: float = 0.3, noise: float = 0.01,
                   seed: int = 0) -> CSR:
    """Diagonal blocks of size ``block`` filled with probability ``fill`` plus
    uniform noise entries amounting to ``noise`` of the block entries."""
    rng = np.random.Generator(np.random.PCG64(seed))
    per_row = max(1, int(round(block * fill)))
    rows = np.repeat(np.arange(num_rows, dtype=np.int64), per_row)
    base = (rows // block) * block
    span = np.minimum(block, num_rows - base)
    cols = base + (rng.random(rows.shape[0]) * span).astype(np.int64)
    m_noise = int(rows.shape[0] * noise)
    rows_n = rng.integers(0, num_rows, m_noise)
    cols_n = rng.integers(0, num_rows, m_noise)
    return _dedup_pairs(np.concatenate([rows, rows_n]), np.concatenate([cols, cols_n]),
                        num_rows, num_rows)


def dense_b(num_rows: int, n: int, seed: int, kind: str = "normal") -> np.ndarray:
    """Dense B (num_rows x n float32): N(0,1) or U[-1,1] from PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if kind == "normal":
        return rng.standard_normal((num_rows, n), dtype=np.float32)
    if kind == "uniform":
        return rng.uniform(-1.0, 1.0, (num_rows, n)).astype(np.float32)
    if kind == "ones":
        return np.ones((num_rows, n), np.float32)
    raise ValueError(kind)
