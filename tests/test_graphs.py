"""Pins the seeded generators that stand in for the reference's dataset
loader (scripts/DTCSpMM/dataset.py:21-44): the properties each config
promises, checked on the arrays at their full sizes where that is cheap,
with expectations derived from the generating process (not restated from
the generator's code).

R-MAT (a, b, c, d) = (.57, .19, .19, .05), scale 20, edge factor 16:
  * pattern symmetric, no self loops, no repeated (row, column);
  * seed-reproducible, and a different seed gives a different graph;
  * at most 2 * 16 * 2^20 entries;
  * degree of node 0 against its closed form: a draw reaches node 0 from
    either end with probability (a + b)^s + (a + c)^s; given that, the other
    end's bits are 0 with probability a / (a + b) each, so with K such
    draws E[deg 0] = sum_j C(s, j) (1 - (1 - p0^(s-j) p1^j)^K), j >= 1;
  * the expected number of empty rows from the same per-node rates;
  * values 1 / sqrt(deg_i deg_j), as scipy's D^-1/2 A D^-1/2.
Block model (232,965 rows, 50 communities, Zipf(1.1), ~490 per row, 5%
cross, 200 hubs of 20k-50k):
  * community blocks recovered from the data (jumps in the rows' median
    column) have sizes proportional to k^-1.1: log-log slope and each size;
  * entries per non-hub row 490 +- 5%; cross fraction 5% +- 1% and equal to
    its expectation given the recovered blocks;
  * exactly 200 hub rows, each with 20,000-50,000 + (its community share)
    entries, extra counts averaging (20k + 50k) / 2;
  * values U[0, 1]; seed reproducibility on a smaller instance.
"""
from math import comb

import numpy as np
import pytest
import scipy.sparse as sp

from dtc.graphs import banded, block_diagonal, rmat, sbm_hubs, uniform_rows

A_, B_, C_ = 0.57, 0.19, 0.19


@pytest.fixture(scope="module")
def rmat20():
    return rmat(20, 16, seed=1)


@pytest.fixture(scope="module")
def sbm_full():
    return sbm_hubs()


def _keys(csr):
    rows = np.repeat(np.arange(csr.num_rows, dtype=np.int64), np.diff(csr.rowptr))
    return rows, csr.col.astype(np.int64)


def _check_simple_pattern(csr):
    rows, cols = _keys(csr)
    assert csr.rowptr[0] == 0 and csr.rowptr[-1] == csr.nnz
    assert np.all(np.diff(csr.rowptr) >= 0)
    # strictly increasing columns inside every row: sorted, no repeats
    same_row = rows[1:] == rows[:-1]
    assert np.all(np.diff(cols)[same_row] > 0)
    assert cols.min(initial=0) >= 0 and cols.max(initial=0) < csr.num_cols
    return rows, cols


def test_rmat_pattern_symmetric_simple(rmat20):
    a = rmat20
    rows, cols = _check_simple_pattern(a)
    assert not np.any(rows == cols), "self loop"
    n = np.int64(a.num_rows)
    fwd = np.sort(rows * n + cols)
    bwd = np.sort(cols * n + rows)
    assert np.array_equal(fwd, bwd), "pattern not symmetric"
    assert a.nnz <= 2 * 16 * (1 << 20)


def test_rmat_degree_closed_forms(rmat20):
    a = rmat20
    s, m = 20, 16 * (1 << 20)
    deg = np.diff(a.rowptr)
    K = m * ((A_ + B_) ** s + (A_ + C_) ** s)
    p0, p1 = A_ / (A_ + B_), B_ / (A_ + B_)
    e_deg0 = sum(comb(s, j) * (1 - (1 - p0 ** (s - j) * p1 ** j) ** K) for j in range(1, s + 1))
    assert abs(deg[0] / e_deg0 - 1) < 0.01, (deg[0], e_deg0)
    assert deg[0] == deg.max()  # node 0 is the hub of an R-MAT with a > b, c > d
    # empty rows: a node with z zero bits is hit by draws at rate
    # m ((a+b)^z (c+d)^(s-z) + (a+c)^z (b+d)^(s-z)); self loops (both ends v)
    # do not count, at rate m a^z d^(s-z).
    d_ = 1 - A_ - B_ - C_
    e_empty = sum(comb(s, z) * np.exp(-m * ((A_ + B_) ** z * (C_ + d_) ** (s - z)
                                            + (A_ + C_) ** z * (B_ + d_) ** (s - z)
                                            - 2 * A_ ** z * d_ ** (s - z)))
                  for z in range(s + 1))
    assert abs((deg == 0).sum() / e_empty - 1) < 0.01, ((deg == 0).sum(), e_empty)
    # heavy tail: the largest degree is thousands of times the mean
    assert deg.max() > 1000 * deg.mean()


def test_rmat_gcn_values(rmat20):
    a = rmat20
    A = sp.csr_matrix((np.ones(a.nnz), a.col, a.rowptr), shape=(a.num_rows, a.num_cols))
    d = np.asarray(A.sum(axis=1)).ravel()
    dinv = sp.diags(1.0 / np.sqrt(np.maximum(d, 1.0)))
    N = (dinv @ A @ dinv).tocsr()
    N.sort_indices()
    assert np.array_equal(N.indices, a.col)
    assert np.allclose(a.values.astype(np.float64), N.data, rtol=1e-6, atol=0)


def test_rmat_seeds():
    x, y, z = rmat(12, 16, seed=7), rmat(12, 16, seed=7), rmat(12, 16, seed=8)
    assert np.array_equal(x.rowptr, y.rowptr) and np.array_equal(x.col, y.col)
    assert np.array_equal(x.values, y.values)
    assert x.nnz != z.nnz or not np.array_equal(x.col, z.col)


def test_rmat_quadrant_special_cases():
    """Degenerate quadrant probabilities fix the graph: b = 1 sends every
    draw from node 0 to node n-1; a = b = 0.5 from node 0 to a uniform node;
    a = 1 gives only self loops, which are dropped."""
    g = rmat(6, 4, a=0.0, b=1.0, c=0.0, seed=0, symmetric=False, values=None)
    assert g.nnz == 1 and g.col[0] == 63 and g.rowptr[1] == 1
    g = rmat(6, 4, a=0.0, b=1.0, c=0.0, seed=0, values=None)
    assert g.nnz == 2 and set(_keys(g)[0].tolist()) == {0, 63}
    g = rmat(8, 64, a=0.5, b=0.5, c=0.0, seed=0, symmetric=False, values=None)
    assert np.all(np.diff(g.rowptr)[1:] == 0) and g.nnz == 255  # every column but 0 (self loop)
    g = rmat(8, 4, a=1.0, b=0.0, c=0.0, seed=0, values=None)
    assert g.nnz == 0


def _recover_blocks(csr, hub_mask):
    """Community boundaries from the data: rows of one community share the
    middle of their block as median column; it jumps between communities."""
    med = np.empty(csr.num_rows)
    for lo in range(0, csr.num_rows, 4096):  # chunked: medians of ragged rows
        hi = min(csr.num_rows, lo + 4096)
        for r in range(lo, hi):
            c = csr.col[csr.rowptr[r]:csr.rowptr[r + 1]]
            med[r] = np.median(c) if c.size else np.nan
    ok = ~hub_mask & ~np.isnan(med)
    idx = np.nonzero(ok)[0]
    d = np.abs(np.diff(med[idx]))
    # inside a block of S rows the medians of ~490 columns scatter by ~S/60;
    # between blocks they move by ~S: compare each step with the local spread
    from scipy.ndimage import median_filter
    spread = median_filter(d, size=257, mode="nearest")
    jumps = d > 15 * np.maximum(spread, 1.0)
    cuts = idx[1:][jumps]
    return np.concatenate([[0], cuts, [csr.num_rows]])


def test_sbm_communities_zipf(sbm_full):
    a = sbm_full
    deg = np.diff(a.rowptr)
    hub = deg >= 20000
    bounds = _recover_blocks(a, hub)
    sizes = np.diff(bounds)
    assert sizes.shape[0] == 50, sizes.shape
    k = np.arange(1, 51)
    slope = np.polyfit(np.log(k), np.log(sizes), 1)[0]
    assert abs(slope + 1.1) < 0.02, slope
    expect = a.num_rows * k ** -1.1 / (k ** -1.1).sum()
    assert np.all(np.abs(sizes - expect) <= 0.01 * expect + 3), np.max(np.abs(sizes - expect))
    # entries per ordinary row and the cross fraction
    assert abs(deg[~hub].mean() / 490.0 - 1) < 0.05, deg[~hub].mean()
    rows = np.repeat(np.arange(a.num_rows), deg)
    comm = np.searchsorted(bounds, np.arange(a.num_rows), side="right") - 1
    lo, hi = bounds[comm], bounds[comm + 1]
    keep = ~hub[rows]
    outside = ((a.col < lo[rows]) | (a.col >= hi[rows]))[keep]
    frac = outside.mean()
    assert 0.04 <= frac <= 0.06, frac
    # cross draws are uniform over all columns: the share landing outside a
    # row's own block of size S is 1 - S / n
    own = (sizes[comm] / a.num_rows)[~hub]
    e_frac = 0.05 * (1 - own).mean()
    assert abs(frac - e_frac) < 0.003, (frac, e_frac)


def test_sbm_hubs_and_values(sbm_full):
    a = sbm_full
    deg = np.diff(a.rowptr)
    hub = deg >= 20000
    assert hub.sum() == 200
    assert deg[hub].min() >= 20000 and deg[hub].max() <= 50000 + 1000
    assert deg[~hub].max() < 2000
    assert 33000 <= deg[hub].mean() - 490 <= 37500
    v = a.values
    assert v.dtype == np.float32 and v.shape[0] == a.nnz
    assert v.min() >= 0.0 and v.max() <= 1.0
    assert abs(v.mean() - 0.5) < 1e-3 and abs(v.var() - 1 / 12) < 1e-3
    _check_simple_pattern(a)


def test_sbm_seeds_small():
    kw = dict(num_rows=20000, communities=20, avg_per_row=40, hubs=6, hub_min=3000, hub_max=8000)
    x, y, z = sbm_hubs(seed=2, **kw), sbm_hubs(seed=2, **kw), sbm_hubs(seed=5, **kw)
    assert np.array_equal(x.col, y.col) and np.array_equal(x.values, y.values)
    assert not (x.nnz == z.nnz and np.array_equal(x.col, z.col))
    deg = np.diff(x.rowptr)
    assert (deg >= 3000).sum() == 6


def test_uniform_rows_c1():
    """C1: 4096 rows, 16 distinct sorted columns each, U[-1, 1] values."""
    a = uniform_rows(4096, 4096, 16, seed=0)
    _check_simple_pattern(a)
    assert np.all(np.diff(a.rowptr) == 16)
    assert a.values.min() >= -1 and a.values.max() <= 1 and abs(a.values.mean()) < 0.02
    # columns uniform: chi-square over 16 bins of 256 columns
    h = np.bincount(a.col // 256, minlength=16)
    chi2 = ((h - a.nnz / 16) ** 2 / (a.nnz / 16)).sum()
    assert chi2 < 40  # 15 degrees of freedom: p < 1e-3 above 37.7
    b = uniform_rows(4096, 4096, 16, seed=0)
    assert np.array_equal(a.col, b.col) and np.array_equal(a.values, b.values)


def test_banded_and_block_diagonal():
    b = banded(5000, 17, fill=0.5, seed=1)
    rows, cols = _check_simple_pattern(b)
    assert np.all(np.abs(rows - cols) <= 17)
    assert abs(b.nnz / (5000 * 35 - 17 * 18) - 0.5) < 0.02
    d = block_diagonal(4096, 64, fill=0.3, noise=0.01, seed=1)
    rows, cols = _check_simple_pattern(d)
    assert (rows // 64 == cols // 64).mean() > 0.98
