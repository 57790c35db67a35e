"""Pins the CPU oracle (oracle/metcf_oracle.c) before anything is checked
against it: the hand-derived fixture, an independent brute-force
restatement, structural invariants, round trips and library products."""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import (GOLDEN, assert_metcf_equal, decode_metcf, load_fixture_18, naive_metcf,
                     random_csr)


def test_fixture_18_oracle():
    rowptr, col, M, fx = load_fixture_18()
    m = oracle.metcf(rowptr, col, M, fx["blockSize_h"], fx["blockSize_w"])
    exp = fx["expected"]
    assert m["block_count"] == exp["block_count"]
    for k in ["blockPartition", "RowWindowOffset", "TCblockRowid", "TCblockoffset", "TCblocktileId",
              "SparseAToXindex", "edgeToRow", "edgeToColumn"]:
        assert list(m[k]) == exp[k], k


def test_fixture_18_naive():
    rowptr, col, M, fx = load_fixture_18()
    m = naive_metcf(rowptr, col, M)
    exp = fx["expected"]
    for k in ["blockPartition", "RowWindowOffset", "TCblockRowid", "TCblockoffset", "TCblocktileId",
              "SparseAToXindex"]:
        assert list(m[k]) == exp[k], k


def test_fixture_18_degree_known_answer():
    rowptr, col, M, fx = load_fixture_18()
    B = np.ones((M, 5), np.float32)
    C = oracle.spmm_csr_f64(rowptr, col, None, B, M)
    assert np.array_equal(C[:, 0], np.array(fx["expected"]["degree"], np.float64))
    m = oracle.metcf(rowptr, col, M)
    C2 = oracle.spmm_metcf_f64(m, B, M)
    assert np.array_equal(C, C2)


SHAPES = [(16, 8), (16, 16), (8, 8), (16, 4), (4, 32), (16, 32), (32, 8), (1, 8), (16, 1)]


@pytest.mark.parametrize("bh,bw", SHAPES)
def test_oracle_vs_naive_random(bh, bw):
    rng = np.random.default_rng(1000 * bh + bw)
    for case in range(25):
        M = int(rng.integers(0, 70))
        K = int(rng.integers(1, 90))
        avg = float(rng.uniform(0, 12))
        rowptr, col = random_csr(rng, M, K, avg, sorted_rows=bool(case % 2),
                                 empty_frac=0.3 if case % 3 == 0 else 0.0,
                                 dup_frac=0.2 if case % 5 == 0 else 0.0)
        vals = rng.standard_normal(col.shape[0]).astype(np.float32)
        m = oracle.metcf(rowptr, col, M, bh, bw, values=vals)
        n = naive_metcf(rowptr, col, M, bh, bw, values=vals)
        assert_metcf_equal(m, n, ctx=f"bh={bh} bw={bw} case={case} M={M}")
        assert np.array_equal(m["TCblockValues"], n["TCblockValues"])


def _c1():
    from dtc.graphs import uniform_rows
    return uniform_rows(4096, 4096, 16, seed=0)


def test_invariants_c1():
    a = _c1()
    m = oracle.metcf(a.rowptr, a.col, a.num_rows)
    T, W = m["block_count"], (a.num_rows + 15) // 16
    off, rwo, atox = m["TCblockoffset"], m["RowWindowOffset"], m["SparseAToXindex"]
    assert off[0] == 0 and off[-1] == a.nnz and np.all(np.diff(off) >= 1)
    assert rwo[0] == 0 and rwo[-1] == T and rwo.shape[0] == W + 1
    assert np.all(np.diff(m["TCblockRowid"]) >= 0)
    sum_u = 0
    for w in range(W):
        s, e = a.rowptr[16 * w], a.rowptr[min(16 * w + 16, a.num_rows)]
        U = np.unique(a.col[s:e]).shape[0]
        sum_u += U
        assert m["blockPartition"][w] == -(-U // 8)
        slots = atox[rwo[w] * 8: rwo[w + 1] * 8]
        real, pad = slots[:U], slots[U:]
        assert np.all(np.diff(real) > 0) and np.all(pad == 0)
    for b in range(T):
        t = m["TCblocktileId"][off[b]:off[b + 1]]
        assert np.unique(t).shape[0] == t.shape[0]
    # figures quoted for this configuration: sum U_w = 63,695, T = 8,081
    assert sum_u == 63695 and T == 8081


def test_roundtrip_random():
    rng = np.random.default_rng(7)
    for case in range(20):
        M = int(rng.integers(1, 200))
        rowptr, col = random_csr(rng, M, 300, 6.0, sorted_rows=bool(case % 2), empty_frac=0.2)
        m = oracle.metcf(rowptr, col, M)
        got = sorted(decode_metcf(m, M))
        rows = np.repeat(np.arange(M), np.diff(rowptr))
        assert got == sorted(zip(rows.tolist(), col.tolist()))


def test_spmm_csr_vs_scipy():
    rng = np.random.default_rng(3)
    M, K, N = 300, 250, 19
    rowptr, col = random_csr(rng, M, K, 7.0)
    vals = rng.standard_normal(col.shape[0]).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C, Cabs = oracle.spmm_csr_f64(rowptr, col, vals, B, M, with_abs=True)
    A = sp.csr_matrix((vals.astype(np.float64), col, rowptr), shape=(M, K))
    assert np.allclose(C, A @ B.astype(np.float64), rtol=1e-12, atol=1e-12)
    Aabs = sp.csr_matrix((np.abs(vals).astype(np.float64), col, rowptr), shape=(M, K))
    assert np.allclose(Cabs, Aabs @ np.abs(B).astype(np.float64), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dup_frac", [0.0, 0.7])
def test_spmm_metcf_matches_csr(dup_frac):
    """The ME-TCF products equal the CSR product, also when (row, column)
    pairs repeat (their elements share a tile id and are added)."""
    rng = np.random.default_rng(4)
    M, N = 333, 24
    rowptr, col = random_csr(rng, M, M, 9.0, sorted_rows=False, empty_frac=0.1, dup_frac=dup_frac)
    vals = rng.uniform(-1, 1, col.shape[0]).astype(np.float32)
    B = rng.standard_normal((M, N)).astype(np.float32)
    m = oracle.metcf(rowptr, col, M, values=vals)
    C = oracle.spmm_csr_f64(rowptr, col, vals, B, M)
    A = sp.coo_matrix((vals.astype(np.float64), (np.repeat(np.arange(M), np.diff(rowptr)), col)),
                      shape=(M, M)).tocsr()  # scipy sums duplicates
    assert np.allclose(C, A @ B.astype(np.float64), rtol=1e-12, atol=1e-12)
    C64 = oracle.spmm_metcf_f64(m, B, M, tc_values=m["TCblockValues"])
    assert np.allclose(C64, C, rtol=1e-12, atol=1e-12)
    C32 = oracle.spmm_metcf_f32(m, B, M, tc_values=m["TCblockValues"])
    assert np.allclose(C32, C, rtol=1e-5, atol=1e-5)


def test_sddmm_vs_numpy():
    rng = np.random.default_rng(5)
    M, K, N = 120, 90, 13
    rowptr, col = random_csr(rng, M, K, 5.0)
    X = rng.standard_normal((M, N)).astype(np.float32)
    Y = rng.standard_normal((K, N)).astype(np.float32)
    out = oracle.sddmm_csr_f64(rowptr, col, X, Y, M)
    dense = X.astype(np.float64) @ Y.astype(np.float64).T
    rows = np.repeat(np.arange(M), np.diff(rowptr))
    assert np.allclose(out, dense[rows, col], rtol=1e-12, atol=1e-12)


def test_transpose_vs_scipy():
    rng = np.random.default_rng(6)
    M, K = 150, 170
    rowptr, col = random_csr(rng, M, K, 4.0, sorted_rows=True, empty_frac=0.2)
    vals = np.arange(col.shape[0], dtype=np.float32)
    rpt, ct, perm = oracle.csr_transpose(rowptr, col, M, K)
    At = sp.csr_matrix((vals, col, rowptr), shape=(M, K)).T.tocsr()
    At.sort_indices()
    assert np.array_equal(rpt, At.indptr) and np.array_equal(ct, At.indices)
    assert np.array_equal(vals[perm], At.data)


def test_inconsistent_block_partition_raises():
    rowptr, col, M, _ = load_fixture_18()
    M2 = 40  # add an empty third window
    rowptr2 = np.concatenate([rowptr, np.full(M2 - M, rowptr[-1], np.int32)])
    bp = np.zeros(3, np.int32)
    bp[2] = 5  # caller garbage in an empty window
    with pytest.raises(RuntimeError):
        oracle.metcf(rowptr2, col, M2, block_partition=bp)


def _digest(m):
    return {k: hashlib.sha256(np.ascontiguousarray(m[k]).tobytes()).hexdigest()
            for k in ["RowWindowOffset", "TCblockRowid", "TCblocktileId", "TCblockoffset",
                      "SparseAToXindex", "edgeToColumn", "edgeToRow", "blockPartition"]}


def test_c1_digest_regression():
    """SHA-256 of the oracle's C1 arrays (written by tests/golden/make_golden.py);
    a regression pin for the oracle and the exact-parity target for the GPU."""
    with open(os.path.join(GOLDEN, "c1_metcf_digest.json")) as f:
        exp = json.load(f)
    a = _c1()
    m = oracle.metcf(a.rowptr, a.col, a.num_rows)
    assert m["block_count"] == exp["block_count"]
    assert _digest(m) == exp["sha256"]
