"""GPU ME-TCF build vs the CPU oracle: bit-exact on every output array.

Most tests run under every window routing of the build (DTC_METCF_PATH):
"auto" (by entry count: one LDS workgroup per window up to 16,384 entries,
the device-wide radix sort beyond), "global" (every window through the
radix sort) and "lds0" ... "lds5" (every window that fits the one-wave
bitonic kernel for at most 64 entries, or one LDS radix-sort shape, waves x
entries per lane 1 x 8, 4 x 8, 8 x 8, 8 x 16 or 16 x 16, through it; the
rest through the radix sort)."""
import os

import numpy as np
import pytest
import torch

import oracle
from helpers import assert_metcf_equal, gpu_metcf, load_fixture_18, random_csr, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ext():
    from dtc import load_extension
    return load_extension()


PATHS = ["auto", "global", "lds0", "lds1", "lds2", "lds3", "lds4", "lds5"]


@pytest.fixture(params=PATHS)
def path(request, monkeypatch):
    monkeypatch.setenv("DTC_METCF_PATH", request.param)
    return request.param


def check(ext, rowptr, col, M, bh=16, bw=8, values=None, ctx=""):
    say("metcf", ctx, f"M={M} nnz={col.shape[0]} bh={bh} bw={bw}",
        "path=" + os.environ.get("DTC_METCF_PATH", "auto"))
    g = gpu_metcf(ext, rowptr, col, M, bh, bw, values=values)
    o = oracle.metcf(rowptr, col, M, bh, bw, values=values)
    assert_metcf_equal(g, o, ctx=ctx)
    if values is not None:
        assert np.array_equal(g["TCblockValues"], o["TCblockValues"]), ctx
    # edge order: the CSR edge stored at each TC position
    assert np.array_equal(np.sort(g["TCblockEdge"]), np.arange(col.shape[0])), ctx
    return g


def test_fixture_18(ext, path):
    rowptr, col, M, fx = load_fixture_18()
    g = check(ext, rowptr, col, M, ctx="fixture18")
    for k, v in fx["expected"].items():
        if k == "block_count":
            assert g[k] == v
        elif k in g:
            assert list(g[k]) == v, k


@pytest.mark.parametrize("bh,bw", [(16, 8), (16, 16), (8, 8), (16, 4), (16, 32), (1, 8), (32, 8)])
def test_random_small(ext, path, bh, bw):
    rng = np.random.default_rng(50 + bh * 7 + bw)
    for case in range(12):
        M = int(rng.integers(1, 90))
        rowptr, col = random_csr(rng, M, int(rng.integers(1, 120)), float(rng.uniform(0, 14)),
                                 sorted_rows=bool(case % 2), empty_frac=0.3 if case % 3 == 0 else 0.0,
                                 dup_frac=0.2 if case % 4 == 0 else 0.0)
        vals = rng.standard_normal(col.shape[0]).astype(np.float32)
        check(ext, rowptr, col, M, bh, bw, values=vals, ctx=f"random bh={bh} bw={bw} case={case}")


def test_c1_uniform(ext):
    from dtc.graphs import uniform_rows
    a = uniform_rows(4096, 4096, 16, seed=0)
    g = check(ext, a.rowptr, a.col, a.num_rows, values=a.values, ctx="C1")
    assert g["block_count"] == 8081


@pytest.mark.parametrize("scale", [10, 14])
def test_rmat(ext, path, scale):
    from dtc.graphs import rmat
    a = rmat(scale, 16, seed=1)
    check(ext, a.rowptr, a.col, a.num_rows, values=a.values, ctx=f"rmat{scale}")


def test_sbm_hubs_small(ext, path):
    from dtc.graphs import sbm_hubs
    a = sbm_hubs(20000, communities=20, avg_per_row=40, hubs=6, hub_min=3000, hub_max=8000, seed=2)
    check(ext, a.rowptr, a.col, a.num_rows, values=a.values, ctx="sbm")


def test_banded_and_block_diagonal(ext, path):
    from dtc.graphs import banded, block_diagonal
    a = banded(3000, 40, seed=1)
    check(ext, a.rowptr, a.col, a.num_rows, ctx="banded")
    b = block_diagonal(5000, 100, seed=2)
    check(ext, b.rowptr, b.col, b.num_rows, ctx="blockdiag")


def test_window_beyond_reference_smem_limit(ext, path):
    # one window with more than 16,383 TC blocks overflows the reference's
    # 128 KB shared-memory opt-in (DTCSpMM_kernel.cu:311-325); no limit here.
    rng = np.random.default_rng(9)
    M, K = 16, 1 << 20
    rows = [np.sort(rng.choice(K, 12000, replace=False)) for _ in range(M)]
    rowptr = np.concatenate([[0], np.cumsum([r.shape[0] for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    g = check(ext, rowptr, col, M, ctx="huge-window")
    assert g["block_count"] > 16383


def test_unsorted_rows_and_duplicates(ext, path):
    rng = np.random.default_rng(11)
    rowptr, col = random_csr(rng, 1000, 500, 20.0, sorted_rows=False, dup_frac=0.5)
    check(ext, rowptr, col, 1000, ctx="unsorted+dups")


@pytest.mark.parametrize("kind", ["sorted", "unsorted", "dups", "one_swap"])
def test_dense_large_windows(ext, path, kind):
    """Windows of 5,120 to 14,400 entries over at most 1,200 columns (TC
    blocks of 34 or more entries on average), for the 8- and 16-wave LDS
    shapes: their TC-block positions come from per-block occupancy bitmaps
    when every row of the window is strictly increasing, and from per-entry
    counting otherwise (unsorted rows, duplicates, or one out-of-order pair
    in a single row)."""
    rng = np.random.default_rng(31 + len(kind))
    rows = []
    for w in range(6):
        per_row = [320, 600, 900][w % 3]
        pool = rng.choice(1 << 18, 1200, replace=False)
        for _ in range(16):
            rows.append(np.sort(rng.choice(pool, per_row, replace=False)))
    if kind == "unsorted":
        rows = [rng.permutation(r) for r in rows]
    elif kind == "dups":
        rows[5] = np.sort(np.concatenate([rows[5], rows[5][:3]]))
        rows[16 * 4 + 9] = np.sort(np.concatenate([rows[16 * 4 + 9], rows[16 * 4 + 9][-1:]]))
    elif kind == "one_swap":
        for r in (16 + 7, 16 * 4 + 15):
            rows[r] = rows[r].copy()
            rows[r][[-2, -1]] = rows[r][[-1, -2]]
    M = len(rows)
    rowptr = np.concatenate([[0], np.cumsum([r.shape[0] for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    check(ext, rowptr, col, M, ctx=f"dense large windows {kind}")


def test_one_stream(ext, monkeypatch):
    """DTC_METCF_STREAMS=1: the whole build on the caller's stream, bit-exact
    on matrices with windows for every LDS class and for the device-wide
    sort, unsorted rows and duplicates included."""
    from dtc.graphs import rmat, sbm_hubs
    monkeypatch.setenv("DTC_METCF_STREAMS", "1")
    r = rmat(14, 16, seed=3)
    check(ext, r.rowptr, r.col, r.num_rows, ctx="one stream rmat14")
    b = sbm_hubs(20000, communities=20, avg_per_row=40, hubs=6, hub_min=3000, hub_max=20000, seed=4)
    check(ext, b.rowptr, b.col, b.num_rows, ctx="one stream sbm")
    rng = np.random.default_rng(13)
    rows = [rng.permutation(rng.choice(1 << 16, 1500, replace=False)) for _ in range(16)]
    rows += [rng.choice(1 << 16, int(rng.integers(1, 40)), replace=True) for _ in range(300)]
    rowptr = np.concatenate([[0], np.cumsum([x.shape[0] for x in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    check(ext, rowptr, col, len(rows), ctx="one stream unsorted+dups")


def test_empty_and_degenerate(ext, path):
    for M in [0, 1, 15, 16, 17, 100]:
        rowptr = np.zeros(M + 1, np.int32)
        col = np.zeros(0, np.int32)
        check(ext, rowptr, col, M, ctx=f"empty M={M}")
    rowptr = np.array([0, 0, 0, 1], np.int32)
    check(ext, rowptr, np.array([2], np.int32), 3, ctx="single edge")


def test_inconsistent_block_partition_raises(ext):
    rowptr, col, M, _ = load_fixture_18()
    M2 = 40
    rowptr2 = np.concatenate([rowptr, np.full(M2 - M, rowptr[-1], np.int32)])
    bp = np.zeros(3, np.int32)
    bp[2] = 5
    say("metcf inconsistent blockPartition")
    with pytest.raises(RuntimeError, match="blockPartition"):
        gpu_metcf(ext, rowptr2, col, M2, block_partition=bp)


def test_invalid_csr_raises(ext):
    say("metcf invalid CSR")
    with pytest.raises(RuntimeError, match="invalid CSR"):
        gpu_metcf(ext, np.array([0, 3, 2, 4], np.int32), np.array([0, 1, 2, 3], np.int32), 3)
    with pytest.raises(RuntimeError, match="invalid CSR"):
        gpu_metcf(ext, np.array([0, 2, 4], np.int32), np.array([0, -1, 2, 3], np.int32), 2)
    with pytest.raises(RuntimeError, match="invalid CSR"):
        gpu_metcf(ext, np.array([0, 2, 3], np.int32), np.array([0, 1, 2, 3], np.int32), 2)


@pytest.mark.parametrize("where", ["tiny", "medium", "large"])
def test_negative_column_every_path(ext, path, where):
    """A negative column in a window of 48 entries, of 992 entries, or of
    20,000 entries (the bitonic kernel, an LDS radix-sort shape, or the
    device-wide sort under the automatic routing; the others under the
    forced ones) is reported as an invalid CSR whichever kernel reads it."""
    rng = np.random.default_rng(len(where))
    sizes = [3] * 16 + [62] * 16 + [1250] * 16 + [5] * 16
    rows = [np.sort(rng.choice(1 << 16, n, replace=False)) for n in sizes]
    r = {"tiny": 3, "medium": 16 + 7, "large": 32 + 11}[where]
    rows[r] = rows[r].copy()
    rows[r][0] = -5
    rowptr = np.concatenate([[0], np.cumsum([x.shape[0] for x in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    say("metcf negative column", where, "path=" + path)
    with pytest.raises(RuntimeError, match="invalid CSR"):
        gpu_metcf(ext, rowptr, col, len(rows))


def test_wrong_sizes_raise(ext):
    dev = "cuda"
    i32 = torch.int32
    col = torch.zeros(4, dtype=i32, device=dev)
    rpt = torch.tensor([0, 2, 4], dtype=i32, device=dev)
    with pytest.raises(RuntimeError, match="blockPartition"):
        ext.preprocess_gpu(col, rpt, 2, 16, 8, torch.zeros(3, dtype=i32, device=dev),
                           torch.zeros(4, dtype=i32, device=dev), torch.zeros(4, dtype=i32, device=dev))
    with pytest.raises(RuntimeError, match="dtype"):
        ext.preprocess_gpu(col.long(), rpt, 2, 16, 8, torch.zeros(1, dtype=i32, device=dev),
                           torch.zeros(4, dtype=i32, device=dev), torch.zeros(4, dtype=i32, device=dev))


def test_metcf_cache_roundtrip(ext, tmp_path):
    from dtc.graphs import rmat
    from dtc.io import load_metcf, save_metcf
    from dtc.matrix import DTCMatrix
    a = rmat(10, 8, seed=3)
    say("metcf cache")
    A = DTCMatrix(a.rowptr, a.col, a.values, num_rows=a.num_rows, num_cols=a.num_cols)
    p = str(tmp_path / "a.pt")
    save_metcf(p, A)
    A2 = load_metcf(p)
    B = torch.randn((a.num_rows, 32), device="cuda")
    assert torch.equal(A.spmm(B, schedule="windows"), A2.spmm(B, schedule="windows"))
    assert A2.block_count == A.block_count and torch.equal(A2.values_tc, A.values_tc)
    # the default schedule ("auto") and the differentiable product work on a
    # matrix read back from the cache, with the same results
    assert A2.max_window_blocks == A.max_window_blocks
    assert torch.equal(A.spmm(B), A2.spmm(B))
    from dtc.matrix import spmm_autograd
    v = torch.from_numpy(a.values).cuda()
    grads = []
    for M_ in (A, A2):
        vv = v.clone().requires_grad_(True)
        Bb = B.clone().requires_grad_(True)
        spmm_autograd(vv, Bb, M_).square().sum().backward()
        grads.append((vv.grad, Bb.grad))
    assert torch.equal(grads[0][0], grads[1][0]) and torch.equal(grads[0][1], grads[1][1])


@pytest.mark.parametrize("bw", [1, 8])
def test_mixed_window_sizes(ext, path, bw):
    """Windows of very different sizes side by side: one of 65,000 edges over
    columns up to 3,000,000 (with 1-column blocks, about 60,000 blocks in one
    window), one of 6,000 edges over 50 columns, one of 3,000 edges, and many
    small and empty windows."""
    rng = np.random.default_rng(bw)
    M = 200
    rows = [[] for _ in range(M)]
    rows[3] = list(rng.choice(3_000_000, 60_000, replace=False))          # huge window
    rows[4] = list(rng.integers(0, 3_000_000, 5_000)) + [7, 7, 7]          # same window, repeats
    rows[40] = list(rng.integers(0, 50, 6_000))                            # big window, few columns
    rows[81] = list(rng.integers(0, 4_000, 3_000))                         # a few thousand edges
    for r in range(100, 200, 3):
        rows[r] = list(rng.integers(0, 2_999_999, int(rng.integers(0, 40))))
    rowptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate([np.array(r, np.int64) for r in rows]).astype(np.int32)
    check(ext, rowptr, col, M, bw=bw, ctx=f"mixed window sizes bw={bw}")


@pytest.mark.parametrize("per_window", [4, 8])
def test_many_short_windows(ext, path, per_window):
    """Tens of thousands of windows of a few edges each (random rows of the
    window, duplicates dropped), so that thousands of consecutive windows
    share every stretch of 2,048 edges."""
    rng = np.random.default_rng(per_window)
    M = 16 * (2048 // per_window) * 12
    rows = [set() for _ in range(M)]
    for w in range(M // 16):
        for _ in range(per_window):
            rows[16 * w + int(rng.integers(0, 16))].add(int(rng.integers(0, 5000)))
    rowptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate([np.array(sorted(r), np.int64) for r in rows]).astype(np.int32)
    check(ext, rowptr, col, M, ctx=f"short windows per_window={per_window}")
