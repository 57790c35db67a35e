"""Outputs beyond 2**32 elements: B and C of 4,194,304 x 1,024 float32 (16 GiB
each), so every C row past 2,097,152 lies beyond a 32-bit element offset and
the last ones beyond an unsigned 32-bit one.  All rows but the last window
hold one entry (a permutation, values in [0.5, 1.5)), so C[r] = v_r B[perm[r]]
to one rounding; the last 16 rows hold 512 entries each, a window of ~1,000
TC blocks that the balanced schedule cuts into chunks and merges with atomics
at the far end of C.  Both schedules, checked in row slices against the
single products and, for the last window, the float64 sums."""
import numpy as np
import pytest
import torch

from helpers import say

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

M = 1 << 22
N = 1024
DENSE = 512   # entries per row of the last window
TOL = 4e-6    # float32 componentwise bound used for every fp32 SpMM test


@pytest.fixture(scope="module")
def big():
    from dtc.matrix import DTCMatrix
    rng = np.random.default_rng(31)
    perm = rng.permutation(M).astype(np.int32)
    counts = np.ones(M, np.int64)
    counts[-16:] = DENSE
    rowptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    tail = [np.sort(rng.choice(M, DENSE, replace=False)).astype(np.int32) for _ in range(16)]
    col = np.concatenate([perm[:M - 16]] + tail)
    vals = rng.uniform(0.5, 1.5, col.shape[0]).astype(np.float32)
    say("max-size build", M, col.shape[0])
    A = DTCMatrix(rowptr, col, vals, num_rows=M, num_cols=M)
    assert A.block_count > 2 * (M // 16 - 1) + 900  # the last window spans many chunks of 64
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    B = torch.randn((M, N), device="cuda", generator=g)
    yield A, B, perm, col, vals, rowptr
    del A, B
    torch.cuda.empty_cache()


@pytest.mark.parametrize("schedule", ["balanced", "windows"])
def test_beyond_32bit_offsets(big, schedule):
    A, B, perm, col, vals, rowptr = big
    say("max-size spmm", schedule, M, N)
    C = A.spmm(B, schedule=schedule, precision="fp32", chunk_blocks=64)
    torch.cuda.synchronize()
    assert C.shape == (M, N)
    perm_d = torch.from_numpy(perm).cuda()
    v1 = torch.from_numpy(vals[:M - 16]).cuda()
    step = 1 << 19
    for r0 in range(0, M - 16, step):
        r1 = min(r0 + step, M - 16)
        exp = v1[r0:r1, None] * B[perm_d[r0:r1].long()]
        bad = (C[r0:r1] - exp).abs() > TOL * exp.abs()
        assert not bool(bad.any()), (schedule, r0, int(bad.nonzero()[0, 0]) + r0)
        del exp, bad
    for r in range(M - 16, M):
        lo, hi = int(rowptr[r]), int(rowptr[r + 1])
        cols = torch.from_numpy(col[lo:hi]).cuda().long()
        v = torch.from_numpy(vals[lo:hi]).cuda().double()
        rows = B[cols].double()
        ref = (v[:, None] * rows).sum(0)
        den = (v.abs()[:, None] * rows.abs()).sum(0)
        err = float(((C[r].double() - ref).abs() / den).max())
        assert err <= TOL, (schedule, r, err)
    del C
    torch.cuda.empty_cache()
