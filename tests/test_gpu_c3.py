"""The dense-community block model at full size: 232,965 rows, 50 Zipf(1.1)
communities, ~490 entries per row (5% across communities), 200 hub rows of
20k-50k entries, U[0, 1] values -- about 121M entries.  ME-TCF bit-exact
against the oracle on every array; both schedules at N = 128 and 512 against
float64 on a row sample holding every hub row; the all-ones known answer
C[r, :] = deg(r) exactly; balanced against unbalanced; and the window-split
threshold (balanced chunk) swept over 16 ... 512 blocks plus the
deterministic merge.

Tolerances (componentwise max |C - C64| / (|A| |B|), float64 denominators):
fp32 and bf16x6 4e-6; bf16x3 1e-4 (its dropped limb products are bounded by
~3 * 2^-16 per product)."""
import numpy as np
import pytest
import torch

import oracle
from helpers import say

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

TOL = {"fp32": 4e-6, "bf16x6": 4e-6, "bf16x3": 1e-4}


@pytest.fixture(scope="module")
def sbm():
    from dtc.graphs import sbm_hubs
    from dtc.matrix import DTCMatrix
    a = sbm_hubs()
    say("block model", a.num_rows, a.nnz)
    A = DTCMatrix(a.rowptr, a.col, a.values, num_rows=a.num_rows, num_cols=a.num_cols)
    deg = np.diff(a.rowptr)
    hubs = np.nonzero(deg >= 20000)[0]
    rows = np.unique(np.concatenate([hubs, np.arange(0, 2048), np.arange(a.num_rows - 2048, a.num_rows),
                                     np.random.default_rng(0).integers(0, a.num_rows, 4096)]))
    return a, A, hubs, rows


def _sample_ref(a, rows, Bn):
    sub_ptr = np.concatenate([[0], np.cumsum(np.diff(a.rowptr)[rows])]).astype(np.int32)
    idx = np.concatenate([np.arange(a.rowptr[r], a.rowptr[r + 1]) for r in rows])
    return oracle.spmm_csr_f64(sub_ptr, a.col[idx], a.values[idx], Bn, rows.shape[0], with_abs=True)


def _err(C, rows, C64, Cabs):
    d = np.abs(C[rows].astype(np.float64) - C64)
    m = Cabs > 0
    return max(float((d[m] / Cabs[m]).max()), float(d[~m].max(initial=0.0)))


def test_metcf_bit_exact(sbm):
    a, A, hubs, _ = sbm
    assert hubs.shape[0] == 200
    ref = oracle.metcf(a.rowptr, a.col, a.num_rows, values=a.values)
    assert A.block_count == ref["block_count"]
    for k, v in [("RowWindowOffset", A.row_window_offset), ("TCblockRowid", A.tcblock_rowid),
                 ("TCblocktileId", A.tcblocktile_id), ("TCblockoffset", A.tcblock_offset),
                 ("SparseAToXindex", A.sparse_atox), ("edgeToColumn", A.edge_to_column),
                 ("edgeToRow", A.edge_to_row), ("blockPartition", A.block_partition)]:
        assert np.array_equal(v.cpu().numpy(), ref[k]), k
    assert np.array_equal(A.values_tc.cpu().numpy(), ref["TCblockValues"])


@pytest.mark.parametrize("N", [128, 512])
def test_error_vs_fp64_with_hubs(sbm, N):
    a, A, hubs, rows = sbm
    g = torch.Generator(device="cuda")
    g.manual_seed(N)
    B = torch.randn((a.num_rows, N), device="cuda", generator=g)
    C64, Cabs = _sample_ref(a, rows, B.cpu().numpy())
    results = {}
    for schedule in ("windows", "balanced"):
        for precision in ("fp32", "bf16x6", "bf16x3"):
            say("block model error", schedule, precision, N)
            C = A.spmm(B, schedule=schedule, precision=precision)
            err = _err(C.cpu().numpy(), rows, C64, Cabs)
            assert err <= TOL[precision], (schedule, precision, N, err)
            results[(schedule, precision)] = C
    # the two schedules agree (same products, different merge order)
    Cw, Cb = results[("windows", "fp32")], results[("balanced", "fp32")]
    scale = A.spmm(B.abs(), schedule="windows")
    assert float(((Cw - Cb).abs() / scale.clamp_min(1e-30)).max()) <= 4e-6


@pytest.mark.parametrize("schedule", ["windows", "balanced"])
def test_degree_known_answer(sbm, schedule):
    a, A, _, _ = sbm
    say("block model degree", schedule)
    ones = torch.ones((a.num_rows, 128), device="cuda")
    deg = torch.from_numpy(np.diff(a.rowptr).astype(np.float32)).cuda()
    for prec in ("fp32", "bf16"):
        C = A.spmm(ones, schedule=schedule, precision=prec, use_values=False)
        assert torch.equal(C, deg[:, None].expand_as(C)), (schedule, prec)


def test_split_threshold_sweep(sbm):
    """Balanced chunk (the window-split threshold) 16 ... 512 blocks at
    N = 128, each against float64; the deterministic merge bitwise stable."""
    a, A, _, rows = sbm
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    B = torch.randn((a.num_rows, 128), device="cuda", generator=g)
    C64, Cabs = _sample_ref(a, rows, B.cpu().numpy())
    for chunk in (16, 32, 64, 128, 256, 512):
        say("block model chunk", chunk)
        C = A.spmm(B, schedule="balanced", chunk_blocks=chunk)
        assert _err(C.cpu().numpy(), rows, C64, Cabs) <= TOL["fp32"], chunk
    say("block model deterministic")
    D1 = A.spmm(B, schedule="balanced", chunk_blocks=64, deterministic=True)
    D2 = A.spmm(B, schedule="balanced", chunk_blocks=64, deterministic=True)
    assert torch.equal(D1, D2)
    assert _err(D1.cpu().numpy(), rows, C64, Cabs) <= TOL["fp32"]
