"""The power-law configuration at full size (R-MAT scale 20, 31.4M edges,
N up to 256): ME-TCF bit-exact against the oracle, the all-ones known answer
C[r, :] = deg(r) exactly, linearity A(B1 + 2 B2) = A B1 + 2 A B2, and the
float32 / bf16x6 componentwise error against float64 on a row sample."""
import numpy as np
import pytest
import torch

import oracle
from helpers import say

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


@pytest.fixture(scope="module")
def rmat20():
    from dtc.graphs import rmat
    from dtc.matrix import DTCMatrix
    a = rmat(20, 16, seed=1)
    say("full-size R-MAT scale 20", a.nnz)
    A = DTCMatrix(a.rowptr, a.col, a.values, num_rows=a.num_rows, num_cols=a.num_cols)
    return a, A


def test_metcf_bit_exact_full_size(rmat20):
    a, A = rmat20
    ref = oracle.metcf(a.rowptr, a.col, a.num_rows, values=a.values)
    assert A.block_count == ref["block_count"]
    for k, v in [("RowWindowOffset", A.row_window_offset), ("TCblockRowid", A.tcblock_rowid),
                 ("TCblocktileId", A.tcblocktile_id), ("TCblockoffset", A.tcblock_offset),
                 ("SparseAToXindex", A.sparse_atox), ("edgeToColumn", A.edge_to_column),
                 ("edgeToRow", A.edge_to_row), ("blockPartition", A.block_partition)]:
        assert np.array_equal(v.cpu().numpy(), ref[k]), k
    assert np.array_equal(A.values_tc.cpu().numpy(), ref["TCblockValues"])
    assert A.sum_unique_columns() == 27196090 and A.block_count == 3428580


@pytest.mark.parametrize("schedule", ["windows", "balanced"])
def test_degree_known_answer_full_size(rmat20, schedule):
    a, A = rmat20
    say("full-size degree", schedule)
    ones = torch.ones((a.num_rows, 256), device="cuda")
    deg = torch.from_numpy(np.diff(a.rowptr).astype(np.float32)).cuda()
    for prec in ("fp32", "bf16x6", "bf16"):
        C = A.spmm(ones, schedule=schedule, precision=prec, use_values=False, chunk_blocks=256)
        assert torch.equal(C, deg[:, None].expand_as(C)), (schedule, prec)


def test_linearity_full_size(rmat20):
    a, A = rmat20
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    B1 = torch.randn((a.num_rows, 64), device="cuda", generator=g)
    B2 = torch.randn((a.num_rows, 64), device="cuda", generator=g)
    say("full-size linearity")
    lhs = A.spmm(B1 + 2 * B2, chunk_blocks=256)
    rhs = A.spmm(B1, chunk_blocks=256) + 2 * A.spmm(B2, chunk_blocks=256)
    scale = A.spmm(B1.abs() + 2 * B2.abs(), chunk_blocks=256)
    assert float(((lhs - rhs).abs() / scale.clamp_min(1e-30)).max()) < 1e-5


# Componentwise tolerances on this graph (max |C - C64| / (|A| |B|)):
#   fp32, bf16x6  4e-6   float32-accurate products
#   bf16x3        1e-4   dropped limb products ~3 * 2^-16 per product; rows
#                        with one entry show a single product's error
#   bf16          3e-2   one bf16 pass: unit roundoff 2^-8 per operand
#   fp16          1.05   one fp16 pass: B elements below 2^-14 (N(0,1) has
#                        many) become fp16 subnormals with an absolute spacing
#                        of 2^-24, so one product's relative error reaches 1
#                        (an element rounded to 0 or to one spacing) times
#                        (1 + A's own rounding error, <= 2^-11 relative for
#                        normal values, ~0.2% for the subnormal hub values);
#                        a one-entry row shows it whole.  The normwise error
#                        max |C - C64| / max |C64| stays below 2e-3.
PRECISION_TOL = [("fp32", 4e-6), ("bf16x6", 4e-6), ("bf16x3", 1e-4), ("bf16", 3e-2), ("fp16", 1.05)]


@pytest.fixture(scope="module")
def sample_rows(rmat20):
    a, _ = rmat20
    rows = np.concatenate([np.arange(0, 4096), np.random.default_rng(0).integers(0, a.num_rows, 8192)])
    return np.unique(rows)


@pytest.mark.parametrize("N", [32, 128, 256])
def test_error_vs_fp64_row_sample(rmat20, sample_rows, N):
    a, A = rmat20
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    B = torch.randn((a.num_rows, N), device="cuda", generator=g)
    rows = sample_rows
    sub_ptr = np.concatenate([[0], np.cumsum(np.diff(a.rowptr)[rows])]).astype(np.int32)
    idx = np.concatenate([np.arange(a.rowptr[r], a.rowptr[r + 1]) for r in rows])
    C64, Cabs = oracle.spmm_csr_f64(sub_ptr, a.col[idx], a.values[idx], B.cpu().numpy(), rows.shape[0],
                                    with_abs=True)
    m = Cabs > 0
    cmax = float(np.abs(C64).max())
    runs = [("balanced", p, t) for p, t in PRECISION_TOL] + [("windows", "fp32", 4e-6)]
    for schedule, precision, tol in runs:
        say("full-size error", schedule, precision, N)
        C = A.spmm(B, schedule=schedule, precision=precision, chunk_blocks=256).cpu().numpy()
        d = np.abs(C[rows].astype(np.float64) - C64)
        comp = float((d[m] / Cabs[m]).max())
        assert comp <= tol, (schedule, precision, N, comp)
        assert float(d[~m].max(initial=0.0)) == 0.0
        if precision == "fp16":
            assert float(d.max()) / cmax <= 2e-3, (N, float(d.max()) / cmax)
