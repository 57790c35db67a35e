"""TC-GNN's SpMM (the comparison baseline without TC-block compaction) over
the phase-1 arrays of the ME-TCF build: against the float64 CSR product, the
all-ones known answer through the reference's call sequence (CPU
`preprocess`, then `run_TCGNNSpMM`), and against the ME-TCF kernels."""
import numpy as np
import pytest
import torch

import oracle
from helpers import random_csr, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ext():
    from dtc import load_extension
    return load_extension()


def phase1(ext, rowptr, col, M):
    W = (M + 15) // 16
    bp = torch.zeros(W, dtype=torch.int32, device="cuda")
    e2c = torch.zeros(col.shape[0], dtype=torch.int32, device="cuda")
    e2r = torch.zeros(col.shape[0], dtype=torch.int32, device="cuda")
    rp, cc = torch.from_numpy(rowptr).cuda(), torch.from_numpy(col).cuda()
    ext.preprocess_gpu(cc, rp, M, 16, 8, bp, e2c, e2r)
    return rp, cc, bp, e2c, e2r


@pytest.mark.parametrize("N", [1, 3, 8, 64, 130, 256])
def test_tcgnn_vs_float64(ext, N):
    rng = np.random.default_rng(N)
    M, K = 1200, 900
    # odd N: repeated (row, column) pairs, whose values add up
    rowptr, col = random_csr(rng, M, K, 10.0, empty_frac=0.2, dup_frac=0.5 if N % 2 else 0.0)
    vals = rng.uniform(-1, 1, col.shape[0]).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    say("tcgnn", N)
    rp, cc, bp, e2c, e2r = phase1(ext, rowptr, col, M)
    X = torch.zeros((M, N), device="cuda")  # reference shape: output like input (M rows)
    X[:K] = torch.from_numpy(B).cuda()
    C = ext.run_TCGNNSpMM(X, rp, cc, bp, e2c, e2r, values=torch.from_numpy(vals).cuda())[0].cpu().numpy()
    Bpad = X.cpu().numpy()
    C64, Cabs = oracle.spmm_csr_f64(rowptr, col, vals, Bpad, M, with_abs=True)
    d = np.abs(C.astype(np.float64) - C64)
    m = Cabs > 0
    assert float((d[m] / Cabs[m]).max()) <= 4e-6
    assert float(d[~m].max(initial=0.0)) == 0.0


def test_tcgnn_reference_call_sequence(ext):
    """run_TCGNN_SpMM.py:30-55: CPU preprocess into CPU tensors, then the
    SpMM with X = ones: every row sums to its degree."""
    from dtc.graphs import rmat
    a = rmat(12, 8, seed=4)
    M = a.num_rows
    W = (M + 15) // 16
    bp = torch.zeros(W, dtype=torch.int32)
    e2c = torch.zeros(a.nnz, dtype=torch.int32)
    e2r = torch.zeros(a.nnz, dtype=torch.int32)
    rp, cc = torch.from_numpy(a.rowptr), torch.from_numpy(a.col)
    T = ext.preprocess(cc, rp, M, 16, 8, bp, e2c, e2r)
    assert int(bp.sum()) == T
    assert np.array_equal(e2r.numpy(), np.repeat(np.arange(M), np.diff(a.rowptr)))
    for N in (128, 256):
        X = torch.ones((M, N), device="cuda")
        C = ext.run_TCGNNSpMM(X, rp.cuda(), cc.cuda(), bp.cuda(), e2c.cuda(), e2r.cuda())[0]
        deg = torch.from_numpy(np.diff(a.rowptr).astype(np.float32)).cuda()
        assert torch.equal(C, deg[:, None].expand_as(C))


def test_tcgnn_matches_metcf_kernels(ext):
    from dtc.graphs import rmat
    from dtc.matrix import DTCMatrix
    a = rmat(13, 16, seed=2)
    M = a.num_rows
    A = DTCMatrix(a.rowptr, a.col, a.values, num_rows=M, num_cols=M)
    B = torch.randn((M, 64), device="cuda")
    ref = A.spmm(B, precision="fp32")
    C = ext.run_TCGNNSpMM(B, A.rowptr, A.col, A.block_partition, A.edge_to_column, A.edge_to_row,
                          values=A.values)[0]
    assert torch.allclose(C, ref, rtol=1e-5, atol=1e-6)
