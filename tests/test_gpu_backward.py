"""Backward pass: GPU CSR transpose and SDDMM vs the oracle, and the
autograd gradients of C = A(values) * B against float64 references (the
oracle, a torch.sparse float64 product, and finite differences of the
float64 function on a 2,048-node R-MAT instance)."""
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


def test_transpose_vs_oracle(ext):
    rng = np.random.default_rng(31)
    for case, (M, K) in enumerate([(1, 1), (50, 70), (333, 100), (1000, 4000)]):
        rowptr, col = random_csr(rng, M, K, 6.0, sorted_rows=bool(case % 2), empty_frac=0.2)
        say("transpose", M, K)
        rpt, ct, perm = ext.csr_transpose(torch.from_numpy(rowptr).cuda(), torch.from_numpy(col).cuda(), M, K)
        r2, c2, p2 = oracle.csr_transpose(rowptr, col, M, K)
        assert np.array_equal(rpt.cpu().numpy(), r2)
        assert np.array_equal(ct.cpu().numpy(), c2)
        assert np.array_equal(perm.cpu().numpy(), p2)


@pytest.mark.parametrize("N", [1, 3, 16, 64, 130, 256, 300, 512])
@pytest.mark.parametrize("with_rows", [False, True])
@pytest.mark.parametrize("y_bf16", [False, True])
def test_sddmm_vs_oracle(ext, N, with_rows, y_bf16):
    """Edge-parallel SDDMM vs float64, rows from edgeToRow or from the row
    pointer's binary search, Y float32 or bfloat16 (taken exactly); a
    3,000-entry hub row, empty rows, nnz not a multiple of 8."""
    rng = np.random.default_rng(N)
    M, K = 700, 3500
    rowptr, col = random_csr(rng, M, K, 8.0, empty_frac=0.1)
    hub = np.sort(rng.choice(K, 3000, replace=False)).astype(np.int32)  # row 5 becomes a hub
    cols = [col[rowptr[r]:rowptr[r + 1]] for r in range(M)]
    cols[5] = hub
    if sum(len(c) for c in cols) % 8 == 0:
        cols[5] = hub[:-1]
    rowptr = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int32)
    col = np.concatenate(cols).astype(np.int32)
    X = rng.standard_normal((M, N)).astype(np.float32)
    Yt = torch.from_numpy(rng.standard_normal((K, N)).astype(np.float32))
    if y_bf16:
        Yt = Yt.to(torch.bfloat16)
    Y = Yt.float().numpy()
    say("sddmm", N, with_rows, y_bf16)
    e2r = torch.from_numpy(np.repeat(np.arange(M), np.diff(rowptr)).astype(np.int32)).cuda()
    out = ext.sddmm_csr(torch.from_numpy(rowptr).cuda(), torch.from_numpy(col).cuda(),
                        torch.from_numpy(X).cuda(), Yt.cuda(),
                        e2r if with_rows else None).cpu().numpy()
    ref = oracle.sddmm_csr_f64(rowptr, col, X, Y, M)
    rows = np.repeat(np.arange(M), np.diff(rowptr))
    scale = (np.abs(X[rows]) * np.abs(Y[col])).sum(1)
    assert np.all(np.abs(out - ref) <= 1e-6 * scale + 1e-30)


def _grads_f64(rowptr, col, vals, B, dC, M, K):
    """dvalues and dB of <C, dC> with C = A B, via a torch.sparse float64 product."""
    rows = np.repeat(np.arange(M), np.diff(rowptr))
    idx = torch.from_numpy(np.stack([rows, col]).astype(np.int64))
    v = torch.from_numpy(vals.astype(np.float64)).requires_grad_(True)
    Bt = torch.from_numpy(B.astype(np.float64)).requires_grad_(True)
    A = torch.sparse_coo_tensor(idx, v, (M, K))
    C = torch.sparse.mm(A, Bt)
    (C * torch.from_numpy(dC.astype(np.float64))).sum().backward()
    return v.grad.numpy(), Bt.grad.numpy(), C.detach().numpy()


@pytest.mark.parametrize("schedule", ["windows", "balanced"])
def test_autograd_square_and_rect(ext, schedule):
    from dtc.matrix import DTCMatrix, spmm_autograd
    rng = np.random.default_rng(41)
    for M, K, N in [(300, 300, 32), (257, 400, 20), (400, 130, 64)]:
        rowptr, col = random_csr(rng, M, K, 7.0, sorted_rows=False, empty_frac=0.15)
        vals = rng.uniform(-1, 1, col.shape[0]).astype(np.float32)
        B = rng.standard_normal((K, N)).astype(np.float32)
        dC = rng.standard_normal((M, N)).astype(np.float32)
        say("autograd", schedule, M, K, N)
        A = DTCMatrix(rowptr, col, None, num_rows=M, num_cols=K)
        v = torch.from_numpy(vals).cuda().requires_grad_(True)
        Bt = torch.from_numpy(B).cuda().requires_grad_(True)
        C = spmm_autograd(v, Bt, A, schedule=schedule)
        C.backward(torch.from_numpy(dC).cuda())
        gv, gB, C64 = _grads_f64(rowptr, col, vals, B, dC, M, K)
        assert np.allclose(C.detach().cpu().numpy(), C64, rtol=1e-4, atol=1e-5)
        assert np.allclose(v.grad.cpu().numpy(), gv, rtol=1e-4, atol=1e-4)
        assert np.allclose(Bt.grad.cpu().numpy(), gB, rtol=1e-4, atol=1e-5)


def test_gradients_rmat_2048_finite_differences(ext):
    """GCN-style instance: R-MAT scale 11 (2,048 nodes), symmetric
    normalisation.  Analytic float32 gradients vs central finite differences
    of the float64 function along random directions."""
    from dtc.graphs import rmat
    from dtc.matrix import DTCMatrix, spmm_autograd
    a = rmat(11, 16, seed=1)
    M = a.num_rows
    N = 64
    rng = np.random.default_rng(5)
    B = rng.standard_normal((M, N)).astype(np.float32)
    dC = rng.standard_normal((M, N)).astype(np.float32)
    say("rmat2048 gradients")
    A = DTCMatrix(a.rowptr, a.col, None, num_rows=M, num_cols=M)
    v = torch.from_numpy(a.values).cuda().requires_grad_(True)
    Bt = torch.from_numpy(B).cuda().requires_grad_(True)
    C = spmm_autograd(v, Bt, A)
    C.backward(torch.from_numpy(dC).cuda())
    gv = v.grad.double().cpu().numpy()
    gB = Bt.grad.double().cpu().numpy()

    def f(vals, Bm):
        C64 = oracle.spmm_csr_f64(a.rowptr, a.col, vals.astype(np.float32), Bm.astype(np.float32), M)
        return float((C64 * dC).sum())

    # f is bilinear, so central differences are exact up to float64 rounding.
    for trial in range(3):
        dv = rng.standard_normal(a.nnz)
        dB = rng.standard_normal(B.shape)
        h = 1e-3
        vals64 = a.values.astype(np.float64)
        B64 = B.astype(np.float64)
        fd_v = (f(vals64 + h * dv, B64) - f(vals64 - h * dv, B64)) / (2 * h)
        fd_B = (f(vals64, B64 + h * dB) - f(vals64, B64 - h * dB)) / (2 * h)
        an_v = float(gv @ dv)
        an_B = float((gB * dB).sum())
        assert abs(fd_v - an_v) <= 1e-3 * max(1.0, abs(fd_v)), (trial, fd_v, an_v)
        assert abs(fd_B - an_B) <= 1e-3 * max(1.0, abs(fd_B)), (trial, fd_B, an_B)


@pytest.mark.parametrize("schedule", ["windows", "balanced"])
def test_autograd_b_rows_differ_from_num_cols(ext, schedule):
    """num_cols inferred as max column + 1 (fewer rows than B), and a B with
    fewer rows than num_cols (columns past B's rows read as zero): dB always
    has B's shape, rows past the last column of A get zero gradient."""
    from dtc.matrix import DTCMatrix, spmm_autograd
    rng = np.random.default_rng(43)
    M, K, N = 200, 300, 24
    rowptr, col = random_csr(rng, M, 250, 5.0, empty_frac=0.1)  # columns < 250 <= K
    vals = rng.uniform(-1, 1, col.shape[0]).astype(np.float32)
    dC = rng.standard_normal((M, N)).astype(np.float32)
    for num_cols, b_rows in [(None, K), (K, 180)]:
        say("autograd b_rows", schedule, num_cols, b_rows)
        B = rng.standard_normal((b_rows, N)).astype(np.float32)
        A = DTCMatrix(rowptr, col, None, num_rows=M, num_cols=num_cols)
        v = torch.from_numpy(vals).cuda().requires_grad_(True)
        Bt = torch.from_numpy(B).cuda().requires_grad_(True)
        spmm_autograd(v, Bt, A, schedule=schedule).backward(torch.from_numpy(dC).cuda())
        assert Bt.grad.shape == Bt.shape
        keep = col < b_rows  # entries whose column exists in B
        kc = int(max(b_rows, col.max() + 1))
        Bpad = np.zeros((kc, N), np.float32)
        Bpad[:b_rows] = B
        rp = np.concatenate([[0], np.cumsum([keep[rowptr[r]:rowptr[r + 1]].sum() for r in range(M)])])
        gv, gB, _ = _grads_f64(rp.astype(np.int32), col[keep], vals[keep], Bpad, dC, M, kc)
        assert np.allclose(Bt.grad.cpu().numpy(), gB[:b_rows], rtol=1e-4, atol=1e-5)
        assert np.allclose(v.grad.cpu().numpy()[keep], gv, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("schedule", ["windows", "balanced"])
def test_autograd_duplicate_entries(ext, schedule):
    """Repeated (row, column) pairs: the forward adds them, as the CSR
    product does, and each copy gets its own gradient."""
    from dtc.matrix import DTCMatrix, spmm_autograd
    rng = np.random.default_rng(44)
    M, N = 300, 40
    rowptr, col = random_csr(rng, M, M, 6.0, sorted_rows=False, dup_frac=0.6)
    vals = rng.uniform(-1, 1, col.shape[0]).astype(np.float32)
    B = rng.standard_normal((M, N)).astype(np.float32)
    dC = rng.standard_normal((M, N)).astype(np.float32)
    say("autograd duplicates", schedule, col.shape[0])
    A = DTCMatrix(rowptr, col, None, num_rows=M, num_cols=M)
    v = torch.from_numpy(vals).cuda().requires_grad_(True)
    Bt = torch.from_numpy(B).cuda().requires_grad_(True)
    C = spmm_autograd(v, Bt, A, schedule=schedule)
    C.backward(torch.from_numpy(dC).cuda())
    gv, gB, C64 = _grads_f64(rowptr, col, vals, B, dC, M, M)  # torch.sparse coalesces by adding
    assert np.allclose(C.detach().cpu().numpy(), C64, rtol=1e-4, atol=1e-5)
    assert np.allclose(v.grad.cpu().numpy(), gv, rtol=1e-4, atol=1e-4)
    assert np.allclose(Bt.grad.cpu().numpy(), gB, rtol=1e-4, atol=1e-5)


def test_reference_backward_names(ext):
    """The reference's backward names (DTCSpMM.cpp:386-393) on a non-symmetric
    matrix with values: without `transposed` they return A dY, bitwise the
    forward call's result, as the reference does; with A^T's operands
    (DTCMatrix.transposed_operands) they return the true gradient
    dB = A^T dY, against the float64 product with the transposed CSR (float32
    tolerance 4e-6 componentwise)."""
    from dtc.matrix import DTCMatrix
    rng = np.random.default_rng(41)
    M = 900
    rowptr, col = random_csr(rng, M, M, 8.0, sorted_rows=True, empty_frac=0.1)
    rows = [list(col[rowptr[r]:rowptr[r + 1]]) for r in range(M)]
    rows[5] = sorted(set(rows[5]) | set(range(0, M, 2)))  # a long row: windows of many blocks
    rowptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate([np.array(r, np.int32) for r in rows]).astype(np.int32)
    vals = rng.uniform(-1, 1, col.shape[0]).astype(np.float32)
    A = DTCMatrix(rowptr, col, vals, num_rows=M, num_cols=M)
    dYn = rng.standard_normal((M, 48)).astype(np.float32)
    dY = torch.from_numpy(dYn).cuda()
    # float64 A^T dY through the transposed CSR
    rpt, ct, perm = oracle.csr_transpose(rowptr, col, M, M)
    want, wabs = oracle.spmm_csr_f64(rpt, ct, vals[perm], dYn, M, with_abs=True)

    def comp(C):
        C = C.cpu().numpy().astype(np.float64)
        mask = wabs > 0
        return float((np.abs(C - want)[mask] / wabs[mask]).max())

    nnz = A.nnz
    w_args = (A.row_window_offset, A.tcblocktile_id, A.tcblock_offset, A.sparse_atox, M, nnz)
    b_args = (A.tcblock_rowid, A.tcblocktile_id, A.tcblock_offset, A.sparse_atox, M)
    say("backward names: aliases")
    (fwd,) = ext.run_DTCSpMM(dY, *w_args, "float4_split", values=A.values_tc)
    (alias,) = ext.backward_ptx_uint8_improved(dY, *w_args, "float4_split", values=A.values_tc)
    assert torch.equal(fwd, alias)
    (fwd_b,) = ext.run_DTCSpMM_balance(dY, *b_args, "float4_split", values=A.values_tc, chunk_blocks=8,
                                       deterministic=True)
    (alias_b,) = ext.balance_backward_ptx_uint8_prefetch(dY, *b_args, "float4_split", values=A.values_tc,
                                                         chunk_blocks=8, deterministic=True)
    assert torch.equal(fwd_b, alias_b)
    say("backward names: true gradients")
    tw = A.transposed_operands("windows")
    tb = A.transposed_operands("balanced")
    (g1,) = ext.backward_ptx_uint8_improved(dY, *w_args, "float4_split", values=A.values_tc, transposed=tw)
    (g2,) = ext.backward_ptx_uint8_improved_gcn(dY, *w_args, values=A.values_tc, transposed=tw)
    (g3,) = ext.balance_backward_ptx_uint8_prefetch(dY, *b_args, "float4_split", values=A.values_tc,
                                                    chunk_blocks=8, transposed=tb)
    for g in (g1, g2, g3):
        assert g.shape == (M, 48)
        assert comp(g) <= 4e-6
    # the forward product is not the gradient here (A is not symmetric)
    assert comp(fwd) > 1e-2
    with pytest.raises(ValueError, match="transposed"):
        ext.backward_ptx_uint8_improved(dY, *w_args, "float4_split", transposed=(tw[0],))
