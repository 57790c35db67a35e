"""The allocation-free balanced SpMM (dtc_spmm_balanced_noalloc) through the
C ABI: its output against dtc_spmm_balanced's with a caller workspace, on two
calls in a row (the second with a smaller matrix and N, so stale scratch
contents from the first would show), with the atomic merge (both within
4e-6 componentwise of float64 and of each other) and the deterministic merge
(bitwise equal), plain and persistent; and a request beyond the fixed
scratch rejected before any work."""
import numpy as np
import pytest
import torch

import oracle
from helpers import random_csr, say

pytestmark = pytest.mark.gpu


def _build(rowptr, col, M, vals):
    from dtc.matrix import DTCMatrix
    return DTCMatrix(rowptr, col, vals, num_rows=M, num_cols=M)


def _cases():
    from dtc.graphs import rmat
    r = rmat(12, 16, seed=1)
    rng = np.random.default_rng(5)
    rp, col = random_csr(rng, 1003, 1003, 9.0, sorted_rows=False, empty_frac=0.3)
    return [("rmat12", r.rowptr, r.col, r.num_rows, r.values, 96),
            ("ragged", rp, col, 1003, rng.uniform(-1, 1, col.shape[0]).astype(np.float32), 17)]


def test_noalloc_matches_workspace_entry():
    from dtc.capi import Library
    lib = Library()
    s = torch.cuda.current_stream().cuda_stream
    for name, rowptr, col, M, vals, N in _cases():  # second call: smaller input
        A = _build(rowptr, col, M, vals)
        T, nnz = A.block_count, A.nnz
        Bn = np.random.default_rng(N).standard_normal((M, N)).astype(np.float32)
        B = torch.from_numpy(Bn).cuda()
        C64, Cabs = oracle.spmm_csr_f64(rowptr, col, vals, Bn, M, with_abs=True)
        m = Cabs > 0

        def err(C):
            d = np.abs(C.cpu().numpy().astype(np.float64) - C64)
            return max(float((d[m] / Cabs[m]).max(initial=0.0)), float(d[~m].max(initial=0.0)))

        args = (A.tcblock_rowid.data_ptr(), A.tcblocktile_id.data_ptr(), A.tcblock_offset.data_ptr(),
                A.sparse_atox.data_ptr(), A.values_tc.data_ptr(), T, nnz, M, B.data_ptr(), M, N)
        for det in (False, True):
            for persistent in (False, True):
                # chunks whose workspace fits the 8 MiB scratch (the deterministic
                # merge's partial tiles take 2 x 16 x 128 floats per chunk here)
                for chunk in ((16, 64) if det else (2, 16)):
                    say("noalloc", name, det, persistent, chunk)
                    opt = lib.options(chunk_blocks=chunk, deterministic=det, persistent=persistent)
                    wsb = lib.spmm_balanced_workspace_size(T, N, opt)
                    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
                    C_ws = torch.full((M, N), float("nan"), device="cuda")
                    lib.spmm_balanced(*args, C_ws.data_ptr(), ws.data_ptr(), wsb, options=opt, stream=s)
                    C_na = torch.full((M, N), float("nan"), device="cuda")
                    lib.spmm_balanced_noalloc(*args, C_na.data_ptr(), options=opt, stream=s)
                    torch.cuda.synchronize()
                    assert err(C_na) <= 4e-6 and err(C_ws) <= 4e-6, (name, det, persistent, chunk)
                    if det:
                        assert torch.equal(C_na, C_ws), (name, persistent, chunk)
                    else:
                        scale = torch.from_numpy(Cabs.astype(np.float32)).cuda()
                        assert bool(((C_na - C_ws).abs() <= 4e-6 * scale + 1e-30).all()), (name, chunk)


def test_noalloc_rejects_oversized_request():
    """The deterministic merge of a large SpMM needs more than the 8 MiB
    scratch: refused with DTC_ERROR_UNSUPPORTED on the host, nothing launched
    (the output keeps its contents)."""
    from dtc.capi import DTCError, Library
    from dtc.graphs import rmat
    lib = Library()
    r = rmat(12, 16, seed=1)
    A = _build(r.rowptr, r.col, r.num_rows, r.values)
    N = 512
    opt = lib.options(chunk_blocks=1, deterministic=True)
    assert lib.spmm_balanced_workspace_size(A.block_count, N, opt) > (8 << 20)
    B = torch.zeros((r.num_rows, N), device="cuda")
    C = torch.full((r.num_rows, N), 7.0, device="cuda")
    say("noalloc oversized")
    with pytest.raises(DTCError) as e:
        lib.spmm_balanced_noalloc(A.tcblock_rowid.data_ptr(), A.tcblocktile_id.data_ptr(),
                                  A.tcblock_offset.data_ptr(), A.sparse_atox.data_ptr(), None, A.block_count,
                                  A.nnz, r.num_rows, B.data_ptr(), r.num_rows, N, C.data_ptr(), options=opt,
                                  stream=torch.cuda.current_stream().cuda_stream)
    assert e.value.status == 2 and "8388608" in e.value.detail
    torch.cuda.synchronize()
    assert bool((C == 7.0).all())
