"""The C ABI on the GPU, without the DTCSpMM extension: libdtc.so bound with
ctypes (dtc.capi, the binding a caller outside Python's torch layer uses) on
device buffers from torch's allocator.  The call order is preprocess_gpu's
(DTC-SpMM/DTCSpMM.cpp:313-359): workspace query, phase 1 (window arrays and
the block count on the host), block-sized allocations, phase 2; then
dtc_spmm_windows and dtc_spmm_balanced (DTCSpMM.cpp:117-178).  ME-TCF must be
bit-exact against the oracle, the SpMMs within 4e-6 componentwise of
float64, and the deterministic balanced merge bitwise repeatable."""
import numpy as np
import pytest
import torch

import oracle
from helpers import load_fixture_18, random_csr, say

pytestmark = pytest.mark.gpu


def _cases():
    from dtc.graphs import rmat, uniform_rows
    rng = np.random.default_rng(12)
    rp18, c18, m18, _ = load_fixture_18()
    out = [("fixture18", rp18, c18, m18, rng.uniform(-1, 1, c18.shape[0]).astype(np.float32))]
    a = uniform_rows(4096, 4096, 16, seed=0)
    out.append(("C1", a.rowptr, a.col, a.num_rows, a.values))
    r = rmat(12, 16, seed=1)
    out.append(("rmat12", r.rowptr, r.col, r.num_rows, r.values))
    rp, col = random_csr(rng, 1003, 1003, 9.0, sorted_rows=False, empty_frac=0.3, dup_frac=0.2)
    out.append(("ragged", rp, col, 1003, rng.uniform(-1, 1, col.shape[0]).astype(np.float32)))
    out.append(("no entries", np.zeros(18, np.int32), np.zeros(0, np.int32), 17, np.zeros(0, np.float32)))
    return out


def _dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t if dtype is None else t.to(dtype)).cuda()


def test_capi_metcf_and_spmm():
    from dtc.capi import Library
    lib = Library()
    s = torch.cuda.current_stream().cuda_stream
    i32, u8, f32 = torch.int32, torch.uint8, torch.float32
    for name, rowptr, col, M, vals in _cases():
        nnz = int(col.shape[0])
        W = (M + 15) // 16
        say("capi", name, M, nnz)
        col_d, rp_d, vals_d = _dev(col), _dev(rowptr), _dev(vals)
        ws_bytes = lib.metcf_workspace_size(M, nnz)
        ws = torch.empty(ws_bytes, dtype=u8, device="cuda")
        bp = torch.zeros(W, dtype=i32, device="cuda")
        e2c = torch.empty(nnz, dtype=i32, device="cuda")
        e2r = torch.empty(nnz, dtype=i32, device="cuda")
        rwo = torch.empty(W + 1, dtype=i32, device="cuda")
        tile = torch.empty(nnz, dtype=u8, device="cuda")
        tcv = torch.empty(nnz, dtype=f32, device="cuda")
        T = lib.metcf_build_windows(col_d.data_ptr(), rp_d.data_ptr(), M, nnz, bp.data_ptr(), e2c.data_ptr(),
                                    e2r.data_ptr(), rwo.data_ptr(), tile.data_ptr(), ws.data_ptr(), ws_bytes,
                                    values=vals_d.data_ptr(), tcblock_values=tcv.data_ptr(), stream=s)
        rowid = torch.empty(T, dtype=i32, device="cuda")
        off = torch.empty(T + 1, dtype=i32, device="cuda")
        atox = torch.empty(8 * T, dtype=i32, device="cuda")
        lib.metcf_build_blocks(rp_d.data_ptr(), rwo.data_ptr(), M, nnz, T, rowid.data_ptr(), off.data_ptr(),
                               atox.data_ptr(), ws.data_ptr(), ws_bytes, stream=s)
        # phase 2 again into a SparseAToX 4 bytes past a 16-byte boundary (the
        # build takes any int32-aligned buffer; the SpMMs need 16 bytes)
        atox_odd = torch.full((8 * T + 1,), -7, dtype=i32, device="cuda")
        lib.metcf_build_blocks(rp_d.data_ptr(), rwo.data_ptr(), M, nnz, T, rowid.data_ptr(), off.data_ptr(),
                               atox_odd.data_ptr() + 4, ws.data_ptr(), ws_bytes, stream=s)
        torch.cuda.synchronize()
        assert atox_odd[0].item() == -7 and torch.equal(atox_odd[1:], atox), name
        ref = oracle.metcf(rowptr, col, M, values=vals)
        assert T == ref["block_count"], name
        for k, v in [("RowWindowOffset", rwo), ("TCblockRowid", rowid), ("TCblocktileId", tile),
                     ("TCblockoffset", off), ("SparseAToXindex", atox), ("blockPartition", bp),
                     ("edgeToColumn", e2c), ("edgeToRow", e2r), ("TCblockValues", tcv)]:
            assert np.array_equal(v.cpu().numpy(), ref[k]), (name, k)
        for N in (1, 40, 128):
            Bn = np.random.default_rng(N).standard_normal((M, N)).astype(np.float32)
            B = _dev(Bn)
            C64, Cabs = oracle.spmm_csr_f64(rowptr, col, vals, Bn, M, with_abs=True)
            m = Cabs > 0

            def err(C):
                d = np.abs(C.cpu().numpy().astype(np.float64) - C64)
                return max(float((d[m] / Cabs[m]).max(initial=0.0)), float(d[~m].max(initial=0.0)))
            C = torch.full((M, N), float("nan"), device="cuda")
            lib.spmm_windows(rwo.data_ptr(), tile.data_ptr(), off.data_ptr(), atox.data_ptr(), tcv.data_ptr(),
                             W, T, nnz, M, B.data_ptr(), M, N, C.data_ptr(), stream=s)
            torch.cuda.synchronize()
            assert err(C) <= 4e-6, (name, N, "windows")
            for det in (False, True):
                opt = lib.options(chunk_blocks=3, deterministic=det)
                wsb = lib.spmm_balanced_workspace_size(T, N, opt)
                wsp = torch.empty(max(wsb, 1), dtype=u8, device="cuda")
                outs = []
                for _ in range(2):
                    C = torch.full((M, N), float("nan"), device="cuda")
                    lib.spmm_balanced(rowid.data_ptr(), tile.data_ptr(), off.data_ptr(), atox.data_ptr(),
                                      tcv.data_ptr(), T, nnz, M, B.data_ptr(), M, N, C.data_ptr(), wsp.data_ptr(),
                                      wsb, options=opt, stream=s)
                    torch.cuda.synchronize()
                    outs.append(C)
                assert err(outs[0]) <= 4e-6, (name, N, "balanced", det)
                if det:
                    assert torch.equal(outs[0], outs[1]), (name, N)


def test_capi_errors_are_status_codes():
    """Invalid CSR and a too-small workspace come back as status codes with a
    detail message (the reference would exit or write out of bounds)."""
    from dtc.capi import DTCError, Library
    lib = Library()
    s = torch.cuda.current_stream().cuda_stream
    rowptr = np.array([0, 2, 3, 9], np.int32)  # last entry != nnz
    col = np.array([0, 1, 2], np.int32)
    rp_d, col_d = _dev(rowptr), _dev(col)
    ws_bytes = lib.metcf_workspace_size(3, 3)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    bufs = [torch.zeros(n, dtype=torch.int32, device="cuda") for n in (1, 3, 3, 2)]
    tile = torch.zeros(3, dtype=torch.uint8, device="cuda")
    say("capi invalid CSR")
    with pytest.raises(DTCError) as e:
        lib.metcf_build_windows(col_d.data_ptr(), rp_d.data_ptr(), 3, 3, *[b.data_ptr() for b in bufs],
                                tile.data_ptr(), ws.data_ptr(), ws_bytes, stream=s)
    assert e.value.status == 4 and "nodePointer" in e.value.detail
    with pytest.raises(DTCError) as e:
        lib.metcf_build_windows(col_d.data_ptr(), rp_d.data_ptr(), 3, 3, *[b.data_ptr() for b in bufs],
                                tile.data_ptr(), ws.data_ptr(), ws_bytes - 1, stream=s)
    assert e.value.status == 3


def _capi_build(lib, rowptr, col, M, stream):
    """Phase 1 and phase 2 through the C ABI on `stream` (torch allocations
    on the current stream); returns the ME-TCF arrays on the host."""
    i32, u8 = torch.int32, torch.uint8
    nnz = int(col.shape[0])
    W = (M + 15) // 16
    col_d, rp_d = _dev(col), _dev(rowptr)
    ws_bytes = lib.metcf_workspace_size(M, nnz)
    ws = torch.empty(ws_bytes, dtype=u8, device="cuda")
    bp = torch.zeros(W, dtype=i32, device="cuda")
    e2c = torch.empty(nnz, dtype=i32, device="cuda")
    e2r = torch.empty(nnz, dtype=i32, device="cuda")
    rwo = torch.empty(W + 1, dtype=i32, device="cuda")
    tile = torch.empty(nnz, dtype=u8, device="cuda")
    T = lib.metcf_build_windows(col_d.data_ptr(), rp_d.data_ptr(), M, nnz, bp.data_ptr(), e2c.data_ptr(),
                                e2r.data_ptr(), rwo.data_ptr(), tile.data_ptr(), ws.data_ptr(), ws_bytes,
                                stream=stream)
    rowid = torch.empty(T, dtype=i32, device="cuda")
    off = torch.empty(T + 1, dtype=i32, device="cuda")
    atox = torch.empty(8 * T, dtype=i32, device="cuda")
    lib.metcf_build_blocks(rp_d.data_ptr(), rwo.data_ptr(), M, nnz, T, rowid.data_ptr(), off.data_ptr(),
                           atox.data_ptr(), ws.data_ptr(), ws_bytes, stream=stream)
    torch.cuda.current_stream().synchronize()
    return T, {"RowWindowOffset": rwo, "TCblockRowid": rowid, "TCblocktileId": tile,
               "TCblockoffset": off, "SparseAToXindex": atox, "blockPartition": bp,
               "edgeToColumn": e2c, "edgeToRow": e2r}


def test_capi_metcf_concurrent_threads():
    """Two host threads building ME-TCF at the same time through the C ABI
    (ctypes releases the GIL), each on a stream of its own, on matrices with a
    window of 48,000 entries (the device-wide sort, on the library's side
    stream) beside thousands of small ones (the LDS path): the builds share
    the side stream and the pinned read buffer, and every one of them must be
    bit-exact against the oracle."""
    import threading
    from dtc.capi import Library
    lib = Library()
    rng = np.random.default_rng(5)
    cases = []
    for k in range(2):
        rows = [np.sort(rng.choice(1 << 16, 3000, replace=False)) for _ in range(16)]
        rows += [np.sort(rng.choice(1 << 16, int(rng.integers(1, 9)), replace=False))
                 for _ in range(2000 + 16 * k)]
        M = len(rows)
        rowptr = np.concatenate([[0], np.cumsum([r.shape[0] for r in rows])]).astype(np.int32)
        col = np.concatenate(rows).astype(np.int32)
        cases.append((rowptr, col, M, oracle.metcf(rowptr, col, M)))
    errors = []

    def work(idx):
        try:
            rowptr, col, M, ref = cases[idx]
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for rep in range(4):
                    T, out = _capi_build(lib, rowptr, col, M, s.cuda_stream)
                    assert T == ref["block_count"], (idx, rep)
                    for k, v in out.items():
                        assert np.array_equal(v.cpu().numpy(), ref[k]), (idx, rep, k)
        except Exception as e:  # reported by the main thread
            errors.append(f"thread {idx}: {e!r}")

    say("capi", "two threads, four builds each")
    threads = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a build thread did not finish"
    assert not errors, errors
