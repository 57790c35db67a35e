"""numpy front end of the CPU oracle (oracle/metcf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product path
never imports this module.

Each function mirrors one restated piece of the reference; the C file's
header lists the reference lines each follows and the checks that pin it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DTC_ORACLE_LIBRARY selects another build of the same source (the sanitized
# one under _build/san/, see the Makefile's `sanitize` target).
_LIB_PATH = os.environ.get("DTC_ORACLE_LIBRARY") or os.path.join(_HERE, "_build", "libmetcf_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p


def build() -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    i32, i64, c_int = ctypes.c_int32, ctypes.c_int64, ctypes.c_int
    lib.oracle_metcf_windows.argtypes = [i32, i32, _i32p, _i32p, i32, i32, _i32p, _i32p,
                                         _i32p, _i32p, ctypes.POINTER(i64), c_int]
    lib.oracle_metcf_windows.restype = c_int
    lib.oracle_metcf_blocks.argtypes = [i32, i32, _i32p, _i32p, _i32p, _i32p, _i32p, i32, i32,
                                        i64, _vp, _i32p, _u8p, _i32p, _i32p, _vp, c_int]
    lib.oracle_metcf_blocks.restype = c_int
    lib.oracle_spmm_csr_f64.argtypes = [i32, _i32p, _i32p, _vp, _f32p, i32, i32, _f64p, _vp, c_int]
    lib.oracle_spmm_csr_f64.restype = None
    lib.oracle_spmm_metcf_f64.argtypes = [i32, i32, _i32p, _u8p, _i32p, _i32p, _vp, i32, i32,
                                          _f32p, i32, i32, _f64p, c_int]
    lib.oracle_spmm_metcf_f64.restype = c_int
    lib.oracle_spmm_metcf_f32.argtypes = [i32, i32, _i32p, _u8p, _i32p, _i32p, _vp, i32, i32,
                                          _f32p, i32, i32, _f32p, c_int]
    lib.oracle_spmm_metcf_f32.restype = None
    lib.oracle_sddmm_csr_f64.argtypes = [i32, _i32p, _i32p, _f32p, _f32p, i32, i32, _f64p, c_int]
    lib.oracle_sddmm_csr_f64.restype = None
    lib.oracle_csr_transpose.argtypes = [i32, i32, _i32p, _i32p, _i32p, _i32p, _i32p]
    lib.oracle_csr_transpose.restype = c_int
    _lib = lib
    return lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def metcf(rowptr, col, num_nodes, blk_h=16, blk_w=8, values=None, block_partition=None,
          nthreads=0):
    """Full ME-TCF build (both phases) as the reference's ``preprocess_gpu``.

    Returns a dict with RowWindowOffset, TCblockRowid, TCblocktileId,
    TCblockoffset, SparseAToXindex, block_count, the in/out buffers
    blockPartition, edgeToColumn, edgeToRow and, when ``values`` is given,
    TCblockValues (values in TC-block order).
    """
    lib = _load()
    rowptr, col = _i32(rowptr), _i32(col)
    M = int(num_nodes)
    nnz = int(col.shape[0])
    W = (M + blk_h - 1) // blk_h
    bp = np.zeros(W, np.int32) if block_partition is None else _i32(block_partition).copy()
    e2c = np.zeros(nnz, np.int32)
    e2r = np.zeros(nnz, np.int32)
    rwo = np.zeros(W + 1, np.int32)
    T = ctypes.c_int64(0)
    rc = lib.oracle_metcf_windows(M, nnz, rowptr, col, blk_h, blk_w, bp, e2c, e2r, rwo,
                                  ctypes.byref(T), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_metcf_windows failed ({rc})")
    T = int(T.value)
    rowid = np.zeros(T, np.int32)
    tile = np.zeros(nnz, np.uint8)
    off = np.zeros(T + 1, np.int32)
    atox = np.zeros(T * blk_w, np.int32)
    vals = None if values is None else _f32(values)
    tcv = None if values is None else np.zeros(nnz, np.float32)
    rc = lib.oracle_metcf_blocks(M, nnz, rowptr, col, e2c, e2r, rwo, blk_h, blk_w, T,
                                 _ptr(vals), rowid, tile, off, atox, _ptr(tcv), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_metcf_blocks failed ({rc})")
    out = dict(RowWindowOffset=rwo, TCblockRowid=rowid, TCblocktileId=tile, TCblockoffset=off,
               SparseAToXindex=atox, block_count=T, blockPartition=bp, edgeToColumn=e2c,
               edgeToRow=e2r)
    if tcv is not None:
        out["TCblockValues"] = tcv
    return out


def metcf_windows_only(rowptr, col, num_nodes, blk_h=16, blk_w=8, nthreads=0):
    """Phase 1 only (timed separately by the CPU baseline)."""
    lib = _load()
    rowptr, col = _i32(rowptr), _i32(col)
    M, nnz = int(num_nodes), int(col.shape[0])
    W = (M + blk_h - 1) // blk_h
    bp = np.zeros(W, np.int32)
    e2c = np.zeros(nnz, np.int32)
    e2r = np.zeros(nnz, np.int32)
    rwo = np.zeros(W + 1, np.int32)
    T = ctypes.c_int64(0)
    rc = lib.oracle_metcf_windows(M, nnz, rowptr, col, blk_h, blk_w, bp, e2c, e2r, rwo,
                                  ctypes.byref(T), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_metcf_windows failed ({rc})")
    return rwo, int(T.value)


def spmm_csr_f64(rowptr, col, values, B, num_nodes, with_abs=False, nthreads=0):
    """float64 C = A * B for CSR A (values None means all ones)."""
    lib = _load()
    rowptr, col, B = _i32(rowptr), _i32(col), _f32(B)
    M = int(num_nodes)
    K, N = B.shape
    vals = None if values is None else _f32(values)
    C = np.empty((M, N), np.float64)
    Cabs = np.empty((M, N), np.float64) if with_abs else None
    lib.oracle_spmm_csr_f64(M, rowptr, col, _ptr(vals), B, K, N, C, _ptr(Cabs), nthreads)
    return (C, Cabs) if with_abs else C


def spmm_metcf_f64(m, B, num_nodes, blk_h=16, blk_w=8, tc_values=None, nthreads=0):
    """float64 C = A * B with A read from an ME-TCF dict (see ``metcf``)."""
    lib = _load()
    B = _f32(B)
    K, N = B.shape
    M = int(num_nodes)
    W = m["RowWindowOffset"].shape[0] - 1
    C = np.empty((M, N), np.float64)
    tcv = None if tc_values is None else _f32(tc_values)
    rc = lib.oracle_spmm_metcf_f64(M, W, _i32(m["RowWindowOffset"]),
                                   np.ascontiguousarray(m["TCblocktileId"], np.uint8),
                                   _i32(m["TCblockoffset"]), _i32(m["SparseAToXindex"]), _ptr(tcv),
                                   blk_h, blk_w, B, K, N, C, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle_spmm_metcf_f64 failed ({rc})")
    return C


def spmm_metcf_f32(m, B, num_nodes, blk_h=16, blk_w=8, tc_values=None, nthreads=0, out=None):
    """float32 CPU port of the ME-TCF SpMM (the CPU baseline)."""
    lib = _load()
    B = _f32(B)
    K, N = B.shape
    M = int(num_nodes)
    W = m["RowWindowOffset"].shape[0] - 1
    C = np.empty((M, N), np.float32) if out is None else out
    tcv = None if tc_values is None else _f32(tc_values)
    lib.oracle_spmm_metcf_f32(M, W, _i32(m["RowWindowOffset"]),
                              np.ascontiguousarray(m["TCblocktileId"], np.uint8),
                              _i32(m["TCblockoffset"]), _i32(m["SparseAToXindex"]), _ptr(tcv),
                              blk_h, blk_w, B, K, N, C, nthreads)
    return C


def sddmm_csr_f64(rowptr, col, X, Y, num_nodes, nthreads=0):
    """out[e] = X[row(e)] . Y[col(e)] in float64."""
    lib = _load()
    rowptr, col, X, Y = _i32(rowptr), _i32(col), _f32(X), _f32(Y)
    K, N = Y.shape
    out = np.empty(col.shape[0], np.float64)
    lib.oracle_sddmm_csr_f64(int(num_nodes), rowptr, col, X, Y, K, N, out, nthreads)
    return out


def csr_transpose(rowptr, col, num_rows, num_cols):
    """Transpose CSR; returns (rowptr_t, col_t, perm) with vals_t = vals[perm]."""
    lib = _load()
    rowptr, col = _i32(rowptr), _i32(col)
    rpt = np.empty(int(num_cols) + 1, np.int32)
    ct = np.empty(col.shape[0], np.int32)
    perm = np.empty(col.shape[0], np.int32)
    rc = lib.oracle_csr_transpose(int(num_rows), int(num_cols), rowptr, col, rpt, ct, perm)
    if rc != 0:
        raise RuntimeError(f"oracle_csr_transpose failed ({rc})")
    return rpt, ct, perm
