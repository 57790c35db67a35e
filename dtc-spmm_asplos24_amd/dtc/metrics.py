"""Error metrics and the algorithmic byte/FLOP model of the ME-TCF SpMM.

Byte model (per launch, fixed per matrix and N, independent of the variant):
    4*N*sum_w U_w   gathered B rows (each window reads each unique column once)
  + 4*M*N           C written once
  + 5*nnz           uint8 tile id + float32 value per stored element
  + 36*T            8 SparseAToX entries + 1 block offset (int32) per TC block
  + 4*(W+1)         row-window offsets
FLOPs: 2*nnz*N, the reference's throughput numerator (DTCSpMM_kernel.cu:934).
"""
from __future__ import annotations

import numpy as np


def gather_model_bytes(sum_unique: int, M: int, N: int, nnz: int, T: int, W: int) -> int:
    return 4 * N * int(sum_unique) + 4 * M * N + 5 * nnz + 36 * int(T) + 4 * (W + 1)


def compulsory_bytes(M: int, K: int, N: int, nnz: int) -> int:
    """Ideal-cache traffic: CSR once, B once, C once."""
    return 8 * nnz + 4 * (M + 1) + 4 * K * N + 4 * M * N


def spmm_flops(nnz: int, N: int) -> int:
    return 2 * int(nnz) * int(N)


def error_report(C, C64, Cabs64) -> dict:
    """Errors of C against the float64 product C64.

    componentwise: max |C - C64| / (|A| |B|) over outputs with a non-zero
        denominator (|A| |B| is the float64 sum of |a| |b| per output);
    normwise:      max |C - C64| / max |C64|.
    """
    C = np.asarray(C, dtype=np.float64)
    diff = np.abs(C - C64)
    den = np.asarray(Cabs64, dtype=np.float64)
    mask = den > 0
    comp = float((diff[mask] / den[mask]).max()) if mask.any() else 0.0
    if (~mask).any():
        comp = max(comp, float(diff[~mask].max()) if diff[~mask].size else 0.0)
    mx = float(np.abs(C64).max()) if C64.size else 0.0
    norm = float(diff.max() / mx) if mx > 0 else float(diff.max() if diff.size else 0.0)
    return {"componentwise": comp, "normwise": norm}


SPMM_SOURCES = ("csrc/spmm.hip", "csrc/dtc_internal.h")


def spmm_source_sha16() -> str:
    """First 16 hex digits of the SHA-256 over the SpMM kernel's sources:
    the stamp that ties a counter measurement (HBM traffic per launch) to the
    kernels it was taken on."""
    import hashlib
    import os
    pkg = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    h = hashlib.sha256()
    for rel in SPMM_SOURCES:
        with open(os.path.join(pkg, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]
