"""ctypes binding of libdtc's C ABI (include/dtc.h) for callers without
torch or the DTCSpMM extension.

Device buffers are passed as integer addresses from any HIP allocator
(hipMalloc, or ``tensor.data_ptr()`` of a torch tensor); streams as integer
hipStream_t handles (0 = the default stream).  Every call raises
``DTCError`` carrying the library's status and detail message when the
status is not DTC_SUCCESS.

This is the binding a reference-side maintainer would add: it replaces the
pybind operators of DTC-SpMM/DTCSpMM.cpp:313-359 (preprocess_gpu) and
:117-178 (run_DTCSpMM, run_DTCSpMM_balance) with plain C calls.
"""
from __future__ import annotations

import ctypes
import os

# DTC_LIBRARY selects another build of the library (e.g. the sanitized host
# build that build.py --sanitize writes under _build_san/).
LIB_PATH = os.environ.get("DTC_LIBRARY") or os.path.join(
    os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "libdtc.so")

i32, i64, vp, sz = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t

PRECISIONS = {"fp32": 0, "bf16": 1, "fp16": 2, "bf16x3": 3, "bf16x6": 4}


class SpmmOptions(ctypes.Structure):
    """dtc_spmm_options_t"""
    _fields_ = [("precision", i32), ("vec", i32), ("subtiles", i32), ("chunk_blocks", i32),
                ("persistent", i32), ("b_dtype", i32), ("deterministic", i32),
                ("cut_span_blocks", i32)]


class DTCError(RuntimeError):
    def __init__(self, fn: str, status: int, detail: str):
        super().__init__(f"{fn}: status {status} ({detail})")
        self.status = status
        self.detail = detail


_SIGS = {
    "dtc_status_string": ([i32], ctypes.c_char_p),
    "dtc_last_error": ([], ctypes.c_char_p),
    "dtc_version": ([], i32),
    "dtc_metcf_workspace_size": ([i32, i32, i32, i32, ctypes.POINTER(sz)], i32),
    "dtc_metcf_build_windows": ([vp, vp, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp,
                                 ctypes.POINTER(i32), vp, sz, vp], i32),
    "dtc_metcf_build_blocks": ([vp, vp, i32, i32, i32, i32, i32, vp, vp, vp, vp, sz, vp], i32),
    "dtc_spmm_windows": ([vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, i32, i32, vp,
                          ctypes.POINTER(SpmmOptions), vp], i32),
    "dtc_spmm_balanced_workspace_size": ([i32, ctypes.POINTER(sz)], i32),
    "dtc_spmm_balanced_workspace_size_ex": ([i32, i32, ctypes.POINTER(SpmmOptions), ctypes.POINTER(sz)], i32),
    "dtc_spmm_balanced": ([vp, vp, vp, vp, vp, i32, i32, i32, vp, i32, i32, vp,
                           ctypes.POINTER(SpmmOptions), vp, sz, vp], i32),
    "dtc_spmm_balanced_noalloc": ([vp, vp, vp, vp, vp, i32, i32, i32, vp, i32, i32, vp,
                                   ctypes.POINTER(SpmmOptions), vp], i32),
}


class Library:
    """The loaded libdtc.so with typed entry points (see include/dtc.h)."""

    def __init__(self, path: str = LIB_PATH):
        self.lib = ctypes.CDLL(path)
        for name, (args, res) in _SIGS.items():
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = res

    def _check(self, name: str, status: int) -> None:
        if status != 0:
            raise DTCError(name, status, self.lib.dtc_last_error().decode())

    def call(self, name: str, *args) -> None:
        """Calls an entry point by name and raises on a non-success status."""
        self._check(name, getattr(self.lib, name)(*args))

    # -------------------------------------------------------- ME-TCF build
    def metcf_workspace_size(self, num_nodes: int, nnz: int, blk_h: int = 16, blk_w: int = 8) -> int:
        out = sz(0)
        self.call("dtc_metcf_workspace_size", num_nodes, nnz, blk_h, blk_w, ctypes.byref(out))
        return out.value

    def metcf_build_windows(self, edge_list, node_pointer, num_nodes, nnz, block_partition, edge_to_column,
                            edge_to_row, row_window_offset, tcblocktile_id, workspace, workspace_bytes,
                            values=None, tcblock_values=None, tcblock_edge=None, stream=0,
                            blk_h=16, blk_w=8) -> int:
        """Phase 1 (every output sized by the CSR); returns the TC-block count."""
        t = i32(0)
        self.call("dtc_metcf_build_windows", edge_list, node_pointer, num_nodes, nnz, blk_h, blk_w, values,
                  block_partition, edge_to_column, edge_to_row, row_window_offset, tcblocktile_id,
                  tcblock_values, tcblock_edge, ctypes.byref(t), workspace, workspace_bytes, stream)
        return t.value

    def metcf_build_blocks(self, node_pointer, row_window_offset, num_nodes, nnz, block_count,
                           tcblock_rowid, tcblock_offset, sparse_atox, workspace, workspace_bytes, stream=0,
                           blk_h=16, blk_w=8) -> None:
        """Phase 2 (the outputs sized by the block count)."""
        self.call("dtc_metcf_build_blocks", node_pointer, row_window_offset, num_nodes, nnz, blk_h, blk_w,
                  block_count, tcblock_rowid, tcblock_offset, sparse_atox, workspace, workspace_bytes, stream)

    # ----------------------------------------------------------------- SpMM
    @staticmethod
    def options(precision="fp32", vec=0, subtiles=0, chunk_blocks=0, persistent=False, b_dtype=0,
                deterministic=False, cut_span_blocks=0) -> SpmmOptions:
        return SpmmOptions(PRECISIONS[precision], vec, subtiles, chunk_blocks, int(persistent), b_dtype,
                           int(deterministic), cut_span_blocks)

    def spmm_windows(self, row_window_offset, tcblocktile_id, tcblock_offset, sparse_atox, tcblock_values,
                     num_windows, num_blocks, nnz, num_nodes, b, b_rows, n, c, options=None, stream=0) -> None:
        opt = options or self.options()
        self.call("dtc_spmm_windows", row_window_offset, tcblocktile_id, tcblock_offset, sparse_atox,
                  tcblock_values, num_windows, num_blocks, nnz, num_nodes, b, b_rows, n, c, ctypes.byref(opt),
                  stream)

    def spmm_balanced_workspace_size(self, num_blocks: int, n: int = 0, options=None) -> int:
        out = sz(0)
        if options is None and n == 0:
            self.call("dtc_spmm_balanced_workspace_size", num_blocks, ctypes.byref(out))
        else:
            opt = options or self.options()
            self.call("dtc_spmm_balanced_workspace_size_ex", num_blocks, n, ctypes.byref(opt), ctypes.byref(out))
        return out.value

    def spmm_balanced(self, tcblock_rowid, tcblocktile_id, tcblock_offset, sparse_atox, tcblock_values,
                      num_blocks, nnz, num_nodes, b, b_rows, n, c, workspace, workspace_bytes, options=None,
                      stream=0) -> None:
        opt = options or self.options()
        self.call("dtc_spmm_balanced", tcblock_rowid, tcblocktile_id, tcblock_offset, sparse_atox,
                  tcblock_values, num_blocks, nnz, num_nodes, b, b_rows, n, c, ctypes.byref(opt), workspace,
                  workspace_bytes, stream)

    def spmm_balanced_noalloc(self, tcblock_rowid, tcblocktile_id, tcblock_offset, sparse_atox, tcblock_values,
                              num_blocks, nnz, num_nodes, b, b_rows, n, c, options=None, stream=0) -> None:
        """The balanced SpMM on the library's fixed 8 MiB scratch (no allocation)."""
        opt = options or self.options()
        self.call("dtc_spmm_balanced_noalloc", tcblock_rowid, tcblocktile_id, tcblock_offset, sparse_atox,
                  tcblock_values, num_blocks, nnz, num_nodes, b, b_rows, n, c, ctypes.byref(opt), stream)
