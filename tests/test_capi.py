"""The C ABI library and the DTCSpMM extension without a GPU: every symbol
declared in include/dtc.h is exported, host-side argument checks return
status codes / raise before any device work, and the extension exposes the
reference's operator names."""
import ctypes
import os
import re
import subprocess

import pytest
import torch

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "dtc-spmm_asplos24_amd")
# DTC_LIBRARY: another build of the library (the sanitized host build)
LIB = os.environ.get("DTC_LIBRARY") or os.path.join(PKG, "libdtc.so")
HEADER = os.path.join(REPO, "include", "dtc.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dtc_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libdtc.so not built")
    return ctypes.CDLL(LIB)


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ["dtc_metcf_build_windows", "dtc_metcf_build_blocks", "dtc_spmm_windows",
              "dtc_spmm_balanced", "dtc_csr_transpose", "dtc_sddmm_csr"]:
        assert s in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_status_strings(lib):
    lib.dtc_status_string.restype = ctypes.c_char_p
    assert lib.dtc_status_string(0) == b"success"
    assert b"invalid" in lib.dtc_status_string(1)
    assert lib.dtc_version() >= 10000


class Opts(ctypes.Structure):  # dtc_spmm_options_t
    _fields_ = [("precision", ctypes.c_int32), ("vec", ctypes.c_int32), ("subtiles", ctypes.c_int32),
                ("chunk_blocks", ctypes.c_int32), ("persistent", ctypes.c_int32), ("b_dtype", ctypes.c_int32),
                ("deterministic", ctypes.c_int32), ("cut_span_blocks", ctypes.c_int32)]


class TcaOpts(ctypes.Structure):  # dtc_tca_options_t
    _fields_ = [("thres", ctypes.c_int32), ("cluster_thres", ctypes.c_int32), ("bands", ctypes.c_int32),
                ("rows_per_band", ctypes.c_int32), ("max_neighbors", ctypes.c_int32), ("seed", ctypes.c_uint32),
                ("max_pairs", ctypes.c_int64)]


def test_struct_sizes_match_header(tmp_path):
    # sizes as the C compiler lays the header's structs out
    src = tmp_path / "sizes.c"
    src.write_text('#include "dtc.h"\n#include <stdio.h>\nint main(void) {\n'
                   '  printf("%zu %zu", sizeof(dtc_spmm_options_t), sizeof(dtc_tca_options_t));\n'
                   '  return 0;\n}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c11", f"-I{os.path.dirname(HEADER)}", str(src), "-o", str(exe)], check=True)
    spmm_size, tca_size = map(int, subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.split())
    assert ctypes.sizeof(Opts) == spmm_size and ctypes.sizeof(TcaOpts) == tca_size
    from dtc.capi import SpmmOptions
    assert [f[0] for f in SpmmOptions._fields_] == [f[0] for f in Opts._fields_]


def test_reorder_and_sddmm_argument_errors(lib):
    lib.dtc_last_error.restype = ctypes.c_char_p
    size = ctypes.c_size_t(0)
    good = TcaOpts(16, 128, 28, 2, 8, 2022, 1000)
    assert lib.dtc_tca_reorder_workspace_size(1000, 5000, ctypes.byref(good), ctypes.byref(size)) == 0
    assert size.value > 0
    for bad in (TcaOpts(0, 128, 28, 2, 8, 1, 1000),      # thres < 1
                TcaOpts(16, 128, 70, 2, 8, 1, 1000),     # 140 hashes > 128
                TcaOpts(16, 128, 28, 2, 17, 1, 1000),    # max_neighbors > 16
                TcaOpts(16, 128, 28, 2, 8, 1, -1)):      # negative capacity
        assert lib.dtc_tca_reorder_workspace_size(1000, 5000, ctypes.byref(bad), ctypes.byref(size)) == 1
    assert lib.dtc_tca_reorder_workspace_size(1000, 5000, None, ctypes.byref(size)) == 1
    assert lib.dtc_lsh_pairs_workspace_size(-1, 10, ctypes.byref(size)) == 1
    assert lib.dtc_minhash_signatures(None, None, 10, 0, 1, None, None) == 1  # zero hashes
    # greedy: thres < 1, and a pair naming an item out of range (host memory only)
    ptr = (ctypes.c_int32 * 3)(0, 0, 0)
    root = (ctypes.c_int32 * 2)()
    pair = (ctypes.c_uint64 * 1)((0 << 32) | 7)
    sim = (ctypes.c_float * 1)(0.5)
    assert lib.dtc_tca_greedy(2, pair, sim, 1, ptr, None, 0, root) == 1
    assert lib.dtc_tca_greedy(2, pair, sim, 1, ptr, None, 16, root) == 1
    assert b"out of range" in lib.dtc_last_error()
    # sddmm: unknown y dtype is rejected before any launch
    dummy = ctypes.c_void_p(16)
    assert lib.dtc_sddmm_csr(dummy, dummy, None, 4, 4, dummy, dummy, 7, 4, 8, dummy, None) == 1


def test_host_side_argument_errors(lib):
    lib.dtc_last_error.restype = ctypes.c_char_p
    size = ctypes.c_size_t(0)
    assert lib.dtc_metcf_workspace_size(-1, 10, 16, 8, ctypes.byref(size)) == 1
    assert lib.dtc_metcf_workspace_size(10, 10, 0, 8, ctypes.byref(size)) == 1
    assert lib.dtc_metcf_workspace_size(100, 1000, 16, 8, ctypes.byref(size)) == 0
    assert size.value >= 1000 * 28  # two key (8 B) and two value (4 B) arrays and the scan (4 B) per entry
    dummy = ctypes.c_void_p(16)
    # N = 0 is rejected before anything touches the device.
    st = lib.dtc_spmm_windows(dummy, dummy, dummy, dummy, None, 1, 1, 4, 16, dummy, 16, 0, dummy, None, None)
    assert st == 1 and b"N" in lib.dtc_last_error()
    # unknown precision
    o = Opts(9, 0, 0, 0, 0)
    st = lib.dtc_spmm_windows(dummy, dummy, dummy, dummy, None, 1, 1, 4, 16, dummy, 16, 8, dummy, ctypes.byref(o), None)
    assert st == 1
    # too few windows for the rows
    st = lib.dtc_spmm_windows(dummy, dummy, dummy, dummy, None, 1, 1, 4, 100, dummy, 16, 8, dummy, None, None)
    assert st == 1
    # missing workspace for the ME-TCF build
    st = lib.dtc_metcf_build_windows(dummy, dummy, 16, 4, 16, 8, None, dummy, dummy, dummy, dummy,
                                     dummy, None, None, ctypes.byref(ctypes.c_int32()), None,
                                     ctypes.c_size_t(0), None)
    assert st == 3
    # phase 2 with a block count no phase 1 can return (more blocks than entries)
    st = lib.dtc_metcf_build_blocks(dummy, dummy, 16, 4, 16, 8, 5, dummy, dummy, dummy, dummy,
                                    ctypes.c_size_t(1 << 20), None)
    assert st == 1 and b"block_count" in lib.dtc_last_error()
    # balanced launches without their workspace (persistent or not), or with
    # one smaller than the size query reports
    size = ctypes.c_size_t()
    assert lib.dtc_spmm_balanced_workspace_size(4, ctypes.byref(size)) == 0
    assert size.value == 256 + 8 * 4
    assert lib.dtc_spmm_balanced_workspace_size(-1, ctypes.byref(size)) == 1
    # exact sizes for a given chunk: one int2 run per chunk, and with the
    # deterministic merge two 16 x 128 float32 slots per chunk (N = 8 -> 128)
    o = Opts(0, 0, 0, 2, 0, 0, 0)
    assert lib.dtc_spmm_balanced_workspace_size_ex(4, 8, ctypes.byref(o), ctypes.byref(size)) == 0
    assert size.value == 256 + 8 * 2
    o.deterministic = 1
    assert lib.dtc_spmm_balanced_workspace_size_ex(4, 8, ctypes.byref(o), ctypes.byref(size)) == 0
    assert size.value == 512 + 2 * 2 * 16 * 128 * 4
    assert lib.dtc_spmm_balanced_workspace_size_ex(4, 200, ctypes.byref(o), ctypes.byref(size)) == 0
    assert size.value == 512 + 2 * 2 * 16 * 256 * 4
    assert lib.dtc_spmm_balanced_workspace_size_ex(4, 8, None, ctypes.byref(size)) == 0
    assert size.value == 256 + 8 * 1  # default chunk for 4 blocks: 32
    for pers in (0, 1):
        for det in (0, 1):
            o = Opts(0, 0, 0, 2, pers, 0, det)
            need = ctypes.c_size_t()
            assert lib.dtc_spmm_balanced_workspace_size_ex(4, 8, ctypes.byref(o), ctypes.byref(need)) == 0
            # sizes as size_t: a plain int would leave the upper half of the
            # 64-bit argument slot undefined (no argtypes on this handle)
            st = lib.dtc_spmm_balanced(dummy, dummy, dummy, dummy, None, 4, 8, 16, dummy, 16, 8, dummy,
                                       ctypes.byref(o), None, ctypes.c_size_t(0), None)
            assert st == 3
            st = lib.dtc_spmm_balanced(dummy, dummy, dummy, dummy, None, 4, 8, 16, dummy, 16, 8, dummy,
                                       ctypes.byref(o), dummy, ctypes.c_size_t(need.value - 1), None)
            assert st == 3 and b"workspace" in lib.dtc_last_error()
    o = Opts(0, 0, 0, 2, 0, 0, 2)  # deterministic must be 0 or 1
    st = lib.dtc_spmm_balanced(dummy, dummy, dummy, dummy, None, 4, 8, 16, dummy, 16, 8, dummy,
                               ctypes.byref(o), dummy, ctypes.c_size_t(1 << 20), None)
    assert st == 1
    # the cut span is a non-negative multiple of the chunk
    for span in (-4, 3):
        o = Opts(0, 0, 0, 2, 0, 0, 0, span)
        st = lib.dtc_spmm_balanced(dummy, dummy, dummy, dummy, None, 4, 8, 16, dummy, 16, 8, dummy,
                                   ctypes.byref(o), dummy, ctypes.c_size_t(1 << 20), None)
        assert st == 1 and b"cut_span_blocks" in lib.dtc_last_error()


@pytest.fixture(scope="module")
def ext():
    from dtc import load_extension
    return load_extension()


def test_extension_exposes_reference_names(ext):
    for name in ["preprocess_gpu", "run_DTCSpMM", "run_DTCSpMM_balance", "DTCSpMM_gcn",
                 "backward_ptx_uint8_improved", "backward_ptx_uint8_improved_gcn",
                 "balance_backward_ptx_uint8_prefetch", "spmm", "csr_transpose", "sddmm_csr"]:
        assert hasattr(ext, name), name
    assert ext.BLK_H == 16 and ext.BLK_W == 8


def _cpu_operands():
    i32 = torch.int32
    return (torch.ones(32, 8), torch.tensor([0, 1, 2], dtype=i32), torch.zeros(4, dtype=torch.uint8),
            torch.tensor([0, 2, 4], dtype=i32), torch.zeros(16, dtype=i32))


def test_extension_rejects_cpu_tensors(ext):
    X, rwo, tile, off, atox = _cpu_operands()
    with pytest.raises(RuntimeError, match="CUDA tensor"):
        ext.run_DTCSpMM(X, rwo, tile, off, atox, 32, 4, "float_nonsplit")
    with pytest.raises(RuntimeError, match="CUDA tensor"):
        ext.run_DTCSpMM_balance(X, torch.zeros(2, dtype=torch.int32), tile, off, atox, 32, "float4_split")
    with pytest.raises(RuntimeError, match="CUDA tensor"):
        ext.preprocess_gpu(torch.zeros(4, dtype=torch.int32), torch.zeros(33, dtype=torch.int32), 32, 16, 8,
                           torch.zeros(2, dtype=torch.int32), torch.zeros(4, dtype=torch.int32),
                           torch.zeros(4, dtype=torch.int32))


def test_extension_rejects_bad_plan_and_precision(ext):
    X, rwo, tile, off, atox = _cpu_operands()
    with pytest.raises(ValueError, match="exe plan"):
        ext.run_DTCSpMM(X, rwo, tile, off, atox, 32, 4, "float8_split")
    with pytest.raises(ValueError, match="precision"):
        ext.run_DTCSpMM(X, rwo, tile, off, atox, 32, 4, "float_nonsplit", precision="tf32")
    with pytest.raises(ValueError, match="schedule"):
        ext.spmm(X, rwo, tile, off, atox, 32, schedule="diagonal")


def test_gather_ceiling_probe_library():
    """The measurement aid beside libdtc: exported entry points and host-side
    size checks (no device work)."""
    path = os.path.join(PKG, "libdtc_probe.so")
    if not os.path.exists(path):
        pytest.skip("libdtc_probe.so not built")
    probe = ctypes.CDLL(path)
    probe.dtc_gather_ceiling_out_size.restype = ctypes.c_int64
    # 1000 blocks in 4 chunks of 256 -> one workgroup per column tile
    assert probe.dtc_gather_ceiling_out_size(1000, 256, 256, 4) == 4 * 1 * 256
    assert probe.dtc_gather_ceiling_out_size(1000, 256, 256, 2) == 8 * 1 * 256
    assert probe.dtc_gather_ceiling_out_size(1000, 256, 96, 2) == 3 * 1 * 256
    assert probe.dtc_gather_ceiling_out_size(1000, 256, 96, 4) == -1  # N not a multiple of 64
    assert probe.dtc_gather_ceiling_out_size(1000, 256, 256, 1) == -1  # vec 1 not modelled
    assert probe.dtc_gather_ceiling_out_size(0, 256, 256, 4) == -1
    assert probe.dtc_gather_ceiling(None, 0, 256, None, 256, 4, None, None) == 1


def test_tcgnn_entry_points_without_gpu(ext, lib):
    """TC-GNN's preprocess needs the GPU build and says so; the C entry point
    rejects a non-float32 precision before any launch."""
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    i32 = torch.int32
    with pytest.raises(RuntimeError, match="GPU"):
        ext.preprocess(torch.zeros(4, dtype=i32), torch.tensor([0, 4], dtype=i32), 1, 16, 8,
                       torch.zeros(1, dtype=i32), torch.zeros(4, dtype=i32), torch.zeros(4, dtype=i32))
    dummy = ctypes.c_void_p(16)
    o = Opts(4, 0, 0, 0, 0, 0)  # bf16x6
    assert lib.dtc_spmm_tcgnn(dummy, dummy, dummy, dummy, dummy, None, 16, 4, dummy, 16, 8, dummy,
                              ctypes.byref(o), None) == 2
