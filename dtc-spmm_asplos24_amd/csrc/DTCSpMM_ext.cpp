// The `DTCSpMM` Python module: the reference's operator surface
// (DTC-SpMM/DTCSpMM.cpp:361-399) over the libdtc C ABI (include/dtc.h).
//
// Tensors are validated here (device, contiguity, dtype, sizes) before any
// GPU work, device buffers come from torch's caching allocator, and every
// launch goes to the current HIP stream of the input's device.  Library
// failures raise RuntimeError (argument checks, as the reference's
// TORCH_CHECKs) or ValueError (unknown plan or precision names), never exit
// the process.
#include <torch/extension.h>

#include <ATen/hip/impl/HIPStreamMasqueradingAsCUDA.h>
#include <hip/hip_runtime_api.h>

#include <cstdlib>
#include <fstream>
#include <string>

#include "dtc.h"

namespace py = pybind11;
using torch::Tensor;

namespace {

constexpr int kBlkH = 16;
constexpr int kBlkW = 8;
constexpr int kExeTime = 1000;  // timed launches under DTC_BENCH=1 (DTCSpMM_kernel.cu:18)

#define CHECK_CUDA(x) TORCH_CHECK((x).is_cuda(), #x " must be a CUDA tensor")
#define CHECK_CONTIGUOUS(x) TORCH_CHECK((x).is_contiguous(), #x " must be contiguous")
#define CHECK_INPUT(x) \
  CHECK_CUDA(x);       \
  CHECK_CONTIGUOUS(x)
#define CHECK_DTYPE(x, dt) \
  TORCH_CHECK((x).scalar_type() == (dt), #x " must have dtype " #dt ", got ", (x).scalar_type())
#define CHECK_SAME_DEVICE(x, ref) \
  TORCH_CHECK((x).device() == (ref).device(), #x " must be on the same device as " #ref)

dtc_stream_t current_stream() {
  return reinterpret_cast<dtc_stream_t>(c10::hip::getCurrentHIPStreamMasqueradingAsCUDA().stream());
}

void check_status(dtc_status_t st, const char *where) {
  if (st == DTC_SUCCESS) return;
  const std::string msg = std::string(where) + ": " + dtc_status_string(st) + " (" + dtc_last_error() + ")";
  TORCH_CHECK(false, msg);
}

bool env_flag(const char *name) {
  const char *v = std::getenv(name);
  return v && v[0] && std::string(v) != "0";
}

int32_t parse_precision(const std::string &p) {
  if (p == "fp32" || p == "float32") return DTC_PRECISION_FP32;
  if (p == "bf16" || p == "bfloat16") return DTC_PRECISION_BF16;
  if (p == "fp16" || p == "float16") return DTC_PRECISION_FP16;
  if (p == "bf16x3") return DTC_PRECISION_BF16X3;
  if (p == "bf16x6") return DTC_PRECISION_BF16X6;
  TORCH_CHECK_VALUE(false, "unknown precision '", p, "' (fp32, bf16, fp16, bf16x3, bf16x6)");
  return 0;
}

// Reference exeplan names (DTCSpMM_kernel.cu:840-930) -> vector width and
// column subtiles per wavefront.  The names only steer performance here:
// every plan computes the full product for any N.
void parse_exeplan(const std::string &plan, dtc_spmm_options_t &o) {
  if (plan == "float_nonsplit" || plan == "float_split") { o.vec = 1; o.subtiles = 1; return; }
  if (plan == "float2_nonsplit" || plan == "float2_split") { o.vec = 2; o.subtiles = 1; return; }
  if (plan == "float4_nonsplit") { o.vec = 4; o.subtiles = 2; return; }
  if (plan == "float4_split") { o.vec = 4; o.subtiles = 1; return; }
  if (plan == "auto") { o.vec = 0; o.subtiles = 0; return; }
  TORCH_CHECK_VALUE(false, "Not supported exe plan '", plan,
                    "' (float_nonsplit, float2_nonsplit, float2_split, float4_nonsplit, float4_split, auto)");
}

// The widest vector width N and the base addresses allow, capped at `want`
// (0: at the library's default width for the schedule, `dflt`).
int32_t fit_vec(int32_t want, int64_t n, const Tensor &b, const Tensor &c, int32_t dflt) {
  const int32_t cap = want != 0 ? want : dflt;
  for (int v : {4, 2, 1}) {
    if (v > cap) continue;
    const bool ok = n % v == 0 && reinterpret_cast<uintptr_t>(b.data_ptr()) % (4 * v) == 0 &&
                    reinterpret_cast<uintptr_t>(c.data_ptr()) % (4 * v) == 0;
    if (ok) return v;
  }
  return 1;
}

struct Operands {
  Tensor input, tile, off, atox;
  c10::optional<Tensor> values;
};

void check_operands(const Operands &o, int64_t num_nodes) {
  CHECK_INPUT(o.input);
  CHECK_INPUT(o.tile);
  CHECK_INPUT(o.off);
  CHECK_INPUT(o.atox);
  TORCH_CHECK(o.input.scalar_type() == torch::kFloat32 || o.input.scalar_type() == torch::kBFloat16,
              "input must be float32 (or bfloat16 for spmm), got ", o.input.scalar_type());
  CHECK_DTYPE(o.tile, torch::kUInt8);
  CHECK_DTYPE(o.off, torch::kInt32);
  CHECK_DTYPE(o.atox, torch::kInt32);
  CHECK_SAME_DEVICE(o.tile, o.input);
  CHECK_SAME_DEVICE(o.off, o.input);
  CHECK_SAME_DEVICE(o.atox, o.input);
  TORCH_CHECK(o.input.dim() == 2, "input must be 2-D [rows, N]");
  TORCH_CHECK(o.input.size(1) >= 1, "input must have at least one column");
  TORCH_CHECK(num_nodes >= 0, "num_nodes must be >= 0");
  TORCH_CHECK(o.off.dim() == 1 && o.off.size(0) >= 1, "TCblock_offset must be 1-D with >= 1 entry");
  const int64_t T = o.off.size(0) - 1;
  TORCH_CHECK(o.atox.numel() >= T * kBlkW, "sparse_AToX_idx must hold 8 entries per TC block");
  if (o.values.has_value()) {
    const Tensor &v = *o.values;
    CHECK_INPUT(v);
    CHECK_DTYPE(v, torch::kFloat32);
    CHECK_SAME_DEVICE(v, o.input);
    TORCH_CHECK(v.numel() == o.tile.numel(), "values must hold one value per TCblocktile_id entry");
  }
}

// Reference-named entries return a tensor shaped like input, as
// zeros_like(input) does (DTCSpMM_kernel.cu:823), with rows the kernels do not
// cover left at zero.  The general entry returns [num_nodes, N] or writes the
// first num_nodes rows of `out`.
Tensor output_for(const Tensor &input, int64_t num_nodes, bool ref_shape, const c10::optional<Tensor> &out,
                  int64_t *m_eff) {
  if (out.has_value()) {
    const Tensor &c = *out;
    CHECK_INPUT(c);
    CHECK_DTYPE(c, torch::kFloat32);
    CHECK_SAME_DEVICE(c, input);
    TORCH_CHECK(c.dim() == 2 && c.size(1) == input.size(1), "out must be 2-D with N columns");
    *m_eff = std::min<int64_t>(num_nodes, c.size(0));
    return c;
  }
  if (!ref_shape) {
    *m_eff = num_nodes;
    return torch::empty({num_nodes, input.size(1)}, input.options().dtype(torch::kFloat32));
  }
  const int64_t rows = input.size(0);
  *m_eff = std::min<int64_t>(num_nodes, rows);
  if (*m_eff == rows) return torch::empty_like(input);
  return torch::zeros_like(input);
}

template <class Launch>
void maybe_bench(const char *tag, int64_t num_edges, int64_t n, Launch launch) {
  // DTC_BENCH=1 reproduces the reference's timed loop and CSV line
  // (DTCSpMM_kernel.cu:840-937): 1000 launches, mean time, appended as
  // "DTCSpMM:,<ms>,<GFLOP/s>," to ./DTCSpMM_exe_time_and_throughput.csv.
  if (!env_flag("DTC_BENCH")) {
    launch();
    return;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(current_stream());
  hipEvent_t e0, e1;
  TORCH_CHECK(hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess, "hipEventCreate failed");
  TORCH_CHECK(hipEventRecord(e0, s) == hipSuccess, "hipEventRecord failed");
  for (int i = 0; i < kExeTime; ++i) launch();
  TORCH_CHECK(hipEventRecord(e1, s) == hipSuccess, "hipEventRecord failed");
  TORCH_CHECK(hipEventSynchronize(e1) == hipSuccess, "hipEventSynchronize failed");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double t = ms / kExeTime;
  const double gflops = 2.0 * (double)num_edges * (double)n / (t * 1e6);
  std::ofstream f("DTCSpMM_exe_time_and_throughput.csv", std::ios::app);
  f << "DTCSpMM:" << "," << t << "," << gflops << ",";
  (void)tag;
}

// ------------------------------------------------------------ preprocess

py::tuple preprocess_gpu(Tensor edgeList, Tensor nodePointer, int64_t num_nodes, int64_t blockSize_h,
                         int64_t blockSize_w, Tensor blockPartition, Tensor edgeToColumn,
                         Tensor edgeToRow, c10::optional<Tensor> values, bool return_edge_order) {
  CHECK_INPUT(edgeList);
  CHECK_INPUT(nodePointer);
  CHECK_INPUT(blockPartition);
  CHECK_INPUT(edgeToColumn);
  CHECK_INPUT(edgeToRow);
  for (const Tensor *t : {&edgeList, &nodePointer, &blockPartition, &edgeToColumn, &edgeToRow})
    CHECK_DTYPE(*t, torch::kInt32);
  CHECK_SAME_DEVICE(nodePointer, edgeList);
  CHECK_SAME_DEVICE(blockPartition, edgeList);
  CHECK_SAME_DEVICE(edgeToColumn, edgeList);
  CHECK_SAME_DEVICE(edgeToRow, edgeList);
  TORCH_CHECK(num_nodes >= 0 && num_nodes < INT32_MAX, "num_nodes out of range");
  TORCH_CHECK(blockSize_h >= 1 && blockSize_w >= 1, "block sizes must be >= 1");
  const int64_t nnz = edgeList.numel();
  TORCH_CHECK(nnz < INT32_MAX, "more than 2^31-1 edges");
  const int64_t W = (num_nodes + blockSize_h - 1) / blockSize_h;
  TORCH_CHECK(nodePointer.numel() == num_nodes + 1, "nodePointer must have num_nodes + 1 entries, got ",
              nodePointer.numel());
  TORCH_CHECK(blockPartition.numel() == W, "blockPartition must have ceil(num_nodes / blockSize_h) = ", W,
              " entries, got ", blockPartition.numel());
  TORCH_CHECK(edgeToColumn.numel() == nnz, "edgeToColumn must have one entry per edge");
  TORCH_CHECK(edgeToRow.numel() == nnz, "edgeToRow must have one entry per edge");
  if (values.has_value()) {
    CHECK_INPUT(*values);
    CHECK_DTYPE(*values, torch::kFloat32);
    CHECK_SAME_DEVICE(*values, edgeList);
    TORCH_CHECK(values->numel() == nnz, "values must have one entry per edge");
  }
  const c10::OptionalDeviceGuard guard(edgeList.device());
  const auto i32 = torch::TensorOptions().dtype(torch::kInt32).device(edgeList.device());
  const auto u8 = torch::TensorOptions().dtype(torch::kUInt8).device(edgeList.device());
  const auto f32 = torch::TensorOptions().dtype(torch::kFloat32).device(edgeList.device());
  size_t ws_bytes = 0;
  check_status(dtc_metcf_workspace_size((int32_t)num_nodes, (int32_t)nnz, (int32_t)blockSize_h,
                                        (int32_t)blockSize_w, &ws_bytes),
               "preprocess_gpu");
  Tensor ws = torch::empty({(int64_t)ws_bytes}, u8);
  Tensor rwo = torch::empty({W + 1}, i32);
  Tensor tile = torch::empty({nnz}, u8);
  Tensor tcv, tce;
  if (values.has_value()) tcv = torch::empty({nnz}, f32);
  if (return_edge_order) tce = torch::empty({nnz}, i32);
  dtc_stream_t s = current_stream();
  int32_t T = 0;
  check_status(dtc_metcf_build_windows(edgeList.data_ptr<int32_t>(), nodePointer.data_ptr<int32_t>(),
                                       (int32_t)num_nodes, (int32_t)nnz, (int32_t)blockSize_h,
                                       (int32_t)blockSize_w,
                                       values.has_value() ? values->data_ptr<float>() : nullptr,
                                       blockPartition.data_ptr<int32_t>(),
                                       edgeToColumn.data_ptr<int32_t>(), edgeToRow.data_ptr<int32_t>(),
                                       rwo.data_ptr<int32_t>(), tile.data_ptr<uint8_t>(),
                                       values.has_value() ? tcv.data_ptr<float>() : nullptr,
                                       return_edge_order ? tce.data_ptr<int32_t>() : nullptr, &T,
                                       ws.data_ptr(), ws_bytes, s),
               "preprocess_gpu");
  Tensor rowid = torch::empty({(int64_t)T}, i32);
  Tensor off = torch::empty({(int64_t)T + 1}, i32);
  Tensor atox = torch::empty({(int64_t)T * blockSize_w}, i32);
  check_status(dtc_metcf_build_blocks(nodePointer.data_ptr<int32_t>(), rwo.data_ptr<int32_t>(),
                                      (int32_t)num_nodes, (int32_t)nnz, (int32_t)blockSize_h,
                                      (int32_t)blockSize_w, T, rowid.data_ptr<int32_t>(),
                                      off.data_ptr<int32_t>(), atox.data_ptr<int32_t>(),
                                      ws.data_ptr(), ws_bytes, s),
               "preprocess_gpu");
  if (env_flag("DTC_VERBOSE"))
    std::printf("TC_Blocks:\t%d\nExp_Edges:\t%lld\n", T, (long long)T * blockSize_w * blockSize_h);
  py::list out;
  out.append(rwo);
  out.append(rowid);
  out.append(tile);
  out.append(off);
  out.append(atox);
  out.append(py::int_(T));
  if (values.has_value()) out.append(tcv);
  if (return_edge_order) out.append(tce);
  return py::tuple(out);
}

void check_csr_pair(const Tensor &ptr, const Tensor &idx);

// TC-GNN's format (DTCSpMM.cpp:237-311): the reference builds it on the CPU
// into caller-allocated CPU tensors.  Here the ME-TCF build's phase 1 computes
// the same arrays on the GPU and they are copied back into the caller's
// tensors.  One difference, deliberate: an empty row window gets 0 blocks
// (the reference's CPU loop gives it 1, and writes blockPartition[W] out of
// bounds when num_nodes % 16 == 0, DTCSpMM.cpp:281).
int64_t preprocess(Tensor edgeList, Tensor nodePointer, int64_t num_nodes, int64_t blockSize_h,
                   int64_t blockSize_w, Tensor blockPartition, Tensor edgeToColumn, Tensor edgeToRow) {
  for (const Tensor *t : {&edgeList, &nodePointer, &blockPartition, &edgeToColumn, &edgeToRow}) {
    CHECK_CONTIGUOUS(*t);
    CHECK_DTYPE(*t, torch::kInt32);
  }
  TORCH_CHECK(torch::cuda::is_available(), "preprocess runs the ME-TCF build on the GPU; no GPU is visible");
  const auto dev = torch::Device(torch::kCUDA);  // the current device
  Tensor el = edgeList.to(dev), np = nodePointer.to(dev);
  Tensor bp = torch::zeros({blockPartition.numel()}, el.options());
  Tensor e2c = torch::empty({edgeToColumn.numel()}, el.options());
  Tensor e2r = torch::empty({edgeToRow.numel()}, el.options());
  py::tuple r = preprocess_gpu(el, np, num_nodes, blockSize_h, blockSize_w, bp, e2c, e2r, c10::nullopt, false);
  blockPartition.copy_(bp);
  edgeToColumn.copy_(e2c);
  edgeToRow.copy_(e2r);
  return r[5].cast<int64_t>();
}

std::vector<Tensor> run_TCGNNSpMM(Tensor input, Tensor nodePointer, Tensor edgeList, Tensor blockPartition,
                                  Tensor edgeToColumn, Tensor edgeToRow, c10::optional<Tensor> values) {
  CHECK_INPUT(input);
  CHECK_DTYPE(input, torch::kFloat32);
  check_csr_pair(nodePointer, edgeList);
  for (const Tensor *t : {&blockPartition, &edgeToColumn, &edgeToRow}) {
    CHECK_INPUT(*t);
    CHECK_DTYPE(*t, torch::kInt32);
    CHECK_SAME_DEVICE(*t, input);
  }
  CHECK_SAME_DEVICE(nodePointer, input);
  TORCH_CHECK(input.dim() == 2, "input must be 2-D [rows, N]");
  const int64_t M = nodePointer.numel() - 1, nnz = edgeList.numel();
  TORCH_CHECK(blockPartition.numel() == (M + 15) / 16, "blockPartition must have one entry per row window");
  TORCH_CHECK(edgeToColumn.numel() == nnz && edgeToRow.numel() == nnz, "edgeToColumn / edgeToRow need one entry per edge");
  const float *vals = nullptr;
  if (values.has_value()) {
    CHECK_INPUT(*values);
    CHECK_DTYPE(*values, torch::kFloat32);
    TORCH_CHECK(values->numel() == nnz, "values must have one entry per edge");
    vals = values->data_ptr<float>();
  }
  const c10::OptionalDeviceGuard guard(input.device());
  // the reference returns zeros_like(input) filled by the kernel (DTCSpMM_kernel.cu:715)
  Tensor out = M == input.size(0) ? torch::empty_like(input) : torch::zeros_like(input);
  const int64_t m_eff = std::min<int64_t>(M, input.size(0));
  maybe_bench("TCGNNSpMM", nnz, input.size(1), [&] {
    check_status(dtc_spmm_tcgnn(nodePointer.data_ptr<int32_t>(), edgeList.data_ptr<int32_t>(),
                                blockPartition.data_ptr<int32_t>(), edgeToColumn.data_ptr<int32_t>(),
                                edgeToRow.data_ptr<int32_t>(), vals, (int32_t)m_eff, (int32_t)nnz,
                                input.data_ptr<float>(), (int32_t)input.size(0), (int32_t)input.size(1),
                                out.data_ptr<float>(), nullptr, current_stream()),
                 "run_TCGNNSpMM");
  });
  return {out};
}

// --------------------------------------------------------------- SpMM

std::vector<Tensor> run_windows(const Operands &o, const Tensor &rwo, int64_t num_nodes, int64_t num_edges,
                                dtc_spmm_options_t opt, bool bench, bool ref_shape, c10::optional<Tensor> out) {
  CHECK_INPUT(rwo);
  CHECK_DTYPE(rwo, torch::kInt32);
  CHECK_SAME_DEVICE(rwo, o.input);
  check_operands(o, num_nodes);
  const int64_t W = rwo.numel() - 1;
  TORCH_CHECK(W >= 0, "Rowwindow_offset must have >= 1 entry");
  TORCH_CHECK(W * kBlkH >= num_nodes, "Rowwindow_offset covers ", W * kBlkH, " rows, fewer than num_nodes");
  const c10::OptionalDeviceGuard guard(o.input.device());
  int64_t m_eff = 0;
  Tensor c = output_for(o.input, num_nodes, ref_shape, out, &m_eff);
  const int64_t n = o.input.size(1);
  if (o.input.scalar_type() == torch::kBFloat16) opt.b_dtype = DTC_B_BFLOAT16;
  else opt.vec = fit_vec(opt.vec, n, o.input, c, 4);
  // The reference bounds B rows by num_nodes (dense_bound,
  // DTCSpMM_kernel.cu:2687); the general entry takes B's own row count.
  const int64_t b_rows = ref_shape ? std::min<int64_t>(num_nodes, o.input.size(0)) : o.input.size(0);
  const float *vals = o.values.has_value() ? o.values->data_ptr<float>() : nullptr;
  auto launch = [&]() {
    check_status(dtc_spmm_windows(rwo.data_ptr<int32_t>(), o.tile.data_ptr<uint8_t>(), o.off.data_ptr<int32_t>(),
                                  o.atox.data_ptr<int32_t>(), vals, (int32_t)W, (int32_t)(o.off.numel() - 1),
                                  (int32_t)o.tile.numel(), (int32_t)m_eff,
                                  o.input.data_ptr(), (int32_t)b_rows, (int32_t)n, c.data_ptr<float>(),
                                  &opt, current_stream()),
                 "run_DTCSpMM");
  };
  if (bench) maybe_bench("DTCSpMM", num_edges, n, launch);
  else launch();
  return {c};
}

std::vector<Tensor> run_balanced(const Operands &o, const Tensor &rowid, int64_t num_nodes, dtc_spmm_options_t opt,
                                 bool bench, bool ref_shape, c10::optional<Tensor> out) {
  CHECK_INPUT(rowid);
  CHECK_DTYPE(rowid, torch::kInt32);
  CHECK_SAME_DEVICE(rowid, o.input);
  check_operands(o, num_nodes);
  const int64_t T = rowid.numel();
  TORCH_CHECK(o.off.numel() == T + 1, "TCblock_offset must have one entry more than TCblock_rowid");
  const c10::OptionalDeviceGuard guard(o.input.device());
  int64_t m_eff = 0;
  Tensor c = output_for(o.input, num_nodes, ref_shape, out, &m_eff);
  const int64_t n = o.input.size(1);
  if (o.input.scalar_type() == torch::kBFloat16) opt.b_dtype = DTC_B_BFLOAT16;
  else opt.vec = fit_vec(opt.vec, n, o.input, c, 2);
  // The reference bounds B rows by num_nodes (dense_bound,
  // DTCSpMM_kernel.cu:2687); the general entry takes B's own row count.
  const int64_t b_rows = ref_shape ? std::min<int64_t>(num_nodes, o.input.size(0)) : o.input.size(0);
  size_t ws_bytes = 0;
  check_status(dtc_spmm_balanced_workspace_size_ex((int32_t)T, (int32_t)n, &opt, &ws_bytes), "run_DTCSpMM_balance");
  Tensor ws = torch::empty({(int64_t)ws_bytes}, o.input.options().dtype(torch::kUInt8));
  const float *vals = o.values.has_value() ? o.values->data_ptr<float>() : nullptr;
  auto launch = [&]() {
    check_status(dtc_spmm_balanced(rowid.data_ptr<int32_t>(), o.tile.data_ptr<uint8_t>(), o.off.data_ptr<int32_t>(),
                                   o.atox.data_ptr<int32_t>(), vals, (int32_t)T, (int32_t)o.tile.numel(),
                                   (int32_t)m_eff,
                                   o.input.data_ptr(), (int32_t)b_rows, (int32_t)n, c.data_ptr<float>(), &opt,
                                   ws.data_ptr(), ws_bytes, current_stream()),
                 "run_DTCSpMM_balance");
  };
  if (bench) maybe_bench("DTCSpMM", o.tile.numel(), n, launch);
  else launch();
  return {c};
}

// run_DTCSpMM(input, Rowwindow_offset, TCblocktile_id, TCblock_offset,
// sparse_AToX_idx, num_nodes, num_edges, exeplan)  -- DTCSpMM.cpp:117-131
std::vector<Tensor> run_DTCSpMM(Tensor input, Tensor Rowwindow_offset, Tensor TCblocktile_id, Tensor TCblock_offset,
                                Tensor sparse_AToX_idx, int64_t num_nodes, int64_t num_edges, std::string exeplan,
                                c10::optional<Tensor> values, std::string precision) {
  CHECK_DTYPE(input, torch::kFloat32);
  dtc_spmm_options_t opt{};
  parse_exeplan(exeplan, opt);
  opt.precision = parse_precision(precision);
  Operands o{input, TCblocktile_id, TCblock_offset, sparse_AToX_idx, values};
  return run_windows(o, Rowwindow_offset, num_nodes, num_edges, opt, true, true, c10::nullopt);
}

// run_DTCSpMM_balance(input, TCblock_rowid, TCblocktile_id, TCblock_offset,
// sparse_AToX_idx, num_nodes, exeplan)  -- DTCSpMM.cpp:163-178
std::vector<Tensor> run_DTCSpMM_balance(Tensor input, Tensor TCblock_rowid, Tensor TCblocktile_id,
                                        Tensor TCblock_offset, Tensor sparse_AToX_idx, int64_t num_nodes,
                                        std::string exeplan, c10::optional<Tensor> values, std::string precision,
                                        int64_t chunk_blocks, bool persistent, bool deterministic,
                                        int64_t cut_span_blocks) {
  CHECK_DTYPE(input, torch::kFloat32);
  dtc_spmm_options_t opt{};
  parse_exeplan(exeplan, opt);
  opt.precision = parse_precision(precision);
  opt.chunk_blocks = (int32_t)chunk_blocks;
  opt.cut_span_blocks = (int32_t)cut_span_blocks;
  opt.persistent = persistent ? 1 : 0;
  opt.deterministic = deterministic ? 1 : 0;
  Operands o{input, TCblocktile_id, TCblock_offset, sparse_AToX_idx, values};
  return run_balanced(o, TCblock_rowid, num_nodes, opt, true, true, c10::nullopt);
}

// DTCSpMM_gcn(input, Rowwindow_offset, TCblocktile_id, TCblock_offset,
// sparse_AToX_idx, num_nodes, num_edges)  -- DTCSpMM.cpp:147-161: one launch,
// variant picked from N as DTCSpMM_kernel.cu:1044-1108 does.
std::vector<Tensor> DTCSpMM_gcn(Tensor input, Tensor Rowwindow_offset, Tensor TCblocktile_id, Tensor TCblock_offset,
                                Tensor sparse_AToX_idx, int64_t num_nodes, int64_t num_edges,
                                c10::optional<Tensor> values) {
  CHECK_DTYPE(input, torch::kFloat32);
  dtc_spmm_options_t opt{};
  const int64_t n = input.dim() == 2 ? input.size(1) : 0;
  if (n >= 128 && n % 4 == 0) { opt.vec = 4; opt.subtiles = 1; }
  else if (n % 2 == 0) { opt.vec = 2; opt.subtiles = 1; }
  else { opt.vec = 1; opt.subtiles = 1; }
  opt.precision = DTC_PRECISION_FP32;
  Operands o{input, TCblocktile_id, TCblock_offset, sparse_AToX_idx, values};
  return run_windows(o, Rowwindow_offset, num_nodes, num_edges, opt, false, true, c10::nullopt);
}

// The reference binds its backward names to the forward functions
// (DTCSpMM.cpp:386-393): called with dY they return A dY, the gradient with
// respect to B only when A is symmetric.  Here they also take, as an optional
// last argument `transposed`, A^T's ME-TCF operands -- (Rowwindow_offset,
// TCblocktile_id, TCblock_offset, sparse_AToX_idx, values or None, rows of
// A^T) for the window schedules, TCblock_rowid in place of Rowwindow_offset
// for the balanced one (dtc.matrix.DTCMatrix.transposed_operands builds
// them) -- and with it return the true gradient dB = A^T dY.  Without it they
// behave as the reference's.
struct TransposedOperands {
  Tensor index, tile, off, atox;
  c10::optional<Tensor> values;
  int64_t rows;
};

TransposedOperands unpack_transposed(const py::tuple &t) {
  TORCH_CHECK_VALUE(t.size() == 6,
                    "transposed must be (index, TCblocktile_id, TCblock_offset, sparse_AToX_idx, values or None, "
                    "num_rows) of A^T");
  TransposedOperands o{t[0].cast<Tensor>(), t[1].cast<Tensor>(), t[2].cast<Tensor>(), t[3].cast<Tensor>(),
                       c10::nullopt, t[5].cast<int64_t>()};
  if (!t[4].is_none()) o.values = t[4].cast<Tensor>();
  return o;
}

std::vector<Tensor> backward_windows(Tensor input, Tensor Rowwindow_offset, Tensor TCblocktile_id,
                                     Tensor TCblock_offset, Tensor sparse_AToX_idx, int64_t num_nodes,
                                     int64_t num_edges, std::string exeplan, c10::optional<Tensor> values,
                                     std::string precision, c10::optional<py::tuple> transposed) {
  if (!transposed)
    return run_DTCSpMM(input, Rowwindow_offset, TCblocktile_id, TCblock_offset, sparse_AToX_idx, num_nodes,
                       num_edges, exeplan, values, precision);
  const TransposedOperands t = unpack_transposed(*transposed);
  return run_DTCSpMM(input, t.index, t.tile, t.off, t.atox, t.rows, t.tile.numel(), exeplan, t.values,
                     precision);
}

std::vector<Tensor> backward_windows_gcn(Tensor input, Tensor Rowwindow_offset, Tensor TCblocktile_id,
                                         Tensor TCblock_offset, Tensor sparse_AToX_idx, int64_t num_nodes,
                                         int64_t num_edges, c10::optional<Tensor> values,
                                         c10::optional<py::tuple> transposed) {
  if (!transposed)
    return DTCSpMM_gcn(input, Rowwindow_offset, TCblocktile_id, TCblock_offset, sparse_AToX_idx, num_nodes,
                       num_edges, values);
  const TransposedOperands t = unpack_transposed(*transposed);
  return DTCSpMM_gcn(input, t.index, t.tile, t.off, t.atox, t.rows, t.tile.numel(), t.values);
}

std::vector<Tensor> backward_balanced(Tensor input, Tensor TCblock_rowid, Tensor TCblocktile_id,
                                      Tensor TCblock_offset, Tensor sparse_AToX_idx, int64_t num_nodes,
                                      std::string exeplan, c10::optional<Tensor> values, std::string precision,
                                      int64_t chunk_blocks, bool persistent, bool deterministic,
                                      int64_t cut_span_blocks, c10::optional<py::tuple> transposed) {
  if (!transposed)
    return run_DTCSpMM_balance(input, TCblock_rowid, TCblocktile_id, TCblock_offset, sparse_AToX_idx, num_nodes,
                               exeplan, values, precision, chunk_blocks, persistent, deterministic,
                               cut_span_blocks);
  const TransposedOperands t = unpack_transposed(*transposed);
  return run_DTCSpMM_balance(input, t.index, t.tile, t.off, t.atox, t.rows, exeplan, t.values, precision,
                             chunk_blocks, persistent, deterministic, cut_span_blocks);
}

// General entry (build extension): either schedule, every tuning knob, an
// optional preallocated output, no CSV side effects.
Tensor spmm(Tensor input, Tensor window_or_rowid, Tensor TCblocktile_id, Tensor TCblock_offset, Tensor sparse_AToX_idx,
            int64_t num_nodes, std::string schedule, c10::optional<Tensor> values, std::string precision, int64_t vec,
            int64_t subtiles, int64_t chunk_blocks, bool persistent, c10::optional<Tensor> out,
            bool deterministic, int64_t cut_span_blocks) {
  dtc_spmm_options_t opt{};
  opt.deterministic = deterministic ? 1 : 0;
  opt.cut_span_blocks = (int32_t)cut_span_blocks;
  opt.precision = parse_precision(precision);
  TORCH_CHECK_VALUE(vec == 0 || vec == 1 || vec == 2 || vec == 4 || vec == 8, "vec must be 0, 1, 2, 4 or 8");
  TORCH_CHECK_VALUE(subtiles >= 0 && subtiles <= 2, "subtiles must be 0, 1 or 2");
  opt.vec = (int32_t)vec;
  opt.subtiles = (int32_t)subtiles;
  opt.chunk_blocks = (int32_t)chunk_blocks;
  opt.persistent = persistent ? 1 : 0;
  Operands o{input, TCblocktile_id, TCblock_offset, sparse_AToX_idx, values};
  if (schedule == "windows" || schedule == "unbalanced")
    return run_windows(o, window_or_rowid, num_nodes, TCblocktile_id.numel(), opt, false, false, out)[0];
  if (schedule == "balanced")
    return run_balanced(o, window_or_rowid, num_nodes, opt, false, false, out)[0];
  TORCH_CHECK_VALUE(false, "schedule must be 'windows' or 'balanced', got '", schedule, "'");
  return Tensor();
}

// ------------------------------------------------------------- backward

py::tuple csr_transpose(Tensor nodePointer, Tensor edgeList, int64_t num_rows, int64_t num_cols) {
  CHECK_INPUT(nodePointer);
  CHECK_INPUT(edgeList);
  CHECK_DTYPE(nodePointer, torch::kInt32);
  CHECK_DTYPE(edgeList, torch::kInt32);
  CHECK_SAME_DEVICE(edgeList, nodePointer);
  TORCH_CHECK(num_rows >= 0 && num_cols >= 0, "sizes must be >= 0");
  TORCH_CHECK(nodePointer.numel() == num_rows + 1, "nodePointer must have num_rows + 1 entries");
  const int64_t nnz = edgeList.numel();
  const c10::OptionalDeviceGuard guard(nodePointer.device());
  const auto i32 = nodePointer.options();
  size_t ws_bytes = 0;
  check_status(dtc_csr_transpose_workspace_size((int32_t)num_rows, (int32_t)num_cols, (int32_t)nnz, &ws_bytes),
               "csr_transpose");
  Tensor ws = torch::empty({(int64_t)ws_bytes}, i32.dtype(torch::kUInt8));
  Tensor rpt = torch::empty({num_cols + 1}, i32);
  Tensor ct = torch::empty({nnz}, i32);
  Tensor perm = torch::empty({nnz}, i32);
  check_status(dtc_csr_transpose(nodePointer.data_ptr<int32_t>(), edgeList.data_ptr<int32_t>(), (int32_t)num_rows,
                                 (int32_t)num_cols, (int32_t)nnz, rpt.data_ptr<int32_t>(), ct.data_ptr<int32_t>(),
                                 perm.data_ptr<int32_t>(), ws.data_ptr(), ws_bytes, current_stream()),
               "csr_transpose");
  return py::make_tuple(rpt, ct, perm);
}

Tensor sddmm_csr(Tensor nodePointer, Tensor edgeList, Tensor X, Tensor Y, c10::optional<Tensor> edgeToRow) {
  CHECK_INPUT(nodePointer);
  CHECK_INPUT(edgeList);
  CHECK_INPUT(X);
  CHECK_INPUT(Y);
  CHECK_DTYPE(nodePointer, torch::kInt32);
  CHECK_DTYPE(edgeList, torch::kInt32);
  CHECK_DTYPE(X, torch::kFloat32);
  TORCH_CHECK(Y.scalar_type() == torch::kFloat32 || Y.scalar_type() == torch::kBFloat16,
              "Y must be float32 or bfloat16, got ", Y.scalar_type());
  TORCH_CHECK(X.dim() == 2 && Y.dim() == 2 && X.size(1) == Y.size(1), "X and Y must be 2-D with equal widths");
  TORCH_CHECK(nodePointer.numel() == X.size(0) + 1, "X must have one row per CSR row");
  const int32_t *e2r = nullptr;
  if (edgeToRow.has_value()) {
    CHECK_INPUT(*edgeToRow);
    CHECK_DTYPE(*edgeToRow, torch::kInt32);
    TORCH_CHECK(edgeToRow->numel() == edgeList.numel(), "edgeToRow must have one entry per edge");
    e2r = edgeToRow->data_ptr<int32_t>();
  }
  const c10::OptionalDeviceGuard guard(X.device());
  Tensor out = torch::empty({edgeList.numel()}, X.options());
  const int32_t y_dtype = Y.scalar_type() == torch::kBFloat16 ? DTC_B_BFLOAT16 : DTC_B_FLOAT32;
  check_status(dtc_sddmm_csr(nodePointer.data_ptr<int32_t>(), edgeList.data_ptr<int32_t>(), e2r,
                             (int32_t)X.size(0), (int32_t)edgeList.numel(), X.data_ptr<float>(), Y.data_ptr(),
                             y_dtype, (int32_t)Y.size(0), (int32_t)X.size(1), out.data_ptr<float>(),
                             current_stream()),
               "sddmm_csr");
  return out;
}

Tensor gather(Tensor values, Tensor index) {
  CHECK_INPUT(values);
  CHECK_INPUT(index);
  CHECK_DTYPE(values, torch::kFloat32);
  CHECK_DTYPE(index, torch::kInt32);
  const c10::OptionalDeviceGuard guard(values.device());
  Tensor out = torch::empty({index.numel()}, values.options());
  check_status(dtc_gather_f32(values.data_ptr<float>(), index.data_ptr<int32_t>(), (int32_t)index.numel(),
                              out.data_ptr<float>(), current_stream()),
               "gather");
  return out;
}

// ------------------------------------------------------ GCN feature transform

// W [K x F] bf16 -> the fragment order dtc_gcn_transform reads (K * F bf16).
Tensor gcn_pack_weights(Tensor W) {
  CHECK_INPUT(W);
  CHECK_DTYPE(W, torch::kBFloat16);
  TORCH_CHECK_VALUE(W.dim() == 2, "W must be 2-D [K x F]");
  const c10::OptionalDeviceGuard guard(W.device());
  Tensor out = torch::empty({W.numel()}, W.options());
  check_status(dtc_gcn_pack_weights(W.data_ptr(), (int32_t)W.size(0), (int32_t)W.size(1), out.data_ptr(),
                                    current_stream()),
               "gcn_pack_weights");
  return out;
}

// B = bf16(act(X) W): X [M x K] float32 or bf16, W packed by gcn_pack_weights
// for K x F; act = ReLU when relu.  Returns bf16 [M x F].
Tensor gcn_transform(Tensor X, Tensor W_packed, int64_t F, bool relu) {
  CHECK_INPUT(X);
  CHECK_INPUT(W_packed);
  CHECK_DTYPE(W_packed, torch::kBFloat16);
  CHECK_SAME_DEVICE(X, W_packed);
  TORCH_CHECK_VALUE(X.dim() == 2, "X must be 2-D [M x K]");
  TORCH_CHECK_VALUE(X.scalar_type() == torch::kFloat32 || X.scalar_type() == torch::kBFloat16,
                    "X must be float32 or bfloat16");
  const int64_t M = X.size(0), K = X.size(1);
  TORCH_CHECK_VALUE(W_packed.numel() == K * F, "packed W must hold K * F elements");
  TORCH_CHECK_VALUE(M < (int64_t(1) << 31) && K < (int64_t(1) << 31) && F < (int64_t(1) << 31), "sizes too large");
  const c10::OptionalDeviceGuard guard(X.device());
  Tensor out = torch::empty({M, F}, X.options().dtype(torch::kBFloat16));
  const int32_t xdt = X.scalar_type() == torch::kFloat32 ? DTC_GCN_X_FLOAT32 : DTC_GCN_X_BFLOAT16;
  check_status(dtc_gcn_transform(X.data_ptr(), xdt, relu ? 1 : 0, W_packed.data_ptr(), (int32_t)M, (int32_t)K,
                                 (int32_t)F, out.data_ptr(), current_stream()),
               "gcn_transform");
  return out;
}

// ------------------------------------------------------------- selector

void check_f32_2d(const Tensor &t, const char *name) {
  TORCH_CHECK(t.is_cuda() && t.is_contiguous() && t.scalar_type() == torch::kFloat32 && t.dim() == 2,
              name, " must be a contiguous 2-D float32 CUDA tensor");
}

Tensor mlp_linear_forward(Tensor x, Tensor w, c10::optional<Tensor> bias, bool relu) {
  check_f32_2d(x, "x");
  check_f32_2d(w, "w");
  TORCH_CHECK(x.size(1) == w.size(0), "x and w shapes do not chain");
  if (bias.has_value()) {
    CHECK_INPUT(*bias);
    TORCH_CHECK(bias->numel() == w.size(1), "bias must have out_dim entries");
  }
  const c10::OptionalDeviceGuard guard(x.device());
  Tensor y = torch::empty({x.size(0), w.size(1)}, x.options());
  check_status(dtc_mlp_linear_forward(x.data_ptr<float>(), w.data_ptr<float>(),
                                      bias.has_value() ? bias->data_ptr<float>() : nullptr, y.data_ptr<float>(),
                                      (int32_t)x.size(0), (int32_t)x.size(1), (int32_t)w.size(1), relu ? 1 : 0,
                                      current_stream()),
               "mlp_linear_forward");
  return y;
}

py::tuple mlp_linear_backward(Tensor x, Tensor w, Tensor y, Tensor dy, bool relu, bool need_dx) {
  check_f32_2d(x, "x");
  check_f32_2d(w, "w");
  check_f32_2d(y, "y");
  check_f32_2d(dy, "dy");
  TORCH_CHECK(y.sizes() == dy.sizes() && y.size(0) == x.size(0) && y.size(1) == w.size(1) && x.size(1) == w.size(0),
              "inconsistent shapes");
  const c10::OptionalDeviceGuard guard(x.device());
  Tensor dw = torch::empty_like(w);
  Tensor db = torch::empty({w.size(1)}, w.options());
  Tensor dx;
  if (need_dx) dx = torch::empty_like(x);
  check_status(dtc_mlp_linear_backward(x.data_ptr<float>(), w.data_ptr<float>(), y.data_ptr<float>(),
                                       dy.data_ptr<float>(), need_dx ? dx.data_ptr<float>() : nullptr,
                                       dw.data_ptr<float>(), db.data_ptr<float>(), (int32_t)x.size(0),
                                       (int32_t)x.size(1), (int32_t)w.size(1), relu ? 1 : 0, current_stream()),
               "mlp_linear_backward");
  return py::make_tuple(need_dx ? py::cast(dx) : py::none(), dw, db);
}

py::tuple mlp_mse(Tensor pred, Tensor target, Tensor mask, double inv_count) {
  check_f32_2d(pred, "pred");
  check_f32_2d(target, "target");
  check_f32_2d(mask, "mask");
  TORCH_CHECK(pred.sizes() == target.sizes() && pred.sizes() == mask.sizes(), "shapes differ");
  const c10::OptionalDeviceGuard guard(pred.device());
  Tensor d = torch::empty_like(pred);
  Tensor loss = torch::empty({1}, pred.options());
  check_status(dtc_mlp_mse(pred.data_ptr<float>(), target.data_ptr<float>(), mask.data_ptr<float>(), pred.numel(),
                           (float)inv_count, d.data_ptr<float>(), loss.data_ptr<float>(), current_stream()),
               "mlp_mse");
  return py::make_tuple(d, loss);
}

void adam_step(Tensor param, Tensor grad, Tensor m, Tensor v, int64_t step, double lr, double beta1, double beta2,
               double eps) {
  for (const Tensor *t : {&param, &grad, &m, &v}) {
    CHECK_INPUT(*t);
    CHECK_DTYPE(*t, torch::kFloat32);
    TORCH_CHECK(t->numel() == param.numel(), "Adam tensors must have the parameter's size");
  }
  const c10::OptionalDeviceGuard guard(param.device());
  check_status(dtc_adam_step(param.data_ptr<float>(), grad.data_ptr<float>(), m.data_ptr<float>(),
                             v.data_ptr<float>(), param.numel(), (float)lr, (float)beta1, (float)beta2, (float)eps,
                             (int32_t)step, current_stream()),
               "adam_step");
}

// ------------------------------------------------------------ row reordering

void check_csr_pair(const Tensor &ptr, const Tensor &idx) {
  CHECK_INPUT(ptr);
  CHECK_INPUT(idx);
  CHECK_DTYPE(ptr, torch::kInt32);
  CHECK_DTYPE(idx, torch::kInt32);
  CHECK_SAME_DEVICE(idx, ptr);
  TORCH_CHECK(ptr.dim() == 1 && ptr.numel() >= 1, "row pointer must be 1-D with at least one entry");
  // The stage entry points trust their CSR; check it here (one host sync).
  TORCH_CHECK(ptr[0].item<int32_t>() == 0 && ptr[-1].item<int64_t>() == idx.numel(),
              "row pointer must start at 0 and end at the number of entries");
  if (ptr.numel() > 1)
    TORCH_CHECK((ptr.slice(0, 1) >= ptr.slice(0, 0, -1)).all().item<bool>(), "row pointer must not decrease");
  if (idx.numel() > 0) TORCH_CHECK(idx.min().item<int32_t>() >= 0, "column indices must be >= 0");
}

// uint32 signatures held in an int32 tensor [num_rows, num_hashes].
Tensor minhash(Tensor nodePointer, Tensor edgeList, int64_t num_hashes, int64_t seed) {
  check_csr_pair(nodePointer, edgeList);
  const int64_t M = nodePointer.numel() - 1;
  const c10::OptionalDeviceGuard guard(nodePointer.device());
  Tensor sig = torch::empty({M, num_hashes}, nodePointer.options());
  check_status(dtc_minhash_signatures(nodePointer.data_ptr<int32_t>(), edgeList.data_ptr<int32_t>(), (int32_t)M,
                                      (int32_t)num_hashes, (uint32_t)seed,
                                      reinterpret_cast<uint32_t *>(sig.data_ptr<int32_t>()), current_stream()),
               "minhash");
  return sig;
}

int64_t default_max_pairs(int64_t n, int64_t bands, int64_t neighbors) {
  return std::max<int64_t>(1, std::min<int64_t>(n * bands * neighbors, int64_t(1) << 28));
}

// Packed candidate pairs (u << 32 | v) as int64.
Tensor lsh_pairs(Tensor signatures, Tensor set_ptr, int64_t bands, int64_t rows_per_band, int64_t max_neighbors,
                 int64_t max_pairs) {
  CHECK_INPUT(signatures);
  CHECK_INPUT(set_ptr);
  CHECK_DTYPE(signatures, torch::kInt32);
  CHECK_DTYPE(set_ptr, torch::kInt32);
  CHECK_SAME_DEVICE(set_ptr, signatures);
  TORCH_CHECK(signatures.dim() == 2 && set_ptr.numel() == signatures.size(0) + 1,
              "signatures must be [n, num_hashes] and set_ptr [n + 1]");
  const int64_t n = signatures.size(0);
  if (max_pairs <= 0) max_pairs = default_max_pairs(n, bands, max_neighbors);
  const c10::OptionalDeviceGuard guard(signatures.device());
  size_t ws_bytes = 0;
  check_status(dtc_lsh_pairs_workspace_size((int32_t)n, max_pairs, &ws_bytes), "lsh_pairs");
  Tensor ws = torch::empty({(int64_t)ws_bytes}, signatures.options().dtype(torch::kUInt8));
  Tensor pairs = torch::empty({max_pairs}, signatures.options().dtype(torch::kInt64));
  int64_t count = 0;
  check_status(dtc_lsh_pairs(reinterpret_cast<const uint32_t *>(signatures.data_ptr<int32_t>()),
                             set_ptr.data_ptr<int32_t>(), (int32_t)n, (int32_t)signatures.size(1), (int32_t)bands,
                             (int32_t)rows_per_band, (int32_t)max_neighbors,
                             reinterpret_cast<uint64_t *>(pairs.data_ptr<int64_t>()), max_pairs, &count,
                             ws.data_ptr(), ws_bytes, current_stream()),
               "lsh_pairs");
  return pairs.slice(0, 0, count).clone();
}

Tensor jaccard_pairs(Tensor set_ptr, Tensor set_idx, Tensor pairs) {
  check_csr_pair(set_ptr, set_idx);
  CHECK_INPUT(pairs);
  CHECK_DTYPE(pairs, torch::kInt64);
  CHECK_SAME_DEVICE(pairs, set_ptr);
  const int64_t n = set_ptr.numel() - 1;
  if (pairs.numel() > 0) {
    const Tensor u = torch::bitwise_right_shift(pairs, 32), v = torch::bitwise_and(pairs, 0xFFFFFFFFLL);
    TORCH_CHECK(u.min().item<int64_t>() >= 0 && torch::maximum(u, v).max().item<int64_t>() < n,
                "pairs name sets out of range");
  }
  const c10::OptionalDeviceGuard guard(pairs.device());
  Tensor sims = torch::empty({pairs.numel()}, set_ptr.options().dtype(torch::kFloat32));
  check_status(dtc_jaccard_pairs(set_ptr.data_ptr<int32_t>(), set_idx.data_ptr<int32_t>(),
                                 reinterpret_cast<const uint64_t *>(pairs.data_ptr<int64_t>()), pairs.numel(),
                                 sims.data_ptr<float>(), current_stream()),
               "jaccard_pairs");
  return sims;
}

// Host tensors in, host roots out.
Tensor tca_greedy(Tensor pairs, Tensor similarity, Tensor list_ptr, Tensor list_idx, int64_t thres) {
  for (const Tensor *t : {&pairs, &similarity, &list_ptr, &list_idx}) {
    TORCH_CHECK(t->device().is_cpu(), "tca_greedy takes host tensors");
    CHECK_CONTIGUOUS(*t);
  }
  CHECK_DTYPE(pairs, torch::kInt64);
  CHECK_DTYPE(similarity, torch::kFloat32);
  CHECK_DTYPE(list_ptr, torch::kInt32);
  CHECK_DTYPE(list_idx, torch::kInt32);
  TORCH_CHECK(pairs.numel() == similarity.numel(), "one similarity per pair");
  const int64_t n = list_ptr.numel() - 1;
  TORCH_CHECK(n >= 0, "list_ptr must have at least one entry");
  Tensor root = torch::empty({n}, list_ptr.options());
  check_status(dtc_tca_greedy((int32_t)n, reinterpret_cast<const uint64_t *>(pairs.data_ptr<int64_t>()),
                              similarity.data_ptr<float>(), pairs.numel(), list_ptr.data_ptr<int32_t>(),
                              list_idx.data_ptr<int32_t>(), (int32_t)thres, root.data_ptr<int32_t>()),
               "tca_greedy");
  return root;
}

py::tuple tca_reorder(Tensor nodePointer, Tensor edgeList, int64_t thres, int64_t cluster_thres, int64_t bands,
                      int64_t rows_per_band, int64_t max_neighbors, int64_t seed, int64_t max_pairs) {
  check_csr_pair(nodePointer, edgeList);
  const int64_t M = nodePointer.numel() - 1;
  if (max_pairs <= 0) max_pairs = default_max_pairs(M, bands, max_neighbors);
  dtc_tca_options_t o{};
  o.thres = (int32_t)thres;
  o.cluster_thres = (int32_t)cluster_thres;
  o.bands = (int32_t)bands;
  o.rows_per_band = (int32_t)rows_per_band;
  o.max_neighbors = (int32_t)max_neighbors;
  o.seed = (uint32_t)seed;
  o.max_pairs = max_pairs;
  const c10::OptionalDeviceGuard guard(nodePointer.device());
  size_t ws_bytes = 0;
  check_status(dtc_tca_reorder_workspace_size((int32_t)M, (int32_t)edgeList.numel(), &o, &ws_bytes), "tca_reorder");
  Tensor ws = torch::empty({(int64_t)ws_bytes}, nodePointer.options().dtype(torch::kUInt8));
  Tensor order = torch::empty({M}, torch::TensorOptions().dtype(torch::kInt32));
  dtc_tca_stats_t stats{};
  check_status(dtc_tca_reorder(nodePointer.data_ptr<int32_t>(), edgeList.data_ptr<int32_t>(), (int32_t)M,
                               (int32_t)edgeList.numel(), &o, order.data_ptr<int32_t>(), &stats, ws.data_ptr(),
                               ws_bytes, current_stream()),
               "tca_reorder");
  py::dict d;
  d["row_pairs"] = stats.row_pairs;
  d["cluster_pairs"] = stats.cluster_pairs;
  d["clusters"] = stats.clusters;
  d["super_clusters"] = stats.super_clusters;
  return py::make_tuple(order, d);
}

}  // namespace

PYBIND11_MODULE(DTCSpMM, m) {
  m.def("preprocess", &preprocess, "TC-GNN format (blockPartition, edgeToColumn, edgeToRow) into CPU tensors",
        py::arg("edgeList"), py::arg("nodePointer"), py::arg("num_nodes"), py::arg("blockSize_h"),
        py::arg("blockSize_w"), py::arg("blockPartition"), py::arg("edgeToColumn"), py::arg("edgeToRow"));
  m.def("run_TCGNNSpMM", &run_TCGNNSpMM, "TC-GNN SpMM (comparison baseline)", py::arg("input"),
        py::arg("nodePointer"), py::arg("edgeList"), py::arg("blockPartition"), py::arg("edgeToColumn"),
        py::arg("edgeToRow"), py::arg("values") = py::none());
  m.def("minhash", &minhash, "MinHash signatures of CSR rows (uint32 bits in int32)", py::arg("nodePointer"),
        py::arg("edgeList"), py::arg("num_hashes") = 56, py::arg("seed") = 2022);
  m.def("lsh_pairs", &lsh_pairs, "LSH candidate pairs, packed u << 32 | v", py::arg("signatures"),
        py::arg("set_ptr"), py::arg("bands") = 28, py::arg("rows_per_band") = 2, py::arg("max_neighbors") = 8,
        py::arg("max_pairs") = 0);
  m.def("jaccard_pairs", &jaccard_pairs, "exact Jaccard similarity of set pairs", py::arg("set_ptr"),
        py::arg("set_idx"), py::arg("pairs"));
  m.def("tca_greedy", &tca_greedy, "greedy pair-merge clustering on the host -> roots", py::arg("pairs"),
        py::arg("similarity"), py::arg("list_ptr"), py::arg("list_idx"), py::arg("thres"));
  m.def("tca_reorder", &tca_reorder, "two-level TCA row reordering -> (reorder_id, stats)", py::arg("nodePointer"),
        py::arg("edgeList"), py::arg("thres") = 16, py::arg("cluster_thres") = 128, py::arg("bands") = 28,
        py::arg("rows_per_band") = 2, py::arg("max_neighbors") = 8, py::arg("seed") = 2022,
        py::arg("max_pairs") = 0);
  m.def("mlp_linear_forward", &mlp_linear_forward, "y = act(x W + b)", py::arg("x"), py::arg("w"),
        py::arg("bias") = py::none(), py::arg("relu") = false);
  m.def("mlp_linear_backward", &mlp_linear_backward, "(dx, dw, db) of one layer", py::arg("x"), py::arg("w"),
        py::arg("y"), py::arg("dy"), py::arg("relu") = false, py::arg("need_dx") = true);
  m.def("mlp_mse", &mlp_mse, "masked MSE -> (dpred, loss)", py::arg("pred"), py::arg("target"), py::arg("mask"),
        py::arg("inv_count"));
  m.def("adam_step", &adam_step, "in-place Adam update", py::arg("param"), py::arg("grad"), py::arg("m"),
        py::arg("v"), py::arg("step"), py::arg("lr"), py::arg("beta1") = 0.9, py::arg("beta2") = 0.999,
        py::arg("eps") = 1e-8);
  m.doc() = "DTC-SpMM for AMD MI355X (gfx950): ME-TCF build and SpMM over libdtc";
  m.def("preprocess_gpu", &preprocess_gpu, "ME-TCF build on the GPU", py::arg("edgeList"), py::arg("nodePointer"),
        py::arg("num_nodes"), py::arg("blockSize_h"), py::arg("blockSize_w"), py::arg("blockPartition"),
        py::arg("edgeToColumn"), py::arg("edgeToRow"), py::arg("values") = py::none(),
        py::arg("return_edge_order") = false);
  m.def("run_DTCSpMM", &run_DTCSpMM, "DTC-SpMM forward, one work unit per row window", py::arg("input"),
        py::arg("Rowwindow_offset"), py::arg("TCblocktile_id"), py::arg("TCblock_offset"), py::arg("sparse_AToX_idx"),
        py::arg("num_nodes"), py::arg("num_edges"), py::arg("exeplan"), py::arg("values") = py::none(),
        py::arg("precision") = "fp32");
  m.def("run_DTCSpMM_balance", &run_DTCSpMM_balance, "DTC-SpMM forward, balanced TC-block chunks", py::arg("input"),
        py::arg("TCblock_rowid"), py::arg("TCblocktile_id"), py::arg("TCblock_offset"), py::arg("sparse_AToX_idx"),
        py::arg("num_nodes"), py::arg("exeplan"), py::arg("values") = py::none(), py::arg("precision") = "fp32",
        py::arg("chunk_blocks") = 0, py::arg("persistent") = false, py::arg("deterministic") = false,
        py::arg("cut_span_blocks") = 0);
  m.def("DTCSpMM_gcn", &DTCSpMM_gcn, "DTC-SpMM forward, single launch, variant chosen from N", py::arg("input"),
        py::arg("Rowwindow_offset"), py::arg("TCblocktile_id"), py::arg("TCblock_offset"), py::arg("sparse_AToX_idx"),
        py::arg("num_nodes"), py::arg("num_edges"), py::arg("values") = py::none());
  // The reference binds these names to the forward functions
  // (DTCSpMM.cpp:386-393): without `transposed` they are the same aliases;
  // with A^T's operands they return the true gradient A^T dY.
  m.def("backward_ptx_uint8_improved", &backward_windows,
        "alias of run_DTCSpMM (reference behaviour); with transposed=A^T's operands, A^T dY", py::arg("input"),
        py::arg("Rowwindow_offset"), py::arg("TCblocktile_id"), py::arg("TCblock_offset"), py::arg("sparse_AToX_idx"),
        py::arg("num_nodes"), py::arg("num_edges"), py::arg("exeplan"), py::arg("values") = py::none(),
        py::arg("precision") = "fp32", py::arg("transposed") = py::none());
  m.def("backward_ptx_uint8_improved_gcn", &backward_windows_gcn,
        "alias of DTCSpMM_gcn (reference behaviour); with transposed=A^T's operands, A^T dY", py::arg("input"),
        py::arg("Rowwindow_offset"), py::arg("TCblocktile_id"), py::arg("TCblock_offset"), py::arg("sparse_AToX_idx"),
        py::arg("num_nodes"), py::arg("num_edges"), py::arg("values") = py::none(), py::arg("transposed") = py::none());
  m.def("balance_backward_ptx_uint8_prefetch", &backward_balanced,
        "alias of run_DTCSpMM_balance (reference behaviour); with transposed=A^T's operands, A^T dY",
        py::arg("input"), py::arg("TCblock_rowid"), py::arg("TCblocktile_id"), py::arg("TCblock_offset"),
        py::arg("sparse_AToX_idx"), py::arg("num_nodes"), py::arg("exeplan"), py::arg("values") = py::none(),
        py::arg("precision") = "fp32", py::arg("chunk_blocks") = 0, py::arg("persistent") = false,
        py::arg("deterministic") = false, py::arg("cut_span_blocks") = 0, py::arg("transposed") = py::none());
  m.def("spmm", &spmm, "SpMM with explicit schedule and tuning knobs", py::arg("input"), py::arg("window_or_rowid"),
        py::arg("TCblocktile_id"), py::arg("TCblock_offset"), py::arg("sparse_AToX_idx"), py::arg("num_nodes"),
        py::arg("schedule") = "balanced", py::arg("values") = py::none(), py::arg("precision") = "fp32",
        py::arg("vec") = 0, py::arg("subtiles") = 0, py::arg("chunk_blocks") = 0, py::arg("persistent") = false,
        py::arg("out") = py::none(), py::arg("deterministic") = false, py::arg("cut_span_blocks") = 0);
  m.def("csr_transpose", &csr_transpose, "CSR transpose on the GPU -> (rowptr_t, col_t, perm)",
        py::arg("nodePointer"), py::arg("edgeList"), py::arg("num_rows"), py::arg("num_cols"));
  m.def("sddmm_csr", &sddmm_csr, "out[e] = X[row(e)] . Y[col(e)]", py::arg("nodePointer"), py::arg("edgeList"),
        py::arg("X"), py::arg("Y"), py::arg("edgeToRow") = py::none());
  m.def("gather", &gather, "out[p] = values[index[p]]", py::arg("values"), py::arg("index"));
  m.def("gcn_pack_weights", &gcn_pack_weights, "W [K x F] bf16 -> the GCN transform's fragment order",
        py::arg("W"));
  m.def("gcn_transform", &gcn_transform, "bf16(act(X) W) as the bf16 B of the SpMM", py::arg("X"),
        py::arg("W_packed"), py::arg("F"), py::arg("relu") = false);
  m.attr("BLK_H") = kBlkH;
  m.attr("BLK_W") = kBlkW;
  m.attr("library_version") = dtc_version();
}
