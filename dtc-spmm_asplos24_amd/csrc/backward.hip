// Gradient building blocks for C = A * B on gfx950.
//
// The reference binds its backward names to the forwards
// (DTC-SpMM/DTCSpMM.cpp:384-399) and so returns A * dY; these kernels supply
// the true gradients instead:
//   dB = A^T * dC  -> the CSR transpose here, then the ME-TCF build and the
//                     SpMM of the transpose;
//   dA = SDDMM     -> dA[e] = dC[row(e), :] . B[col(e), :] on A's pattern.
#include <algorithm>
#include <string>
#include <type_traits>

#include "dtc_internal.h"

namespace dtc {
namespace {

struct TransposeWs {
  int32_t *flags;
  uint64_t *keys0, *keys1;
  int32_t *vals0, *vals1;
  int32_t *e2r;
  void *tmp;
  size_t bytes;
};

TransposeWs carve_t(void *ws, int64_t nnz) {
  Carver c(ws);
  TransposeWs t;
  t.flags = c.take<int32_t>(64);
  const size_t n = (size_t)std::max<int64_t>(nnz, 1);
  t.keys0 = c.take<uint64_t>(n);
  t.keys1 = c.take<uint64_t>(n);
  t.vals0 = c.take<int32_t>(n);
  t.vals1 = c.take<int32_t>(n);
  t.e2r = c.take<int32_t>(n);
  t.tmp = c.take<char>(radix_sort_workspace_bytes(nnz));
  t.bytes = c.used();
  return t;
}

__global__ void k_flags_init(int32_t *f) {
  if (threadIdx.x == 0 && blockIdx.x == 0) f[0] = 0;
}

__global__ __launch_bounds__(256) void k_check_csr(const int32_t *rowptr, const int32_t *col,
                                                   int32_t M, int32_t K, int32_t nnz,
                                                   int32_t *flags) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int f = 0;
  for (int64_t r = tid; r < M; r += stride)
    if (rowptr[r] > rowptr[r + 1]) f |= 1;
  for (int64_t e = tid; e < nnz; e += stride) {
    const int c = col[e];
    if (c < 0 || c >= K) f |= 2;
  }
  if (tid == 0 && (rowptr[0] != 0 || rowptr[M] != nnz)) f |= 4;
  if (f) atomicOr(flags, f);
}

// Row, key (the column) and value (the CSR index) of every edge, edge-balanced.
__global__ __launch_bounds__(256) void k_transpose_keys(const int32_t *rowptr,
                                                        const int32_t *col, int32_t M,
                                                        int32_t nnz, int32_t *e2r,
                                                        uint64_t *keys, int32_t *vals) {
  for_each_edge_row(rowptr, M, nnz, [&](int idx, int row) {
    e2r[idx] = row;
    keys[idx] = (uint32_t)col[idx];
    vals[idx] = idx;
  });
}

__global__ __launch_bounds__(256) void k_transpose_fill(const uint64_t *skeys,
                                                        const int32_t *svals,
                                                        const int32_t *e2r, int32_t nnz,
                                                        int32_t K, int32_t *rowptr_t,
                                                        int32_t *col_t, int32_t *perm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnz) {
    const int src = svals[i];
    perm[i] = src;
    col_t[i] = e2r[src];
  }
  if (i <= K) rowptr_t[i] = lower_bound_dev<uint64_t>(skeys, 0, nnz, (uint64_t)i);
}

// out[e] = x[row(e)] . y[col(e)], edge-parallel: every wavefront takes 8
// consecutive edges at a time (CSR order), all 64 lanes reading VEC floats of
// each edge's two rows (one 1 KB row per edge at n = 256 and VEC = 4), so a
// hub row is spread over many wavefronts.  The row of an edge comes from
// edge_to_row when given, else from a binary search of rowptr (lanes 0-7,
// one edge each).  The 8 partial dot products are reduced together: three
// halving exchanges (4, 2, 1 values) then three plain butterfly steps.
constexpr int kSddmmEdges = 8;

// Y as bfloat16 (YBF = true): VEC bf16 per lane (8 bytes at VEC = 4), taken
// exactly and widened to float32 before the product.
template <int VEC, bool YBF>
struct YVec;
template <int VEC>
struct YVec<VEC, false> {
  typedef float type __attribute__((ext_vector_type(VEC)));
  __device__ static float at(const type &v, int i) { return v[i]; }
};
template <int VEC>
struct YVec<VEC, true> {
  typedef unsigned short type __attribute__((ext_vector_type(VEC)));
  __device__ static float at(const type &v, int i) {
    return __uint_as_float((uint32_t)v[i] << 16);
  }
};

template <int VEC, bool YBF>
__global__ __launch_bounds__(256) void k_sddmm(const int32_t *rowptr, const int32_t *col,
                                               const int32_t *e2r, int32_t M, int32_t nnz,
                                               const float *x, const void *y_, int32_t y_rows,
                                               int32_t n, float *out) {
  typedef float vfloat __attribute__((ext_vector_type(VEC)));
  typedef YVec<VEC, YBF> YV;
  typedef typename YV::type yvec;
  typedef typename std::conditional<YBF, unsigned short, float>::type yelem;
  const yelem *y = static_cast<const yelem *>(y_);
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t groups = ((int64_t)nnz + kSddmmEdges - 1) / kSddmmEdges;
  for (int64_t g = wave; g < groups; g += nwaves) {
    const int64_t e0 = g * kSddmmEdges;
    // rows and columns of the 8 edges (lane k holds edge e0 + k)
    int my_row = 0, my_col = -1;
    if (lane < kSddmmEdges && e0 + lane < nnz) {
      const int e = (int)(e0 + lane);
      my_col = col[e];
      if (e2r) {
        my_row = e2r[e];
      } else {
        int lo = 0, hi = M;  // last row r with rowptr[r] <= e
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (rowptr[mid] <= e) lo = mid;
          else hi = mid;
        }
        my_row = lo;
      }
      if ((unsigned)my_row >= (unsigned)M) my_col = -1;  // defensive: never read outside x
    }
    // Branch-free loads: an edge without a valid column reads row 0 of x and
    // y (both exist: M >= 1 and y_rows >= 1 here) and is masked afterwards,
    // so all 16 row loads of a lane are in flight together.
    int rk[kSddmmEdges], jk[kSddmmEdges];
    bool ok[kSddmmEdges];
#pragma unroll
    for (int k = 0; k < kSddmmEdges; ++k) {
      rk[k] = __builtin_amdgcn_readlane(my_row, k);
      jk[k] = __builtin_amdgcn_readlane(my_col, k);
      ok[k] = (unsigned)jk[k] < (unsigned)y_rows;
      if (!ok[k]) rk[k] = jk[k] = 0;
    }
    float p[kSddmmEdges];
#pragma unroll
    for (int k = 0; k < kSddmmEdges; ++k) p[k] = 0.f;
    // bfloat16 y, all 8 edges in one row (the common case inside long rows):
    // x is loaded once, so a lane moves 16 + 8*8 bytes instead of 8*(16 + 8).
    // rk[] is wave-uniform, so this is a scalar branch; the general path is
    // left as it was rather than copying registers whose loads are still in
    // flight, which would make the loads wait on one another.  Not for
    // float32 y: there the second path costs registers (one wave fewer per
    // SIMD) and measured slower.
    bool one_row = YBF;
#pragma unroll
    for (int k = 1; k < kSddmmEdges; ++k) one_row = one_row && rk[k] == rk[0];
    if (one_row) {
      for (int c = lane * VEC; c < n; c += kWave * VEC) {
        const vfloat xv = *reinterpret_cast<const vfloat *>(x + (int64_t)rk[0] * n + c);
        yvec yv[kSddmmEdges];
#pragma unroll
        for (int k = 0; k < kSddmmEdges; ++k)
          yv[k] = *reinterpret_cast<const yvec *>(y + (int64_t)jk[k] * n + c);
#pragma unroll
        for (int k = 0; k < kSddmmEdges; ++k)
#pragma unroll
          for (int v = 0; v < VEC; ++v) p[k] += xv[v] * YV::at(yv[k], v);
      }
    } else {
      for (int c = lane * VEC; c < n; c += kWave * VEC) {
        vfloat xv[kSddmmEdges];
        yvec yv[kSddmmEdges];
#pragma unroll
        for (int k = 0; k < kSddmmEdges; ++k) {
          xv[k] = *reinterpret_cast<const vfloat *>(x + (int64_t)rk[k] * n + c);
          yv[k] = *reinterpret_cast<const yvec *>(y + (int64_t)jk[k] * n + c);
        }
#pragma unroll
        for (int k = 0; k < kSddmmEdges; ++k)
#pragma unroll
          for (int v = 0; v < VEC; ++v) p[k] += xv[k][v] * YV::at(yv[k], v);
      }
    }
#pragma unroll
    for (int k = 0; k < kSddmmEdges; ++k)
      if (!ok[k]) p[k] = 0.f;
    // halving exchanges: lanes keep the half of the values their bit selects
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
    float q[4], t[2];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float send = b5 ? p[i] : p[i + 4];
      q[i] = (b5 ? p[i + 4] : p[i]) + __shfl_xor(send, 32, kWave);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const float send = b4 ? q[i] : q[i + 2];
      t[i] = (b4 ? q[i + 2] : q[i]) + __shfl_xor(send, 16, kWave);
    }
    float v = (b3 ? t[1] : t[0]) + __shfl_xor(b3 ? t[0] : t[1], 8, kWave);
    v += __shfl_xor(v, 4, kWave);
    v += __shfl_xor(v, 2, kWave);
    v += __shfl_xor(v, 1, kWave);
    const int k = (lane >> 3) & 7;  // edge index this lane's sum belongs to
    if ((lane & 7) == 0 && e0 + k < nnz) out[e0 + k] = v;
  }
}

__global__ void k_zero_f32(float *p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.f;
}

__global__ __launch_bounds__(256) void k_gather(const float *in, const int32_t *idx,
                                                int32_t count, float *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = in[idx[i]];
}

}  // namespace
}  // namespace dtc

using namespace dtc;

extern "C" {

dtc_status_t dtc_csr_transpose_workspace_size(int32_t num_rows, int32_t num_cols, int32_t nnz,
                                              size_t *bytes) {
  clear_error();
  if (!bytes) return fail(DTC_ERROR_INVALID_VALUE, "bytes is NULL");
  if (num_rows < 0 || num_cols < 0 || nnz < 0)
    return fail(DTC_ERROR_INVALID_VALUE, "sizes must be >= 0");
  *bytes = carve_t(nullptr, nnz).bytes;
  return DTC_SUCCESS;
}

dtc_status_t dtc_csr_transpose(const int32_t *node_pointer, const int32_t *edge_list,
                               int32_t num_rows, int32_t num_cols, int32_t nnz,
                               int32_t *node_pointer_t, int32_t *edge_list_t, int32_t *perm,
                               void *workspace, size_t workspace_bytes, dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (num_rows < 0 || num_cols < 0 || nnz < 0)
    return fail(DTC_ERROR_INVALID_VALUE, "sizes must be >= 0");
  if (!node_pointer || !node_pointer_t)
    return fail(DTC_ERROR_INVALID_VALUE, "row pointers are required");
  if (nnz > 0 && (!edge_list || !edge_list_t || !perm))
    return fail(DTC_ERROR_INVALID_VALUE, "edge arrays are required when nnz > 0");
  TransposeWs ws = carve_t(workspace, nnz);
  if (!workspace || workspace_bytes < ws.bytes)
    return fail(DTC_ERROR_WORKSPACE, "transpose workspace needs " + std::to_string(ws.bytes) + " bytes");
  k_flags_init<<<1, 64, 0, stream>>>(ws.flags);
  DTC_LAUNCH_CHECK("flags_init");
  const int64_t work = std::max<int64_t>(std::max<int64_t>(num_rows, nnz), 1);
  k_check_csr<<<(unsigned)std::min<int64_t>(4096, ceil_div(work, 256)), 256, 0, stream>>>(
      node_pointer, edge_list, num_rows, num_cols, nnz, ws.flags);
  DTC_LAUNCH_CHECK("check_csr");
  int32_t flags = 0;
  DTC_HIP_CHECK(hipMemcpyAsync(&flags, ws.flags, 4, hipMemcpyDeviceToHost, stream));
  DTC_HIP_CHECK(hipStreamSynchronize(stream));
  if (flags)
    return fail(DTC_ERROR_INVALID_CSR, "invalid CSR for transpose (flags " + std::to_string(flags) + ")");
  if (nnz > 0) {
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(65536, ceil_div(nnz, 4096)));
    k_transpose_keys<<<grid, 256, 0, stream>>>(node_pointer, edge_list, num_rows, nnz, ws.e2r,
                                               ws.keys0, ws.vals0);
    DTC_LAUNCH_CHECK("transpose_keys");
  }
  int key_bits = 1;
  while (key_bits < 31 && (int64_t(1) << key_bits) < num_cols) ++key_bits;
  bool in_1 = false;
  dtc_status_t st = radix_sort_pairs(ws.keys0, ws.vals0, ws.keys1, ws.vals1, nnz, key_bits,
                                     &in_1, ws.tmp, stream);
  if (st != DTC_SUCCESS) return st;
  const int64_t n = std::max<int64_t>(nnz, (int64_t)num_cols + 1);
  k_transpose_fill<<<(unsigned)ceil_div(n, 256), 256, 0, stream>>>(
      in_1 ? ws.keys1 : ws.keys0, in_1 ? ws.vals1 : ws.vals0, ws.e2r, nnz, num_cols,
      node_pointer_t, edge_list_t, perm);
  DTC_LAUNCH_CHECK("transpose_fill");
  return DTC_SUCCESS;
}

dtc_status_t dtc_sddmm_csr(const int32_t *node_pointer, const int32_t *edge_list,
                           const int32_t *edge_to_row, int32_t num_rows, int32_t nnz,
                           const float *x, const void *y, int32_t y_dtype, int32_t y_rows,
                           int32_t n, float *out, dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (num_rows < 0 || nnz < 0 || y_rows < 0 || n <= 0)
    return fail(DTC_ERROR_INVALID_VALUE, "sizes must be >= 0 and N >= 1");
  if (y_dtype != DTC_B_FLOAT32 && y_dtype != DTC_B_BFLOAT16)
    return fail(DTC_ERROR_INVALID_VALUE, "y_dtype must be DTC_B_FLOAT32 or DTC_B_BFLOAT16");
  if (!node_pointer) return fail(DTC_ERROR_INVALID_VALUE, "node_pointer is NULL");
  if (nnz == 0 || num_rows == 0) return DTC_SUCCESS;
  if (!edge_list || !x || !y || !out) return fail(DTC_ERROR_INVALID_VALUE, "NULL argument");
  if (y_rows == 0) {  // every column is out of range: all zeros
    k_zero_f32<<<(unsigned)std::min<int64_t>(4096, ceil_div(nnz, 256)), 256, 0, stream>>>(out, nnz);
    DTC_LAUNCH_CHECK("zero");
    return DTC_SUCCESS;
  }
  const int64_t groups = ceil_div(nnz, kSddmmEdges);
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1 << 16, ceil_div(groups, 4)));
  const bool ybf = y_dtype == DTC_B_BFLOAT16;
  const bool v4 = n % 4 == 0 && (uintptr_t)x % 16 == 0 && (uintptr_t)y % (ybf ? 8 : 16) == 0;
  if (v4 && ybf)
    k_sddmm<4, true><<<grid, 256, 0, stream>>>(node_pointer, edge_list, edge_to_row, num_rows, nnz,
                                               x, y, y_rows, n, out);
  else if (v4)
    k_sddmm<4, false><<<grid, 256, 0, stream>>>(node_pointer, edge_list, edge_to_row, num_rows, nnz,
                                                x, y, y_rows, n, out);
  else if (ybf)
    k_sddmm<1, true><<<grid, 256, 0, stream>>>(node_pointer, edge_list, edge_to_row, num_rows, nnz,
                                               x, y, y_rows, n, out);
  else
    k_sddmm<1, false><<<grid, 256, 0, stream>>>(node_pointer, edge_list, edge_to_row, num_rows, nnz,
                                                x, y, y_rows, n, out);
  DTC_LAUNCH_CHECK("sddmm");
  return DTC_SUCCESS;
}

dtc_status_t dtc_gather_f32(const float *in, const int32_t *idx, int32_t count, float *out,
                            dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (count < 0) return fail(DTC_ERROR_INVALID_VALUE, "count must be >= 0");
  if (count == 0) return DTC_SUCCESS;
  if (!in || !idx || !out) return fail(DTC_ERROR_INVALID_VALUE, "NULL argument");
  k_gather<<<(unsigned)ceil_div(count, 256), 256, 0, stream>>>(in, idx, count, out);
  DTC_LAUNCH_CHECK("gather");
  return DTC_SUCCESS;
}

}  // extern "C"
