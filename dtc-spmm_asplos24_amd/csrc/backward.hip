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

__global__ __launch_bounds__(256) void k_transpose_keys(const int32_t *rowptr,
                                                        const int32_t *col, int32_t M,
                                                        int32_t *e2r, uint64_t *keys,
                                                        int32_t *vals) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t r = wave; r < M; r += nwaves) {
    const int s = rowptr[r];
    const int e = rowptr[r + 1];
    for (int idx = s + lane; idx < e; idx += 64) {
      e2r[idx] = (int32_t)r;
      keys[idx] = (uint32_t)col[idx];
      vals[idx] = idx;
    }
  }
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

// out[e] = x[row(e)] . y[col(e)]: one wavefront per row, four 16-lane groups
// each taking one edge at a time, VEC floats per lane per step.
template <int VEC>
__global__ __launch_bounds__(256) void k_sddmm(const int32_t *rowptr, const int32_t *col,
                                               int32_t M, const float *x, const float *y,
                                               int32_t y_rows, int32_t n, float *out) {
  typedef float vfloat __attribute__((ext_vector_type(VEC)));
  const int lane = threadIdx.x & 63;
  const int grp = lane >> 4;
  const int q = lane & 15;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t r = wave; r < M; r += nwaves) {
    const int s = rowptr[r];
    const int e = rowptr[r + 1];
    const float *xr = x + r * (int64_t)n;
    for (int base = s; base < e; base += 4) {
      const int ed = base + grp;
      float acc = 0.f;
      if (ed < e) {
        const int j = col[ed];
        if ((unsigned)j < (unsigned)y_rows) {
          const float *yr = y + (int64_t)j * n;
          for (int c = q * VEC; c < n; c += 16 * VEC) {
            const vfloat xv = *reinterpret_cast<const vfloat *>(xr + c);
            const vfloat yv = *reinterpret_cast<const vfloat *>(yr + c);
#pragma unroll
            for (int v = 0; v < VEC; ++v) acc += xv[v] * yv[v];
          }
        }
      }
#pragma unroll
      for (int d = 8; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
      if (q == 0 && ed < e) out[ed] = acc;
    }
  }
}

template <>
__global__ __launch_bounds__(256) void k_sddmm<1>(const int32_t *rowptr, const int32_t *col,
                                                  int32_t M, const float *x, const float *y,
                                                  int32_t y_rows, int32_t n, float *out) {
  const int lane = threadIdx.x & 63;
  const int grp = lane >> 4;
  const int q = lane & 15;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t r = wave; r < M; r += nwaves) {
    const int s = rowptr[r];
    const int e = rowptr[r + 1];
    const float *xr = x + r * (int64_t)n;
    for (int base = s; base < e; base += 4) {
      const int ed = base + grp;
      float acc = 0.f;
      if (ed < e) {
        const int j = col[ed];
        if ((unsigned)j < (unsigned)y_rows) {
          const float *yr = y + (int64_t)j * n;
          for (int c = q; c < n; c += 16) acc += xr[c] * yr[c];
        }
      }
#pragma unroll
      for (int d = 8; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
      if (q == 0 && ed < e) out[ed] = acc;
    }
  }
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
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(65536, ceil_div(num_rows, 4)));
    k_transpose_keys<<<grid, 256, 0, stream>>>(node_pointer, edge_list, num_rows, ws.e2r,
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
                           int32_t num_rows, int32_t nnz, const float *x, const float *y,
                           int32_t y_rows, int32_t n, float *out, dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (num_rows < 0 || nnz < 0 || y_rows < 0 || n <= 0)
    return fail(DTC_ERROR_INVALID_VALUE, "sizes must be >= 0 and N >= 1");
  if (!node_pointer) return fail(DTC_ERROR_INVALID_VALUE, "node_pointer is NULL");
  if (nnz == 0 || num_rows == 0) return DTC_SUCCESS;
  if (!edge_list || !x || !y || !out) return fail(DTC_ERROR_INVALID_VALUE, "NULL argument");
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(65536, ceil_div(num_rows, 4)));
  const bool v4 = n % 4 == 0 && (uintptr_t)x % 16 == 0 && (uintptr_t)y % 16 == 0;
  if (v4)
    k_sddmm<4><<<grid, 256, 0, stream>>>(node_pointer, edge_list, num_rows, x, y, y_rows, n, out);
  else
    k_sddmm<1><<<grid, 256, 0, stream>>>(node_pointer, edge_list, num_rows, x, y, y_rows, n, out);
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
