// Dense half of a GCN layer on gfx950: B = act(X) * W in bf16, written as
// the row-major bf16 B the SpMM gathers (b_dtype bfloat16).
//
// The reference's GCN forward (DTCSpMM_gcn, DTC-SpMM/DTCSpMM_kernel.cu:1033-1116)
// takes X W as a float32 tensor that the caller computed with a torch GEMM;
// here the product is one MFMA kernel whose prologue applies the previous
// layer's activation (ReLU) and the float32 -> bf16 cast while it loads X, and
// whose epilogue rounds to bf16 in the layout and precision the bf16-B SpMM
// reads -- no separate activation, cast or conversion pass over the
// M x F intermediates.
//
// The product is HBM-bound (X read once, B written once: 128 flop per byte at
// K = F = 256, against MI355X's 2.5 PFLOP/s over 8 TB/s = 312), so the
// kernel is built to read X once and keep W in L2:
//   * a workgroup computes 128 rows x 256 columns: 4 waves, 64 columns each
//     (4 MFMA column tiles), 8 row tiles of 16 each -- 128 accumulator VGPRs;
//   * X's 128 rows are staged in LDS as bf16, 128 columns of K at a time
//     (32 KB, rows padded by 16 bytes), shared by the four waves;
//   * W is packed once per call into MFMA fragment order (K/32 x F/16 groups
//     of 64 lanes x 8 bf16), so each B fragment is one contiguous 1 KB read,
//     served by L2 (W is 128 KB at K = F = 256);
//   * v_mfma_f32_16x16x32_bf16: lane l holds A[m = l % 16][k = 8 (l / 16) + j],
//     B[k = 8 (l / 16) + j][n = l % 16] and D[m = 4 (l / 16) + i][n = l % 16];
//   * the epilogue stages each wave's 16 x 64 tile through LDS and writes
//     whole 128-byte row segments.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "dtc.h"
#include "dtc_internal.h"

namespace dtc {
namespace {

typedef __bf16 gbf16x8 __attribute__((ext_vector_type(8)));
typedef float gfloatx4 __attribute__((ext_vector_type(4)));

constexpr int kGemmRows = 64;         // rows per workgroup (128 measured slower: one wave per SIMD)
constexpr int kGemmCols = 256;        // columns per workgroup (4 waves x 64)
constexpr int kWaveCols = 64;         // columns per wave (4 MFMA column tiles)
constexpr int kKChunk = 128;          // K columns of X staged per LDS pass
constexpr int kXStride = kKChunk + 8; // LDS row stride in bf16 (16 bytes of padding)
constexpr int kColTiles = kWaveCols / 16;

// Packs W [K x F] (row-major bf16) into fragment order: group (s, t) for
// k-step s (32 rows of W) and column tile t (16 columns) holds, for lane l,
// W[32 s + 8 (l / 16) + j][16 t + l % 16], j = 0..7, as 8 consecutive bf16.
__global__ __launch_bounds__(256) void k_pack_w(const __bf16 *__restrict__ w, int32_t K, int32_t F,
                                                __bf16 *__restrict__ wp) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one (s, t, lane)
  const int64_t groups = (int64_t)(K / 32) * (F / 16);
  if (gid >= groups * 64) return;
  const int lane = (int)(gid & 63);
  const int64_t grp = gid >> 6;
  const int t = (int)(grp % (F / 16));
  const int s = (int)(grp / (F / 16));
  const int n = 16 * t + (lane & 15);
  const int k0 = 32 * s + 8 * (lane >> 4);
  gbf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = w[(int64_t)(k0 + j) * F + n];
  *reinterpret_cast<gbf16x8 *>(wp + gid * 8) = v;
}

template <bool XF32, bool RELU>
__device__ __forceinline__ __bf16 act(float x) {
  return (__bf16)(RELU ? fmaxf(x, 0.0f) : x);
}

// X[r0 .. r0 + ROWS)[kc .. kc + kw) (kw a multiple of 32, at most 128) is
// staged in two steps, so that the next chunk's loads are in flight while the
// current one is multiplied: load_x reads thread tid's pieces (8 consecutive
// elements of one row; piece idx = tid + 256 q, row idx / (kw / 8)) into
// registers, rows past M as zeros; store_x applies act, rounds to bf16 and
// writes them to LDS.
template <bool XF32, int ROWS>
struct XPieces {
  static constexpr int kPieces = ROWS * kKChunk / 8 / 256;  // per thread, at the widest chunk
  typename std::conditional<XF32, gfloatx4, gbf16x8>::type v[kPieces][XF32 ? 2 : 1];
};

template <bool XF32, int ROWS>
__device__ __forceinline__ void load_x(const void *__restrict__ x, int64_t r0, int32_t M, int32_t K,
                                       int kc, int kw, XPieces<XF32, ROWS> &xp) {
  const int per_row = kw / 8;
#pragma unroll
  for (int q = 0; q < XPieces<XF32, ROWS>::kPieces; ++q) {
    const int idx = threadIdx.x + 256 * q;
    const int r = idx / per_row;
    const int p = idx - r * per_row;
    const int64_t row = r0 + r;
    const bool ok = idx < ROWS * per_row && row < M;
    if constexpr (XF32) {
      const float *src = static_cast<const float *>(x) + (ok ? row * K + kc + 8 * p : 0);
      xp.v[q][0] = ok ? *reinterpret_cast<const gfloatx4 *>(src) : gfloatx4{0.f, 0.f, 0.f, 0.f};
      xp.v[q][1] = ok ? *reinterpret_cast<const gfloatx4 *>(src + 4) : gfloatx4{0.f, 0.f, 0.f, 0.f};
    } else {
      const __bf16 *src = static_cast<const __bf16 *>(x) + (ok ? row * K + kc + 8 * p : 0);
      gbf16x8 z;
#pragma unroll
      for (int j = 0; j < 8; ++j) z[j] = (__bf16)0.0f;
      xp.v[q][0] = ok ? *reinterpret_cast<const gbf16x8 *>(src) : z;
    }
  }
}

template <bool XF32, bool RELU, int ROWS>
__device__ __forceinline__ void store_x(int kw, const XPieces<XF32, ROWS> &xp, __bf16 *s_x) {
  const int per_row = kw / 8;
#pragma unroll
  for (int q = 0; q < XPieces<XF32, ROWS>::kPieces; ++q) {
    const int idx = threadIdx.x + 256 * q;
    if (idx >= ROWS * per_row) continue;  // a narrower last chunk (a guard, not a loop exit)
    const int r = idx / per_row;
    const int p = idx - r * per_row;
    gbf16x8 v;
    if constexpr (XF32) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v[j] = act<XF32, RELU>(xp.v[q][0][j]);
        v[4 + j] = act<XF32, RELU>(xp.v[q][1][j]);
      }
    } else {
      v = xp.v[q][0];
      if constexpr (RELU) {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = act<XF32, RELU>((float)v[j]);
      }
    }
    *reinterpret_cast<gbf16x8 *>(s_x + r * kXStride + 8 * p) = v;
  }
}

// ROWS rows (ROWS / 16 row tiles) per workgroup.
template <bool XF32, bool RELU, int ROWS>
__global__ __launch_bounds__(256) void k_gcn_transform(const void *__restrict__ x,
                                                       const __bf16 *__restrict__ wp, int32_t M,
                                                       int32_t K, int32_t F,
                                                       __bf16 *__restrict__ out) {
  constexpr int kRowTiles = ROWS / 16;
  __shared__ __attribute__((aligned(16))) __bf16 s_x[ROWS * kXStride];  // 34,816 bytes at 128 rows
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int64_t r0 = (int64_t)blockIdx.x * ROWS;
  const int col0 = blockIdx.y * kGemmCols + wave * kWaveCols;  // this wave's first column
  const bool active = col0 < F;
  const int ntile_f = F / 16;
  gfloatx4 acc[kRowTiles][kColTiles];
#pragma unroll
  for (int i = 0; i < kRowTiles; ++i)
#pragma unroll
    for (int c = 0; c < kColTiles; ++c) acc[i][c] = gfloatx4{0.f, 0.f, 0.f, 0.f};
  XPieces<XF32, ROWS> xp;
  load_x<XF32, ROWS>(x, r0, M, K, 0, min(kKChunk, K), xp);
  for (int kc = 0; kc < K; kc += kKChunk) {
    const int kw = min(kKChunk, K - kc);
    __syncthreads();  // the previous chunk's fragments are read
    store_x<XF32, RELU, ROWS>(kw, xp, s_x);
    __syncthreads();
    if (kc + kKChunk < K) load_x<XF32, ROWS>(x, r0, M, K, kc + kKChunk, min(kKChunk, K - kc - kKChunk), xp);
    if (active) {
      for (int ks = 0; ks < kw; ks += 32) {
        const int s = (kc + ks) / 32;
        gbf16x8 bfrag[kColTiles];
#pragma unroll
        for (int c = 0; c < kColTiles; ++c) {
          const int t = col0 / 16 + c;
          bfrag[c] = *reinterpret_cast<const gbf16x8 *>(wp + (((int64_t)s * ntile_f + t) * 64 + lane) * 8);
        }
#pragma unroll
        for (int i = 0; i < kRowTiles; ++i) {
          const gbf16x8 afrag = *reinterpret_cast<const gbf16x8 *>(
              s_x + (16 * i + (lane & 15)) * kXStride + ks + 8 * (lane >> 4));
#pragma unroll
          for (int c = 0; c < kColTiles; ++c)
            acc[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag, bfrag[c], acc[i][c], 0, 0, 0);
        }
      }
    }
  }
  // Epilogue: each wave stages a 16 x 64 bf16 tile (2 KB) in its own part of
  // s_x and writes it out as 16 rows of 128 bytes.
  __syncthreads();  // every wave is done reading X from s_x
  if (!active) return;
  __bf16 *s_o = s_x + wave * (16 * kWaveCols);
#pragma unroll
  for (int i = 0; i < kRowTiles; ++i) {
#pragma unroll
    for (int c = 0; c < kColTiles; ++c)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        s_o[(4 * (lane >> 4) + e) * kWaveCols + 16 * c + (lane & 15)] = (__bf16)acc[i][c][e];
    wave_lds_sync();
    // 16 rows x 64 columns = 128 pieces of 8 bf16: two per lane
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int piece = h * 64 + lane;
      const int r = piece >> 3;
      const int p = piece & 7;
      const int64_t row = r0 + 16 * i + r;
      const gbf16x8 v = *reinterpret_cast<const gbf16x8 *>(s_o + r * kWaveCols + 8 * p);
      if (row < M) *reinterpret_cast<gbf16x8 *>(out + row * F + col0 + 8 * p) = v;
    }
    wave_lds_sync();
  }
}

dtc_status_t check_gcn_shape(int32_t m, int32_t k, int32_t f) {
  if (m < 0 || k <= 0 || f <= 0) return fail(DTC_ERROR_INVALID_VALUE, "M must be >= 0, K and F > 0");
  if (k % 32 != 0) return fail(DTC_ERROR_UNSUPPORTED, "K must be a multiple of 32");
  if (f % 64 != 0) return fail(DTC_ERROR_UNSUPPORTED, "F must be a multiple of 64");
  return DTC_SUCCESS;
}

}  // namespace
}  // namespace dtc

using namespace dtc;

extern "C" {

dtc_status_t dtc_gcn_pack_weights(const void *w, int32_t k, int32_t f, void *w_packed,
                                  dtc_stream_t stream_) {
  clear_error();
  dtc_status_t st = check_gcn_shape(0, k, f);
  if (st != DTC_SUCCESS) return st;
  if (!w || !w_packed) return fail(DTC_ERROR_INVALID_VALUE, "W and the packed buffer are required");
  if (((uintptr_t)w_packed & 15) != 0) return fail(DTC_ERROR_INVALID_VALUE, "the packed buffer must be 16-byte aligned");
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  const int64_t threads = (int64_t)(k / 32) * (f / 16) * 64;
  k_pack_w<<<(unsigned)ceil_div(threads, 256), 256, 0, stream>>>(static_cast<const __bf16 *>(w), k, f,
                                                                static_cast<__bf16 *>(w_packed));
  DTC_LAUNCH_CHECK("gcn_pack_weights");
  return DTC_SUCCESS;
}

dtc_status_t dtc_gcn_transform(const void *x, int32_t x_dtype, int32_t relu, const void *w_packed,
                               int32_t m, int32_t k, int32_t f, void *out, dtc_stream_t stream_) {
  clear_error();
  dtc_status_t st = check_gcn_shape(m, k, f);
  if (st != DTC_SUCCESS) return st;
  if (x_dtype != DTC_GCN_X_FLOAT32 && x_dtype != DTC_GCN_X_BFLOAT16)
    return fail(DTC_ERROR_INVALID_VALUE, "x_dtype must be DTC_GCN_X_FLOAT32 or DTC_GCN_X_BFLOAT16");
  if (relu != 0 && relu != 1) return fail(DTC_ERROR_INVALID_VALUE, "relu must be 0 or 1");
  if (m == 0) return DTC_SUCCESS;
  if (!x || !w_packed || !out) return fail(DTC_ERROR_INVALID_VALUE, "X, packed W and the output are required");
  if (((uintptr_t)x & 15) || ((uintptr_t)w_packed & 15) || ((uintptr_t)out & 15))
    return fail(DTC_ERROR_INVALID_VALUE, "X, packed W and the output must be 16-byte aligned");
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  const __bf16 *wp = static_cast<const __bf16 *>(w_packed);
  __bf16 *o = static_cast<__bf16 *>(out);
  const dim3 grid((unsigned)ceil_div(m, kGemmRows), (unsigned)ceil_div(f, kGemmCols));
  auto launch = [&](auto kernel) { kernel<<<grid, 256, 0, stream>>>(x, wp, m, k, f, o); };
  if (x_dtype == DTC_GCN_X_FLOAT32)
    relu ? launch(k_gcn_transform<true, true, kGemmRows>) : launch(k_gcn_transform<true, false, kGemmRows>);
  else
    relu ? launch(k_gcn_transform<false, true, kGemmRows>) : launch(k_gcn_transform<false, false, kGemmRows>);
  DTC_LAUNCH_CHECK("gcn_transform");
  return DTC_SUCCESS;
}

}  // extern "C"
