// Measurement aid, not part of the libdtc ABI: a kernel that issues exactly
// the B-row loads of the balanced float32 SpMM (same chunk / column-tile
// mapping, 4 TC blocks per group, each 16-lane group reading the 256 bytes of
// one row slice per SparseAToX slot) and nothing else -- no A tiles, no MFMA,
// no C stores.  Its time is the practical ceiling of the SpMM's access
// pattern on the memory system (L2 and MALL hits included), against which
// the SpMM's own time is compared.  One float per lane is written so that
// the loads are kept.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__global__ __launch_bounds__(256) void k_gather_ceiling(const int32_t *__restrict__ atox,
                                                        int32_t T, int32_t chunk,
                                                        const float *__restrict__ B, int32_t N,
                                                        int32_t n_ct, int32_t num_items,
                                                        float *__restrict__ out) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int ct = blockIdx.x % n_ct;
  const int item = (blockIdx.x / n_ct) * 4 + wave;
  if (item >= num_items) return;
  const int b0 = item * chunk;
  const int b1 = min(T, b0 + chunk);
  const int g = lane >> 4, q = lane & 15;
  const float *Bc = B + ct * 64 + q * 4;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int b = b0; b < b1; b += 4) {
    const int blk = min(b + g, b1 - 1);
    const int4 r0 = *reinterpret_cast<const int4 *>(atox + (int64_t)blk * 8);
    const int4 r1 = *reinterpret_cast<const int4 *>(atox + (int64_t)blk * 8 + 4);
    const int rows[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
    float4 v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = *reinterpret_cast<const float4 *>(Bc + (int64_t)rows[c] * N);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      acc.x += v[c].x;
      acc.y += v[c].y;
      acc.z += v[c].z;
      acc.w += v[c].w;
    }
  }
  out[((int64_t)blockIdx.x * 4 + wave) * 64 + lane] = acc.x + acc.y + acc.z + acc.w;
}

}  // namespace

// Entries of atox must index rows of B ([b_rows x N], N a multiple of 64,
// 16-byte aligned); the caller checks that before the launch.  out holds
// dtc_gather_ceiling_out_size(T, chunk, N) floats.
extern "C" int64_t dtc_gather_ceiling_out_size(int32_t T, int32_t chunk, int32_t N) {
  if (T <= 0 || chunk <= 0 || N <= 0 || N % 64) return -1;
  const int64_t items = ((int64_t)T + chunk - 1) / chunk;
  return (int64_t)(N / 64) * ((items + 3) / 4) * 256;
}

extern "C" int dtc_gather_ceiling(const int32_t *atox, int32_t T, int32_t chunk, const float *B,
                                  int32_t N, float *out, hipStream_t stream) {
  if (dtc_gather_ceiling_out_size(T, chunk, N) < 0) return 1;
  const int n_ct = N / 64;
  const int items = (T + chunk - 1) / chunk;
  const int grid = n_ct * ((items + 3) / 4);
  k_gather_ceiling<<<grid, 256, 0, stream>>>(atox, T, chunk, B, N, n_ct, items, out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
