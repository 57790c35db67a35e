// Measurement aid, not part of the libdtc ABI: a kernel that issues exactly
// the B-row loads of the balanced float32 SpMM (same chunk / column-tile
// mapping, 4 TC blocks per group, each 16-lane group reading the 16 * vec
// floats of one row slice per SparseAToX slot) and nothing else -- no A
// tiles, no MFMA, no C stores.  Its time is the practical ceiling of the
// SpMM's access pattern on the memory system (L2 and MALL hits included),
// against which the SpMM's own time is compared.  One float per lane is
// written so that the loads are kept.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

template <int VEC>
struct Vec {
  typedef float type __attribute__((ext_vector_type(VEC)));
};

template <int VEC>
__global__ __launch_bounds__(256) void k_gather_ceiling(const int32_t *__restrict__ atox,
                                                        int32_t T, int32_t chunk,
                                                        const float *__restrict__ B, int32_t N,
                                                        int32_t n_ct, int32_t num_items,
                                                        float *__restrict__ out) {
  typedef typename Vec<VEC>::type vt;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int ct = blockIdx.x % n_ct;
  const int item = (blockIdx.x / n_ct) * 4 + wave;
  if (item >= num_items) return;
  const int b0 = item * chunk;
  const int b1 = min(T, b0 + chunk);
  const int g = lane >> 4, q = lane & 15;
  const float *Bc = B + ct * (16 * VEC) + q * VEC;
  vt acc;
#pragma unroll
  for (int k = 0; k < VEC; ++k) acc[k] = 0.f;
  for (int b = b0; b < b1; b += 4) {
    const int blk = min(b + g, b1 - 1);
    const int4 r0 = *reinterpret_cast<const int4 *>(atox + (int64_t)blk * 8);
    const int4 r1 = *reinterpret_cast<const int4 *>(atox + (int64_t)blk * 8 + 4);
    const int rows[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
    vt v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = *reinterpret_cast<const vt *>(Bc + (int64_t)rows[c] * N);
#pragma unroll
    for (int c = 0; c < 8; ++c) acc += v[c];
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < VEC; ++k) s += acc[k];
  out[((int64_t)blockIdx.x * 4 + wave) * 64 + lane] = s;
}

int64_t items_of(int32_t T, int32_t chunk) { return ((int64_t)T + chunk - 1) / chunk; }

}  // namespace

// vec is 2 or 4 (floats per lane per row, a column tile of 16 * vec).
// Entries of atox must index rows of B ([b_rows x N], N a multiple of
// 16 * vec, rows 4 * vec-byte aligned); the caller checks that before the
// launch.  out holds dtc_gather_ceiling_out_size(T, chunk, N, vec) floats.
extern "C" int64_t dtc_gather_ceiling_out_size(int32_t T, int32_t chunk, int32_t N, int32_t vec) {
  if (T <= 0 || chunk <= 0 || N <= 0 || (vec != 2 && vec != 4) || N % (16 * vec)) return -1;
  return (int64_t)(N / (16 * vec)) * ((items_of(T, chunk) + 3) / 4) * 256;
}

extern "C" int dtc_gather_ceiling(const int32_t *atox, int32_t T, int32_t chunk, const float *B,
                                  int32_t N, int32_t vec, float *out, hipStream_t stream) {
  if (dtc_gather_ceiling_out_size(T, chunk, N, vec) < 0) return 1;
  const int n_ct = N / (16 * vec);
  const int items = (int)items_of(T, chunk);
  const int grid = n_ct * ((items + 3) / 4);
  if (vec == 4) k_gather_ceiling<4><<<grid, 256, 0, stream>>>(atox, T, chunk, B, N, n_ct, items, out);
  else k_gather_ceiling<2><<<grid, 256, 0, stream>>>(atox, T, chunk, B, N, n_ct, items, out);
  return hipGetLastError() == hipSuccess ? 0 : 2;
}
