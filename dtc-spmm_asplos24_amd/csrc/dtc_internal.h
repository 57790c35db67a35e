// Internal helpers shared by the libdtc HIP sources (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "dtc.h"

namespace dtc {

constexpr int kWave = 64;

// ---------------------------------------------------------------- status

// Records `msg` as the calling thread's last error detail and returns `st`.
dtc_status_t fail(dtc_status_t st, const std::string &msg);
void clear_error();

#define DTC_HIP_CHECK(expr)                                                        \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return ::dtc::fail(DTC_ERROR_HIP, std::string(#expr " failed: ") +           \
                                            hipGetErrorString(_e));                \
  } while (0)

#define DTC_LAUNCH_CHECK(what)                                                     \
  do {                                                                             \
    hipError_t _e = hipGetLastError();                                             \
    if (_e != hipSuccess)                                                          \
      return ::dtc::fail(DTC_ERROR_HIP, std::string(what " launch failed: ") +     \
                                            hipGetErrorString(_e));                \
  } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Carves aligned sub-buffers out of a caller workspace.
struct Carver {
  char *base;
  size_t off = 0;
  explicit Carver(void *p) : base(static_cast<char *>(p)) {}
  template <class T>
  T *take(size_t count) {
    off = align_up(off, 256);
    T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
  size_t used() const { return align_up(off, 256); }
};

// ------------------------------------------------------ wave64 primitives

__device__ __forceinline__ int lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Orders LDS accesses of one wavefront (lanes exchange data through LDS
// without a workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Inclusive prefix sum across the 64 lanes of a wavefront.
__device__ __forceinline__ int wave_inclusive_sum(int x) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int t = __shfl_up(x, d, kWave);
    if (lane >= d) x += t;
  }
  return x;
}

__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) x = max(x, __shfl_xor(x, d, kWave));
  return x;
}

// Inclusive prefix sum across a 256-thread workgroup; `total` receives the
// workgroup sum.  `smem` needs 4 ints.  Contains __syncthreads().
__device__ __forceinline__ int block256_inclusive_sum(int x, int *smem, int &total) {
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  int inc = wave_inclusive_sum(x);
  if (lane == kWave - 1) smem[wave] = inc;
  __syncthreads();
  int prefix = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const int s = smem[w];
    if (w < wave) prefix += s;
    total += s;
  }
  __syncthreads();
  return inc + prefix;
}

// First index in [lo, hi) with a[idx] >= key (a ascending).
template <class T>
__device__ __forceinline__ int lower_bound_dev(const T *a, int lo, int hi, T key) {
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// First index in [lo, hi) with a[idx] >= key (a ascending), computed by the
// whole wavefront: each round probes 64 evenly spaced positions and keeps the
// one interval the ballot points at, so a search over n elements takes
// ceil(log64 n) rounds of one coalesced load instead of log2 n dependent
// loads.  Must be called by all 64 lanes with the same arguments.
template <class T>
__device__ __forceinline__ int wave_lower_bound(const T *a, int lo, int hi, T key) {
  const int lane = lane_id();
  while (hi - lo > kWave) {
    const int step = (hi - lo + kWave - 1) / kWave;
    const int p = lo + lane * step;
    const bool less = p < hi && a[p] < key;
    const int k = __popcll(__ballot(less));
    if (k == 0) return lo;
    const int nlo = lo + (k - 1) * step + 1;
    const int nhi = min(hi, lo + k * step);
    lo = nlo;
    hi = nhi;
  }
  const int p = lo + lane;
  const bool less = p < hi && a[p] < key;
  return lo + __popcll(__ballot(less));
}

// Calls emit(idx, row) once for every edge idx of a CSR with M rows, the
// edges cut into equal ranges over all waves of the grid (a long row is
// shared by many waves instead of walked by one).  The wave finds the row of
// its first edge by one search, then takes up to 64 edges per step, each
// lane locating its edge's row among the next 64 row ends held in lanes.
// Must be called by whole waves with the same arguments.
// As for_each_edge_row, with a per-edge value loaded at the start of each
// step together with the row ends (one memory round trip per step instead of
// two): load(e) for the step's edge e = p + lane (clamped to the range), and
// emit(idx, row, value) gets the value of its own edge.
template <class Load, class Emit>
__device__ __forceinline__ void for_each_edge_row_load(const int32_t *rowptr, int32_t M, int32_t nnz,
                                                       Load load, Emit emit) {
  const int lane = lane_id();
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave_uniform(threadIdx.x >> 6);
  const int e_begin = (int)((int64_t)nnz * w / nw);
  const int e_end = (int)((int64_t)nnz * (w + 1) / nw);
  if (e_begin >= e_end) return;
  // the row of e_begin: the last r with rowptr[r] <= e_begin
  int r = wave_uniform(wave_lower_bound<int32_t>(rowptr, 0, M + 1, e_begin + 1) - 1);
  int p = e_begin;
  while (p < e_end) {
    const auto val = load(min(p + lane, e_end - 1));
    const int64_t bi = (int64_t)r + 1 + lane;
    const int bnd = bi <= M ? rowptr[bi] : INT32_MAX;
    const uint64_t past = __ballot(bnd > p);
    if (past == 0) {  // rows r .. r+63 all end at or before p (empty rows)
      r += kWave;
      continue;
    }
    const int last = __builtin_amdgcn_readlane(bnd, kWave - 1);  // > p
    const int cnt = min(min(kWave, e_end - p), last - p);
    const int idx = p + min(lane, cnt - 1);
    int lo = 0, hi = kWave - 1;  // first lane whose row end exceeds idx (lane 63's does)
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (__shfl(bnd, mid, kWave) > idx) hi = mid;
      else lo = mid + 1;
    }
    const int row = r + lo;
    if (lane < cnt) emit(idx, row, val);  // idx = p + lane here: val is its own edge's
    r = wave_uniform(__shfl(row, cnt - 1, kWave));  // row of the step's last edge
    p += cnt;
  }
}

template <class Emit>
__device__ __forceinline__ void for_each_edge_row(const int32_t *rowptr, int32_t M, int32_t nnz,
                                                  Emit emit) {
  for_each_edge_row_load(
      rowptr, M, nnz, [](int) { return 0; }, [&](int idx, int row, int) { emit(idx, row); });
}

// ----------------------------------------------------- device-wide pieces
// (implemented in primitives.hip)

// Workspace bytes for a scan over n int32 elements.
size_t scan_workspace_bytes(int64_t n);
// out[i] = sum_{j <= i} in[j] (inclusive) or sum_{j < i} (exclusive).
// `in` and `out` may alias.
dtc_status_t scan_i32(const int32_t *in, int32_t *out, int64_t n, bool inclusive,
                      void *ws, hipStream_t stream);
// Inclusive scan of (keys[i] != keys[i-1]) with keys[-1] treated as different.
dtc_status_t scan_key_changes(const uint64_t *keys, int32_t *out, int64_t n, void *ws,
                              hipStream_t stream);

// Stable LSD radix sort of (key, value) pairs by bits [0, key_bits).
// Result lands in (keys0, vals0) if `*result_in_1` is false, else in
// (keys1, vals1).
size_t radix_sort_workspace_bytes(int64_t n);
dtc_status_t radix_sort_pairs(uint64_t *keys0, int32_t *vals0, uint64_t *keys1,
                              int32_t *vals1, int64_t n, int key_bits, bool *result_in_1,
                              void *ws, hipStream_t stream);

}  // namespace dtc
