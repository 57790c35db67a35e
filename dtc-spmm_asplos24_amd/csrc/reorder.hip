// TCA-style row reordering (reordering/TCA_reorder.py): rows with similar
// column sets are clustered so that a 16-row window holds fewer distinct
// columns (fewer TC blocks, fewer B rows gathered).
//
// GPU, data-parallel over rows, edges and pairs:
//   * MinHash signatures: sig[r][j] = min over the row's columns c of
//     fmix32(c ^ s_j), s_j = fmix32(j * 0x9E3779B9 + seed); each wavefront
//     takes an equal range of edges, lane j owns hash j (and j + 64), columns
//     are broadcast with readlane.
//   * LSH: per band, a stable radix sort of rows by band key; rows with equal
//     keys at most D positions apart become candidate pairs (one atomic
//     append per workgroup), then one sort + unique over all bands.
//   * Exact Jaccard of each pair over sorted rows, one wavefront per pair
//     (binary search of the shorter row's columns in the longer row, staged
//     in LDS when it is short).
//   * Cluster unions for the second level: entrywise-min signatures and the
//     union column lists by a sort + unique of (cluster, column) keys.
//   * The pairs' pop order (a stable sort by decreasing similarity).
// Host: the greedy merge itself, which pops the most similar pair first and
// re-queues stale pairs (TCA_reorder.py:152-187, :250-286) -- a sequential
// priority-queue algorithm, as in the reference, kept in C++ here.
#include <algorithm>
#include <queue>
#include <vector>

#include "dtc_internal.h"

namespace dtc {
namespace {

__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

__host__ __device__ __forceinline__ uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xFF51AFD7ED558CCDull;
  k ^= k >> 33;
  k *= 0xC4CEB9FE1A85EC53ull;
  k ^= k >> 33;
  return k;
}

constexpr int kMaxHashes = 128;

enum : int { kBadFirst = 1, kBadLast = 2, kBadOrder = 4, kBadColumn = 8, kUnsorted = 16 };

int bits_of(int64_t max_value) {
  int b = 1;
  while (b < 63 && (int64_t(1) << b) <= max_value) ++b;
  return b;
}

unsigned blocks_for(int64_t n, int per_block) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(1 << 20, ceil_div(n, per_block)));
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_zero_u64(unsigned long long *p) {
  if (threadIdx.x == 0) *p = 0ull;
}

// flags[0]: CSR faults; flags[1]: largest column.  One wavefront per row:
// columns must be strictly increasing inside a row (sorted, no duplicates).
// Results are combined per workgroup first, so the grid issues one atomic pair
// per workgroup rather than per wavefront.
__global__ __launch_bounds__(256) void k_check_sorted_csr(const int32_t *rowptr,
                                                         const int32_t *col, int32_t M,
                                                         int32_t nnz, int32_t *flags) {
  __shared__ int s_f[4], s_m[4];
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  const int64_t first = (int64_t)blockIdx.x * 4 + wave;
  const int64_t stride = (int64_t)gridDim.x * 4;
  int f = 0, mx = 0;
  if (first == 0 && lane == 0) {
    if (rowptr[0] != 0) f |= kBadFirst;
    if (rowptr[M] != nnz) f |= kBadLast;
  }
  for (int64_t r = first; r < M; r += stride) {
    const int b = rowptr[r], e = rowptr[r + 1];
    if (b > e || b < 0 || e > nnz) {
      f |= kBadOrder;
      continue;
    }
    for (int p = b + lane; p < e; p += kWave) {
      const int c = col[p];
      if (c < 0) f |= kBadColumn;
      if (p > b && col[p - 1] >= c) f |= kUnsorted;
      mx = max(mx, c);
    }
  }
  for (int d = kWave / 2; d >= 1; d >>= 1) f |= __shfl_xor(f, d, kWave);
  mx = wave_max(mx);
  if (lane == 0) {
    s_f[wave] = f;
    s_m[wave] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int bf = s_f[0] | s_f[1] | s_f[2] | s_f[3];
    const int bm = max(max(s_m[0], s_m[1]), max(s_m[2], s_m[3]));
    if (bf) atomicOr(&flags[0], bf);
    atomicMax(&flags[1], bm);
  }
}

// Edge-balanced: wavefront w takes the CSR edges [nnz*w/NW, nnz*(w+1)/NW)
// (nnz = rowptr[M]), walks the rows they belong to, and for each row segment
// folds the segment's columns into lane j's minimum of hash j (and j + 64);
// columns are broadcast 64 at a time with readlane.  A segment that is a
// whole row is stored, a partial one (a row cut by a range boundary, such as
// a hub row spread over many wavefronts) is merged with atomicMin into the
// preset 0xFFFFFFFF.  Empty rows keep the preset.
__global__ __launch_bounds__(256) void k_minhash(const int32_t *rowptr, const int32_t *col,
                                                int32_t M, int K, uint32_t seed,
                                                uint32_t *sig) {
  const int lane = lane_id();
  const int64_t nw = (int64_t)gridDim.x * 4;
  const int64_t w = (int64_t)blockIdx.x * 4 + wave_uniform(threadIdx.x >> 6);
  const int64_t nnz = wave_uniform(rowptr[M]);
  const int e_begin = (int)(nnz * w / nw);
  const int e_end = (int)(nnz * (w + 1) / nw);
  if (e_begin >= e_end) return;
  const uint32_t s0 = fmix32((uint32_t)lane * 0x9E3779B9u + seed);
  const uint32_t s1 = fmix32((uint32_t)(lane + kWave) * 0x9E3779B9u + seed);
  const bool two = K > kWave;
  // the row holding e_begin: the last r with rowptr[r] <= e_begin
  int r = wave_uniform(wave_lower_bound<int32_t>(rowptr, 0, M + 1, e_begin + 1) - 1);
  int p = e_begin;
  while (p < e_end) {
    const int rs = wave_uniform(rowptr[r]);
    const int re = wave_uniform(rowptr[r + 1]);
    const int q = min(re, e_end);
    uint32_t m0 = 0xFFFFFFFFu, m1 = 0xFFFFFFFFu;
    for (int c0 = p; c0 < q; c0 += kWave) {
      const int cnt = min(kWave, q - c0);
      const int c = lane < cnt ? col[c0 + lane] : 0;
      for (int t = 0; t < cnt; ++t) {
        const uint32_t x = (uint32_t)__builtin_amdgcn_readlane(c, t);
        m0 = min(m0, fmix32(x ^ s0));
        if (two) m1 = min(m1, fmix32(x ^ s1));
      }
    }
    uint32_t *out = sig + (int64_t)r * K;
    if (p == rs && q == re) {
      if (lane < K) out[lane] = m0;
      if (lane + kWave < K) out[lane + kWave] = m1;
    } else {
      if (lane < K) atomicMin(&out[lane], m0);
      if (lane + kWave < K) atomicMin(&out[lane + kWave], m1);
    }
    p = q;
    ++r;
    while (r < M && wave_uniform(rowptr[r + 1]) <= p) ++r;  // skip empty rows
  }
}

__global__ void k_band_keys(const uint32_t *sig, int32_t n, int K, int band, int r,
                            uint64_t *keys, int32_t *vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t *s = sig + i * K + (int64_t)band * r;
  uint64_t k;
  if (r == 1) {
    k = s[0];
  } else if (r == 2) {
    k = ((uint64_t)s[0] << 32) | s[1];
  } else {
    k = 0;
    for (int t = 0; t < r; ++t) k = fmix64(k ^ ((uint64_t)s[t] + 0x9E3779B97F4A7C15ull));
  }
  keys[i] = k;
  vals[i] = (int32_t)i;
}

// Candidate pairs of one band: sorted positions i and i + d (d <= D) with
// equal keys, both rows non-empty.  Pairs are packed (min << pb) | max and
// appended at count[0] with one atomic per workgroup (the workgroup's pairs
// are counted from the per-d ballots first); writes past `cap` are dropped
// (the count still grows, so the host sees the overflow).
constexpr int kMaxNeighbors = 16;

__global__ __launch_bounds__(256) void k_emit_pairs(const uint64_t *keys, const int32_t *rows,
                                                   int32_t n, const int32_t *set_ptr, int D,
                                                   int pb, uint64_t *pairs, int64_t cap,
                                                   unsigned long long *count) {
  __shared__ unsigned s_cnt[4];
  __shared__ unsigned long long s_base;
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < n;
  const uint64_t ki = in ? keys[i] : 0ull;
  const int u = in ? rows[i] : 0;
  const bool u_live = in && set_ptr[u + 1] > set_ptr[u];
  uint64_t pk[kMaxNeighbors];
  uint64_t mask[kMaxNeighbors];
  unsigned total = 0;
#pragma unroll
  for (int d = 1; d <= kMaxNeighbors; ++d) {
    bool emit = false;
    pk[d - 1] = 0;
    if (d <= D && u_live && i + d < n && keys[i + d] == ki) {
      const int v = rows[i + d];
      if (set_ptr[v + 1] > set_ptr[v]) {
        emit = true;
        pk[d - 1] = ((uint64_t)min(u, v) << pb) | (uint64_t)max(u, v);
      }
    }
    mask[d - 1] = __ballot(emit);
    total += (unsigned)__popcll(mask[d - 1]);
  }
  if (lane == 0) s_cnt[wave] = total;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned all = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    s_base = all ? atomicAdd(count, (unsigned long long)all) : 0ull;
  }
  __syncthreads();
  int64_t pos = (int64_t)s_base;
  for (int w2 = 0; w2 < wave; ++w2) pos += s_cnt[w2];
#pragma unroll
  for (int d = 0; d < kMaxNeighbors; ++d) {
    const uint64_t m = mask[d];
    const int64_t at = pos + __popcll(m & lt_mask);
    if (((m >> lane) & 1ull) && at < cap) pairs[at] = pk[d];
    pos += __popcll(m);
  }
}

// Unique sorted keys -> packed (u << 32 | v) pairs; G is the inclusive scan
// of key changes.
__global__ void k_unique_pairs(const uint64_t *keys, const int32_t *G, int64_t n, int pb,
                               uint64_t *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  if (i == 0 || keys[i - 1] != k) {
    const uint64_t lo_mask = (1ull << pb) - 1;
    out[G[i] - 1] = ((k >> pb) << 32) | (k & lo_mask);
  }
}

// One wavefront per pair.  The longer list, when it has at most
// kJaccardLds entries, is staged in the wavefront's LDS slice and the shorter
// list's entries (one per lane, 64 at a time) are binary-searched there;
// longer lists are searched in global memory.
constexpr int kJaccardLds = 512;

__global__ __launch_bounds__(256) void k_jaccard(const int32_t *set_ptr, const int32_t *col,
                                                const uint64_t *pairs, int64_t P,
                                                float *sims) {
  __shared__ int32_t s_b[4][kJaccardLds];
  const int lane = lane_id();
  const int wave = wave_uniform(threadIdx.x >> 6);
  const int64_t p = (int64_t)blockIdx.x * 4 + wave;
  if (p >= P) return;
  const uint64_t pk = pairs[p];
  int u = (int)(pk >> 32), v = (int)(pk & 0xFFFFFFFFu);
  int ua = set_ptr[u], la = set_ptr[u + 1] - ua;
  int va = set_ptr[v], lb = set_ptr[v + 1] - va;
  if (la > lb) {
    const int t0 = ua, t1 = la;
    ua = va;
    la = lb;
    va = t0;
    lb = t1;
  }
  la = wave_uniform(la);
  lb = wave_uniform(lb);
  int hits = 0;
  if (lb <= kJaccardLds) {
    int32_t *sb = s_b[wave];
    for (int k = lane; k < lb; k += kWave) sb[k] = col[va + k];
    wave_lds_sync();
    for (int k = lane; k < la; k += kWave) {
      const int x = col[ua + k];
      int lo = 0, hi = lb;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sb[mid] < x) lo = mid + 1;
        else hi = mid;
      }
      hits += (lo < lb && sb[lo] == x) ? 1 : 0;
    }
  } else {
    for (int k = lane; k < la; k += kWave) {
      const int x = col[ua + k];
      int lo = 0, hi = lb;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[va + mid] < x) lo = mid + 1;
        else hi = mid;
      }
      hits += (lo < lb && col[va + lo] == x) ? 1 : 0;
    }
  }
  const int inter = wave_sum(hits);
  if (lane == 0)
    sims[p] = la == 0 ? 0.0f : (float)inter / (float)(la + lb - inter);
}

// sig_c[group[r]][j] = min over rows r of the group (sig_c preset to ~0).
__global__ void k_group_min(const uint32_t *sig, int32_t M, int K, const int32_t *group,
                            uint32_t *sig_c) {
  const int64_t total = (int64_t)M * K;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / K;
    const int j = (int)(t - r * K);
    const uint32_t v = sig[t];
    if (v != 0xFFFFFFFFu) atomicMin(&sig_c[(int64_t)group[r] * K + j], v);
  }
}

// (group, column) keys of every edge; one wavefront per row.
__global__ __launch_bounds__(256) void k_union_keys(const int32_t *rowptr, const int32_t *col,
                                                   int32_t M, const int32_t *group, int cb,
                                                   uint64_t *keys, int32_t *vals) {
  const int lane = lane_id();
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= M) return;
  const uint64_t g = (uint64_t)group[r] << cb;
  const int b = rowptr[r], e = rowptr[r + 1];
  for (int p = b + lane; p < e; p += kWave) {
    keys[p] = g | (uint32_t)col[p];
    vals[p] = 0;
  }
}

// Unique (group, column) keys -> column list and group of every entry.
__global__ void k_union_compact(const uint64_t *keys, const int32_t *G, int64_t n, int cb,
                                int32_t *ucol, int32_t *ugroup) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  if (i == 0 || keys[i - 1] != k) {
    ucol[G[i] - 1] = (int32_t)(k & ((1ull << cb) - 1));
    ugroup[G[i] - 1] = (int32_t)(k >> cb);
  }
}

// uptr[g] = first entry of group g (ugroup is non-decreasing), g = 0..C.
__global__ void k_union_ptr(const int32_t *ugroup, int32_t U, int32_t C, int32_t *uptr) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g > C) return;
  uptr[g] = lower_bound_dev<int32_t>(ugroup, 0, U, (int32_t)g);
}

// ---------------------------------------------------------------- host

struct LshWs {
  uint64_t *bkeys0, *bkeys1;
  int32_t *bvals0, *bvals1;
  uint64_t *pkeys0, *pkeys1;
  int32_t *pvals0, *pvals1;
  int32_t *G;
  unsigned long long *count;
  void *tmp;
  size_t bytes;
};

LshWs carve_lsh(void *base, int64_t n, int64_t cap) {
  Carver c(base);
  LshWs w{};
  w.bkeys0 = c.take<uint64_t>(n);
  w.bkeys1 = c.take<uint64_t>(n);
  w.bvals0 = c.take<int32_t>(n);
  w.bvals1 = c.take<int32_t>(n);
  w.pkeys0 = c.take<uint64_t>(cap);
  w.pkeys1 = c.take<uint64_t>(cap);
  w.pvals0 = c.take<int32_t>(cap);
  w.pvals1 = c.take<int32_t>(cap);
  w.G = c.take<int32_t>(cap);
  w.count = c.take<unsigned long long>(1);
  const int64_t m = std::max<int64_t>(n, cap);
  w.tmp = c.take<char>(std::max(radix_sort_workspace_bytes(m), scan_workspace_bytes(m)));
  w.bytes = c.used();
  return w;
}

dtc_status_t check_lsh_args(int32_t n, int32_t K, int32_t bands, int32_t r, int32_t D,
                            int64_t max_pairs) {
  if (n < 0) return fail(DTC_ERROR_INVALID_VALUE, "number of sets must be >= 0");
  if (K < 1 || K > kMaxHashes)
    return fail(DTC_ERROR_INVALID_VALUE, "num_hashes must be in [1, 128]");
  if (bands < 1 || r < 1 || (int64_t)bands * r > K)
    return fail(DTC_ERROR_INVALID_VALUE, "bands * rows_per_band must be in [1, num_hashes]");
  if (D < 1 || D > kMaxNeighbors) return fail(DTC_ERROR_INVALID_VALUE, "max_neighbors must be in [1, 16]");
  if (max_pairs < 0 || max_pairs > INT32_MAX)
    return fail(DTC_ERROR_INVALID_VALUE, "max_pairs must be in [0, 2^31)");
  return DTC_SUCCESS;
}

// LSH candidates of n sets (signatures sig[n][K], set sizes from set_ptr)
// into out[0, *num_pairs).
dtc_status_t lsh_pairs(const uint32_t *sig, const int32_t *set_ptr, int32_t n, int32_t K,
                       int32_t bands, int32_t r, int32_t D, uint64_t *out, int64_t max_pairs,
                       int64_t *num_pairs, const LshWs &w, hipStream_t stream) {
  *num_pairs = 0;
  if (n < 2 || max_pairs == 0) return DTC_SUCCESS;
  const int pb = bits_of(n - 1);
  k_zero_u64<<<1, 64, 0, stream>>>(w.count);
  DTC_LAUNCH_CHECK("zero");
  for (int b = 0; b < bands; ++b) {
    k_band_keys<<<blocks_for(n, 256), 256, 0, stream>>>(sig, n, K, b, r, w.bkeys0, w.bvals0);
    DTC_LAUNCH_CHECK("band_keys");
    bool in_1 = false;
    dtc_status_t st = radix_sort_pairs(w.bkeys0, w.bvals0, w.bkeys1, w.bvals1, n, 64, &in_1,
                                       w.tmp, stream);
    if (st != DTC_SUCCESS) return st;
    k_emit_pairs<<<blocks_for(n, 256), 256, 0, stream>>>(
        in_1 ? w.bkeys1 : w.bkeys0, in_1 ? w.bvals1 : w.bvals0, n, set_ptr, D, pb, w.pkeys0,
        max_pairs, w.count);
    DTC_LAUNCH_CHECK("emit_pairs");
  }
  unsigned long long emitted = 0;
  DTC_HIP_CHECK(hipMemcpyAsync(&emitted, w.count, sizeof(emitted), hipMemcpyDeviceToHost, stream));
  DTC_HIP_CHECK(hipStreamSynchronize(stream));
  if ((int64_t)emitted > max_pairs)
    return fail(DTC_ERROR_WORKSPACE, "LSH produced " + std::to_string(emitted) +
                                         " candidate pairs before deduplication; max_pairs is " +
                                         std::to_string(max_pairs));
  const int64_t E = (int64_t)emitted;
  if (E == 0) return DTC_SUCCESS;
  bool in_1 = false;
  dtc_status_t st = radix_sort_pairs(w.pkeys0, w.pvals0, w.pkeys1, w.pvals1, E, 2 * pb, &in_1,
                                     w.tmp, stream);
  if (st != DTC_SUCCESS) return st;
  const uint64_t *sorted = in_1 ? w.pkeys1 : w.pkeys0;
  st = scan_key_changes(sorted, w.G, E, w.tmp, stream);
  if (st != DTC_SUCCESS) return st;
  k_unique_pairs<<<blocks_for(E, 256), 256, 0, stream>>>(sorted, w.G, E, pb, out);
  DTC_LAUNCH_CHECK("unique_pairs");
  int32_t U = 0;
  DTC_HIP_CHECK(hipMemcpyAsync(&U, w.G + (E - 1), 4, hipMemcpyDeviceToHost, stream));
  DTC_HIP_CHECK(hipStreamSynchronize(stream));
  *num_pairs = U;
  return DTC_SUCCESS;
}

// Stream order of the pairs: a stable sort of the key-ordered pairs by
// decreasing similarity (similarities are >= 0, so ~bits orders them).
__global__ void k_sim_keys(const float *sims, int64_t P, uint64_t *keys, int32_t *vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  keys[i] = (uint64_t)(0xFFFFFFFFu - __float_as_uint(sims[i]));
  vals[i] = (int32_t)i;
}

// --- greedy merge (host).  The queue pops the most similar pair first and,
// among equal similarities, the smaller (p1, p2) first: a total order, so the
// result does not depend on how the queue is built.  The initial pairs are
// consumed as one stream pre-sorted in that order (sorted on the GPU by the
// composite entry point); only re-queued pairs go through a binary heap.
// "Is this pair queued" covers both: an initial pair is queued until the
// stream passes it (its stream position is found by binary search of the
// pairs in key order), a re-queued one while it sits in the heap.

struct QItem {
  double simi;
  int32_t p1, p2;
};

// a before b in pop order
inline bool q_first(const QItem &a, const QItem &b) {
  if (a.simi != b.simi) return a.simi > b.simi;
  if (a.p1 != b.p1) return a.p1 < b.p1;
  return a.p2 < b.p2;
}

struct QLater {  // std::priority_queue keeps the element no other one precedes on top
  bool operator()(const QItem &a, const QItem &b) const { return q_first(b, a); }
};

// |S1 ∩ S2| / |S1 ∪ S2| of sorted duplicate-free lists (TCA_reorder.py:95-100).
double jaccard_host(const int32_t *a, int la, const int32_t *b, int lb) {
  if (la == 0 || lb == 0) return 0.0;
  int i = 0, j = 0, inter = 0;
  while (i < la && j < lb) {
    if (a[i] < b[j]) ++i;
    else if (b[j] < a[i]) ++j;
    else {
      ++inter;
      ++i;
      ++j;
    }
  }
  return (double)inter / (double)(la + lb - inter);
}

inline uint64_t pair_key(int32_t a, int32_t b) {
  return ((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
}

// Open-addressing hash map from pair keys (< 2^63) to int64, linear probing,
// tombstones on erase: one or two cache lines per lookup where a sorted-array
// search of ten million keys costs twenty-odd misses.
class PairMap {
 public:
  explicit PairMap(size_t expected) {
    size_t cap = 16;
    while (cap < 2 * expected + 16) cap <<= 1;
    keys_.assign(cap, kEmpty);
    vals_.assign(cap, 0);
    mask_ = cap - 1;
  }
  void put(uint64_t k, int64_t v) {
    if (2 * (used_ + 1) > keys_.size()) grow();
    size_t i = slot(k), tomb = SIZE_MAX;
    while (keys_[i] != kEmpty) {
      if (keys_[i] == k) {
        vals_[i] = v;
        return;
      }
      if (keys_[i] == kTomb && tomb == SIZE_MAX) tomb = i;
      i = (i + 1) & mask_;
    }
    if (tomb != SIZE_MAX) i = tomb;
    else ++used_;
    keys_[i] = k;
    vals_[i] = v;
  }
  const int64_t *get(uint64_t k) const {
    for (size_t i = slot(k); keys_[i] != kEmpty; i = (i + 1) & mask_)
      if (keys_[i] == k) return &vals_[i];
    return nullptr;
  }
  void erase(uint64_t k) {
    for (size_t i = slot(k); keys_[i] != kEmpty; i = (i + 1) & mask_)
      if (keys_[i] == k) {
        keys_[i] = kTomb;
        return;
      }
  }

 private:
  static constexpr uint64_t kEmpty = ~0ull, kTomb = ~0ull - 1;
  size_t slot(uint64_t k) const { return (size_t)fmix64(k) & mask_; }
  void grow() {  // rehash without tombstones, at least 4x the live keys
    std::vector<uint64_t> ok;
    std::vector<int64_t> ov;
    ok.swap(keys_);
    ov.swap(vals_);
    size_t live = 0;
    for (uint64_t k : ok) live += (k != kEmpty && k != kTomb) ? 1 : 0;
    size_t cap = 16;
    while (cap < 4 * live + 16) cap <<= 1;
    keys_.assign(cap, kEmpty);
    vals_.assign(cap, 0);
    mask_ = keys_.size() - 1;
    used_ = 0;
    for (size_t i = 0; i < ok.size(); ++i)
      if (ok[i] != kEmpty && ok[i] != kTomb) put(ok[i], ov[i]);
  }
  std::vector<uint64_t> keys_;
  std::vector<int64_t> vals_;
  size_t mask_ = 0, used_ = 0;  // used_ counts live keys and tombstones
};

// The initial pairs in stream (pop) order, and each pair's stream position.
struct GreedyInput {
  std::vector<QItem> stream;
  PairMap rank{0};
};

// From key-sorted unique pairs (u < v), their similarities and the stream
// order perm (stream position j holds key index perm[j]).
GreedyInput greedy_input_sorted(const uint64_t *keys, const float *sims, const int32_t *perm,
                                int64_t P) {
  GreedyInput g;
  g.stream.resize((size_t)P);
  g.rank = PairMap((size_t)P);
  for (int64_t j = 0; j < P; ++j) {
    const int64_t k = perm[j];
    g.stream[j] = QItem{(double)sims[k], (int32_t)(keys[k] >> 32), (int32_t)(keys[k] & 0xFFFFFFFFu)};
    g.rank.put(keys[k], j);
  }
  return g;
}

// From pairs in any order (host sorts).  Pairs with a == b are dropped, and
// of repeated pairs the first one given is kept (the reference's queue skips
// later copies, TCA_reorder.py:139).
GreedyInput greedy_input(const uint64_t *pairs, const float *sims, int64_t P) {
  std::vector<int64_t> idx;
  idx.reserve((size_t)P);
  for (int64_t i = 0; i < P; ++i) {
    const int32_t a = (int32_t)(pairs[i] >> 32), b = (int32_t)(pairs[i] & 0xFFFFFFFFu);
    if (a != b) idx.push_back(i);
  }
  auto key_of = [&](int64_t i) {
    return pair_key((int32_t)(pairs[i] >> 32), (int32_t)(pairs[i] & 0xFFFFFFFFu));
  };
  std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
    const uint64_t ka = key_of(a), kb = key_of(b);
    return ka != kb ? ka < kb : a < b;
  });
  std::vector<uint64_t> keys;
  std::vector<float> ksims;
  for (size_t t = 0; t < idx.size(); ++t) {
    const uint64_t k = key_of(idx[t]);
    if (!keys.empty() && keys.back() == k) continue;
    keys.push_back(k);
    ksims.push_back(sims[idx[t]]);
  }
  const int64_t U = (int64_t)keys.size();
  std::vector<int32_t> perm((size_t)U);
  for (int64_t k = 0; k < U; ++k) perm[k] = (int32_t)k;
  std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return ksims[a] > ksims[b]; });
  return greedy_input_sorted(keys.data(), ksims.data(), perm.data(), U);
}

void greedy(int32_t n, const GreedyInput &in, const int32_t *lptr, const int32_t *lidx,
            int32_t thres, int32_t *root_out) {
  const int64_t P = (int64_t)in.stream.size();
  int64_t next = 0;  // stream position of the next initial pair
  std::priority_queue<QItem, std::vector<QItem>, QLater> heap;
  PairMap in_heap(1024);
  auto queued = [&](uint64_t key) {
    if (in_heap.get(key)) return true;
    const int64_t *r = in.rank.get(key);
    return r && *r >= next;
  };
  std::vector<int32_t> cid(n), csz(n, 1);
  std::vector<uint8_t> deleted(n, 0);
  for (int32_t i = 0; i < n; ++i) cid[i] = i;
  int64_t num_cluster = n;
  auto root = [&cid](int32_t i) {
    while (i != cid[i]) {
      cid[i] = cid[cid[i]];
      i = cid[i];
    }
    return i;
  };
  while ((next < P || !heap.empty()) && num_cluster > 0) {
    QItem it;
    if (heap.empty() || (next < P && q_first(in.stream[next], heap.top()))) {
      it = in.stream[next++];
    } else {
      it = heap.top();
      heap.pop();
      in_heap.erase(pair_key(it.p1, it.p2));
    }
    int32_t p1 = it.p1, p2 = it.p2;
    if (p1 == cid[p1] && p2 == cid[p2]) {
      if (deleted[p1] || deleted[p2]) continue;
      if (csz[p1] < csz[p2]) std::swap(p1, p2);  // p1 absorbs p2 (ties: p1 keeps its root)
      cid[p2] = p1;
      --num_cluster;
      csz[p1] += csz[p2];
      if (csz[p1] >= thres) {
        deleted[p1] = 1;
        --num_cluster;
      }
    } else {
      p1 = root(p1);
      p2 = root(p2);
      if (deleted[p1] || deleted[p2]) continue;
      const uint64_t key = pair_key(p1, p2);
      if (p1 != p2 && !queued(key)) {
        heap.push(QItem{jaccard_host(lidx + lptr[p1], lptr[p1 + 1] - lptr[p1], lidx + lptr[p2],
                                     lptr[p2 + 1] - lptr[p2]),
                        p1, p2});
        in_heap.put(key, 1);
      }
    }
  }
  for (int32_t i = 0; i < n; ++i) root_out[i] = root(i);
}

// Stream order of P key-ordered pairs whose similarities are on the device
// (sorted there; the LSH workspace's pair buffers are free again by now).
dtc_status_t stream_order(const float *sims, int64_t P, const LshWs &w, hipStream_t stream,
                          std::vector<int32_t> &perm) {
  perm.resize((size_t)P);
  if (P == 0) return DTC_SUCCESS;
  k_sim_keys<<<blocks_for(P, 256), 256, 0, stream>>>(sims, P, w.pkeys0, w.pvals0);
  DTC_LAUNCH_CHECK("sim_keys");
  bool in_1 = false;
  dtc_status_t st = radix_sort_pairs(w.pkeys0, w.pvals0, w.pkeys1, w.pvals1, P, 32, &in_1, w.tmp,
                                     stream);
  if (st != DTC_SUCCESS) return st;
  DTC_HIP_CHECK(hipMemcpyAsync(perm.data(), in_1 ? w.pvals1 : w.pvals0, (size_t)P * 4,
                               hipMemcpyDeviceToHost, stream));
  DTC_HIP_CHECK(hipStreamSynchronize(stream));
  return DTC_SUCCESS;
}

// Items grouped by root, groups in first-occurrence order, members
// ascending: index[i] = group of item i, members in CSR form.
int32_t group_in_order(const std::vector<int32_t> &roots, std::vector<int32_t> &index,
                       std::vector<int32_t> &gptr, std::vector<int32_t> &members) {
  const int32_t n = (int32_t)roots.size();
  std::vector<int32_t> slot(n, -1);
  index.assign(n, 0);
  int32_t G = 0;
  for (int32_t i = 0; i < n; ++i) {
    int32_t &s = slot[roots[i]];
    if (s < 0) s = G++;
    index[i] = s;
  }
  gptr.assign(G + 1, 0);
  for (int32_t i = 0; i < n; ++i) ++gptr[index[i] + 1];
  for (int32_t g = 0; g < G; ++g) gptr[g + 1] += gptr[g];
  members.assign(n, 0);
  std::vector<int32_t> fill(gptr.begin(), gptr.end() - 1);
  for (int32_t i = 0; i < n; ++i) members[fill[index[i]]++] = i;
  return G;
}

struct TcaWs {
  int32_t *flags;
  uint32_t *sig, *sig_c;
  uint64_t *pairs;
  float *sims;
  int32_t *group, *ucount, *ucol, *ugroup;
  uint64_t *ukeys0, *ukeys1;
  int32_t *uvals0, *uvals1, *uG;
  void *utmp;
  void *lsh;
  size_t lsh_bytes;
  size_t bytes;
};

TcaWs carve_tca(void *base, int64_t M, int64_t nnz, int K, int64_t cap) {
  Carver c(base);
  TcaWs w{};
  w.flags = c.take<int32_t>(2);
  w.sig = c.take<uint32_t>(M * K);
  w.sig_c = c.take<uint32_t>(M * K);
  w.pairs = c.take<uint64_t>(cap);
  w.sims = c.take<float>(cap);
  w.group = c.take<int32_t>(M);
  w.ucount = c.take<int32_t>(M + 1);
  w.ucol = c.take<int32_t>(nnz);
  w.ugroup = c.take<int32_t>(nnz);
  w.ukeys0 = c.take<uint64_t>(nnz);
  w.ukeys1 = c.take<uint64_t>(nnz);
  w.uvals0 = c.take<int32_t>(nnz);
  w.uvals1 = c.take<int32_t>(nnz);
  w.uG = c.take<int32_t>(nnz);
  w.utmp = c.take<char>(std::max(radix_sort_workspace_bytes(nnz), scan_workspace_bytes(std::max<int64_t>(nnz, M + 1))));
  w.lsh_bytes = carve_lsh(nullptr, M, cap).bytes;
  w.lsh = c.take<char>(w.lsh_bytes);
  w.bytes = c.used();
  return w;
}

dtc_status_t check_options(const dtc_tca_options_t *o) {
  if (!o) return fail(DTC_ERROR_INVALID_VALUE, "options are required");
  if (o->thres < 1 || o->cluster_thres < 1)
    return fail(DTC_ERROR_INVALID_VALUE, "cluster thresholds must be >= 1");
  return check_lsh_args(0, o->bands * o->rows_per_band, o->bands, o->rows_per_band,
                        o->max_neighbors, o->max_pairs);
}

}  // namespace
}  // namespace dtc

using namespace dtc;

extern "C" {

dtc_status_t dtc_minhash_signatures(const int32_t *node_pointer, const int32_t *edge_list,
                                    int32_t num_rows, int32_t num_hashes, uint32_t seed,
                                    uint32_t *signatures, dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (num_rows < 0) return fail(DTC_ERROR_INVALID_VALUE, "num_rows must be >= 0");
  if (num_hashes < 1 || num_hashes > kMaxHashes)
    return fail(DTC_ERROR_INVALID_VALUE, "num_hashes must be in [1, 128]");
  if (num_rows == 0) return DTC_SUCCESS;
  if (!node_pointer || !edge_list || !signatures)
    return fail(DTC_ERROR_INVALID_VALUE, "node_pointer, edge_list and signatures are required");
  const int64_t cells = (int64_t)num_rows * num_hashes;
  k_fill_u32<<<blocks_for(cells, 256), 256, 0, stream>>>(signatures, cells, 0xFFFFFFFFu);
  DTC_LAUNCH_CHECK("fill");
  const unsigned grid = (unsigned)std::min<int64_t>(16384, ceil_div(num_rows, 4));
  k_minhash<<<grid, 256, 0, stream>>>(node_pointer, edge_list, num_rows, num_hashes, seed,
                                      signatures);
  DTC_LAUNCH_CHECK("minhash");
  return DTC_SUCCESS;
}

dtc_status_t dtc_lsh_pairs_workspace_size(int32_t num_sets, int64_t max_pairs, size_t *bytes) {
  clear_error();
  if (!bytes) return fail(DTC_ERROR_INVALID_VALUE, "bytes is NULL");
  if (num_sets < 0 || max_pairs < 0) return fail(DTC_ERROR_INVALID_VALUE, "sizes must be >= 0");
  *bytes = carve_lsh(nullptr, num_sets, max_pairs).bytes;
  return DTC_SUCCESS;
}

dtc_status_t dtc_lsh_pairs(const uint32_t *signatures, const int32_t *set_ptr, int32_t num_sets,
                           int32_t num_hashes, int32_t bands, int32_t rows_per_band,
                           int32_t max_neighbors, uint64_t *pairs, int64_t max_pairs,
                           int64_t *num_pairs, void *workspace, size_t workspace_bytes,
                           dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (!num_pairs) return fail(DTC_ERROR_INVALID_VALUE, "num_pairs is NULL");
  *num_pairs = 0;
  dtc_status_t st = check_lsh_args(num_sets, num_hashes, bands, rows_per_band, max_neighbors,
                                   max_pairs);
  if (st != DTC_SUCCESS) return st;
  if (num_sets < 2) return DTC_SUCCESS;
  if (!signatures || !set_ptr || (max_pairs > 0 && !pairs))
    return fail(DTC_ERROR_INVALID_VALUE, "signatures, set_ptr and pairs are required");
  const LshWs w = carve_lsh(workspace, num_sets, max_pairs);
  if (!workspace || workspace_bytes < w.bytes)
    return fail(DTC_ERROR_WORKSPACE, "LSH workspace needs " + std::to_string(w.bytes) + " bytes");
  return lsh_pairs(signatures, set_ptr, num_sets, num_hashes, bands, rows_per_band,
                   max_neighbors, pairs, max_pairs, num_pairs, w, stream);
}

dtc_status_t dtc_jaccard_pairs(const int32_t *set_ptr, const int32_t *set_idx,
                               const uint64_t *pairs, int64_t num_pairs, float *similarity,
                               dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (num_pairs < 0) return fail(DTC_ERROR_INVALID_VALUE, "num_pairs must be >= 0");
  if (num_pairs == 0) return DTC_SUCCESS;
  if (!set_ptr || !set_idx || !pairs || !similarity)
    return fail(DTC_ERROR_INVALID_VALUE, "set_ptr, set_idx, pairs and similarity are required");
  k_jaccard<<<(unsigned)ceil_div(num_pairs, 4), 256, 0, stream>>>(set_ptr, set_idx, pairs,
                                                                  num_pairs, similarity);
  DTC_LAUNCH_CHECK("jaccard");
  return DTC_SUCCESS;
}

dtc_status_t dtc_tca_greedy(int32_t num_items, const uint64_t *pairs, const float *similarity,
                            int64_t num_pairs, const int32_t *list_ptr, const int32_t *list_idx,
                            int32_t thres, int32_t *root) {
  clear_error();
  if (num_items < 0 || num_pairs < 0) return fail(DTC_ERROR_INVALID_VALUE, "sizes must be >= 0");
  if (thres < 1) return fail(DTC_ERROR_INVALID_VALUE, "thres must be >= 1");
  if (num_items == 0) return DTC_SUCCESS;
  if (!root || !list_ptr || (num_pairs > 0 && (!pairs || !similarity)))
    return fail(DTC_ERROR_INVALID_VALUE, "pairs, similarity, list_ptr and root are required");
  for (int64_t i = 0; i < num_pairs; ++i) {
    const uint64_t a = pairs[i] >> 32, b = pairs[i] & 0xFFFFFFFFu;
    if (a >= (uint64_t)num_items || b >= (uint64_t)num_items)
      return fail(DTC_ERROR_INVALID_VALUE, "pair " + std::to_string(i) + " names an item out of range");
  }
  greedy(num_items, greedy_input(pairs, similarity, num_pairs), list_ptr, list_idx, thres, root);
  return DTC_SUCCESS;
}

dtc_status_t dtc_tca_reorder_workspace_size(int32_t num_rows, int32_t nnz,
                                            const dtc_tca_options_t *options, size_t *bytes) {
  clear_error();
  if (!bytes) return fail(DTC_ERROR_INVALID_VALUE, "bytes is NULL");
  dtc_status_t st = check_options(options);
  if (st != DTC_SUCCESS) return st;
  if (num_rows < 0 || nnz < 0) return fail(DTC_ERROR_INVALID_VALUE, "sizes must be >= 0");
  *bytes = carve_tca(nullptr, num_rows, nnz, options->bands * options->rows_per_band,
                     options->max_pairs).bytes;
  return DTC_SUCCESS;
}

dtc_status_t dtc_tca_reorder(const int32_t *node_pointer, const int32_t *edge_list,
                             int32_t num_rows, int32_t nnz, const dtc_tca_options_t *options,
                             int32_t *reorder_id, dtc_tca_stats_t *stats, void *workspace,
                             size_t workspace_bytes, dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  dtc_status_t st = check_options(options);
  if (st != DTC_SUCCESS) return st;
  if (num_rows < 0 || nnz < 0) return fail(DTC_ERROR_INVALID_VALUE, "sizes must be >= 0");
  if (!reorder_id && num_rows > 0) return fail(DTC_ERROR_INVALID_VALUE, "reorder_id is NULL");
  if (stats) *stats = dtc_tca_stats_t{};
  if (num_rows == 0) return DTC_SUCCESS;
  if (!node_pointer || (nnz > 0 && !edge_list))
    return fail(DTC_ERROR_INVALID_VALUE, "node_pointer and edge_list are required");
  const dtc_tca_options_t o = *options;
  const int K = o.bands * o.rows_per_band;
  const TcaWs w = carve_tca(workspace, num_rows, nnz, K, o.max_pairs);
  if (!workspace || workspace_bytes < w.bytes)
    return fail(DTC_ERROR_WORKSPACE, "reorder workspace needs " + std::to_string(w.bytes) + " bytes");
  const LshWs lw = carve_lsh(w.lsh, num_rows, o.max_pairs);
  const int64_t M = num_rows;

  // 1. CSR check: sorted, duplicate-free rows.
  k_fill_i32<<<1, 64, 0, stream>>>(w.flags, 2, 0);
  DTC_LAUNCH_CHECK("fill");
  k_check_sorted_csr<<<(unsigned)std::min<int64_t>(4096, ceil_div(M, 4)), 256, 0, stream>>>(
      node_pointer, edge_list, num_rows, nnz, w.flags);
  DTC_LAUNCH_CHECK("check_sorted_csr");
  int32_t flags[2] = {0, 0};
  DTC_HIP_CHECK(hipMemcpyAsync(flags, w.flags, 8, hipMemcpyDeviceToHost, stream));
  DTC_HIP_CHECK(hipStreamSynchronize(stream));
  if (flags[0] & kUnsorted)
    return fail(DTC_ERROR_INVALID_CSR, "rows must hold strictly increasing column indices");
  if (flags[0]) return fail(DTC_ERROR_INVALID_CSR, "invalid CSR (flags " + std::to_string(flags[0]) + ")");

  // 2. Row signatures, LSH pairs and their Jaccard similarity.
  st = dtc_minhash_signatures(node_pointer, edge_list, num_rows, K, o.seed, w.sig, stream_);
  if (st != DTC_SUCCESS) return st;
  int64_t P = 0;
  st = lsh_pairs(w.sig, node_pointer, num_rows, K, o.bands, o.rows_per_band, o.max_neighbors,
                 w.pairs, o.max_pairs, &P, lw, stream);
  if (st != DTC_SUCCESS) return st;
  st = dtc_jaccard_pairs(node_pointer, edge_list, w.pairs, P, w.sims, stream_);
  if (st != DTC_SUCCESS) return st;
  std::vector<uint64_t> h_pairs((size_t)P);
  std::vector<float> h_sims((size_t)P);
  std::vector<int32_t> h_ptr((size_t)M + 1), h_col((size_t)nnz);
  if (P > 0) {
    DTC_HIP_CHECK(hipMemcpyAsync(h_pairs.data(), w.pairs, P * 8, hipMemcpyDeviceToHost, stream));
    DTC_HIP_CHECK(hipMemcpyAsync(h_sims.data(), w.sims, P * 4, hipMemcpyDeviceToHost, stream));
  }
  DTC_HIP_CHECK(hipMemcpyAsync(h_ptr.data(), node_pointer, (M + 1) * 4, hipMemcpyDeviceToHost, stream));
  if (nnz > 0)
    DTC_HIP_CHECK(hipMemcpyAsync(h_col.data(), edge_list, (size_t)nnz * 4, hipMemcpyDeviceToHost, stream));
  DTC_HIP_CHECK(hipStreamSynchronize(stream));

  // 3. Row clusters (TCU-aware level).
  std::vector<int32_t> roots((size_t)M), group, gptr, members;
  std::vector<int32_t> perm;
  st = stream_order(w.sims, P, lw, stream, perm);
  if (st != DTC_SUCCESS) return st;
  greedy(num_rows, greedy_input_sorted(h_pairs.data(), h_sims.data(), perm.data(), P), h_ptr.data(),
         h_col.data(), o.thres, roots.data());
  const int32_t C = group_in_order(roots, group, gptr, members);

  // 4. Cluster unions: signatures and column lists.
  DTC_HIP_CHECK(hipMemcpyAsync(w.group, group.data(), M * 4, hipMemcpyHostToDevice, stream));
  k_fill_u32<<<blocks_for((int64_t)C * K, 256), 256, 0, stream>>>(w.sig_c, (int64_t)C * K, 0xFFFFFFFFu);
  DTC_LAUNCH_CHECK("fill");
  k_group_min<<<blocks_for(M * K, 256), 256, 0, stream>>>(w.sig, num_rows, K, w.group, w.sig_c);
  DTC_LAUNCH_CHECK("group_min");
  int64_t U = 0;
  if (nnz > 0) {
    const int cb = bits_of(std::max(flags[1], 1));
    const int gb = bits_of(std::max(C - 1, 1));
    k_union_keys<<<(unsigned)ceil_div(M, 4), 256, 0, stream>>>(node_pointer, edge_list,
                                                               num_rows, w.group, cb, w.ukeys0,
                                                               w.uvals0);
    DTC_LAUNCH_CHECK("union_keys");
    bool in_1 = false;
    st = radix_sort_pairs(w.ukeys0, w.uvals0, w.ukeys1, w.uvals1, nnz, cb + gb, &in_1, w.utmp,
                          stream);
    if (st != DTC_SUCCESS) return st;
    const uint64_t *sorted = in_1 ? w.ukeys1 : w.ukeys0;
    st = scan_key_changes(sorted, w.uG, nnz, w.utmp, stream);
    if (st != DTC_SUCCESS) return st;
    k_union_compact<<<blocks_for(nnz, 256), 256, 0, stream>>>(sorted, w.uG, nnz, cb, w.ucol,
                                                              w.ugroup);
    DTC_LAUNCH_CHECK("union_compact");
    int32_t u32 = 0;
    DTC_HIP_CHECK(hipMemcpyAsync(&u32, w.uG + (nnz - 1), 4, hipMemcpyDeviceToHost, stream));
    DTC_HIP_CHECK(hipStreamSynchronize(stream));
    U = u32;
  }
  k_union_ptr<<<blocks_for((int64_t)C + 1, 256), 256, 0, stream>>>(w.ugroup, (int32_t)U, C,
                                                                   w.ucount);
  DTC_LAUNCH_CHECK("union_ptr");

  // 5. Cluster pairs and the cache-aware level.
  int64_t Pc = 0;
  st = lsh_pairs(w.sig_c, w.ucount, C, K, o.bands, o.rows_per_band, o.max_neighbors, w.pairs,
                 o.max_pairs, &Pc, lw, stream);
  if (st != DTC_SUCCESS) return st;
  st = dtc_jaccard_pairs(w.ucount, w.ucol, w.pairs, Pc, w.sims, stream_);
  if (st != DTC_SUCCESS) return st;
  std::vector<uint64_t> c_pairs((size_t)Pc);
  std::vector<float> c_sims((size_t)Pc);
  std::vector<int32_t> c_ptr((size_t)C + 1), c_col((size_t)U);
  if (Pc > 0) {
    DTC_HIP_CHECK(hipMemcpyAsync(c_pairs.data(), w.pairs, Pc * 8, hipMemcpyDeviceToHost, stream));
    DTC_HIP_CHECK(hipMemcpyAsync(c_sims.data(), w.sims, Pc * 4, hipMemcpyDeviceToHost, stream));
  }
  DTC_HIP_CHECK(hipMemcpyAsync(c_ptr.data(), w.ucount, ((size_t)C + 1) * 4, hipMemcpyDeviceToHost, stream));
  if (U > 0)
    DTC_HIP_CHECK(hipMemcpyAsync(c_col.data(), w.ucol, (size_t)U * 4, hipMemcpyDeviceToHost, stream));
  DTC_HIP_CHECK(hipStreamSynchronize(stream));
  std::vector<int32_t> c_roots((size_t)C), sgroup, sptr, smembers;
  st = stream_order(w.sims, Pc, lw, stream, perm);
  if (st != DTC_SUCCESS) return st;
  greedy(C, greedy_input_sorted(c_pairs.data(), c_sims.data(), perm.data(), Pc), c_ptr.data(),
         c_col.data(), o.cluster_thres, c_roots.data());
  const int32_t S = group_in_order(c_roots, sgroup, sptr, smembers);

  // 6. Rows cluster by cluster, clusters super-cluster by super-cluster
  // (TCA_reorder.py:300-305).
  int64_t out = 0;
  for (int32_t s = 0; s < S; ++s)
    for (int32_t q = sptr[s]; q < sptr[s + 1]; ++q) {
      const int32_t k = smembers[q];
      for (int32_t t = gptr[k]; t < gptr[k + 1]; ++t) reorder_id[out++] = members[t];
    }
  if (stats) {
    stats->row_pairs = P;
    stats->cluster_pairs = Pc;
    stats->clusters = C;
    stats->super_clusters = S;
  }
  return DTC_SUCCESS;
}

}  // extern "C"
