// Device-wide building blocks for the ME-TCF build on gfx950 (wave64):
// reduce-then-scan prefix sums and a stable LSD radix sort of
// (uint64 key, int32 value) pairs whose per-tile ranking uses wave ballots.
#include <string>

#include "seg_sort.h"

namespace dtc {

// ------------------------------------------------------------------ status

static thread_local std::string g_last_error;

dtc_status_t fail(dtc_status_t st, const std::string &msg) {
  g_last_error = msg;
  return st;
}
void clear_error() { g_last_error.clear(); }

// -------------------------------------------------------------------- scan

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

struct LoadI32 {
  const int32_t *p;
  __device__ __forceinline__ int operator()(int64_t i) const { return p[i]; }
};

template <class Key>
struct LoadKeyChange {
  const Key *k;
  __device__ __forceinline__ int operator()(int64_t i) const {
    return (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
  }
};

template <class Load>
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(Load load, int64_t n,
                                                                    int32_t *partials) {
  __shared__ int smem[4];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k * kScanThreads + threadIdx.x;
    if (i < n) sum += load(i);
  }
  sum = wave_sum(sum);
  if ((threadIdx.x & 63) == 0) smem[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = smem[0] + smem[1] + smem[2] + smem[3];
}

// Exclusive scan of partials[0..m) in place, one workgroup.
__global__ __launch_bounds__(kScanThreads) void scan_partials_kernel(int32_t *partials,
                                                                      int64_t m) {
  __shared__ int smem[4];
  int carry = 0;
  for (int64_t base = 0; base < m; base += kScanTile) {
    int v[kScanItems];
    int local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const int64_t i = base + (int64_t)threadIdx.x * kScanItems + k;
      v[k] = i < m ? partials[i] : 0;
      local += v[k];
    }
    int total;
    const int inc = block256_inclusive_sum(local, smem, total);
    int run = carry + inc - local;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const int64_t i = base + (int64_t)threadIdx.x * kScanItems + k;
      if (i < m) partials[i] = run;
      run += v[k];
    }
    carry += total;
  }
}

template <class Load>
__global__ __launch_bounds__(kScanThreads) void scan_downsweep_kernel(
    Load load, int32_t *out, int64_t n, const int32_t *partials, int inclusive) {
  __shared__ int smem[4];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int v[kScanItems];
  int local = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    v[k] = i < n ? load(i) : 0;
    local += v[k];
  }
  int total;
  const int inc = block256_inclusive_sum(local, smem, total);
  int run = partials[blockIdx.x] + inc - local;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    if (inclusive) run += v[k];
    if (i < n) out[i] = run;
    if (!inclusive) run += v[k];
  }
}

size_t scan_workspace_bytes(int64_t n) {
  return align_up(sizeof(int32_t) * (size_t)std::max<int64_t>(1, ceil_div(n, kScanTile)), 256);
}

template <class Load>
static dtc_status_t scan_impl(Load load, int32_t *out, int64_t n, bool inclusive, void *ws,
                              hipStream_t stream) {
  if (n <= 0) return DTC_SUCCESS;
  const int64_t tiles = ceil_div(n, kScanTile);
  if (tiles > INT32_MAX) return fail(DTC_ERROR_UNSUPPORTED, "scan too large");
  int32_t *partials = static_cast<int32_t *>(ws);
  scan_reduce_kernel<Load><<<(unsigned)tiles, kScanThreads, 0, stream>>>(load, n, partials);
  DTC_LAUNCH_CHECK("scan_reduce");
  scan_partials_kernel<<<1, kScanThreads, 0, stream>>>(partials, tiles);
  DTC_LAUNCH_CHECK("scan_partials");
  scan_downsweep_kernel<Load>
      <<<(unsigned)tiles, kScanThreads, 0, stream>>>(load, out, n, partials, inclusive ? 1 : 0);
  DTC_LAUNCH_CHECK("scan_downsweep");
  return DTC_SUCCESS;
}

dtc_status_t scan_i32(const int32_t *in, int32_t *out, int64_t n, bool inclusive, void *ws,
                      hipStream_t stream) {
  return scan_impl(LoadI32{in}, out, n, inclusive, ws, stream);
}

dtc_status_t scan_key_changes(const uint64_t *keys, int32_t *out, int64_t n, void *ws,
                              hipStream_t stream) {
  return scan_impl(LoadKeyChange<uint64_t>{keys}, out, n, true, ws, stream);
}

dtc_status_t scan_key_changes_u32(const uint32_t *keys, int32_t *out, int64_t n, void *ws,
                                  hipStream_t stream) {
  return scan_impl(LoadKeyChange<uint32_t>{keys}, out, n, true, ws, stream);
}

// -------------------------------------------------------------- radix sort

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kRadix = 256;

// Tile maps: tile t covers [start, start + len) of the key array, and the
// count of digit d in it lives at hist[hbase + d * hstride].
//   PlainTiles: tiles of kSortTile keys over the whole array, hist laid out
//     [digit][tile], so one exclusive scan of hist gives every (digit, tile)
//     its first output position;
//   SegTiles: the array is a sequence of segments (segment w' = [csb[w'],
//     csb[w'+1])), each cut into its own tiles (tile t belongs to segment
//     tw[t], segment w' owns tiles [tb[w'], tb[w'+1])), hist laid out
//     [segment][digit][tile of the segment]: the same scan then gives
//     csb[w'] + (entries of w' with a smaller digit) + (entries of digit d in
//     earlier tiles of w'), i.e. each pass sorts every segment on its own.
struct PlainTiles {
  int64_t n;
  int ntiles;
  __device__ __forceinline__ bool valid(int) const { return true; }
  __device__ __forceinline__ void get(int t, int64_t &start, int &len, int64_t &hbase,
                                      int &hstride) const {
    start = (int64_t)t * kSortTile;
    len = (int)min<int64_t>(kSortTile, n - start);
    hbase = t;
    hstride = ntiles;
  }
};

struct SegTiles {
  const int32_t *tw, *tb, *csb;
  const int32_t *ntot;  // number of tiles (the grid may be larger)
  __device__ __forceinline__ bool valid(int t) const { return t < *ntot; }
  __device__ __forceinline__ void get(int t, int64_t &start, int &len, int64_t &hbase,
                                      int &hstride) const {
    const int w = tw[t];
    const int t0 = tb[w];
    const int tl = t - t0;
    start = (int64_t)csb[w] + (int64_t)tl * kSortTile;
    len = (int)min<int64_t>(kSortTile, (int64_t)csb[w + 1] - start);
    hbase = (int64_t)t0 * kRadix + tl;
    hstride = tb[w + 1] - t0;
  }
};

// Digit counts of one tile.
template <class Key, class Tiles>
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const Key *keys, int shift,
                                                                   int32_t *hist, Tiles tiles) {
  __shared__ int cnt[kRadix];
  if (!tiles.valid(blockIdx.x)) {  // past the last tile: the scanned tail reads zeros
    hist[(int64_t)blockIdx.x * kRadix + threadIdx.x] = 0;
    return;
  }
  cnt[threadIdx.x] = 0;
  __syncthreads();
  int64_t start, hbase;
  int len, hstride;
  tiles.get(blockIdx.x, start, len, hbase, hstride);
#pragma unroll 4
  for (int k = 0; k < kSortItems; ++k) {
    const int i = k * kSortThreads + threadIdx.x;
    if (i < len) atomicAdd(&cnt[(int)((keys[start + i] >> shift) & (kRadix - 1))], 1);
  }
  __syncthreads();
  hist[hbase + (int64_t)threadIdx.x * hstride] = cnt[threadIdx.x];
}

// Stable scatter: element i of the tile goes to offsets[d][tile] + (number of
// earlier elements of the tile with the same digit d).  Tiles are 4,096
// keys.  Wave w owns the tile's elements [1024 w, 1024 w + 1024), taken in 16
// rounds of 64 in index order, so index order is wave-major, then round,
// then lane.  Each wave
// ranks its own elements without workgroup barriers: inside a round the rank
// among equal digits comes from eight ballots (one per digit bit) and a
// popcount, added to the wave's running count of the digit (per-wave LDS
// counters).  Then, once per tile, the four waves' counts give each digit's
// per-wave prefixes and the tile's digit-sorted layout (a scan over the 256
// digits), and the tile is laid out in LDS in digit order and written out so
// that consecutive threads store consecutive elements of one digit's run
// (~16 keys on average: runs, not 4,096 scattered 12-byte writes; with
// 2,048-key tiles the runs of ~8 keys made the scatter 1.5x slower than a
// pass whose few digits give long runs).
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kWaveItems = kSortTile / kSortWaves;  // 1,024 elements per wave

template <class Key, class Tiles>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const Key *kin, const int32_t *vin, Key *kout, int32_t *vout, int shift,
    const int32_t *offsets, Tiles tiles) {
  __shared__ int s_off[kRadix];             // global position of the tile's first element of digit d
  __shared__ int s_loc[kRadix];             // local position of the same element
  __shared__ int s_wcnt[kSortWaves][kRadix];  // per-wave digit counts, then per-wave prefixes
  __shared__ int s_scan[kSortWaves];
  __shared__ Key s_keys[kSortTile];
  __shared__ int32_t s_vals[kSortTile];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  if (!tiles.valid(blockIdx.x)) return;  // workgroup-uniform
  int64_t start, hbase;
  int len, hstride;
  tiles.get(blockIdx.x, start, len, hbase, hstride);
  const int wbase = wave * kWaveItems;  // tile-relative
  // All of the wave's keys and values are loaded up front (one memory
  // latency for the whole tile).
  Key kreg[kSortItems];
  int32_t vreg[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int i = wbase + r * 64 + lane;
    kreg[r] = i < len ? kin[start + i] : (Key)0;
    vreg[r] = i < len ? vin[start + i] : 0;
  }
  s_off[tid] = offsets[hbase + (int64_t)tid * hstride];
#pragma unroll
  for (int k = 0; k < kRadix / 64; ++k) s_wcnt[wave][k * 64 + lane] = 0;
  wave_lds_sync();
  int lrank[kSortItems];  // rank among the wave's earlier elements of the same digit
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const bool valid = wbase + r * 64 + lane < len;
    const int d = (int)((kreg[r] >> shift) & (kRadix - 1));
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool b = (d >> bit) & 1;
      const uint64_t m = __ballot(b);
      peers &= b ? m : ~m;
    }
    const int rank = __popcll(peers & lt_mask);
    lrank[r] = valid ? s_wcnt[wave][d] + rank : 0;
    wave_lds_sync();
    if (valid && rank == 0) s_wcnt[wave][d] += __popcll(peers);
    wave_lds_sync();
  }
  __syncthreads();
  // per digit d = tid: the waves' exclusive prefixes, and the tile's count of d
  int cnt = 0;
#pragma unroll
  for (int w = 0; w < kSortWaves; ++w) {
    const int c = s_wcnt[w][tid];
    s_wcnt[w][tid] = cnt;
    cnt += c;
  }
  // exclusive scan of the 256 digit counts: the digit-sorted local layout
  const int inc = wave_inclusive_sum(cnt);
  if (lane == 63) s_scan[wave] = inc;
  __syncthreads();
  int pre = 0;
#pragma unroll
  for (int w = 0; w < kSortWaves; ++w) pre += w < wave ? s_scan[w] : 0;
  s_loc[tid] = pre + inc - cnt;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    if (wbase + r * 64 + lane < len) {
      const int d = (int)((kreg[r] >> shift) & (kRadix - 1));
      const int lp = s_loc[d] + s_wcnt[wave][d] + lrank[r];
      s_keys[lp] = kreg[r];
      s_vals[lp] = vreg[r];
    }
  }
  __syncthreads();
  for (int j = tid; j < len; j += kSortThreads) {
    const Key key = s_keys[j];
    const int d = (int)((key >> shift) & (kRadix - 1));
    const int64_t pos = (int64_t)s_off[d] + (j - s_loc[d]);
    kout[pos] = key;
    vout[pos] = s_vals[j];
  }
}

size_t radix_sort_workspace_bytes(int64_t n) {
  const int64_t ntiles = std::max<int64_t>(1, ceil_div(n, kSortTile));
  return align_up(sizeof(int32_t) * (size_t)(ntiles * kRadix), 256) +
         scan_workspace_bytes(ntiles * kRadix);
}

dtc_status_t radix_sort_pairs(uint64_t *keys0, int32_t *vals0, uint64_t *keys1,
                              int32_t *vals1, int64_t n, int key_bits, bool *result_in_1,
                              void *ws, hipStream_t stream) {
  *result_in_1 = false;
  if (n <= 1 || key_bits <= 0) return DTC_SUCCESS;
  const int64_t ntiles = ceil_div(n, kSortTile);
  if (ntiles * kRadix > INT32_MAX) return fail(DTC_ERROR_UNSUPPORTED, "sort too large");
  Carver c(ws);
  int32_t *hist = c.take<int32_t>((size_t)(ntiles * kRadix));
  void *scan_ws = c.take<char>(scan_workspace_bytes(ntiles * kRadix));
  uint64_t *kin = keys0, *kout = keys1;
  int32_t *vin = vals0, *vout = vals1;
  bool in_1 = false;
  const PlainTiles tiles{n, (int)ntiles};
  for (int shift = 0; shift < key_bits; shift += 8) {
    radix_hist_kernel<uint64_t, PlainTiles>
        <<<(unsigned)ntiles, kSortThreads, 0, stream>>>(kin, shift, hist, tiles);
    DTC_LAUNCH_CHECK("radix_hist");
    dtc_status_t st = scan_i32(hist, hist, ntiles * kRadix, false, scan_ws, stream);
    if (st != DTC_SUCCESS) return st;
    radix_scatter_kernel<uint64_t, PlainTiles>
        <<<(unsigned)ntiles, kSortThreads, 0, stream>>>(kin, vin, kout, vout, shift, hist, tiles);
    DTC_LAUNCH_CHECK("radix_scatter");
    std::swap(kin, kout);
    std::swap(vin, vout);
    in_1 = !in_1;
  }
  *result_in_1 = in_1;
  return DTC_SUCCESS;
}

size_t seg_radix_sort_workspace_bytes(int64_t n, int64_t segments) {
  const int64_t ntiles = std::max<int64_t>(1, ceil_div(n, kSortTile) + segments);
  return align_up(sizeof(int32_t) * (size_t)(ntiles * kRadix), 256) +
         scan_workspace_bytes(ntiles * kRadix);
}

int64_t seg_radix_tile_size() { return kSortTile; }

dtc_status_t seg_radix_sort_pairs_u32(uint32_t *keys0, int32_t *vals0, uint32_t *keys1,
                                      int32_t *vals1, int key_bits, const int32_t *tw,
                                      const int32_t *tb, const int32_t *csb,
                                      const int32_t *ntiles_dev, int64_t ntiles,
                                      bool *result_in_1, void *ws, hipStream_t stream) {
  *result_in_1 = false;
  if (ntiles <= 0 || key_bits <= 0) return DTC_SUCCESS;
  if (ntiles * kRadix > INT32_MAX) return fail(DTC_ERROR_UNSUPPORTED, "sort too large");
  Carver c(ws);
  int32_t *hist = c.take<int32_t>((size_t)(ntiles * kRadix));
  void *scan_ws = c.take<char>(scan_workspace_bytes(ntiles * kRadix));
  uint32_t *kin = keys0, *kout = keys1;
  int32_t *vin = vals0, *vout = vals1;
  bool in_1 = false;
  const SegTiles tiles{tw, tb, csb, ntiles_dev};
  for (int shift = 0; shift < key_bits; shift += 8) {
    radix_hist_kernel<uint32_t, SegTiles>
        <<<(unsigned)ntiles, kSortThreads, 0, stream>>>(kin, shift, hist, tiles);
    DTC_LAUNCH_CHECK("radix_hist");
    dtc_status_t st = scan_i32(hist, hist, ntiles * kRadix, false, scan_ws, stream);
    if (st != DTC_SUCCESS) return st;
    radix_scatter_kernel<uint32_t, SegTiles>
        <<<(unsigned)ntiles, kSortThreads, 0, stream>>>(kin, vin, kout, vout, shift, hist, tiles);
    DTC_LAUNCH_CHECK("radix_scatter");
    std::swap(kin, kout);
    std::swap(vin, vout);
    in_1 = !in_1;
  }
  *result_in_1 = in_1;
  return DTC_SUCCESS;
}

}  // namespace dtc

extern "C" {

const char *dtc_status_string(dtc_status_t status) {
  switch (status) {
    case DTC_SUCCESS: return "success";
    case DTC_ERROR_INVALID_VALUE: return "invalid value";
    case DTC_ERROR_UNSUPPORTED: return "unsupported";
    case DTC_ERROR_WORKSPACE: return "workspace too small";
    case DTC_ERROR_INVALID_CSR: return "invalid CSR";
    case DTC_ERROR_HIP: return "HIP error";
    case DTC_ERROR_INCONSISTENT: return "inconsistent blockPartition";
  }
  return "unknown status";
}

const char *dtc_last_error(void) { return dtc::g_last_error.c_str(); }

int32_t dtc_version(void) { return 10000; }

}  // extern "C"
