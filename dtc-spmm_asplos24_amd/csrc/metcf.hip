// ME-TCF construction on gfx950.
//
// The reference builds the format with a thrust sort of (window, column)
// pairs followed by two kernels that run ONE THREAD PER ROW WINDOW
// (generate_edgetocolumn, DTCSpMM_kernel.cu:205-249, and
// generate_tcoffset_id_atob, :252-335).  The CSR is already grouped by
// window, so the sort only has to order each window's own entries by column:
// the sorted positions of window w are exactly its CSR range
// [rowptr[w*blk_h], rowptr[min(w*blk_h+blk_h, M)]).  Windows are classified
// by entry count E:
//
//   LDS path (16-row windows of 8-column blocks, E <= 16,384): one workgroup
//   per window (five shapes, 1 to 16 waves, up to 512 ... 16,384 entries;
//   windows of at most 64 entries take one wave each and bitonic sorts)
//   sorts the window's (column, local index) pairs in LDS (LSD radix sort
//   over the window's column span, ranks from LDS atomics in the first pass
//   and from wave ballots in the stable later ones), numbers the distinct
//   columns by a workgroup scan and writes every output sized by the CSR:
//   edgeToRow, edgeToColumn, blockPartition, and the tile ids, values and
//   edge indices in TC-block order, each entry at its block's start plus
//   the number of the block's entries with a smaller CSR index (the
//   reference's CSR-order placement inside a block, :291-300).
//
//   Global path (larger windows, and every window of other block shapes),
//   on a side stream beside the LDS path: the windows are compacted and
//   sorted by column, each within its own range (segmented radix sort of
//   32-bit keys) when they average at least one sort tile, else together by
//   (compacted window, column); the inclusive scan G of "key differs from
//   its predecessor" gives every sorted element its unique-column number, so
//   the rank of a column inside its window is G - G(window start), and
//   edge-parallel kernels write the same outputs.
//
// Both paths leave per-window block data in the workspace (distinct columns,
// block starts, |S_w|); after the caller has allocated the block-sized
// arrays, phase 2 writes TCblockoffset, TCblockRowid and SparseAToX from it.
// Both paths write every output of a window and nothing of another window,
// so the classification changes no result.
#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "seg_sort.h"

namespace dtc {
namespace {

// Window classes of the LDS path: (waves, entries per lane), capacity
// 64 * waves * entries.
constexpr int kWinH = 16;  // rows per window on the LDS path (the reference's BLK_H)
constexpr int kWinW = 8;   // columns per TC block on the LDS path (BLK_W)
// Class 0 (at most 64 entries) is one wave per window with bitonic sorts
// (k_window_tiny); classes 1-5 are k_window_sort.
constexpr int kClasses = 6;
// (Measured against 2 x 16 and 4 x 16 for the middle classes, 2 x 4 for
// class 1 and 16 x 8 for class 4 on power-law and block-model matrices.)
constexpr int kClassWaves[kClasses] = {1, 1, 4, 8, 8, 16};
constexpr int kClassItems[kClasses] = {1, 8, 8, 8, 16, 16};
constexpr int kClassCap[kClasses] = {64, 64 * 1 * 8, 64 * 4 * 8, 64 * 8 * 8, 64 * 8 * 16,
                                     64 * 16 * 16};
// Phase 2 takes the TC blocks in chunks of kOutBlocks per workgroup.
constexpr int kOutBlocks = 512;

// Routing of windows (environment DTC_METCF_PATH): by size, every window
// through the global path, or every window that fits class c through c.
constexpr int kPathAuto = 0;
constexpr int kPathGlobal = 1;
constexpr int kPathClass0 = 2;

constexpr int kMaxSlots = 16;  // the column maximum is spread over this many words
struct MetcfHeader {
  int32_t flags;    // CSR check bits, see kBad*
  int32_t nnz_big;  // entries of the global-path windows
  int32_t cls_count[kClasses + 1];  // windows per LDS class; [kClasses]: global-path windows
  int32_t max_col[kMaxSlots];       // largest column index: the maximum of these
  int32_t pad[64 - 2 - (kClasses + 1) - kMaxSlots];
};
constexpr int kHeaderRead = 128;  // bytes of the header read by the host
static_assert(sizeof(MetcfHeader) == 256, "header size");
static_assert(offsetof(MetcfHeader, pad) <= kHeaderRead, "header read size");

constexpr int kBadFirst = 1;     // rowptr[0] != 0
constexpr int kBadLast = 2;      // rowptr[M] != nnz
constexpr int kBadOrder = 4;     // rowptr decreases
constexpr int kBadColumn = 8;    // negative column index
constexpr int kBadClass = 16;    // a window larger than its class (internal)
constexpr int kBadPartition = 32;  // blockPartition of an empty window is not zero

struct MetcfWs {
  MetcfHeader *hdr;
  uint64_t *keys0, *keys1;  // global path: (compacted window, column) keys
  int32_t *vals0, *vals1;   // global path: CSR edge index
  int32_t *gscan;           // global path: inclusive unique-column count
  int32_t *boff;            // staged block data for phase 2 (see WinStage)
  int32_t *dcol;
  int32_t *wU;
  int32_t *lists;           // [kClasses + 1][W] windows of each class; the last row
                            // (global path) maps compacted window -> window
  int32_t *bstart;          // global path: first compacted entry of each of its blocks
  int32_t *cbb;             // global path: first block of each compacted window [n_big + 1]
  int32_t *csb;             // compacted window -> first compacted entry [n_big + 1]
  int32_t *tb;              // global path: first sort tile of each compacted window [n_big + 1]
  int32_t *tw;              // global path: compacted window of each sort tile
  // 32-bit views of keys0 / keys1 for the segmented global-path sort
  uint32_t *k32_0, *k32_1;
  int32_t *wpos;            // global path: compacted window of each compacted position
  int32_t *cw;              // [2][nnz / kOutBlocks + 2] window of the first and of the
                            // last block of each phase-2 chunk
  int64_t cw_stride;
  void *tmp;                // scan / sort scratch
  size_t bytes;
};

MetcfWs carve(void *ws, int64_t nnz, int64_t W) {
  Carver c(ws);
  MetcfWs w;
  const size_t n = (size_t)std::max<int64_t>(nnz, 1);
  const size_t nw = (size_t)std::max<int64_t>(W, 1);
  w.hdr = c.take<MetcfHeader>(1);
  w.keys0 = c.take<uint64_t>(n);
  w.keys1 = c.take<uint64_t>(n);
  w.vals0 = c.take<int32_t>(n);
  w.vals1 = c.take<int32_t>(n);
  w.gscan = c.take<int32_t>(n);
  w.boff = c.take<int32_t>(n);
  w.dcol = c.take<int32_t>(n);
  w.wU = c.take<int32_t>(nw);
  w.lists = c.take<int32_t>((kClasses + 1) * nw);
  w.bstart = c.take<int32_t>(n);
  w.cbb = c.take<int32_t>(nw + 1);
  w.csb = c.take<int32_t>(nw + 1);
  w.tb = c.take<int32_t>(nw + 1);
  w.tw = c.take<int32_t>((size_t)(ceil_div(std::max<int64_t>(nnz, 1), seg_radix_tile_size()) + nw));
  w.k32_0 = reinterpret_cast<uint32_t *>(w.keys0);
  w.k32_1 = reinterpret_cast<uint32_t *>(w.keys1);
  w.wpos = c.take<int32_t>(n);
  w.cw_stride = ceil_div((int64_t)n, kOutBlocks) + 1;
  w.cw = c.take<int32_t>((size_t)(2 * w.cw_stride));
  const size_t tmp = std::max({seg_radix_sort_workspace_bytes(nnz, W), scan_workspace_bytes(nnz),
                               scan_workspace_bytes(W)});
  w.tmp = c.take<char>(tmp);
  w.bytes = c.used();
  return w;
}

int bits_for(int64_t max_value) {
  int b = 0;
  while (b < 63 && (int64_t(1) << b) <= max_value) ++b;
  return b;
}

int path_mode(int32_t blk_h, int32_t blk_w) {
  if (blk_h != kWinH || blk_w != kWinW) return kPathGlobal;
  const char *e = std::getenv("DTC_METCF_PATH");
  if (!e || !*e || !std::strcmp(e, "auto")) return kPathAuto;
  if (!std::strcmp(e, "global")) return kPathGlobal;
  for (int c = 0; c < kClasses; ++c) {
    const std::string name = "lds" + std::to_string(c);
    if (name == e) return kPathClass0 + c;
  }
  return kPathAuto;
}

// DTC_METCF_STREAMS=1 runs the whole build on the caller's stream (no side
// stream for the global path): the same kernels in the same order, one after
// another, as a caller needing single-stream work may want, and to time the
// kernels without overlap.
bool one_stream() {
  const char *e = std::getenv("DTC_METCF_STREAMS");
  return e && !std::strcmp(e, "1");
}

__global__ void k_header_init(MetcfHeader *h) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    h->flags = 0;
    h->nnz_big = 0;
    for (int c = 0; c <= kClasses; ++c) h->cls_count[c] = 0;
    for (int c = 0; c < kMaxSlots; ++c) h->max_col[c] = 0;
  }
}


// Checks rowptr[0] == 0, rowptr[M] == nnz and rowptr non-decreasing, which
// keeps every window's entries inside the edge arrays.  (Columns are checked
// where they are read: negative ones by the window kernels and by
// k_big_maxcol, which also finds the largest column the global path sorts.)
__device__ __forceinline__ void validate_part(const int32_t *rowptr, int32_t M, int32_t nnz,
                                              MetcfHeader *h) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int flags = 0;
  for (int64_t r = tid; r < M; r += stride)
    if (rowptr[r] > rowptr[r + 1]) flags |= kBadOrder;
  if (tid == 0) {
    if (rowptr[0] != 0) flags |= kBadFirst;
    if (rowptr[M] != nnz) flags |= kBadLast;
  }
  if (__ballot(flags != 0) == 0) return;  // wave-uniform: nothing to report
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) flags |= __shfl_xor(flags, d, 64);
  if ((threadIdx.x & 63) == 0) atomicOr(&h->flags, flags);
}

__device__ __forceinline__ int window_begin(const int32_t *rowptr, int w, int blk_h, int M) {
  return rowptr[min((int64_t)w * blk_h, (int64_t)M)];
}
__device__ __forceinline__ int window_end(const int32_t *rowptr, int w, int blk_h, int M) {
  return rowptr[min((int64_t)w * blk_h + blk_h, (int64_t)M)];
}

// Window classification.  Each non-empty window is appended to the list of
// its class (one global atomic per workgroup and class; the order of a list
// does not affect any output): an LDS class, or the global path (list
// kClasses, whose entries are summed into nnz_big).  Empty windows must have
// blockPartition 0 (the caller zero-fills it; empty windows are not written).
// Also writes row_window_offset[0] = 0.
__device__ __forceinline__ void classify_part(const int32_t *rowptr, int32_t M, int32_t W,
                                              int32_t blk_h, int mode, const int32_t *bp,
                                              int32_t *lists, int32_t *rwo, MetcfHeader *h) {
  constexpr int NC = kClasses + 1;
  __shared__ int s_cnt[NC], s_base[NC], s_big;
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  if (threadIdx.x < NC) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) s_big = 0;
  __syncthreads();
  int cls = -1, E = 0;
  bool bad = false;
  if (w < W) {
    E = window_end(rowptr, w, blk_h, M) - window_begin(rowptr, w, blk_h, M);
    if (E > 0) {
      cls = kClasses;
      if (mode == kPathAuto) {
        for (int c = kClasses - 1; c >= 0; --c)
          if (E <= kClassCap[c]) cls = c;
      } else if (mode >= kPathClass0) {
        const int c = mode - kPathClass0;
        if (E <= kClassCap[c]) cls = c;
      }
    } else {
      bad = bp[w] != 0;
    }
  }
  if (w == 0) rwo[0] = 0;
  if (__ballot(bad) && lane == 0) atomicOr(&h->flags, kBadPartition);
  const int big_e = wave_sum(cls == kClasses ? E : 0);
  if (lane == 0 && big_e) atomicAdd(&s_big, big_e);
  uint64_t mask[NC];
  int woff[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    mask[c] = __ballot(cls == c);
    woff[c] = 0;
    if (mask[c]) {
      const int leader = __ffsll((unsigned long long)mask[c]) - 1;
      if (lane == leader) woff[c] = atomicAdd(&s_cnt[c], __popcll(mask[c]));
      woff[c] = __shfl(woff[c], leader, 64);
    }
  }
  __syncthreads();
  if (threadIdx.x < NC)
    s_base[threadIdx.x] = s_cnt[threadIdx.x] ? atomicAdd(&h->cls_count[threadIdx.x], s_cnt[threadIdx.x]) : 0;
  if (threadIdx.x == 0 && s_big) atomicAdd(&h->nnz_big, s_big);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (cls == c) {
      const int below = (int)__builtin_amdgcn_mbcnt_hi(
          (uint32_t)(mask[c] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask[c], 0u));
      lists[(int64_t)c * W + s_base[c] + woff[c] + below] = w;
    }
  }
}

// The row-pointer check and the window classes in one launch: every
// workgroup takes its share of the check (a grid-stride walk over the rows)
// and classes the 256 windows of its own index (none past W).
__global__ __launch_bounds__(256) void k_check_classify(const int32_t *rowptr, int32_t M,
                                                        int32_t nnz, int32_t W, int32_t blk_h,
                                                        int mode, const int32_t *bp,
                                                        int32_t *lists, int32_t *rwo,
                                                        MetcfHeader *h) {
  validate_part(rowptr, M, nnz, h);
  classify_part(rowptr, M, W, blk_h, mode, bp, lists, rwo, h);
}

// Exclusive prefix sum by one 1,024-thread workgroup, 16 items per thread
// per step: store(i, sum of load(k) for k < i) for i in [0, n), in order of i
// for each thread's items; returns the total.  Every thread must call it.
template <class Load, class Store>
__device__ __forceinline__ int wg1024_exclusive_scan(Load load, Store store, int n, int *s_tmp) {
  constexpr int NT = 1024, IT = 16;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int carry = 0;
  for (int base = 0; base < n; base += NT * IT) {
    int v[IT];
    int sum = 0;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int i = base + tid * IT + it;
      v[it] = i < n ? load(i) : 0;
      sum += v[it];
    }
    const int inc = wave_inclusive_sum(sum);
    if (lane == 63) s_tmp[wave] = inc;
    __syncthreads();
    int pre = 0, tot = 0;
#pragma unroll
    for (int j = 0; j < NT / 64; ++j) {
      const int t = s_tmp[j];
      pre += j < wave ? t : 0;
      tot += t;
    }
    __syncthreads();
    int run = carry + pre + inc - sum;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int i = base + tid * IT + it;
      if (i < n) store(i, run);
      run += v[it];
    }
    carry += tot;
  }
  return carry;
}

// First compacted entry of each global-path window: csb[k] = entries of the
// windows listed before k; csb[n_big] = all of them.
// Also cuts each window into sort tiles of `tile` entries: tb[k] = tiles of
// the windows listed before k, tb[n_big] = all of them, tw[t] = window of
// tile t.
//
// Runs before the host knows the window count: reads it from the header and
// leaves the work to the multi-workgroup kernels when it exceeds
// kOneGroupWindows.
constexpr int kOneGroupWindows = 4096;  // up to this many: one-workgroup kernels
__global__ __launch_bounds__(1024) void k_big_offsets(const int32_t *rowptr, int32_t M,
                                                      int32_t blk_h, const int32_t *wmap,
                                                      const MetcfHeader *h, int32_t tile,
                                                      int32_t *csb, int32_t *tb, int32_t *tw) {
  constexpr int IT = kOneGroupWindows / 1024;  // windows per thread
  __shared__ int s_e[16], s_t[16];
  const int n_big = h->cls_count[kClasses];
  if (n_big == 0 || n_big > kOneGroupWindows) return;  // workgroup-uniform
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // entries and sort tiles of this thread's windows, loaded once
  int e[IT], t[IT], se = 0, st = 0;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int i = tid * IT + it;
    const int w = i < n_big ? wmap[i] : 0;
    e[it] = i < n_big ? window_end(rowptr, w, blk_h, M) - window_begin(rowptr, w, blk_h, M) : 0;
    t[it] = (e[it] + tile - 1) / tile;
    se += e[it];
    st += t[it];
  }
  const int ie = wave_inclusive_sum(se), itl = wave_inclusive_sum(st);
  if (lane == 63) {
    s_e[wave] = ie;
    s_t[wave] = itl;
  }
  __syncthreads();
  int re = ie - se, rt = itl - st, te = 0, tt = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    re += j < wave ? s_e[j] : 0;
    rt += j < wave ? s_t[j] : 0;
    te += s_e[j];
    tt += s_t[j];
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int i = tid * IT + it;
    if (i < n_big) {
      csb[i] = re;
      tb[i] = rt;
      for (int k = 0; k < t[it]; ++k) tw[rt + k] = i;
    }
    re += e[it];
    rt += t[it];
  }
  if (tid == 0) {
    csb[n_big] = te;
    tb[n_big] = tt;
  }
}

// Largest column of the global path's windows (the width of its sort keys)
// and negative columns there, read by tiles of `tile` entries as laid out by
// k_big_offsets; with more than kOneGroupWindows such windows (their tiles
// are laid out after the first host read) every column is read.  One atomic
// pair per workgroup.
__global__ __launch_bounds__(256) void k_big_maxcol(const int32_t *rowptr, const int32_t *col,
                                                    int32_t M, int32_t nnz, int32_t blk_h,
                                                    const int32_t *wmap, const int32_t *csb,
                                                    const int32_t *tb, const int32_t *tw,
                                                    int32_t tile, MetcfHeader *h) {
  const int n_big = h->cls_count[kClasses];
  if (n_big == 0) return;  // workgroup-uniform
  int mx = 0, neg = 0;
  if (n_big <= kOneGroupWindows) {
    const int tiles = tb[n_big];
    for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
      const int k = tw[t];
      const int s = window_begin(rowptr, wmap[k], blk_h, M);
      const int off = (t - tb[k]) * tile;
      const int len = min(tile, csb[k + 1] - csb[k] - off);
      for (int i = threadIdx.x; i < len; i += blockDim.x) {
        const int c = col[s + off + i];
        mx = max(mx, c);
        neg |= c;
      }
    }
  } else {
    // every column, 16 bytes per lane when the array is 16-byte aligned
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n4 = ((uintptr_t)col % 16 == 0) ? nnz / 4 : 0;
    const int4 *col4 = reinterpret_cast<const int4 *>(col);
    for (int64_t i = tid; i < n4; i += 4 * stride) {
      int4 c[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t j = i + u * stride;
        c[u] = j < n4 ? col4[j] : make_int4(0, 0, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        neg |= c[u].x | c[u].y | c[u].z | c[u].w;
        mx = max(mx, max(max(c[u].x, c[u].y), max(c[u].z, c[u].w)));
      }
    }
    for (int64_t e = n4 * 4 + tid; e < nnz; e += stride) {
      const int c = col[e];
      neg |= c;
      mx = max(mx, c);
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    neg |= __shfl_xor(neg, d, 64);
    mx = max(mx, __shfl_xor(mx, d, 64));
  }
  __shared__ int s_m[4], s_n[4];
  if ((threadIdx.x & 63) == 0) {
    s_m[threadIdx.x >> 6] = mx;
    s_n[threadIdx.x >> 6] = neg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if ((s_n[0] | s_n[1] | s_n[2] | s_n[3]) < 0) atomicOr(&h->flags, kBadColumn);
    // spread over kMaxSlots words: atomics on one word serialise in L2
    atomicMax(&h->max_col[blockIdx.x % kMaxSlots], max(max(s_m[0], s_m[1]), max(s_m[2], s_m[3])));
  }
}

// Many global-path windows: the same offsets with one thread per window
// and device-wide scans (csb and tb hold the counts, with a trailing 0, until
// scanned), then the tile map.
__global__ __launch_bounds__(256) void k_big_sizes(const int32_t *rowptr, int32_t M,
                                                   int32_t blk_h, const int32_t *wmap,
                                                   int32_t n_big, int32_t tile, int32_t *csb,
                                                   int32_t *tb) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_big) {
    const int w = wmap[k];
    const int E = window_end(rowptr, w, blk_h, M) - window_begin(rowptr, w, blk_h, M);
    csb[k] = E;
    tb[k] = (E + tile - 1) / tile;
  } else if (k == n_big) {
    csb[k] = 0;
    tb[k] = 0;
  }
}

__global__ __launch_bounds__(256) void k_big_tilemap(const int32_t *tb, int32_t n_big,
                                                     int32_t *tw) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_big)
    for (int t = tb[k]; t < tb[k + 1]; ++t) tw[t] = k;
}

// ------------------------------------------------------------- LDS path

// Block data kept in the workspace by phase 1 (both paths) for phase 2, at
// CSR positions of the window: block slots s + kb and distinct-column slots
// s + r stay inside [s, s + E).
struct WinStage {
  int32_t *boff;   // first TC-block position of block kb, window-relative
  int32_t *dcol;   // r-th distinct column of the window
  int32_t *wU;     // [W] distinct columns of the window
};

// Largest r in [0, 16) with s_rp[r] <= q: the window row holding local entry
// q (s_rp[r] = start of row r relative to the window; empty rows share the
// start of the next row, s_rp[0] = 0 <= q).
__device__ __forceinline__ int row_of(const int *s_rp, int q) {
  int r = 0;
#pragma unroll
  for (int step = kWinH / 2; step >= 1; step >>= 1)
    if (s_rp[r + step] <= q) r += step;
  return r;
}

// Number of set bits of m in lanes below this one.
__device__ __forceinline__ int lanes_below(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// cnt[row * NDIG + d] holds the number of entries of digit d in counter row
// `row` (one row per wave, or a single workgroup-wide row); replaced by the
// position of that row's first entry of digit d in the digit-sorted order
// (digit-major, then row).  Every thread of the NT-thread workgroup must call
// it.
template <int NT, int NDIG, int NROWS>
__device__ __forceinline__ void counts_to_offsets(int *cnt, int *s_scan) {
  constexpr int DPT = NDIG >= NT ? NDIG / NT : 1;  // digits per thread
  constexpr int ACTIVE = NDIG / DPT;               // threads holding digits
  const int tid = threadIdx.x;
  int sum = 0;
  if (tid < ACTIVE) {
#pragma unroll 4
    for (int q = 0; q < DPT; ++q)
#pragma unroll 4
      for (int vv = 0; vv < NROWS; ++vv) sum += cnt[vv * NDIG + tid * DPT + q];
  }
  const int inc = wave_inclusive_sum(sum);
  int base = inc - sum;
  if (ACTIVE > 64) {
    if ((tid & 63) == 63 && tid < ACTIVE) s_scan[tid >> 6] = inc;
    __syncthreads();
    const int wv = tid >> 6;
#pragma unroll
    for (int j = 0; j < ACTIVE / 64; ++j) base += j < wv ? s_scan[j] : 0;
  }
  if (tid < ACTIVE) {
#pragma unroll 2
    for (int q = 0; q < DPT; ++q) {
#pragma unroll 4
      for (int vv = 0; vv < NROWS; ++vv) {
        const int i = vv * NDIG + tid * DPT + q;
        const int c = cnt[i];
        cnt[i] = base;
        base += c;
      }
    }
  }
  __syncthreads();
}

// Digit widths of the window sort per workgroup shape: the first pass (not
// stable, one workgroup-wide counter row) takes D1 bits, the later ones
// (stable, one counter row per wave) D2 bits, so that a 20-bit column span
// takes two passes in the 4- to 16-wave shapes while the counters stay small
// enough for four waves per SIMD (the 1-wave shape, for windows of at most
// 512 entries, keeps 8-bit digits: its fixed cost per digit dominates).  The
// shapes of at most 4,096 entries move key and index in one step through a
// 32-bit and a 16-bit buffer; the larger ones through one 16-bit buffer.
constexpr int win_d1(int waves) { return waves == 1 ? 8 : waves <= 4 ? 11 : 12; }
constexpr int win_d2(int waves) { return waves == 1 ? 8 : waves <= 8 ? 9 : 8; }
constexpr int win_cnt(int waves) {
  return (1 << win_d1(waves)) > (waves << win_d2(waves)) ? (1 << win_d1(waves))
                                                          : (waves << win_d2(waves));
}
constexpr bool win_wide(int waves, int items) { return 64 * waves * items <= 4096; }

// Phase 1 of the LDS path, one workgroup per window of the class list.
// Local CSR entry q is held by wave q / (64 ITEMS), slot (q / 64) % ITEMS,
// lane q % 64; after the sort the same layout holds sorted position p.
//   sort   LSD radix sort of (column - window minimum, q) over the window's
//          column span: a first pass of D1 bits ranked by LDS atomics on
//          one workgroup-wide counter row (it need not be stable: there is
//          no earlier order to keep, and no output depends on the order of
//          entries of equal column), then stable passes of D2 bits in which
//          each wave ranks its entries slot by slot (one ballot per digit
//          bit gives the lanes of equal digit, added to per-wave LDS
//          counters); counts_to_offsets turns the counts into positions,
//          and key and index move to them through LDS;
//   ranks  a sorted entry is new when its column differs from its
//          predecessor's; rank = (new entries up to it) - 1 is its
//          edgeToColumn, U = number of distinct columns;
//   blocks block kb holds ranks [8 kb, 8 kb + 8): its entries are the sorted
//          positions [first of rank 8 kb, first of rank 8 kb + 8); an entry's
//          TC-block position is the block start plus the number of the
//          block's entries with a smaller CSR index (the reference's
//          CSR-order placement inside a block).
// Writes edgeToRow, edgeToColumn, blockPartition, the tile ids (and values,
// edge indices) in TC-block order, and the WinStage block data.
template <int WAVES, int ITEMS>
__global__ __launch_bounds__(64 * WAVES) void k_window_sort(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col, int32_t M,
    const int32_t *__restrict__ list, const float *__restrict__ values, int32_t *__restrict__ e2r,
    int32_t *__restrict__ e2c, int32_t *__restrict__ bp, uint8_t *__restrict__ tile,
    float *__restrict__ tcvalues, int32_t *__restrict__ tcedge, WinStage st, MetcfHeader *h) {
  constexpr int NT = 64 * WAVES;
  constexpr int CAP = NT * ITEMS;
  constexpr bool kWide = win_wide(WAVES, ITEMS);
  // the bitmap path for TC-block positions, in the shapes of the largest
  // windows (where blocks fill up); the smaller shapes count per entry
  constexpr bool kBitmap = WAVES >= 8;
  constexpr int D1 = win_d1(WAVES), D2 = win_d2(WAVES);
  constexpr int N1 = 1 << D1, N2 = 1 << D2;
  constexpr int kRaw = kWide ? CAP + CAP / 2 : CAP / 2;
  __shared__ uint32_t s_raw[kRaw];
  __shared__ __align__(16) int s_cnt[win_cnt(WAVES)];
  __shared__ uint16_t s_bs[CAP / kWinW + 1];
  __shared__ int s_rp[kWinH + 1];
  __shared__ int s_min[WAVES], s_max[WAVES], s_wtot[WAVES];
  __shared__ int s_unsorted;  // bitmap path: some row not strictly increasing
  __shared__ uint32_t s_last[WAVES];
  __shared__ int s_scan[16];
  uint32_t *s_k = s_raw;                                                 // kWide only
  uint16_t *s_x = reinterpret_cast<uint16_t *>(kWide ? s_raw + CAP : s_raw);
  const int tid = threadIdx.x, v = tid >> 6, lane = tid & 63;
  const int vbase = v * 64 * ITEMS;
  const int w = list[blockIdx.x];
  const int s = window_begin(rowptr, w, kWinH, M);
  const int E = window_end(rowptr, w, kWinH, M) - s;
  if (E <= 0 || E > CAP) {  // the class list guarantees 0 < E <= CAP
    if (tid == 0) atomicOr(&h->flags, kBadClass);
    return;
  }
  if (tid <= kWinH) s_rp[tid] = rowptr[min((int64_t)w * kWinH + tid, (int64_t)M)] - s;
  __syncthreads();

  // key[k]: column - window minimum.  ix[k]: local index (low 16 bits)
  // and, during a pass, the entry's destination, after the sort its running
  // count of new columns in the wave (high 16 bits).
  uint32_t key[ITEMS];
  uint32_t ix[ITEMS];
  int cmin = INT32_MAX, cmax = 0;
  // Branch-free loads (clamped to the window) so that all of them are in
  // flight together.
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int q = vbase + k * 64 + lane;
    ix[k] = (uint32_t)q;
    const int c = col[s + min(q, E - 1)];
    key[k] = q < E ? (uint32_t)c : 0u;
    cmin = q < E ? min(cmin, c) : cmin;
    cmax = q < E ? max(cmax, c) : cmax;
  }
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int q = vbase + k * 64 + lane;
    if (q < E) e2r[s + q] = w * kWinH + row_of(s_rp, q);
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    cmin = min(cmin, __shfl_xor(cmin, d, 64));
    cmax = max(cmax, __shfl_xor(cmax, d, 64));
  }
  if (lane == 0) {
    s_min[v] = cmin;
    s_max[v] = cmax;
  }
  __syncthreads();
#pragma unroll
  for (int vv = 0; vv < WAVES; ++vv) {
    cmin = min(cmin, s_min[vv]);
    cmax = max(cmax, s_max[vv]);
  }
  if (tid == 0 && cmin < 0) atomicOr(&h->flags, kBadColumn);  // reported after phase 1
  const uint32_t span = (uint32_t)(cmax - cmin);
  const int bits = span ? 32 - __clz(span) : 0;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k)
    if (vbase + k * 64 + lane < E) key[k] -= (uint32_t)cmin;

  // Moves key and index of every entry to the position in the high half of
  // its ix word.
  auto move = [&]() {
    if (kWide) {
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        if (vbase + k * 64 + lane < E) {
          s_k[ix[k] >> 16] = key[k];
          s_x[ix[k] >> 16] = (uint16_t)ix[k];
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        const int q = vbase + k * 64 + lane;
        if (vbase + k * 64 < E && q < E) {
          key[k] = s_k[q];
          ix[k] = s_x[q];
        }
      }
      __syncthreads();
    } else {
      // through the 16-bit buffer: the high key half (if the span needs it),
      // the low key half, then the local index
      uint32_t hi[ITEMS];
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) hi[k] = 0;
      if (bits > 16) {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k)
          if (vbase + k * 64 + lane < E) s_x[ix[k] >> 16] = (uint16_t)(key[k] >> 16);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
          const int q = vbase + k * 64 + lane;
          if (vbase + k * 64 < E && q < E) hi[k] = s_x[q];
        }
        __syncthreads();
      }
#pragma unroll
      for (int k = 0; k < ITEMS; ++k)
        if (vbase + k * 64 + lane < E) s_x[ix[k] >> 16] = (uint16_t)key[k];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        const int q = vbase + k * 64 + lane;
        if (vbase + k * 64 < E && q < E) key[k] = (hi[k] << 16) | s_x[q];
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < ITEMS; ++k)
        if (vbase + k * 64 + lane < E) s_x[ix[k] >> 16] = (uint16_t)ix[k];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        const int q = vbase + k * 64 + lane;
        if (vbase + k * 64 < E && q < E) ix[k] = s_x[q];
      }
      __syncthreads();
    }
  };

  int shift = 0;
  if (bits > 0) {
    // first pass: low D1 bits, ranks from atomics on one counter row
    for (int i = tid; i < N1; i += NT) s_cnt[i] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      if (vbase + k * 64 < E) {  // wave-uniform
        const bool valid = vbase + k * 64 + lane < E;
        const int r = valid ? atomicAdd(&s_cnt[key[k] & (N1 - 1)], 1) : 0;
        ix[k] = (ix[k] & 0xffffu) | ((uint32_t)r << 16);
      }
    }
    __syncthreads();
    counts_to_offsets<NT, N1, 1>(s_cnt, s_scan);
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
      if (vbase + k * 64 + lane < E) ix[k] += (uint32_t)s_cnt[key[k] & (N1 - 1)] << 16;
    move();
    shift = D1;
  }
  for (; shift < bits; shift += D2) {
    // stable pass: D2 bits from `shift`, ranks from ballots, one counter
    // row per wave
    const int nbits = min(D2, bits - shift);
    int *row = s_cnt + v * N2;
#pragma unroll
    for (int j = 0; j < N2 / 64; ++j) row[j * 64 + lane] = 0;
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      if (vbase + k * 64 < E) {  // wave-uniform
        const bool valid = vbase + k * 64 + lane < E;
        const int d = (int)((key[k] >> shift) & (uint32_t)(N2 - 1));
        uint64_t peers = __ballot(valid);
        for (int bit = 0; bit < nbits; ++bit) {
          const bool b = (d >> bit) & 1;
          const uint64_t m = __ballot(b);
          peers &= b ? m : ~m;
        }
        const int r = lanes_below(peers);
        const int c = valid ? row[d] : 0;
        ix[k] = (ix[k] & 0xffffu) | ((uint32_t)(c + r) << 16);
        wave_lds_sync();
        if (valid && r == 0) row[d] = c + __popcll(peers);
        wave_lds_sync();
      }
    }
    __syncthreads();
    counts_to_offsets<NT, N2, WAVES>(s_cnt, s_scan);
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
      if (vbase + k * 64 + lane < E)
        ix[k] += (uint32_t)row[(key[k] >> shift) & (uint32_t)(N2 - 1)] << 16;
    move();
  }
  // s_x[p] = local index of sorted position p (left there by the last move).
  if (bits == 0) {
#pragma unroll
    for (int k = 0; k < ITEMS; ++k)
      if (vbase + k * 64 + lane < E) s_x[vbase + k * 64 + lane] = (uint16_t)ix[k];
  }

  // ranks of the distinct columns: high half of ix = new columns of the wave
  // up to and including the entry
  if (lane == 63) s_last[v] = key[ITEMS - 1];
  __syncthreads();
  const uint32_t prev_wave = v > 0 ? s_last[v - 1] : 0u;
  uint32_t newmask = 0;
  int run = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    if (vbase + k * 64 < E) {  // wave-uniform
      const int q = vbase + k * 64 + lane;
      const uint32_t up = __shfl_up(key[k], 1, 64);
      const uint32_t from_prev =
          k > 0 ? (uint32_t)__builtin_amdgcn_readlane((int)key[k > 0 ? k - 1 : 0], 63) : prev_wave;
      const uint32_t pk = lane == 0 ? from_prev : up;
      const bool isnew = q < E && (q == 0 || key[k] != pk);
      const uint64_t m = __ballot(isnew);
      const int rk = run + lanes_below(m) + (isnew ? 1 : 0);
      run += __popcll(m);
      ix[k] = (ix[k] & 0xffffu) | ((uint32_t)rk << 16);
      newmask |= isnew ? (1u << k) : 0u;
    }
  }
  if (lane == 0) s_wtot[v] = run;
  __syncthreads();
  int pre = 0, U = 0;
#pragma unroll
  for (int vv = 0; vv < WAVES; ++vv) {
    const int t = s_wtot[vv];
    pre += vv < v ? t : 0;
    U += t;
  }
  const int nb = (U + kWinW - 1) / kWinW;
  // Bitmap path: block kb's 16 x 8 (row, slot) occupancy in words
  // s_cnt[4 kb .. 4 kb + 3] (the sort's counters are free now).  With
  // strictly increasing rows, the CSR order of a block's entries is the
  // row-major (row, slot) order, so an entry's place in its block is the
  // number of set bits before its own.  It pays where blocks are full (at
  // least 16 entries on average); otherwise each entry counts the few
  // entries of its block with a smaller CSR index.
  bool bitmap = false;
  if (kBitmap && nb * 4 <= win_cnt(WAVES) && E >= 16 * nb) {  // workgroup-uniform
    for (int i = tid; i < nb * 4; i += NT) s_cnt[i] = 0;
    if (tid == 0) s_unsorted = 0;
    __syncthreads();
    bool unsorted = false;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int q = vbase + k * 64 + lane;
      if (q < E) {
        const int r = row_of(s_rp, q);
        if (q > s_rp[r]) unsorted |= col[s + q] <= col[s + q - 1];  // not a row's first entry
      }
    }
    if (__ballot(unsorted) != 0 && lane == 0) s_unsorted = 1;
    __syncthreads();
    bitmap = s_unsorted == 0;
  }
  if (bitmap) {
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const int q = vbase + k * 64 + lane;
      if (vbase + k * 64 < E && q < E) {
        const int rank = (int)(ix[k] >> 16) + pre - 1;
        const int bit = row_of(s_rp, (int)(ix[k] & 0xffffu)) * kWinW + rank % kWinW;
        atomicOr(reinterpret_cast<unsigned *>(&s_cnt[(rank / kWinW) * 4 + (bit >> 5)]),
                 1u << (bit & 31));
      }
    }
  }
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    if ((newmask >> k) & 1u) {
      const int q = vbase + k * 64 + lane;
      const int rank = (int)(ix[k] >> 16) + pre - 1;
      st.dcol[s + rank] = (int32_t)(key[k] + (uint32_t)cmin);
      if (rank % kWinW == 0) {
        s_bs[rank / kWinW] = (uint16_t)q;
        st.boff[s + rank / kWinW] = q;
      }
    }
  }
  if (tid == 0) {
    s_bs[nb] = (uint16_t)E;
    bp[w] = nb;
    st.wU[w] = U;
  }
  __syncthreads();
  // TC-block position of every entry; its tile id, value and edge index
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int q = vbase + k * 64 + lane;
    if (vbase + k * 64 < E && q < E) {
      const int rank = (int)(ix[k] >> 16) + pre - 1;
      const uint32_t i = ix[k] & 0xffffu;
      const int kb = rank / kWinW, c = rank % kWinW;
      const int lo = s_bs[kb];
      const int rr = row_of(s_rp, (int)i);
      int pos = lo;
      if (bitmap) {
        const int bit = rr * kWinW + c;
        const int wi = bit >> 5;
        const uint32_t below = (1u << (bit & 31)) - 1u;
        const int4 m = *reinterpret_cast<const int4 *>(&s_cnt[kb * 4]);
        pos += __popc(wi == 0 ? ((uint32_t)m.x & below) : (uint32_t)m.x);
        pos += wi >= 1 ? __popc(wi == 1 ? ((uint32_t)m.y & below) : (uint32_t)m.y) : 0;
        pos += wi >= 2 ? __popc(wi == 2 ? ((uint32_t)m.z & below) : (uint32_t)m.z) : 0;
        pos += wi >= 3 ? __popc((uint32_t)m.w & below) : 0;
      } else {
        const int hi = s_bs[kb + 1];
        for (int j = lo; j < hi; ++j) pos += (uint32_t)s_x[j] < i ? 1 : 0;
      }
      if (pos < E) {
        tile[s + pos] = (uint8_t)(rr * kWinW + c);
        if (tcvalues) tcvalues[s + pos] = values ? values[s + i] : 1.0f;
        if (tcedge) tcedge[s + pos] = s + (int)i;
      }
    }
  }
  __syncthreads();
  // edgeToColumn in CSR order through LDS
#pragma unroll
  for (int k = 0; k < ITEMS; ++k)
    if (vbase + k * 64 + lane < E) s_x[ix[k] & 0xffffu] = (uint16_t)((int)(ix[k] >> 16) + pre - 1);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int q = vbase + k * 64 + lane;
    if (vbase + k * 64 < E && q < E) e2c[s + q] = (int)s_x[q];
  }
}

// Ascending bitonic sort of one key per lane across the 64 lanes of a wave.
template <class T>
__device__ __forceinline__ T wave_bitonic_sort(T key) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const T other = __shfl_xor(key, j, 64);
      const bool asc = (lane & k) == 0;  // always true in the last merge (k = 64)
      const bool low = (lane & j) == 0;
      const T lo = key < other ? key : other;
      const T hi = key < other ? other : key;
      key = low == asc ? lo : hi;
    }
  }
  return key;
}

// Windows of at most 64 entries: one wave per window, four windows per
// workgroup, no LDS.  Lane q holds local entry q.  A bitonic sort of
// (column - window minimum, q) gives the sorted order; a sorted entry is new
// when its column differs from its predecessor's, rank = (new entries up to
// it) - 1.  A second bitonic sort of (block, q, row, slot) gives the TC-block
// order directly (inside a block the reference keeps CSR order, which is q
// order), so lane p then writes TC-block position p.
__global__ __launch_bounds__(256) void k_window_tiny(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col, int32_t M,
    const int32_t *__restrict__ list, int32_t count, const float *__restrict__ values,
    int32_t *__restrict__ e2r, int32_t *__restrict__ e2c, int32_t *__restrict__ bp,
    uint8_t *__restrict__ tile, float *__restrict__ tcvalues, int32_t *__restrict__ tcedge,
    WinStage st, MetcfHeader *h) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= count) return;  // wave-uniform
  const int w = list[i];
  const int s = window_begin(rowptr, w, kWinH, M);
  const int E = window_end(rowptr, w, kWinH, M) - s;
  if (E <= 0 || E > 64) {  // the class list guarantees 0 < E <= 64
    if (lane == 0) atomicOr(&h->flags, kBadClass);
    return;
  }
  // lane r (1..15) holds the start of window row r, relative to the window
  const int rs = (lane >= 1 && lane < kWinH)
                     ? rowptr[min((int64_t)w * kWinH + lane, (int64_t)M)] - s
                     : INT32_MAX;
  const int q = lane;
  const bool valid = q < E;
  const int c = col[s + min(q, E - 1)];
  int row = 0;
#pragma unroll
  for (int r = 1; r < kWinH; ++r) row += __shfl(rs, r, 64) <= q ? 1 : 0;
  if (valid) e2r[s + q] = w * kWinH + row;
  int cmin = valid ? c : INT32_MAX, cmax = valid ? c : 0;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    cmin = min(cmin, __shfl_xor(cmin, d, 64));
    cmax = max(cmax, __shfl_xor(cmax, d, 64));
  }
  (void)cmax;
  if (lane == 0 && cmin < 0) atomicOr(&h->flags, kBadColumn);  // reported after phase 1
  // sort 1: (column - window minimum, local index); empty lanes last
  uint64_t k1 = valid ? (((uint64_t)(uint32_t)(c - cmin) << 6) | (uint64_t)q) : ~0ull;
  k1 = wave_bitonic_sort(k1);
  const int p = lane;
  const uint32_t colr = (uint32_t)(k1 >> 6);
  const int idx = (int)(k1 & 63);
  const uint32_t prevc = __shfl_up(colr, 1, 64);
  const bool isnew = p < E && (p == 0 || colr != prevc);
  const uint64_t m = __ballot(isnew);
  const int rank = lanes_below(m) + (isnew ? 1 : 0) - 1;
  const int U = __popcll(m);
  const int kb = rank / kWinW, cc = rank % kWinW;
  if (isnew) {
    st.dcol[s + rank] = (int32_t)(colr + (uint32_t)cmin);
    if (cc == 0) st.boff[s + kb] = p;
  }
  if (lane == 0) {
    bp[w] = (U + kWinW - 1) / kWinW;
    st.wU[w] = U;
  }
  if (p < E) e2c[s + idx] = rank;
  // sort 2: (block, local index, row, slot); empty lanes last
  const int row_s = __shfl(row, idx, 64);  // row of local entry idx
  uint32_t k2 = p < E ? ((uint32_t)kb << 14) | ((uint32_t)idx << 7) | ((uint32_t)row_s << 3) |
                            (uint32_t)cc
                      : ~0u;
  k2 = wave_bitonic_sort(k2);
  if (lane < E) {
    const int i2 = (int)((k2 >> 7) & 63u);
    tile[s + lane] = (uint8_t)(((k2 >> 3) & 15u) * kWinW + (k2 & 7u));
    if (tcvalues) tcvalues[s + lane] = values ? values[s + i2] : 1.0f;
    if (tcedge) tcedge[s + lane] = s + i2;
  }
}

// Phase 2 (every window, both paths): a workgroup takes kOutBlocks
// consecutive TC blocks.  The windows of its first and last block come from
// k_chunk_windows; the row-window offsets and CSR starts of the windows in
// between go to LDS (when they fit), where each block finds its window
// (rwo[w] <= b < rwo[w+1]).  Each block's offset and row id, and then SparseAToX (the window's
// distinct columns from kb * blk_w on, zero past U, consecutive threads on
// consecutive entries) come from the WinStage data.
constexpr int kOutWin = 1024;  // row-window offsets kept in LDS

// Phase 1, after RowWindowOffset: the window of the first and of the last
// block of every phase-2 chunk (blocks [c kOutBlocks, min((c + 1) kOutBlocks,
// T))), so that its workgroups start without a search.  One thread per
// window; a window writes the chunks whose first or last block it holds
// (every block has at least one entry, so T <= nnz and c < cw_stride).
__global__ __launch_bounds__(256) void k_chunk_windows(const int32_t *__restrict__ rwo, int32_t W,
                                                       int32_t *__restrict__ cw_first,
                                                       int32_t *__restrict__ cw_last) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const int64_t b0 = rwo[w], b1 = rwo[w + 1];
  if (b0 >= b1) return;  // empty window: no block
  const int64_t T = rwo[W];
  for (int64_t c = (b0 + kOutBlocks - 1) / kOutBlocks; c * kOutBlocks < b1; ++c)
    cw_first[c] = w;
  for (int64_t c = b0 / kOutBlocks; c <= (b1 - 1) / kOutBlocks; ++c) {
    const int64_t last = std::min(c * kOutBlocks + kOutBlocks, T) - 1;
    if (last >= b0 && last < b1) cw_last[c] = w;
  }
}
__global__ __launch_bounds__(256) void k_blocks_out(const int32_t *__restrict__ rowptr, int32_t M,
                                                    int32_t W, int32_t blk_h, int32_t blk_w,
                                                    int32_t T, int32_t nnz, WinStage st,
                                                    const int32_t *__restrict__ rwo,
                                                    const int32_t *__restrict__ cw_first,
                                                    const int32_t *__restrict__ cw_last,
                                                    int32_t *__restrict__ off,
                                                    int32_t *__restrict__ rowid,
                                                    int32_t *__restrict__ atox) {
  __shared__ int s_rwo[kOutWin + 1], s_ws[kOutWin + 1];
  __shared__ int s_base[kOutBlocks], s_kb[kOutBlocks], s_U[kOutBlocks];
  const int tid = threadIdx.x;
  const int first = blockIdx.x * kOutBlocks;
  if (first >= T) return;  // workgroup-uniform
  if (first == 0 && tid == 0) off[T] = nnz;
  const int nblk = min(kOutBlocks, T - first);
  // (clamped into [0, W) so that the LDS range stays in bounds whatever the
  // workspace holds)
  const int wa = min(max(cw_first[blockIdx.x], 0), W - 1);
  const int wb = min(max(cw_last[blockIdx.x], wa), W - 1);
  const bool in_lds = wb - wa + 1 <= kOutWin;
  if (in_lds)
    for (int k = tid; k <= wb - wa + 1; k += blockDim.x) {
      s_rwo[k] = rwo[wa + k];
      s_ws[k] = rowptr[min((int64_t)(wa + k) * blk_h, (int64_t)M)];
    }
  __syncthreads();
#pragma unroll 2
  for (int t = tid; t < nblk; t += blockDim.x) {
    const int b = first + t;
    int w;
    if (in_lds) {
      int lo = 1, hi = wb - wa + 1;  // first k in [1, wb - wa + 1] with s_rwo[k] > b
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s_rwo[mid] <= b) lo = mid + 1;
        else hi = mid;
      }
      w = wa + lo - 1;
    } else {
      w = lower_bound_dev(rwo, wa + 1, wb + 1, b + 1) - 1;
    }
    const int s = in_lds ? s_ws[w - wa] : rowptr[min((int64_t)w * blk_h, (int64_t)M)];
    const int kb = b - (in_lds ? s_rwo[w - wa] : rwo[w]);
    off[b] = s + st.boff[s + kb];
    rowid[b] = w;
    s_base[t] = s + kb * blk_w;
    s_kb[t] = kb * blk_w;
    s_U[t] = st.wU[w];
  }
  __syncthreads();
  const int n_atox = nblk * blk_w;
  const int64_t a0 = (int64_t)first * blk_w;
  if (blk_w == kWinW && ((uintptr_t)atox & 15) == 0) {
    // 8-column blocks: four consecutive entries (half a block) per thread and
    // one 16-byte store (a0 is a multiple of 8)
#pragma unroll 2
    for (int e = tid * 4; e < n_atox; e += blockDim.x * 4) {
      const int t = e >> 3, c = e & 7;
      const int left = s_U[t] - (s_kb[t] + c);  // distinct columns left from this entry on
      const int32_t *d = st.dcol + s_base[t] + c;
      int4 v;
      v.x = left > 0 ? d[0] : 0;
      v.y = left > 1 ? d[1] : 0;
      v.z = left > 2 ? d[2] : 0;
      v.w = left > 3 ? d[3] : 0;
      *reinterpret_cast<int4 *>(atox + a0 + e) = v;
    }
  } else {
#pragma unroll 4
    for (int e = tid; e < n_atox; e += blockDim.x) {
      const int t = e / blk_w, c = e - t * blk_w;
      atox[a0 + e] = s_kb[t] + c < s_U[t] ? st.dcol[s_base[t] + c] : 0;
    }
  }
}

// ---------------------------------------------------------- global path

// Keys of the global-path windows: compacted entry j of window w'
// (entries [csb[w'], csb[w'+1])) is CSR entry rowptr[w*blk_h] + j - csb[w']
// of window w = wmap[w']; value = CSR edge index; wpos[j] = w'.  Key, for
// the segmented sort (32 bits, each window sorted within its own range):
// column with bit 31 = parity of w', so that neighbouring windows never share
// a key; for the sort of the whole range (64 bits): (w', column).  Either
// way the sorted order keeps every entry inside its window's range.  Also
// writes edgeToRow of those entries.  A workgroup takes 4,096 consecutive
// compacted entries; two of its threads find the windows of the first and
// the last one.  When both are the same window (the usual case: these
// windows hold more than 16,384 entries), its row starts go to LDS once and
// every entry finds its row there; otherwise each entry searches rowptr.
constexpr int kBigChunk = 4096;
constexpr int kBigRowsLds = 256;  // row starts kept in LDS (blk_h <= 255)
template <class Key>
__device__ __forceinline__ Key big_key(int wp, int c, int colbits) {
  if (sizeof(Key) == 4) return (Key)(((uint32_t)(wp & 1) << 31) | (uint32_t)c);
  return (Key)(((uint64_t)wp << colbits) | (uint32_t)c);
}

template <class Key>
__global__ __launch_bounds__(256) void k_big_keys(const int32_t *rowptr, const int32_t *col,
                                                  int32_t M, int32_t blk_h, int colbits,
                                                  int32_t n_big, int32_t nnz_big,
                                                  const int32_t *wmap, const int32_t *csb,
                                                  int32_t *e2r, Key *keys, int32_t *vals,
                                                  int32_t *wpos) {
  __shared__ int s_w[2];
  __shared__ int s_rp[kBigRowsLds];
  const int64_t first = (int64_t)blockIdx.x * kBigChunk;
  if (first >= nnz_big) return;  // workgroup-uniform
  const int last = (int)min<int64_t>(first + kBigChunk - 1, nnz_big - 1);
  // window of entry j: the last w' with csb[w'] <= j
  if (threadIdx.x < 2)
    s_w[threadIdx.x] = lower_bound_dev(csb, 0, n_big + 1, (threadIdx.x ? last : (int)first) + 1) - 1;
  __syncthreads();
  const int wa = s_w[0], wb = s_w[1];
  if (wa == wb && blk_h < kBigRowsLds) {
    const int w = wmap[wa];
    const int64_t r0 = (int64_t)w * blk_h;
    const int nrows = (int)min<int64_t>(blk_h, M - r0);
    if (threadIdx.x <= nrows) s_rp[threadIdx.x] = rowptr[r0 + threadIdx.x];
    __syncthreads();
    const int s = s_rp[0];
    const int delta = s - csb[wa];
#pragma unroll 4
    for (int j = (int)first + threadIdx.x; j <= last; j += blockDim.x) {
      const int e = j + delta;
      // row: the last r in [0, nrows) with s_rp[r] <= e
      int lo = 0, hi = nrows;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_rp[mid] <= e) lo = mid;
        else hi = mid;
      }
      e2r[e] = (int)r0 + lo;
      keys[j] = big_key<Key>(wa, col[e], colbits);
      vals[j] = e;
      wpos[j] = wa;
    }
    return;
  }
  for (int j = (int)first + threadIdx.x; j <= last; j += blockDim.x) {
    const int wp = wa == wb ? wa : lower_bound_dev(csb, wa, wb + 1, j + 1) - 1;
    const int w = wmap[wp];
    const int64_t r0 = (int64_t)w * blk_h;
    const int s = rowptr[r0];
    const int e = s + (j - csb[wp]);
    // row: the last r in [r0, min(r0 + blk_h, M)) with rowptr[r] <= e
    const int rend = (int)min<int64_t>(r0 + blk_h, M);
    const int row = lower_bound_dev(rowptr, (int)r0, rend, e + 1) - 1;
    e2r[e] = row;
    keys[j] = big_key<Key>(wp, col[e], colbits);
    vals[j] = e;
    wpos[j] = wp;
  }
}

struct BigView {
  const uint32_t *keys32;  // segmented sort
  const uint64_t *keys64;  // sort of the whole range
  const int32_t *vals;
  const int32_t *g;
  const int32_t *csb;
  const int32_t *wmap;
  const int32_t *wpos;
  int colbits;
  __device__ __forceinline__ int window(int j) const { return wpos[j]; }
  __device__ __forceinline__ int column(int j) const {
    return keys32 ? (int)(keys32[j] & 0x7fffffffu)
                  : (int)(keys64[j] & ((1ull << colbits) - 1));
  }
  // unique-column count before compacted window wp
  __device__ __forceinline__ int base(int wp) const {
    const int c = csb[wp];
    return c > 0 ? g[c - 1] : 0;
  }
};

// The sorted global-path pairs (key, CSR edge index), their unique-column
// scan and window maps, as the kernels after the sort read them.
BigView big_view(const uint32_t *keys32, const uint64_t *keys64, int colbits, const int32_t *vals,
                 const int32_t *g, const int32_t *csb, const int32_t *wmap, const int32_t *wpos) {
  BigView bv;
  bv.keys32 = keys32;
  bv.keys64 = keys64;
  bv.colbits = colbits;
  bv.vals = vals;
  bv.g = g;
  bv.csb = csb;
  bv.wmap = wmap;
  bv.wpos = wpos;
  return bv;
}

// blockPartition of the global-path windows: ceil(|S_w| / blk_w)
// (DTCSpMM_kernel.cu:215-223), |S_w| (WinStage), and cbb[k] = blocks of the
// windows listed before k (one workgroup).
__global__ __launch_bounds__(1024) void k_big_counts(BigView bv, int32_t n_big, int32_t blk_w,
                                                     int32_t *bp, int32_t *wU, int32_t *cbb) {
  __shared__ int s_tmp[16];
  const int total = wg1024_exclusive_scan(
      [&](int wp) {
        const int c0 = bv.csb[wp], c1 = bv.csb[wp + 1];
        const int U = bv.g[c1 - 1] - (c0 > 0 ? bv.g[c0 - 1] : 0);
        const int nb = (U - 1 + blk_w) / blk_w;
        bp[bv.wmap[wp]] = nb;
        wU[bv.wmap[wp]] = U;
        return nb;
      },
      [&](int wp, int x) { cbb[wp] = x; }, n_big, s_tmp);
  if (threadIdx.x == 0) cbb[n_big] = total;  // as the multi-workgroup path leaves it
}

// The same with one thread per window (cbb then holds the block counts, with
// a trailing 0, until scanned).
__global__ __launch_bounds__(256) void k_big_counts_multi(BigView bv, int32_t n_big, int32_t blk_w,
                                                          int32_t *bp, int32_t *wU,
                                                          int32_t *cbb) {
  const int wp = blockIdx.x * blockDim.x + threadIdx.x;
  if (wp < n_big) {
    const int c0 = bv.csb[wp], c1 = bv.csb[wp + 1];
    const int U = bv.g[c1 - 1] - (c0 > 0 ? bv.g[c0 - 1] : 0);
    const int nb = (U - 1 + blk_w) / blk_w;
    bp[bv.wmap[wp]] = nb;
    wU[bv.wmap[wp]] = U;
    cbb[wp] = nb;
  } else if (wp == n_big) {
    cbb[wp] = 0;
  }
}

// Per sorted entry j of the global path: edgeToColumn[e] = rank of col(e) in
// its window (DTCSpMM_kernel.cu:224-227); the first entry of each distinct
// column stores the column (WinStage dcol), the first of each block its
// window-relative position (WinStage boff) and its compacted position
// (bstart, for the tile pass).
__global__ __launch_bounds__(256) void k_big_stage(const int32_t *rowptr, int32_t M, int32_t blk_h,
                                                   int32_t blk_w, int32_t nnz_big, BigView bv,
                                                   const int32_t *cbb, int32_t *e2c, WinStage st,
                                                   int32_t *bstart) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnz_big) return;
  const int32_t *g = bv.g;
  const int32_t *csb = bv.csb;
  const int32_t *wmap = bv.wmap;
  const int wp = bv.window(j);
  const int r = g[j] - bv.base(wp) - 1;
  e2c[bv.vals[j]] = r;
  if (j > 0 && g[j - 1] == g[j]) return;
  const int w = wmap[wp];
  const int s = window_begin(rowptr, w, blk_h, M);
  st.dcol[s + r] = bv.column(j);
  if (r % blk_w == 0) {
    const int kb = r / blk_w;
    st.boff[s + kb] = j - csb[wp];
    bstart[cbb[wp] + kb] = j;
  }
}

// Tile ids (and the optional values / edge indices) of the global-path
// windows in TC-block order.  A block's entries are one contiguous range of
// the sorted order (its columns are consecutive condensed columns of one
// window), [bstart of the block, bstart of the next or the window end); an
// entry's slot is the block start plus the number of the block's entries
// with a smaller CSR index, one pass over that range (<= blk_h * blk_w
// entries).
__global__ __launch_bounds__(256) void k_big_tile_ids(
    const int32_t *rowptr, const int32_t *e2r, int32_t M, int32_t blk_h, int32_t blk_w,
    int32_t nnz_big, BigView bv, const int32_t *cbb, const int32_t *bstart, const float *values,
    uint8_t *tile, float *tcvalues, int32_t *tcedge) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnz_big) return;
  const int32_t *g = bv.g;
  const int32_t *csb = bv.csb;
  const int32_t *wmap = bv.wmap;
  const int wp = bv.window(j);
  const int w = wmap[wp];
  const int c0 = csb[wp], c1 = csb[wp + 1];
  const int base = bv.base(wp);
  const int U = g[c1 - 1] - base;
  const int r = g[j] - base - 1;
  const int kb = r / blk_w;
  const int c = r % blk_w;
  const int delta = window_begin(rowptr, w, blk_h, M) - c0;  // CSR position - compacted position
  const int lo = bstart[cbb[wp] + kb];
  const int hi = (kb + 1) * blk_w >= U ? c1 : bstart[cbb[wp] + kb + 1];
  const int myv = bv.vals[j];
  int pos = lo;
#pragma unroll 8
  for (int t = lo; t < hi; ++t) pos += bv.vals[t] < myv;
  if (pos < c0 || pos >= c1) return;  // unreachable for a consistent build; keeps writes in bounds
  const int dst = pos + delta;
  const unsigned row_local = (unsigned)e2r[myv] % (unsigned)blk_h;
  tile[dst] = (uint8_t)(row_local * (unsigned)blk_w + (unsigned)c);
  if (tcvalues) tcvalues[dst] = values ? values[myv] : 1.0f;
  if (tcedge) tcedge[dst] = myv;
}

__global__ void k_set_i32(int32_t *p, int32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

unsigned grid_for(int64_t n, int threads) {
  return (unsigned)std::max<int64_t>(1, ceil_div(n, threads));
}

dtc_status_t check_shape(int32_t num_nodes, int32_t nnz, int32_t blk_h, int32_t blk_w) {
  if (num_nodes < 0 || nnz < 0)
    return fail(DTC_ERROR_INVALID_VALUE, "num_nodes and nnz must be >= 0");
  if (blk_h <= 0 || blk_w <= 0)
    return fail(DTC_ERROR_INVALID_VALUE, "blockSize_h and blockSize_w must be >= 1");
  if (blk_h > 4096 || blk_w > 4096)
    return fail(DTC_ERROR_UNSUPPORTED, "block sizes above 4096 are not supported");
  return DTC_SUCCESS;
}

WinStage win_stage(const MetcfWs &ws) {
  WinStage st;
  st.boff = ws.boff;
  st.dcol = ws.dcol;
  st.wU = ws.wU;
  return st;
}

// The global path runs on a library-owned stream of the current device
// (default priority, created once), forked from and joined back to the
// caller's stream with events, so that it overlaps the LDS path: the two
// write disjoint windows.  `mu` is held from the fork to the join of one
// build so that concurrent builds do not interleave their event records.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  std::mutex mu;
  // pinned host words for the build's device-to-host reads (header, block
  // count), guarded by `read_mu`
  char *pinned = nullptr;
  std::mutex read_mu;
};

dtc_status_t side_stream(SideStream **out) {
  static std::mutex reg;
  static SideStream *per_dev[64] = {};
  int dev = 0;
  DTC_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(DTC_ERROR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> g(reg);
  if (!per_dev[dev]) {
    SideStream *x = new SideStream;
    DTC_HIP_CHECK(hipStreamCreateWithFlags(&x->s, hipStreamNonBlocking));
    DTC_HIP_CHECK(hipEventCreateWithFlags(&x->fork, hipEventDisableTiming));
    DTC_HIP_CHECK(hipEventCreateWithFlags(&x->join, hipEventDisableTiming));
    DTC_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&x->pinned), 256, hipHostMallocDefault));
    per_dev[dev] = x;
  }
  *out = per_dev[dev];
  return DTC_SUCCESS;
}

template <int C>
dtc_status_t launch_window_sort(const int32_t *rowptr, const int32_t *col, int32_t M, int count,
                                const MetcfWs &ws, int64_t W, const float *values, int32_t *e2r,
                                int32_t *e2c, int32_t *bp, uint8_t *tile, float *tcvalues,
                                int32_t *tcedge, hipStream_t stream) {
  if (count <= 0) return DTC_SUCCESS;
  k_window_sort<kClassWaves[C], kClassItems[C]><<<(unsigned)count, 64 * kClassWaves[C], 0, stream>>>(
      rowptr, col, M, ws.lists + (int64_t)C * W, values, e2r, e2c, bp, tile, tcvalues, tcedge,
      win_stage(ws), ws.hdr);
  DTC_LAUNCH_CHECK("window_sort");
  return DTC_SUCCESS;
}

}  // namespace
}  // namespace dtc

using namespace dtc;

extern "C" {

dtc_status_t dtc_metcf_workspace_size(int32_t num_nodes, int32_t nnz, int32_t blk_h,
                                      int32_t blk_w, size_t *bytes) {
  clear_error();
  if (!bytes) return fail(DTC_ERROR_INVALID_VALUE, "bytes is NULL");
  dtc_status_t st = check_shape(num_nodes, nnz, blk_h, blk_w);
  if (st != DTC_SUCCESS) return st;
  const int64_t W = ceil_div(num_nodes, blk_h);
  *bytes = carve(nullptr, nnz, W).bytes;
  return DTC_SUCCESS;
}

dtc_status_t dtc_metcf_build_windows(const int32_t *edge_list, const int32_t *node_pointer,
                                     int32_t num_nodes, int32_t nnz, int32_t blk_h,
                                     int32_t blk_w, const float *values,
                                     int32_t *block_partition, int32_t *edge_to_column,
                                     int32_t *edge_to_row, int32_t *row_window_offset,
                                     uint8_t *tcblocktile_id, float *tcblock_values,
                                     int32_t *tcblock_edge, int32_t *block_count,
                                     void *workspace, size_t workspace_bytes,
                                     dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  dtc_status_t st = check_shape(num_nodes, nnz, blk_h, blk_w);
  if (st != DTC_SUCCESS) return st;
  if (!node_pointer || !row_window_offset || !block_count)
    return fail(DTC_ERROR_INVALID_VALUE, "node_pointer, row_window_offset and block_count are required");
  if (nnz > 0 && (!edge_list || !edge_to_column || !edge_to_row || !tcblocktile_id))
    return fail(DTC_ERROR_INVALID_VALUE, "edge arrays and tcblocktile_id are required when nnz > 0");
  const int64_t W = ceil_div(num_nodes, blk_h);
  if (W > 0 && !block_partition) return fail(DTC_ERROR_INVALID_VALUE, "block_partition is NULL");
  MetcfWs ws = carve(workspace, nnz, W);
  if (!workspace || workspace_bytes < ws.bytes)
    return fail(DTC_ERROR_WORKSPACE, "ME-TCF workspace needs " + std::to_string(ws.bytes) + " bytes");
  const int mode = path_mode(blk_h, blk_w);

  // 1. Row-pointer check, window classes, the global path's largest column
  //    (negative columns are flagged where they are read) (first host
  //    synchronisation).
  k_header_init<<<1, 64, 0, stream>>>(ws.hdr);
  DTC_LAUNCH_CHECK("header_init");
  {
    // up to 1,024 workgroups for the row check, and one per 256 windows for
    // the classes
    const unsigned grid = (unsigned)std::max<int64_t>(
        std::min<int64_t>(1024, grid_for(num_nodes, 256)), grid_for(W, 256));
    k_check_classify<<<grid, 256, 0, stream>>>(node_pointer, num_nodes, nnz, (int32_t)W, blk_h,
                                               mode, block_partition, ws.lists,
                                               row_window_offset, ws.hdr);
    DTC_LAUNCH_CHECK("check_classify");
  }
  if (W > 0) {
    // offsets and sort tiles of up to kOneGroupWindows global-path windows
    k_big_offsets<<<1, 1024, 0, stream>>>(node_pointer, num_nodes, blk_h,
                                          ws.lists + (int64_t)kClasses * W, ws.hdr,
                                          (int32_t)seg_radix_tile_size(), ws.csb, ws.tb, ws.tw);
    DTC_LAUNCH_CHECK("big_offsets");
    k_big_maxcol<<<1024, 256, 0, stream>>>(node_pointer, edge_list, num_nodes, nnz, blk_h,
                                           ws.lists + (int64_t)kClasses * W, ws.csb, ws.tb,
                                           ws.tw, (int32_t)seg_radix_tile_size(), ws.hdr);
    DTC_LAUNCH_CHECK("big_maxcol");
  }
  SideStream *dev_state = nullptr;
  st = side_stream(&dev_state);
  if (st != DTC_SUCCESS) return st;
  // The side stream is forked here, before the first host synchronisation,
  // so that the global path's first kernel starts as soon as it is issued;
  // the lock keeps other builds on this device off the side stream until
  // the join.
  std::unique_lock<std::mutex> side_lock(dev_state->mu);
  if (W > 0) {
    DTC_HIP_CHECK(hipEventRecord(dev_state->fork, stream));
    DTC_HIP_CHECK(hipStreamWaitEvent(dev_state->s, dev_state->fork, 0));
  }
  MetcfHeader host{};
  {
    std::lock_guard<std::mutex> g(dev_state->read_mu);
    DTC_HIP_CHECK(hipMemcpyAsync(dev_state->pinned, ws.hdr, kHeaderRead, hipMemcpyDeviceToHost,
                                 stream));
    DTC_HIP_CHECK(hipStreamSynchronize(stream));
    std::memcpy(&host, dev_state->pinned, kHeaderRead);
  }
  int max_col = 0;
  for (int c = 0; c < kMaxSlots; ++c) max_col = std::max(max_col, host.max_col[c]);
  if (host.flags) {
    std::string why;
    if (host.flags & kBadFirst) why += " nodePointer[0] != 0;";
    if (host.flags & kBadLast) why += " nodePointer[num_nodes] != number of edges;";
    if (host.flags & kBadOrder) why += " nodePointer decreases;";
    if (host.flags & kBadColumn) why += " negative column index;";
    if (why.empty() && (host.flags & kBadPartition))
      return fail(DTC_ERROR_INCONSISTENT,
                  "blockPartition must be zero-filled: an empty row window has a non-zero entry");
    return fail(DTC_ERROR_INVALID_CSR, "invalid CSR:" + why);
  }
  if (W == 0) {
    k_set_i32<<<1, 64, 0, stream>>>(row_window_offset, 0);
    DTC_LAUNCH_CHECK("set");
    *block_count = 0;
    return DTC_SUCCESS;
  }
  const int colbits = std::max(1, bits_for(max_col));
  const int n_big = host.cls_count[kClasses];
  const int nnz_big = host.nnz_big;
  const int32_t *wmap = ws.lists + (int64_t)kClasses * W;

  // 2. The global path goes to the side stream when the LDS path also has
  //    work.  Issue order: the global path's keys, the two largest LDS
  //    classes (the longest kernels), the global path's sort, the third
  //    class, the rest of the global path (many short kernels), then the
  //    other LDS classes, so that neither stream waits for the host's
  //    launches on the other.
  int n_lds = 0;
  for (int c = 0; c < kClasses; ++c) n_lds += host.cls_count[c];
  SideStream *side = nullptr;
  hipStream_t gs = stream;
  if (nnz_big > 0 && n_lds > 0 && !one_stream()) {
    side = dev_state;
    gs = side->s;
  } else {
    side_lock.unlock();
  }
  // 3. Global path, first steps (issued before the LDS path so that the
  //    side stream starts at once): window offsets and sort tiles of many
  //    windows; keys.  Windows averaging at least one sort tile are sorted
  //    each within its own range by column (32-bit keys, a pass per 8 column
  //    bits); many small windows (other block shapes) are sorted as one
  //    range by (window, column) (64-bit keys), where per-window tiles would
  //    be mostly empty.
  if (n_big > kOneGroupWindows) {
    k_big_sizes<<<grid_for(n_big + 1, 256), 256, 0, gs>>>(
        node_pointer, num_nodes, blk_h, wmap, n_big, (int32_t)seg_radix_tile_size(), ws.csb,
        ws.tb);
    DTC_LAUNCH_CHECK("big_sizes");
    st = scan_i32(ws.csb, ws.csb, n_big + 1, false, ws.tmp, gs);
    if (st != DTC_SUCCESS) return st;
    st = scan_i32(ws.tb, ws.tb, n_big + 1, false, ws.tmp, gs);
    if (st != DTC_SUCCESS) return st;
    k_big_tilemap<<<grid_for(n_big, 256), 256, 0, gs>>>(ws.tb, n_big, ws.tw);
    DTC_LAUNCH_CHECK("big_tilemap");
  }

  const bool segmented = n_big > 0 && nnz_big / n_big >= seg_radix_tile_size();
  if (nnz_big > 0 && segmented) {
    k_big_keys<uint32_t><<<grid_for(nnz_big, kBigChunk), 256, 0, gs>>>(
        node_pointer, edge_list, num_nodes, blk_h, colbits, n_big, nnz_big, wmap, ws.csb,
        edge_to_row, ws.k32_0, ws.vals0, ws.wpos);
    DTC_LAUNCH_CHECK("big_keys");
  } else if (nnz_big > 0) {
    k_big_keys<uint64_t><<<grid_for(nnz_big, kBigChunk), 256, 0, gs>>>(
        node_pointer, edge_list, num_nodes, blk_h, colbits, n_big, nnz_big, wmap, ws.csb,
        edge_to_row, ws.keys0, ws.vals0, ws.wpos);
    DTC_LAUNCH_CHECK("big_keys");
  }

  // 4. LDS path (every output of the window sized by its entries, and its
  //    block data for phase 2): the largest class now, the others after the
  //    global path's launches.
  auto launch_class = [&](int c) -> dtc_status_t {
    const int n = host.cls_count[c];
    if (n <= 0) return DTC_SUCCESS;
    switch (c) {
      case 0:
        k_window_tiny<<<grid_for(n, 4), 256, 0, stream>>>(
            node_pointer, edge_list, num_nodes, ws.lists, n, values, edge_to_row,
            edge_to_column, block_partition, tcblocktile_id, tcblock_values, tcblock_edge,
            win_stage(ws), ws.hdr);
        DTC_LAUNCH_CHECK("window_tiny");
        return DTC_SUCCESS;
      case 5: return launch_window_sort<5>(node_pointer, edge_list, num_nodes, n, ws, W, values, edge_to_row, edge_to_column, block_partition, tcblocktile_id, tcblock_values, tcblock_edge, stream);
      case 4: return launch_window_sort<4>(node_pointer, edge_list, num_nodes, n, ws, W, values, edge_to_row, edge_to_column, block_partition, tcblocktile_id, tcblock_values, tcblock_edge, stream);
      case 3: return launch_window_sort<3>(node_pointer, edge_list, num_nodes, n, ws, W, values, edge_to_row, edge_to_column, block_partition, tcblocktile_id, tcblock_values, tcblock_edge, stream);
      case 2: return launch_window_sort<2>(node_pointer, edge_list, num_nodes, n, ws, W, values, edge_to_row, edge_to_column, block_partition, tcblocktile_id, tcblock_values, tcblock_edge, stream);
      default: return launch_window_sort<1>(node_pointer, edge_list, num_nodes, n, ws, W, values, edge_to_row, edge_to_column, block_partition, tcblocktile_id, tcblock_values, tcblock_edge, stream);
    }
  };
  // The two largest classes now (a few hundred µs of work queued on the
  // caller's stream while the host issues the global path), the third after
  // the sort, the rest after the global path.
  bool issued[kClasses] = {};
  auto issue = [&](int c) -> dtc_status_t {
    issued[c] = true;
    return launch_class(c);
  };
  st = issue(kClasses - 1);
  if (st == DTC_SUCCESS) st = issue(kClasses - 2);
  if (st != DTC_SUCCESS) return st;
  // 5. Global path, continued: sort, unique-column numbering;
  //    blockPartition, edgeToColumn, block data, tile ids; then join.
  if (nnz_big > 0) {
    bool in_1 = false;
    BigView bv;
    if (segmented) {
      const int64_t max_tiles = ceil_div(nnz_big, seg_radix_tile_size()) + n_big;
      st = seg_radix_sort_pairs_u32(ws.k32_0, ws.vals0, ws.k32_1, ws.vals1, colbits, ws.tw, ws.tb,
                                    ws.csb, ws.tb + n_big, max_tiles, &in_1, ws.tmp, gs);
      if (st != DTC_SUCCESS) return st;
      const uint32_t *sorted = in_1 ? ws.k32_1 : ws.k32_0;
      st = scan_key_changes_u32(sorted, ws.gscan, nnz_big, ws.tmp, gs);
      if (st != DTC_SUCCESS) return st;
      bv = big_view(sorted, nullptr, colbits, in_1 ? ws.vals1 : ws.vals0, ws.gscan, ws.csb, wmap,
                    ws.wpos);
    } else {
      const int key_bits = colbits + bits_for(n_big - 1);
      st = radix_sort_pairs(ws.keys0, ws.vals0, ws.keys1, ws.vals1, nnz_big, key_bits, &in_1,
                            ws.tmp, gs);
      if (st != DTC_SUCCESS) return st;
      const uint64_t *sorted = in_1 ? ws.keys1 : ws.keys0;
      st = scan_key_changes(sorted, ws.gscan, nnz_big, ws.tmp, gs);
      if (st != DTC_SUCCESS) return st;
      bv = big_view(nullptr, sorted, colbits, in_1 ? ws.vals1 : ws.vals0, ws.gscan, ws.csb, wmap,
                    ws.wpos);
    }
    if (side) {
      st = issue(kClasses - 3);
      if (st != DTC_SUCCESS) return st;
    }
    if (n_big <= kOneGroupWindows) {
      k_big_counts<<<1, 1024, 0, gs>>>(bv, n_big, blk_w, block_partition, ws.wU, ws.cbb);
      DTC_LAUNCH_CHECK("big_counts");
    } else {
      k_big_counts_multi<<<grid_for(n_big + 1, 256), 256, 0, gs>>>(bv, n_big, blk_w,
                                                                       block_partition, ws.wU,
                                                                       ws.cbb);
      DTC_LAUNCH_CHECK("big_counts");
      st = scan_i32(ws.cbb, ws.cbb, n_big + 1, false, ws.tmp, gs);
      if (st != DTC_SUCCESS) return st;
    }
    const unsigned grid = grid_for(nnz_big, 256);
    k_big_stage<<<grid, 256, 0, gs>>>(node_pointer, num_nodes, blk_h, blk_w, nnz_big, bv,
                                          ws.cbb, edge_to_column, win_stage(ws), ws.bstart);
    DTC_LAUNCH_CHECK("big_stage");
    k_big_tile_ids<<<grid, 256, 0, gs>>>(node_pointer, edge_to_row, num_nodes, blk_h, blk_w,
                                             nnz_big, bv, ws.cbb, ws.bstart, values,
                                             tcblocktile_id, tcblock_values, tcblock_edge);
    DTC_LAUNCH_CHECK("big_tile_ids");
  }
  if (side) DTC_HIP_CHECK(hipEventRecord(side->join, gs));
  // 6. The other LDS classes.
  for (int c = kClasses - 1; c >= 0; --c) {
    if (issued[c]) continue;
    st = issue(c);
    if (st != DTC_SUCCESS) return st;
  }
  if (side) {
    DTC_HIP_CHECK(hipStreamWaitEvent(stream, side->join, 0));
    side_lock.unlock();
  }

  // 7. Window offsets (row_window_offset[0] was written by k_check_classify).
  st = scan_i32(block_partition, row_window_offset + 1, W, true, ws.tmp, stream);
  if (st != DTC_SUCCESS) return st;
  k_chunk_windows<<<grid_for(W, 256), 256, 0, stream>>>(row_window_offset, (int32_t)W, ws.cw,
                                                         ws.cw + ws.cw_stride);
  DTC_LAUNCH_CHECK("chunk_windows");

  // 8. Block count (second host synchronisation).
  int32_t T = 0;
  {
    std::lock_guard<std::mutex> g(dev_state->read_mu);
    DTC_HIP_CHECK(hipMemcpyAsync(dev_state->pinned, ws.hdr, kHeaderRead, hipMemcpyDeviceToHost,
                                 stream));
    DTC_HIP_CHECK(hipMemcpyAsync(dev_state->pinned + kHeaderRead, row_window_offset + W,
                                 sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    DTC_HIP_CHECK(hipStreamSynchronize(stream));
    std::memcpy(&host, dev_state->pinned, kHeaderRead);
    std::memcpy(&T, dev_state->pinned + kHeaderRead, sizeof(int32_t));
  }
  if (host.flags & kBadColumn)  // found by the window kernels
    return fail(DTC_ERROR_INVALID_CSR, "invalid CSR: negative column index;");
  if (host.flags & kBadClass)
    return fail(DTC_ERROR_INCONSISTENT, "internal error: a window exceeded its class capacity");
  *block_count = T;
  return DTC_SUCCESS;
}

dtc_status_t dtc_metcf_build_blocks(const int32_t *node_pointer,
                                    const int32_t *row_window_offset, int32_t num_nodes,
                                    int32_t nnz, int32_t blk_h, int32_t blk_w,
                                    int32_t block_count, int32_t *tcblock_rowid,
                                    int32_t *tcblock_offset, int32_t *sparse_atox,
                                    void *workspace, size_t workspace_bytes,
                                    dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  dtc_status_t st = check_shape(num_nodes, nnz, blk_h, blk_w);
  if (st != DTC_SUCCESS) return st;
  if (block_count < 0) return fail(DTC_ERROR_INVALID_VALUE, "block_count must be >= 0");
  if (block_count > nnz)  // every block holds an entry: phase 1 never reports more
    return fail(DTC_ERROR_INVALID_VALUE, "block_count exceeds nnz: pass the count phase 1 returned");
  if (!node_pointer || !row_window_offset || !tcblock_offset)
    return fail(DTC_ERROR_INVALID_VALUE, "node_pointer, row_window_offset and tcblock_offset are required");
  if (block_count > 0 && (!tcblock_rowid || !sparse_atox))
    return fail(DTC_ERROR_INVALID_VALUE, "tcblock_rowid and sparse_atox are required");
  if ((nnz == 0) != (block_count == 0))
    return fail(DTC_ERROR_INVALID_VALUE, "block_count must be 0 exactly when nnz is 0");
  const int64_t W = ceil_div(num_nodes, blk_h);
  MetcfWs ws = carve(workspace, nnz, W);
  if (!workspace || workspace_bytes < ws.bytes)
    return fail(DTC_ERROR_WORKSPACE, "ME-TCF workspace needs " + std::to_string(ws.bytes) + " bytes");
  if (block_count == 0) {
    k_set_i32<<<1, 64, 0, stream>>>(tcblock_offset, nnz);
    DTC_LAUNCH_CHECK("set");
    return DTC_SUCCESS;
  }
  k_blocks_out<<<grid_for(block_count, kOutBlocks), 256, 0, stream>>>(
      node_pointer, num_nodes, (int32_t)W, blk_h, blk_w, block_count, nnz, win_stage(ws),
      row_window_offset, ws.cw, ws.cw + ws.cw_stride, tcblock_offset, tcblock_rowid, sparse_atox);
  DTC_LAUNCH_CHECK("blocks_out");
  return DTC_SUCCESS;
}

}  // extern "C"
