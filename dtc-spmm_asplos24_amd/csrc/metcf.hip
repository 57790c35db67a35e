// ME-TCF construction on gfx950.
//
// The reference builds the format with a thrust sort of (window, column)
// pairs followed by two kernels that run ONE THREAD PER ROW WINDOW
// (generate_edgetocolumn, DTCSpMM_kernel.cu:205-249, and
// generate_tcoffset_id_atob, :252-335).  Here every step is data-parallel
// over edges:
//
//   phase 1  check the CSR, then key every edge by (window, column) and sort
//            the keys stably with the radix sort (value = CSR edge index);
//            the inclusive scan G of "key differs from its predecessor"
//            gives every sorted element its global unique-column number, so
//            the rank of a column inside its window is G - G(window start),
//            |S_w| = G(window end) - G(window start), blockPartition and the
//            window offsets follow with one more scan;
//   phase 2  one thread per sorted element writes SparseAToX, the block
//            offsets (the first element of a block sits exactly at the
//            block's offset in the sorted order), the block row ids and
//            finally the tile ids at offset + (number of elements of the same
//            block with a smaller CSR index), which reproduces the
//            reference's CSR-order placement inside a block (:291-300).
//
// Sorted positions of window w are exactly its CSR range
// [rowptr[w*blk_h], rowptr[min(w*blk_h+blk_h, M)]) because the sort key
// starts with the window.
#include <algorithm>
#include <string>

#include "dtc_internal.h"

namespace dtc {
namespace {

struct MetcfHeader {
  int32_t flags;        // CSR check bits, see kBad*
  int32_t max_col;      // largest column index
  int32_t block_count;  // sum of blocks over non-empty windows
  int32_t rwo_last;     // row_window_offset[W]
  int32_t colbits;      // bits used for the column in the sort key
  int32_t sorted_in_1;  // sorted pairs live in buffer 1
  int32_t pad[58];
};
static_assert(sizeof(MetcfHeader) == 256, "header size");

constexpr int kBadFirst = 1;     // rowptr[0] != 0
constexpr int kBadLast = 2;      // rowptr[M] != nnz
constexpr int kBadOrder = 4;     // rowptr decreases
constexpr int kBadColumn = 8;    // negative column index

struct MetcfWs {
  MetcfHeader *hdr;
  uint64_t *keys0, *keys1;
  int32_t *vals0, *vals1;
  int32_t *gscan;    // inclusive unique-column count per sorted element
  void *tmp;         // scan / sort scratch
  size_t bytes;
};

MetcfWs carve(void *ws, int64_t nnz, int64_t W) {
  Carver c(ws);
  MetcfWs w;
  w.hdr = c.take<MetcfHeader>(1);
  w.keys0 = c.take<uint64_t>((size_t)std::max<int64_t>(nnz, 1));
  w.keys1 = c.take<uint64_t>((size_t)std::max<int64_t>(nnz, 1));
  w.vals0 = c.take<int32_t>((size_t)std::max<int64_t>(nnz, 1));
  w.vals1 = c.take<int32_t>((size_t)std::max<int64_t>(nnz, 1));
  w.gscan = c.take<int32_t>((size_t)std::max<int64_t>(nnz, 1));
  const size_t tmp = std::max({radix_sort_workspace_bytes(nnz), scan_workspace_bytes(nnz),
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

__global__ void k_header_init(MetcfHeader *h) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    h->flags = 0;
    h->max_col = 0;
    h->block_count = 0;
    h->rwo_last = 0;
    h->colbits = 0;
    h->sorted_in_1 = 0;
  }
}

__global__ void k_header_set(MetcfHeader *h, int colbits, int sorted_in_1) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    h->colbits = colbits;
    h->sorted_in_1 = sorted_in_1;
  }
}

// Checks rowptr[0] == 0, rowptr[M] == nnz, rowptr non-decreasing, columns
// >= 0, and finds the largest column.
__global__ __launch_bounds__(256) void k_validate(const int32_t *rowptr, const int32_t *col,
                                                  int32_t M, int32_t nnz, MetcfHeader *h) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int flags = 0;
  int mx = 0;
  for (int64_t r = tid; r < M; r += stride)
    if (rowptr[r] > rowptr[r + 1]) flags |= kBadOrder;
  for (int64_t e = tid; e < nnz; e += stride) {
    const int c = col[e];
    if (c < 0) flags |= kBadColumn;
    mx = max(mx, c);
  }
  if (tid == 0) {
    if (rowptr[0] != 0) flags |= kBadFirst;
    if (rowptr[M] != nnz) flags |= kBadLast;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    flags |= __shfl_xor(flags, d, 64);
    mx = max(mx, __shfl_xor(mx, d, 64));
  }
  // one atomic pair per workgroup
  __shared__ int s_f[4], s_m[4];
  if ((threadIdx.x & 63) == 0) {
    s_f[threadIdx.x >> 6] = flags;
    s_m[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int bf = s_f[0] | s_f[1] | s_f[2] | s_f[3];
    if (bf) atomicOr(&h->flags, bf);
    atomicMax(&h->max_col, max(max(s_m[0], s_m[1]), max(s_m[2], s_m[3])));
  }
}

// edgeToRow (DTCSpMM_kernel.cu:107-120) and the sort keys, edge-parallel:
// wavefront w takes the edges [nnz*w/NW, nnz*(w+1)/NW) and walks them 64 at a
// time.  Lane j holds rowptr[r + 1 + j], the end of row r + j, for the 64 rows
// from the row r containing the next edge; every edge of a step lies before
// the last of those ends, and its row is r + (the first j whose end exceeds
// it), found by a 6-step binary search over the lanes.  A hub row is spread
// over many wavefronts and short rows share one step, so no wavefront walks a
// long row serially.  The step's columns are loaded with its row ends.
__global__ __launch_bounds__(256) void k_rows_keys(const int32_t *rowptr, const int32_t *col,
                                                   int32_t M, int32_t nnz, int32_t blk_h,
                                                   int colbits, int32_t *e2r, uint64_t *keys,
                                                   int32_t *vals) {
  for_each_edge_row_load(
      rowptr, M, nnz, [&](int e) { return col[e]; },
      [&](int idx, int row, int c) {
        e2r[idx] = row;
        keys[idx] = ((uint64_t)(row / blk_h) << colbits) | (uint32_t)c;
        vals[idx] = idx;
      });
}

__device__ __forceinline__ int window_begin(const int32_t *rowptr, int w, int blk_h, int M) {
  return rowptr[min((int64_t)w * blk_h, (int64_t)M)];
}
__device__ __forceinline__ int window_end(const int32_t *rowptr, int w, int blk_h, int M) {
  return rowptr[min((int64_t)w * blk_h + blk_h, (int64_t)M)];
}

// blockPartition[w] = ceil(|S_w| / blk_w) for non-empty windows
// (DTCSpMM_kernel.cu:215-223); the block count sums them.
__global__ __launch_bounds__(256) void k_window_counts(const int32_t *rowptr, int32_t M,
                                                       int32_t W, int32_t blk_h,
                                                       int32_t blk_w, const int32_t *g,
                                                       int32_t *bp, int32_t *rwo,
                                                       MetcfHeader *h) {
  __shared__ int smem[4];
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  int nb = 0;
  if (w < W) {
    const int s = window_begin(rowptr, w, blk_h, M);
    const int e = window_end(rowptr, w, blk_h, M);
    if (e > s) {
      const int base = s > 0 ? g[s - 1] : 0;
      const int U = g[e - 1] - base;
      nb = (U - 1 + blk_w) / blk_w;
      bp[w] = nb;
    }
  }
  int total;
  block256_inclusive_sum(nb, smem, total);
  if (threadIdx.x == 0 && total) atomicAdd(&h->block_count, total);
  if (w == 0) rwo[0] = 0;
}

__global__ void k_copy_last(const int32_t *rwo, int32_t W, MetcfHeader *h) {
  if (threadIdx.x == 0 && blockIdx.x == 0) h->rwo_last = rwo[W];
}

// edgeToColumn[e] = rank of col(e) in its window (DTCSpMM_kernel.cu:224-227).
__global__ __launch_bounds__(256) void k_edge_to_column(const int32_t *rowptr, int32_t M,
                                                        int32_t nnz, int32_t blk_h,
                                                        const MetcfHeader *h,
                                                        const uint64_t *keys0,
                                                        const uint64_t *keys1,
                                                        const int32_t *vals0,
                                                        const int32_t *vals1,
                                                        const int32_t *g, int32_t *e2c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const int colbits = h->colbits;
  const uint64_t *keys = h->sorted_in_1 ? keys1 : keys0;
  const int32_t *vals = h->sorted_in_1 ? vals1 : vals0;
  const int w = (int)(keys[i] >> colbits);
  const int s = window_begin(rowptr, w, blk_h, M);
  const int base = s > 0 ? g[s - 1] : 0;
  e2c[vals[i]] = g[i] - base - 1;
}

// Phase 2a: SparseAToX, block offsets, block row ids, unique positions.
__global__ __launch_bounds__(256) void k_fill_blocks(
    const int32_t *rowptr, const int32_t *rwo, int32_t M, int32_t nnz, int32_t blk_h,
    int32_t blk_w, int32_t T, const MetcfHeader *h, const uint64_t *keys0,
    const uint64_t *keys1, const int32_t *g, int32_t *off, int32_t *rowid, int32_t *atox) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const int colbits = h->colbits;
  const uint64_t *keys = h->sorted_in_1 ? keys1 : keys0;
  const uint64_t key = keys[i];
  const int gi = g[i];
  const bool is_new = (i == 0) || (g[i - 1] != gi);
  if (i == 0) off[T] = nnz;
  if (!is_new) return;
  const int w = (int)(key >> colbits);
  const int colv = (int)(key & ((1ull << colbits) - 1));
  const int s = window_begin(rowptr, w, blk_h, M);
  const int e = window_end(rowptr, w, blk_h, M);
  const int base = s > 0 ? g[s - 1] : 0;
  const int r = gi - base - 1;
  const int64_t b = (int64_t)rwo[w] + r / blk_w;
  const int c = r % blk_w;
  atox[b * blk_w + c] = colv;
  if (c == 0) {
    off[b] = i;
    rowid[b] = w;
  }
  const int U = g[e - 1] - base;
  if (r == U - 1)
    for (int cc = c + 1; cc < blk_w; ++cc) atox[b * blk_w + cc] = 0;
}

// Phase 2b: tile ids (and the optional reordered values / edge ids).  Within
// a block the reference stores elements in CSR order; here an element's slot
// is its block offset plus the number of elements of the same block with a
// smaller CSR index.  A block's elements are one contiguous range of the
// sorted order, [off[b], off[b+1]) (its columns are consecutive condensed
// columns of one window), so that count is one pass over the range: at most
// blk_h * blk_w = 128 loads that the lanes of one block share (the same
// addresses), from lines the neighbouring lanes have just brought in.
__global__ __launch_bounds__(256) void k_tile_ids(
    const int32_t *rowptr, const int32_t *rwo, const int32_t *e2r, int32_t M, int32_t nnz,
    int32_t blk_h, int32_t blk_w, const MetcfHeader *h, const uint64_t *keys0,
    const uint64_t *keys1, const int32_t *vals0, const int32_t *vals1, const int32_t *g,
    const int32_t *off, const float *values, uint8_t *tile, float *tcvalues, int32_t *tcedge) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnz) return;
  const int colbits = h->colbits;
  const uint64_t *keys = h->sorted_in_1 ? keys1 : keys0;
  const int32_t *vals = h->sorted_in_1 ? vals1 : vals0;
  const int w = (int)(keys[i] >> colbits);
  const int s = window_begin(rowptr, w, blk_h, M);
  const int base = s > 0 ? g[s - 1] : 0;
  const int r = g[i] - base - 1;
  const int kb = r / blk_w;
  const int c = r % blk_w;
  const int64_t b = (int64_t)rwo[w] + kb;
  const int myv = vals[i];
  const int lo = off[b], hi = off[b + 1];
  int pos = lo;
#pragma unroll 8
  for (int j = lo; j < hi; ++j) pos += vals[j] < myv;
  if (pos >= hi) return;  // unreachable for a consistent build; keeps writes in bounds
  const unsigned row_local = (unsigned)e2r[myv] % (unsigned)blk_h;
  tile[pos] = (uint8_t)(row_local * (unsigned)blk_w + (unsigned)c);
  if (tcvalues) tcvalues[pos] = values ? values[myv] : 1.0f;
  if (tcedge) tcedge[pos] = myv;
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
                                     int32_t blk_w, int32_t *block_partition,
                                     int32_t *edge_to_column, int32_t *edge_to_row,
                                     int32_t *row_window_offset, int32_t *block_count,
                                     void *workspace, size_t workspace_bytes,
                                     dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  dtc_status_t st = check_shape(num_nodes, nnz, blk_h, blk_w);
  if (st != DTC_SUCCESS) return st;
  if (!node_pointer || !row_window_offset || !block_count)
    return fail(DTC_ERROR_INVALID_VALUE, "node_pointer, row_window_offset and block_count are required");
  if (nnz > 0 && (!edge_list || !edge_to_column || !edge_to_row))
    return fail(DTC_ERROR_INVALID_VALUE, "edge arrays are required when nnz > 0");
  const int64_t W = ceil_div(num_nodes, blk_h);
  if (W > 0 && !block_partition) return fail(DTC_ERROR_INVALID_VALUE, "block_partition is NULL");
  MetcfWs ws = carve(workspace, nnz, W);
  if (!workspace || workspace_bytes < ws.bytes)
    return fail(DTC_ERROR_WORKSPACE, "ME-TCF workspace needs " + std::to_string(ws.bytes) + " bytes");

  // 1. CSR check and largest column (first host synchronisation).
  k_header_init<<<1, 64, 0, stream>>>(ws.hdr);
  DTC_LAUNCH_CHECK("header_init");
  {
    const int64_t work = std::max<int64_t>(num_nodes, nnz);
    const unsigned grid = (unsigned)std::min<int64_t>(4096, grid_for(work, 256));
    k_validate<<<grid, 256, 0, stream>>>(node_pointer, edge_list, num_nodes, nnz, ws.hdr);
    DTC_LAUNCH_CHECK("validate");
  }
  MetcfHeader host{};
  DTC_HIP_CHECK(hipMemcpyAsync(&host, ws.hdr, 16, hipMemcpyDeviceToHost, stream));
  DTC_HIP_CHECK(hipStreamSynchronize(stream));
  if (host.flags) {
    std::string why;
    if (host.flags & kBadFirst) why += " nodePointer[0] != 0;";
    if (host.flags & kBadLast) why += " nodePointer[num_nodes] != number of edges;";
    if (host.flags & kBadOrder) why += " nodePointer decreases;";
    if (host.flags & kBadColumn) why += " negative column index;";
    return fail(DTC_ERROR_INVALID_CSR, "invalid CSR:" + why);
  }
  if (W == 0) {
    k_set_i32<<<1, 64, 0, stream>>>(row_window_offset, 0);
    DTC_LAUNCH_CHECK("set");
    *block_count = 0;
    return DTC_SUCCESS;
  }
  const int colbits = std::max(1, bits_for(host.max_col));
  const int wbits = bits_for(W - 1);
  const int key_bits = colbits + wbits;

  bool in_1 = false;
  if (nnz > 0) {
    // 2. Keys (window, column) with the CSR edge index as value; stable sort.
    const unsigned grid = (unsigned)std::min<int64_t>(65536, grid_for(nnz, 4096));
    k_rows_keys<<<grid, 256, 0, stream>>>(node_pointer, edge_list, num_nodes, nnz, blk_h, colbits,
                                          edge_to_row, ws.keys0, ws.vals0);
    DTC_LAUNCH_CHECK("rows_keys");
    st = radix_sort_pairs(ws.keys0, ws.vals0, ws.keys1, ws.vals1, nnz, key_bits, &in_1, ws.tmp,
                          stream);
    if (st != DTC_SUCCESS) return st;
    // 3. Unique-column numbering of the sorted keys.
    st = scan_key_changes(in_1 ? ws.keys1 : ws.keys0, ws.gscan, nnz, ws.tmp, stream);
    if (st != DTC_SUCCESS) return st;
  }
  k_header_set<<<1, 64, 0, stream>>>(ws.hdr, colbits, in_1 ? 1 : 0);
  DTC_LAUNCH_CHECK("header_set");
  // 4. Blocks per window, window offsets.
  k_window_counts<<<grid_for(W, 256), 256, 0, stream>>>(node_pointer, num_nodes, (int32_t)W, blk_h,
                                                         blk_w, ws.gscan, block_partition,
                                                         row_window_offset, ws.hdr);
  DTC_LAUNCH_CHECK("window_counts");
  st = scan_i32(block_partition, row_window_offset + 1, W, true, ws.tmp, stream);
  if (st != DTC_SUCCESS) return st;
  k_copy_last<<<1, 64, 0, stream>>>(row_window_offset, (int32_t)W, ws.hdr);
  DTC_LAUNCH_CHECK("copy_last");
  // 5. edgeToColumn.
  if (nnz > 0) {
    k_edge_to_column<<<grid_for(nnz, 256), 256, 0, stream>>>(
        node_pointer, num_nodes, nnz, blk_h, ws.hdr, ws.keys0, ws.keys1, ws.vals0, ws.vals1,
        ws.gscan, edge_to_column);
    DTC_LAUNCH_CHECK("edge_to_column");
  }
  // 6. Block count (second host synchronisation).
  DTC_HIP_CHECK(hipMemcpyAsync(&host, ws.hdr, 16, hipMemcpyDeviceToHost, stream));
  DTC_HIP_CHECK(hipStreamSynchronize(stream));
  if (host.block_count != host.rwo_last)
    return fail(DTC_ERROR_INCONSISTENT,
                "blockPartition must be zero-filled: row-window offsets end at " +
                    std::to_string(host.rwo_last) + " but the windows hold " +
                    std::to_string(host.block_count) + " blocks");
  *block_count = host.block_count;
  return DTC_SUCCESS;
}

dtc_status_t dtc_metcf_build_blocks(const int32_t *node_pointer, const int32_t *edge_to_row,
                                    const int32_t *row_window_offset, int32_t num_nodes,
                                    int32_t nnz, int32_t blk_h, int32_t blk_w,
                                    int32_t block_count, const float *values,
                                    int32_t *tcblock_rowid, uint8_t *tcblocktile_id,
                                    int32_t *tcblock_offset, int32_t *sparse_atox,
                                    float *tcblock_values, int32_t *tcblock_edge,
                                    void *workspace, size_t workspace_bytes,
                                    dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  dtc_status_t st = check_shape(num_nodes, nnz, blk_h, blk_w);
  if (st != DTC_SUCCESS) return st;
  if (block_count < 0) return fail(DTC_ERROR_INVALID_VALUE, "block_count must be >= 0");
  if (!node_pointer || !row_window_offset || !tcblock_offset)
    return fail(DTC_ERROR_INVALID_VALUE, "node_pointer, row_window_offset and tcblock_offset are required");
  if (nnz > 0 && (!edge_to_row || !tcblocktile_id))
    return fail(DTC_ERROR_INVALID_VALUE, "edge_to_row and tcblocktile_id are required when nnz > 0");
  if (block_count > 0 && (!tcblock_rowid || !sparse_atox))
    return fail(DTC_ERROR_INVALID_VALUE, "tcblock_rowid and sparse_atox are required");
  if ((nnz == 0) != (block_count == 0))
    return fail(DTC_ERROR_INVALID_VALUE, "block_count must be 0 exactly when nnz is 0");
  const int64_t W = ceil_div(num_nodes, blk_h);
  MetcfWs ws = carve(workspace, nnz, W);
  if (!workspace || workspace_bytes < ws.bytes)
    return fail(DTC_ERROR_WORKSPACE, "ME-TCF workspace needs " + std::to_string(ws.bytes) + " bytes");
  if (nnz == 0) {
    k_set_i32<<<1, 64, 0, stream>>>(tcblock_offset, 0);
    DTC_LAUNCH_CHECK("set");
    return DTC_SUCCESS;
  }
  const unsigned grid = grid_for(nnz, 256);
  k_fill_blocks<<<grid, 256, 0, stream>>>(node_pointer, row_window_offset, num_nodes, nnz, blk_h,
                                          blk_w, block_count, ws.hdr, ws.keys0, ws.keys1,
                                          ws.gscan, tcblock_offset, tcblock_rowid, sparse_atox);
  DTC_LAUNCH_CHECK("fill_blocks");
  k_tile_ids<<<grid, 256, 0, stream>>>(node_pointer, row_window_offset, edge_to_row, num_nodes,
                                       nnz, blk_h, blk_w, ws.hdr, ws.keys0, ws.keys1, ws.vals0,
                                       ws.vals1, ws.gscan, tcblock_offset, values,
                                       tcblocktile_id, tcblock_values, tcblock_edge);
  DTC_LAUNCH_CHECK("tile_ids");
  return DTC_SUCCESS;
}

}  // extern "C"
