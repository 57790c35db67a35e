/*
 * dtc.h -- C ABI of libdtc, the MI355X (gfx950) DTC-SpMM library.
 *
 * Plain pointers and sizes only.  Every device buffer is owned by the caller;
 * the library allocates no device memory.  Scratch memory is passed in as a
 * workspace whose size comes from the matching *_workspace_size query.  All
 * work is enqueued on the given stream (a hipStream_t; NULL is the default
 * stream).  Functions return a dtc_status_t and never exit the process;
 * dtc_last_error() gives a human-readable detail of the calling thread's
 * last failure.
 *
 * Index types follow the reference: int32 CSR, uint8 tile ids, int32 block
 * offsets and SparseAToX (DTC-SpMM/DTCSpMM.cpp:313-359).  Row counts, N,
 * nnz and block counts are below 2^31; the dense operands B and C are
 * addressed with 64-bit element offsets, so M x N may exceed 2^32 elements.
 */
#ifndef DTC_H_
#define DTC_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same type as hipStream_t. */
typedef struct ihipStream_t *dtc_stream_t;

typedef enum dtc_status {
  DTC_SUCCESS = 0,
  DTC_ERROR_INVALID_VALUE = 1, /* bad argument: null pointer, size, shape, plan name */
  DTC_ERROR_UNSUPPORTED = 2,   /* well-formed but not supported (block shape, precision) */
  DTC_ERROR_WORKSPACE = 3,     /* workspace missing or too small */
  DTC_ERROR_INVALID_CSR = 4,   /* row pointer / column index check failed on the device */
  DTC_ERROR_HIP = 5,           /* HIP runtime error */
  DTC_ERROR_INCONSISTENT = 6   /* blockPartition was not zero for an empty row window */
} dtc_status_t;

typedef enum dtc_precision {
  DTC_PRECISION_FP32 = 0,   /* float32 operands, v_mfma_f32_16x16x4_f32 */
  DTC_PRECISION_BF16 = 1,   /* one bf16 pass, v_mfma_f32_16x16x32_bf16 */
  DTC_PRECISION_FP16 = 2,   /* one fp16 pass, v_mfma_f32_16x16x32_f16 */
  DTC_PRECISION_BF16X3 = 3, /* 2-limb bf16 split: hi*hi + hi*lo + lo*hi */
  DTC_PRECISION_BF16X6 = 4  /* 3-limb bf16 split: all limb products of order <= 2 */
} dtc_precision_t;

typedef enum dtc_b_dtype {
  DTC_B_FLOAT32 = 0, /* B is float32 */
  DTC_B_BFLOAT16 = 1 /* B is bfloat16, taken exactly; A enters as three bf16 limbs
                        (one with DTC_PRECISION_BF16); float32 accumulation and C */
} dtc_b_dtype_t;

typedef struct dtc_spmm_options {
  int32_t precision;    /* dtc_precision_t */
  int32_t vec;          /* B elements per lane per gathered row: 1, 2, 4 (float32 B);
                           4, 8 (bfloat16 B); 0 = float32 B: 2 (balanced) or 4
                           (unbalanced), narrower where N or the alignment need it;
                           bfloat16 B: 8 (balanced; 4 where N or the alignment need
                           it) or 4 (unbalanced) */
  int32_t subtiles;     /* 1 or 2 column tiles of 16*vec per wave (0 = 1; float32 B) */
  int32_t chunk_blocks; /* balanced schedule: TC blocks per work chunk (0 = the largest
                           power of two in [32, 512] at most num_blocks / 8192) */
  int32_t persistent;   /* balanced schedule: 1 = persistent waves pulling chunks from a queue */
  int32_t b_dtype;      /* dtc_b_dtype_t */
  int32_t deterministic; /* balanced schedule: 1 = windows cut across chunks are merged in a
                            fixed order (partial tiles in the workspace, then one ordered sum
                            per window) instead of float32 atomics, so repeated calls give
                            bitwise identical results; needs the larger workspace that
                            dtc_spmm_balanced_workspace_size_ex reports */
  int32_t cut_span_blocks; /* balanced schedule: a window of at most this many TC blocks is
                              never cut (it belongs whole to the chunk holding its first
                              block); a longer one is cut at the multiples of this span and
                              its pieces merged.  A multiple of the chunk; 0 = the chunk */
} dtc_spmm_options_t;

/* ------------------------------------------------------------------ misc */

const char *dtc_status_string(dtc_status_t status);
/* Detail of the calling thread's last non-success status ("" if none). */
const char *dtc_last_error(void);
/* Library version, major * 10000 + minor * 100 + patch. */
int32_t dtc_version(void);

/* --------------------------------------------------- ME-TCF construction
 *
 * Replaces DTCSpMM.preprocess_gpu (DTC-SpMM/DTCSpMM.cpp:313-359) and
 * seg_sort_dequ (DTC-SpMM/DTCSpMM_kernel.cu:349-407).  The reference sizes its
 * block-dependent outputs from a device-to-host read of the block count, so
 * the build is split in the same place: phase 1 fills every output sized by
 * the CSR (windows or edges) and returns the block count, the caller
 * allocates the block-sized outputs, phase 2 fills them.  The workspace
 * carries per-window block data from phase 1 to phase 2 and must not be
 * touched in between.
 *
 * Outputs are bit-identical to the reference's for the same inputs,
 * including empty windows (blockPartition left as passed), a partial last
 * window (num_nodes % blk_h != 0), zero padding of SparseAToX, and tile ids
 * computed as (row % blk_h) * blk_w + column slot, stored as uint8.
 */

/* Workspace bytes for both phases. */
dtc_status_t dtc_metcf_workspace_size(int32_t num_nodes, int32_t nnz, int32_t blk_h,
                                      int32_t blk_w, size_t *bytes);

/*
 * Phase 1.  edge_list: CSR column indices [nnz]; node_pointer: CSR row
 * pointer [num_nodes + 1].  In/out (caller-allocated, as the reference's
 * callers do, scripts/DTCSpMM/run_DTC_SpMM.py:49-58): block_partition
 * [ceil(num_nodes / blk_h)], zero-filled by the caller; edge_to_column [nnz];
 * edge_to_row [nnz].  Out: row_window_offset [ceil(num_nodes / blk_h) + 1];
 * tcblocktile_id [nnz] (uint8, TC-block order); *block_count (host) = number
 * of TC blocks T.  Optional (NULL to skip, build extensions): values [nnz]
 * (CSR order) -> tcblock_values [nnz] (TC-block order; 1.0 when values is
 * NULL); tcblock_edge [nnz] receives the CSR edge index stored at each
 * TC-block position.  Synchronises the stream twice (CSR check, then the
 * block count), as the reference does (DTCSpMM_kernel.cu:378-380, :384-387).
 * Windows of more than 16,384 entries are sorted on a stream the library
 * creates once per device (default priority), forked from and joined back to
 * `stream` by events; builds on one device take turns at it under a lock
 * (environment DTC_METCF_STREAMS=1: every kernel on `stream`).
 */
dtc_status_t dtc_metcf_build_windows(const int32_t *edge_list, const int32_t *node_pointer,
                                     int32_t num_nodes, int32_t nnz, int32_t blk_h,
                                     int32_t blk_w, const float *values,
                                     int32_t *block_partition, int32_t *edge_to_column,
                                     int32_t *edge_to_row, int32_t *row_window_offset,
                                     uint8_t *tcblocktile_id, float *tcblock_values,
                                     int32_t *tcblock_edge, int32_t *block_count,
                                     void *workspace, size_t workspace_bytes,
                                     dtc_stream_t stream);

/*
 * Phase 2 (same workspace, after phase 1; no host synchronisation).  Out:
 * tcblock_rowid [T], tcblock_offset [T + 1], sparse_atox [T * blk_w].
 * block_count must be the T phase 1 returned (a count above nnz is rejected:
 * every block holds an entry).
 */
dtc_status_t dtc_metcf_build_blocks(const int32_t *node_pointer,
                                    const int32_t *row_window_offset, int32_t num_nodes,
                                    int32_t nnz, int32_t blk_h, int32_t blk_w,
                                    int32_t block_count, int32_t *tcblock_rowid,
                                    int32_t *tcblock_offset, int32_t *sparse_atox,
                                    void *workspace, size_t workspace_bytes,
                                    dtc_stream_t stream);

/* ------------------------------------------------------------ SpMM forward
 *
 * C[M x N] = A * B with A in ME-TCF at 16 x 8 blocks, B [b_rows x N] row-major
 * (float32, or bfloat16 with options->b_dtype), C [num_nodes x N] float32
 * row-major.  tcblock_values (TC-block
 * order, [nnz]) may be NULL: every stored element of A is then 1, as the
 * reference's launchers pass (DTCSpMM_kernel.cu:837, :1140).  Columns of A
 * that are >= b_rows read as zero (the dense_bound guard of
 * DTCSpMM_kernel.cu:2735).  Any N >= 1 and any num_nodes are handled; rows of
 * a partial last window are not written past num_nodes.  num_blocks (T) and
 * nnz are the lengths of the block arrays and of tcblocktile_id; block ranges
 * and element offsets read from the format are clamped to them.
 */

/*
 * Unbalanced schedule: one work unit per (row window, column tile).
 * Replaces run_DTCSpMM / spmm_forward_improved_ptx_uint8_cuda
 * (DTCSpMM.cpp:117-131, DTCSpMM_kernel.cu:811-944) and DTCSpMM_gcn
 * (DTCSpMM.cpp:147-161, DTCSpMM_kernel.cu:1033-1116).  No workspace; every
 * output element is written (rows of empty windows get zeros).
 */
dtc_status_t dtc_spmm_windows(const int32_t *row_window_offset, const uint8_t *tcblocktile_id,
                              const int32_t *tcblock_offset, const int32_t *sparse_atox,
                              const float *tcblock_values, int32_t num_windows,
                              int32_t num_blocks, int32_t nnz,
                              int32_t num_nodes, const void *b, int32_t b_rows, int32_t n,
                              float *c, const dtc_spmm_options_t *options,
                              dtc_stream_t stream);

/*
 * Balanced schedule: work chunks of options->chunk_blocks consecutive TC
 * blocks, each split at row-window boundaries; windows that span several
 * chunks are merged with float32 atomic adds into rows this call zeroes
 * first (or, with options->deterministic, summed in chunk order from
 * partial tiles in the workspace).  Replaces run_DTCSpMM_balance /
 * spmm_balance_forward_cuda_ptx_unit8_prefetch (DTCSpMM.cpp:163-178,
 * DTCSpMM_kernel.cu:1118-1199).  Every output element is written.  The
 * workspace (required when num_blocks > 0) holds each chunk's block run,
 * worked out by a preparation kernel, the work counter of the persistent
 * variant and, with options->deterministic, the partial tiles; its size is
 * dtc_spmm_balanced_workspace_size_ex for these options and N (exactly
 * 256 + 8 * ceil(num_blocks / chunk) bytes without the deterministic merge),
 * or dtc_spmm_balanced_workspace_size (256 + 8 * num_blocks bytes, enough for
 * any chunk size without the deterministic merge).
 */
dtc_status_t dtc_spmm_balanced_workspace_size(int32_t num_blocks, size_t *bytes);
/* Workspace bytes for the given options and N (NULL options: the defaults):
 * 256 + 8 * ceil(num_blocks / chunk), plus with options->deterministic two
 * partial-tile slots of 16 x N float32 (N rounded up to 128) per chunk. */
dtc_status_t dtc_spmm_balanced_workspace_size_ex(int32_t num_blocks, int32_t n,
                                                 const dtc_spmm_options_t *options, size_t *bytes);
dtc_status_t dtc_spmm_balanced(const int32_t *tcblock_rowid, const uint8_t *tcblocktile_id,
                               const int32_t *tcblock_offset, const int32_t *sparse_atox,
                               const float *tcblock_values, int32_t num_blocks, int32_t nnz,
                               int32_t num_nodes, const void *b, int32_t b_rows, int32_t n,
                               float *c, const dtc_spmm_options_t *options, void *workspace,
                               size_t workspace_bytes, dtc_stream_t stream);

/*
 * dtc_spmm_balanced without a workspace argument, allocating nothing (no
 * hipMalloc, hipFree or allocating library call).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  The workspace is a fixed 8 MiB device array inside the library,
 * one copy per GPU, whose address is looked up on the first call on each
 * device (the stream must belong to the current device) and which is held
 * until the process ends.  A call whose workspace
 * (dtc_spmm_balanced_workspace_size_ex) exceeds 8 MiB -- about a million
 * chunks, or a deterministic merge whose partial tiles do not fit -- returns
 * DTC_ERROR_UNSUPPORTED before any work.  Every call on one GPU uses the same
 * array: calls that overlap in time on one GPU (on different streams, or from
 * different threads without ordering) race and must not be issued.
 */
dtc_status_t dtc_spmm_balanced_noalloc(const int32_t *tcblock_rowid, const uint8_t *tcblocktile_id,
                                       const int32_t *tcblock_offset, const int32_t *sparse_atox,
                                       const float *tcblock_values, int32_t num_blocks, int32_t nnz,
                                       int32_t num_nodes, const void *b, int32_t b_rows, int32_t n,
                                       float *c, const dtc_spmm_options_t *options,
                                       dtc_stream_t stream);

/*
 * TC-GNN's SpMM, the comparison baseline without TC-block compaction
 * (replaces run_TCGNNSpMM / spmm_forward_cuda, DTCSpMM.cpp:97-115,
 * DTCSpMM_kernel.cu:703-808).  Its format is the phase-1 output of the ME-TCF
 * build: block_partition [ceil(num_nodes / 16)], edge_to_column and
 * edge_to_row [nnz] over the CSR (node_pointer, edge_list), which must be a
 * valid CSR of nnz entries.  values (CSR order) may be NULL: A = 1, as the
 * reference's TC-GNN kernel computes.  Float32 only (options may be NULL; a
 * precision other than FP32 is DTC_ERROR_UNSUPPORTED).  Every output element
 * of the first num_nodes rows is written.
 */
dtc_status_t dtc_spmm_tcgnn(const int32_t *node_pointer, const int32_t *edge_list,
                            const int32_t *block_partition, const int32_t *edge_to_column,
                            const int32_t *edge_to_row, const float *values, int32_t num_nodes,
                            int32_t nnz, const float *b, int32_t b_rows, int32_t n, float *c,
                            const dtc_spmm_options_t *options, dtc_stream_t stream);

/* ---------------------------------------------------------------- backward
 *
 * The reference binds its "backward" names to the forward functions
 * (DTCSpMM.cpp:384-399), i.e. it returns A * dY.  These entry points give the
 * true gradients of C = A * B: dB = A^T * dC through the ME-TCF of A^T, and
 * dA = SDDMM(dC, B) on A's pattern.
 */

/* Workspace bytes for dtc_csr_transpose. */
dtc_status_t dtc_csr_transpose_workspace_size(int32_t num_rows, int32_t num_cols, int32_t nnz,
                                              size_t *bytes);
/* CSR [num_rows x num_cols] -> CSR of the transpose [num_cols x num_rows],
 * columns ascending per row; perm[p] = source edge of transposed edge p. */
dtc_status_t dtc_csr_transpose(const int32_t *node_pointer, const int32_t *edge_list,
                               int32_t num_rows, int32_t num_cols, int32_t nnz,
                               int32_t *node_pointer_t, int32_t *edge_list_t, int32_t *perm,
                               void *workspace, size_t workspace_bytes, dtc_stream_t stream);

/* out[e] = sum_j x[row(e), j] * y[col(e), j] for every CSR edge e
 * (x: float32 [num_rows x n], y: [y_rows x n] float32 or bfloat16 per y_dtype
 * (a dtc_b_dtype_t; bf16 taken exactly); columns >= y_rows give 0).
 * edge_to_row (the edgeToRow output of the ME-TCF build) may be NULL: rows
 * are then found by binary search of node_pointer. */
dtc_status_t dtc_sddmm_csr(const int32_t *node_pointer, const int32_t *edge_list,
                           const int32_t *edge_to_row, int32_t num_rows, int32_t nnz,
                           const float *x, const void *y, int32_t y_dtype, int32_t y_rows,
                           int32_t n, float *out, dtc_stream_t stream);

/* out[p] = in[idx[p]] for p < count (moves values between CSR and TC-block order). */
dtc_status_t dtc_gather_f32(const float *in, const int32_t *idx, int32_t count, float *out,
                            dtc_stream_t stream);

/* --------------------------------------------------------- variant selector
 *
 * Training and inference kernels of the MLP that picks the SpMM variant per
 * matrix (the reference's counterpart is the hand-written ExecutionPlan table,
 * scripts/DTCSpMM/run_DTC_SpMM.py:10-30).  Row-major float32 throughout;
 * weights are [in_dim x out_dim] and y = act(x W + bias).
 */

/* y = x W + bias (bias may be NULL); relu != 0 applies max(0, .). */
dtc_status_t dtc_mlp_linear_forward(const float *x, const float *w, const float *bias, float *y,
                                    int32_t batch, int32_t in_dim, int32_t out_dim, int32_t relu,
                                    dtc_stream_t stream);
/* Gradients of a layer from dy and its output y (y gives the ReLU mask):
 * dw = x^T dz, db = sum_b dz, dx = dz W^T (dx may be NULL), dz = dy * act'(y). */
dtc_status_t dtc_mlp_linear_backward(const float *x, const float *w, const float *y,
                                     const float *dy, float *dx, float *dw, float *db,
                                     int32_t batch, int32_t in_dim, int32_t out_dim,
                                     int32_t relu, dtc_stream_t stream);
/* *loss = inv_count * sum mask (pred - target)^2 (device scalar),
 * dpred = 2 inv_count mask (pred - target). */
dtc_status_t dtc_mlp_mse(const float *pred, const float *target, const float *mask, int64_t n,
                         float inv_count, float *dpred, float *loss, dtc_stream_t stream);
/* One Adam update of param[n] (bias-corrected, step >= 1). */
dtc_status_t dtc_adam_step(float *param, const float *grad, float *m, float *v, int64_t n,
                           float lr, float beta1, float beta2, float eps, int32_t step,
                           dtc_stream_t stream);

/* ------------------------------------------------------------ row reordering
 *
 * Replaces the offline TCA reordering script, reordering/TCA_reorder.py
 * (MinHash/LSH candidates :55-80, Jaccard :122-146, greedy clustering
 * :152-187, cluster-of-clusters :203-286, output order :300-315).  Rows (and,
 * at the second level, clusters) are sets of column indices given as CSR with
 * strictly increasing columns per row.  Signature entry j of a set is
 * min over its columns c of fmix32(c ^ s_j), s_j = fmix32(j * 0x9E3779B9 +
 * seed) (fmix32: MurmurHash3's 32-bit finaliser).
 */
typedef struct dtc_tca_options {
  int32_t thres;         /* rows per cluster at which it stops growing (16) */
  int32_t cluster_thres; /* clusters per super-cluster likewise (128) */
  int32_t bands;         /* LSH bands (28 for Jaccard threshold 0.2 at 128 perms) */
  int32_t rows_per_band; /* signature entries per band (2) */
  int32_t max_neighbors; /* pair rows at most this far apart in a band's order */
  uint32_t seed;         /* hash family seed */
  int64_t max_pairs;     /* candidate-pair capacity before deduplication */
} dtc_tca_options_t;

typedef struct dtc_tca_stats {
  int64_t row_pairs;     /* unique candidate row pairs */
  int64_t cluster_pairs; /* unique candidate cluster pairs */
  int32_t clusters;
  int32_t super_clusters;
} dtc_tca_stats_t;

/* signatures[r * num_hashes + j], uint32; empty rows get 0xFFFFFFFF
 * (replaces MinHash / libMHCUDA.minhash_cuda_calc, TCA_reorder.py:55-80). */
dtc_status_t dtc_minhash_signatures(const int32_t *node_pointer, const int32_t *edge_list,
                                    int32_t num_rows, int32_t num_hashes, uint32_t seed,
                                    uint32_t *signatures, dtc_stream_t stream);
/* Candidate pairs (u < v, packed u << 32 | v, sorted, unique): some band of
 * rows_per_band entries agrees and u, v lie at most max_neighbors apart in
 * that band's stable order; empty sets (set_ptr) never pair (replaces
 * MinHashLSH.insert/query, TCA_reorder.py:57, :125). */
dtc_status_t dtc_lsh_pairs_workspace_size(int32_t num_sets, int64_t max_pairs, size_t *bytes);
dtc_status_t dtc_lsh_pairs(const uint32_t *signatures, const int32_t *set_ptr, int32_t num_sets,
                           int32_t num_hashes, int32_t bands, int32_t rows_per_band,
                           int32_t max_neighbors, uint64_t *pairs, int64_t max_pairs,
                           int64_t *num_pairs, void *workspace, size_t workspace_bytes,
                           dtc_stream_t stream);
/* similarity[p] = |S_u ∩ S_v| / |S_u ∪ S_v| in float32 (replaces
 * cugraph.jaccard, TCA_reorder.py:136). */
dtc_status_t dtc_jaccard_pairs(const int32_t *set_ptr, const int32_t *set_idx,
                               const uint64_t *pairs, int64_t num_pairs, float *similarity,
                               dtc_stream_t stream);
/* Host memory: the greedy merge of TCA_reorder.py:152-187 over num_items
 * items: most similar pair popped first (equal similarities in increasing
 * (p1, p2) order), stale pairs re-queued with the Jaccard of the two roots'
 * own lists (list_ptr/list_idx, sorted rows); pairs (a, a) are skipped and of
 * repeated pairs the first given is kept.  root[i] is item i's cluster root. */
dtc_status_t dtc_tca_greedy(int32_t num_items, const uint64_t *pairs, const float *similarity,
                            int64_t num_pairs, const int32_t *list_ptr, const int32_t *list_idx,
                            int32_t thres, int32_t *root);
/* The whole two-level reordering of a device CSR; reorder_id (host,
 * int32[num_rows]) lists the original row ids in their new order, the content
 * of the reference's <dataset>.reorder_id.npz (TCA_reorder.py:317). */
dtc_status_t dtc_tca_reorder_workspace_size(int32_t num_rows, int32_t nnz,
                                            const dtc_tca_options_t *options, size_t *bytes);
dtc_status_t dtc_tca_reorder(const int32_t *node_pointer, const int32_t *edge_list,
                             int32_t num_rows, int32_t nnz, const dtc_tca_options_t *options,
                             int32_t *reorder_id, dtc_tca_stats_t *stats, void *workspace,
                             size_t workspace_bytes, dtc_stream_t stream);

/* ------------------------------------------------------- GCN feature transform
 *
 * The dense half of a GCN layer H = A (act(X) W): the reference's GCN forward
 * (DTCSpMM_gcn, DTC-SpMM/DTCSpMM.cpp:147-161 over
 * DTC-SpMM/DTCSpMM_kernel.cu:1033-1116) takes X W as a float32 tensor its
 * caller computed with a torch GEMM.  Here B = act(X) W is one MFMA kernel
 * that applies act (identity or ReLU) and the cast to bf16 while it loads X,
 * accumulates in float32, and writes B as row-major bf16 [M x F] -- the
 * operand of the SpMM with b_dtype DTC_B_BFLOAT16.  W (bf16, [K x F]
 * row-major) is first packed into the kernel's fragment order (K * F bf16,
 * 16-byte aligned).  K must be a multiple of 32 and F of 64; X, the packed W
 * and the output 16-byte aligned.
 */
typedef enum dtc_gcn_x_dtype {
  DTC_GCN_X_FLOAT32 = 0, /* X is float32: rounded to bf16 (after act) as it is loaded */
  DTC_GCN_X_BFLOAT16 = 1 /* X is bf16 */
} dtc_gcn_x_dtype_t;

dtc_status_t dtc_gcn_pack_weights(const void *w, int32_t k, int32_t f, void *w_packed,
                                  dtc_stream_t stream);
/* out[i][j] = bf16( sum_k act(X[i][k]) W[k][j] ), act = ReLU when relu = 1. */
dtc_status_t dtc_gcn_transform(const void *x, int32_t x_dtype, int32_t relu, const void *w_packed,
                               int32_t m, int32_t k, int32_t f, void *out, dtc_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* DTC_H_ */
