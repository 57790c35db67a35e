/*
 * metcf_oracle.c -- CPU oracle for the DTC-SpMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product path (libdtc.so + the DTCSpMM extension) never links,
 * loads or falls back to it.
 *
 * It restates, in plain C, what the reference computes:
 *   - ME-TCF construction, `preprocess_gpu` (DTC-SpMM/DTCSpMM.cpp:313-359)
 *     and `seg_sort_dequ` with its kernels (DTC-SpMM/DTCSpMM_kernel.cu:107-419);
 *   - the SpMM semantics of the ME-TCF kernels (DTCSpMM_kernel.cu:2663-2866
 *     unbalanced, :4730-4901 balanced): C = A * B with A given by the TC
 *     blocks, in float64 (the parity reference) and float32 (the CPU
 *     baseline "port");
 *   - the CSR product in float64 and the SDDMM / transpose used by the
 *     backward pass.
 *
 * Parity pins (each function is checked by tests/test_oracle.py against):
 *   metcf_windows / metcf_blocks : the hand-derived 18x18 fixture in
 *       tests/golden/metcf_fixture_18.json; a brute-force set-based
 *       restatement on random tiny matrices (M <= 64, empty windows,
 *       M % 16 != 0, unsorted rows, several block shapes); the ME-TCF -> CSR
 *       round trip; the structural invariants (offsets end at nnz, window
 *       offsets end at block_count, ascending SparseAToX, zero padding).
 *   spmm_csr_f64 : scipy.sparse float64 product and the degree known answer
 *       (all-ones A and B give C[r,:] = deg(r), run_DTC_SpMM.py:68).
 *   spmm_metcf_f64 / spmm_metcf_f32 : equality with spmm_csr_f64 on the
 *       same matrix (format semantics), within float32 rounding for f32.
 *   sddmm_csr_f64 : numpy dense (X @ Y^T) sampled on the pattern.
 *   csr_transpose : scipy .T.tocsr() (sorted indices), and A^T^T == A.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_OK 0
#define ORACLE_ERR_ARG 1
#define ORACLE_ERR_ALLOC 2
#define ORACLE_ERR_INCONSISTENT 3

static int imin(int a, int b) { return a < b ? a : b; }

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* First index of `target` in the ascending, duplicate-free array arr[0..n).
 * Follows the device binary search of DTCSpMM_kernel.cu:176-193 (which
 * walks back to the first equal element; with unique keys that is the
 * match itself).  Returns -1 when absent. */
static int find_rank(const int *arr, int n, int target) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = lo + (hi - lo) / 2;
    if (arr[mid] == target) {
      while (mid > 0 && arr[mid - 1] == target) mid--;
      return mid;
    }
    if (arr[mid] < target) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

/*
 * Phase 1 of ME-TCF: row ids, per-window column condensing, block counts
 * and the row-window offsets.
 *
 *   edgeToRow[e]       : row of edge e              (kernel.cu:107-120)
 *   window w            : rows [w*blk_h, min(w*blk_h+blk_h, M))  (:122-134)
 *   S_w                 : sorted unique columns of window w       (:354-362,
 *                         the (segment, column) sort, then the in-place
 *                         dedup of :194-204 / :217-220)
 *   blockPartition[w]   : ceil(|S_w| / blk_w); written only for non-empty
 *                         windows, empty ones keep the caller's value (:215-223)
 *   edgeToColumn[e]     : rank of col(e) in S_w  (:224-227)
 *   block_count         : sum over non-empty windows (atomicAdd at :222)
 *   rowWindowOffset     : [0, inclusive scan of blockPartition]  (:366-373)
 */
int oracle_metcf_windows(int32_t M, int32_t nnz, const int32_t *rowptr,
                         const int32_t *col, int32_t blk_h, int32_t blk_w,
                         int32_t *blockPartition, int32_t *edgeToColumn,
                         int32_t *edgeToRow, int32_t *rowWindowOffset,
                         int64_t *block_count_out, int nthreads) {
  if (M < 0 || nnz < 0 || blk_h <= 0 || blk_w <= 0) return ORACLE_ERR_ARG;
  set_threads(nthreads);
  const int W = (M + blk_h - 1) / blk_h;

  for (int r = 0; r < M; ++r)
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) edgeToRow[e] = r;

  int64_t total = 0;
  int err = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (int w = 0; w < W; ++w) {
    const int s = rowptr[w * blk_h];
    const int e = rowptr[imin(w * blk_h + blk_h, M)];
    const int n = e - s;
    if (n <= 0) continue;
    int *tmp = (int *)malloc(sizeof(int) * (size_t)n);
    if (!tmp) {
#pragma omp atomic write
      err = ORACLE_ERR_ALLOC;
      continue;
    }
    memcpy(tmp, col + s, sizeof(int) * (size_t)n);
    qsort(tmp, (size_t)n, sizeof(int), cmp_int);
    int last = 0; /* index of the last unique element, as `size` at :217-220 */
    for (int cur = 1; cur < n; ++cur)
      if (tmp[cur] != tmp[cur - 1]) tmp[++last] = tmp[cur];
    const int U = last + 1;
    const int nb = (last + blk_w) / blk_w; /* (size + blk_w) / blk_w, :221 */
    blockPartition[w] = nb;
    total += nb;
    for (int idx = s; idx < e; ++idx)
      edgeToColumn[idx] = find_rank(tmp, U, col[idx]);
    free(tmp);
  }
  if (err) return err;

  rowWindowOffset[0] = 0;
  for (int w = 0; w < W; ++w)
    rowWindowOffset[w + 1] = rowWindowOffset[w] + blockPartition[w];
  *block_count_out = total;
  return ORACLE_OK;
}

/*
 * Phase 2 of ME-TCF: TC-block compaction.
 *
 *   tcblockRowid[b]     : window of block b                  (kernel.cu:149-161)
 *   tcblockOffset       : [0, inclusive scan of per-block nnz] (:264-290 count
 *                         into tcblock_offset + 1, then :399-403 scan)
 *   tileId[p]           : (row % blk_h) * blk_w + (rank % blk_w), as uint8,
 *                         placed at window start + exclusive block offset +
 *                         running count, walking edges in CSR order (:291-300)
 *   sparseAToX[b*blk_w+c]: original column of condensed column c of block b,
 *                         0 for padding slots (zero-filled at :390-391, :298)
 *   tcValues[p]         : (build extension) values[e] moved with its tile id
 *
 * Returns ORACLE_ERR_INCONSISTENT when the window offsets do not cover the
 * condensed columns (the reference would index shared memory out of bounds).
 */
int oracle_metcf_blocks(int32_t M, int32_t nnz, const int32_t *rowptr,
                        const int32_t *col, const int32_t *edgeToColumn,
                        const int32_t *edgeToRow,
                        const int32_t *rowWindowOffset, int32_t blk_h,
                        int32_t blk_w, int64_t block_count,
                        const float *values, int32_t *tcblockRowid,
                        uint8_t *tileId, int32_t *tcblockOffset,
                        int32_t *sparseAToX, float *tcValues, int nthreads) {
  if (M < 0 || nnz < 0 || blk_h <= 0 || blk_w <= 0 || block_count < 0)
    return ORACLE_ERR_ARG;
  set_threads(nthreads);
  const int W = (M + blk_h - 1) / blk_h;
  const int64_t T = block_count;

  /* The window offsets must end at the block count; a non-zero
   * blockPartition entry for an empty window breaks that and makes the
   * reference's generate_tcblock_rowid write past TCblockRowid. */
  if ((int64_t)rowWindowOffset[W] != T) return ORACLE_ERR_INCONSISTENT;

  memset(tcblockOffset, 0, sizeof(int32_t) * (size_t)(T + 1));
  memset(sparseAToX, 0, sizeof(int32_t) * (size_t)(T * blk_w));
  memset(tileId, 0, (size_t)nnz);
  if (tcValues) memset(tcValues, 0, sizeof(float) * (size_t)nnz);

  for (int w = 0; w < W; ++w)
    for (int b = rowWindowOffset[w]; b < rowWindowOffset[w + 1]; ++b)
      if (b >= 0 && b < T) tcblockRowid[b] = w;

  int err = 0;
#pragma omp parallel for schedule(dynamic, 64)
  for (int w = 0; w < W; ++w) {
    const int b0 = rowWindowOffset[w];
    const int nb = rowWindowOffset[w + 1] - b0;
    if (nb <= 0) continue;
    const int s = rowptr[w * blk_h];
    const int e = rowptr[imin(w * blk_h + blk_h, M)];
    if (e - s <= 0) continue;
    if (b0 < 0 || (int64_t)b0 + nb > T) {
#pragma omp atomic write
      err = ORACLE_ERR_INCONSISTENT;
      continue;
    }
    int *cnt = (int *)calloc((size_t)nb * 2, sizeof(int));
    if (!cnt) {
#pragma omp atomic write
      err = ORACLE_ERR_ALLOC;
      continue;
    }
    int *pos = cnt + nb;
    int bad = 0;
    for (int idx = s; idx < e; ++idx) {
      const int k = edgeToColumn[idx] / blk_w;
      if (k < 0 || k >= nb) { bad = 1; break; }
      cnt[k]++;
    }
    if (bad) {
      free(cnt);
#pragma omp atomic write
      err = ORACLE_ERR_INCONSISTENT;
      continue;
    }
    for (int k = 0; k < nb; ++k) tcblockOffset[1 + b0 + k] = cnt[k];
    /* exclusive prefix of the per-block counts, window-local */
    int run = 0;
    for (int k = 0; k < nb; ++k) {
      const int c = cnt[k];
      cnt[k] = run;
      run += c;
    }
    for (int idx = s; idx < e; ++idx) {
      const unsigned c = (unsigned)edgeToColumn[idx];
      const unsigned k = c / (unsigned)blk_w;
      const unsigned row_local = (unsigned)edgeToRow[idx] % (unsigned)blk_h;
      const unsigned col_local = c % (unsigned)blk_w;
      const int p = s + cnt[k] + pos[k];
      tileId[p] = (uint8_t)(row_local * (unsigned)blk_w + col_local);
      if (tcValues) tcValues[p] = values ? values[idx] : 1.0f;
      sparseAToX[(size_t)(b0 + (int)k) * blk_w + col_local] = col[idx];
      pos[k]++;
    }
    free(cnt);
  }
  if (err) return err;
  for (int64_t b = 0; b < T; ++b) tcblockOffset[b + 1] += tcblockOffset[b];
  return ORACLE_OK;
}

/* C = A * B in float64, A in CSR (values NULL means every value is 1, as the
 * reference's kernels use, DTCSpMM_kernel.cu:837, :1140).  B is K x N float32
 * row-major; columns >= K contribute nothing, as the kernels' dense_bound
 * guard (DTCSpMM_kernel.cu:2735-2736).  If Cabs is not NULL it receives
 * sum |a| * |b| per output, the denominator of the componentwise error. */
void oracle_spmm_csr_f64(int32_t M, const int32_t *rowptr, const int32_t *col,
                         const float *values, const float *B, int32_t K,
                         int32_t N, double *C, double *Cabs, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
  for (int r = 0; r < M; ++r) {
    double *c = C + (size_t)r * N;
    double *ca = Cabs ? Cabs + (size_t)r * N : NULL;
    for (int n = 0; n < N; ++n) c[n] = 0.0;
    if (ca)
      for (int n = 0; n < N; ++n) ca[n] = 0.0;
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      const int j = col[e];
      if (j < 0 || j >= K) continue;
      const double v = values ? (double)values[e] : 1.0;
      const float *b = B + (size_t)j * N;
      for (int n = 0; n < N; ++n) c[n] += v * (double)b[n];
      if (ca) {
        const double av = v < 0 ? -v : v;
        for (int n = 0; n < N; ++n) {
          const double bv = (double)b[n];
          ca[n] += av * (bv < 0 ? -bv : bv);
        }
      }
    }
  }
}

/* C = A * B in float64 with A decoded from ME-TCF: for every block b of
 * window w and every stored element p of b, tile id t gives row
 * w*blk_h + t / blk_w and column sparseAToX[b*blk_w + t % blk_w]
 * (DTCSpMM_kernel.cu:2716-2722 builds the same dense tile).  Elements with
 * the same tile id (a repeated (row, column) pair in the CSR) are added: the
 * reference's shared-memory store lets one of the racing writes win, and its
 * loader never produces such pairs because it sums duplicates
 * (scripts/DTCSpMM/dataset.py:36-44); adding them makes the ME-TCF product
 * equal to the CSR product for every input.  tcValues NULL means all ones. */
int oracle_spmm_metcf_f64(int32_t M, int32_t num_windows,
                          const int32_t *rowWindowOffset,
                          const uint8_t *tileId, const int32_t *tcblockOffset,
                          const int32_t *sparseAToX, const float *tcValues,
                          int32_t blk_h, int32_t blk_w, const float *B,
                          int32_t K, int32_t N, double *C, int nthreads) {
  if (blk_h * blk_w > 4096) return ORACLE_ERR_ARG;
  set_threads(nthreads);
  memset(C, 0, sizeof(double) * (size_t)M * (size_t)N);
  int err = 0;
#pragma omp parallel for schedule(dynamic, 16)
  for (int w = 0; w < num_windows; ++w) {
    double tile[4096];
    for (int b = rowWindowOffset[w]; b < rowWindowOffset[w + 1]; ++b) {
      for (int i = 0; i < blk_h * blk_w; ++i) tile[i] = 0.0;
      for (int p = tcblockOffset[b]; p < tcblockOffset[b + 1]; ++p) {
        const int t = tileId[p];
        if (t >= blk_h * blk_w) {
#pragma omp atomic write
          err = ORACLE_ERR_INCONSISTENT;
          continue;
        }
        tile[t] += tcValues ? (double)tcValues[p] : 1.0;
      }
      for (int r = 0; r < blk_h; ++r) {
        const int row = w * blk_h + r;
        if (row >= M) break;
        double *c = C + (size_t)row * N;
        for (int cc = 0; cc < blk_w; ++cc) {
          const double v = tile[r * blk_w + cc];
          if (v == 0.0) continue;
          const int j = sparseAToX[(size_t)b * blk_w + cc];
          if (j < 0 || j >= K) continue;
          const float *bb = B + (size_t)j * N;
          for (int n = 0; n < N; ++n) c[n] += v * (double)bb[n];
        }
      }
    }
  }
  return err;
}

/* CPU baseline "port" of the reference's ME-TCF SpMM: windows in parallel,
 * each TC block's stored elements applied as float32 row updates
 * C[row] += a * B[sparseAToX[...]] (the per-element form of the block
 * product of DTCSpMM_kernel.cu:2716-2794).  Rows of a window start at zero. */
__attribute__((target_clones("avx2", "default")))
void oracle_spmm_metcf_f32(int32_t M, int32_t num_windows,
                           const int32_t *rowWindowOffset,
                           const uint8_t *tileId, const int32_t *tcblockOffset,
                           const int32_t *sparseAToX, const float *tcValues,
                           int32_t blk_h, int32_t blk_w, const float *B,
                           int32_t K, int32_t N, float *C, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 8)
  for (int w = 0; w < num_windows; ++w) {
    const int r0 = w * blk_h;
    const int r1 = imin(r0 + blk_h, M);
    for (int r = r0; r < r1; ++r) memset(C + (size_t)r * N, 0, sizeof(float) * (size_t)N);
    for (int b = rowWindowOffset[w]; b < rowWindowOffset[w + 1]; ++b) {
      for (int p = tcblockOffset[b]; p < tcblockOffset[b + 1]; ++p) {
        const int t = tileId[p];
        const int row = r0 + t / blk_w;
        const int j = sparseAToX[(size_t)b * blk_w + t % blk_w];
        if (row >= r1 || j < 0 || j >= K) continue;
        const float v = tcValues ? tcValues[p] : 1.0f;
        float *c = C + (size_t)row * N;
        const float *bb = B + (size_t)j * N;
        for (int n = 0; n < N; ++n) c[n] += v * bb[n];
      }
    }
  }
}

/* out[e] = sum_n X[row(e), n] * Y[col(e), n] in float64: the gradient of
 * the CSR values for C = A * Y with upstream gradient X = dC (SDDMM on A's
 * pattern).  Columns >= K give 0. */
void oracle_sddmm_csr_f64(int32_t M, const int32_t *rowptr, const int32_t *col,
                          const float *X, const float *Y, int32_t K, int32_t N,
                          double *out, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
  for (int r = 0; r < M; ++r) {
    const float *x = X + (size_t)r * N;
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      const int j = col[e];
      double acc = 0.0;
      if (j >= 0 && j < K) {
        const float *y = Y + (size_t)j * N;
        for (int n = 0; n < N; ++n) acc += (double)x[n] * (double)y[n];
      }
      out[e] = acc;
    }
  }
}

/* Transpose of an M x K CSR matrix into K x M CSR with ascending column
 * indices per row (a stable counting sort by column).  perm[p] receives the
 * source edge of transposed edge p, so values move as vals_t[p] = vals[perm[p]]. */
int oracle_csr_transpose(int32_t M, int32_t K, const int32_t *rowptr,
                         const int32_t *col, int32_t *rowptr_t,
                         int32_t *col_t, int32_t *perm) {
  memset(rowptr_t, 0, sizeof(int32_t) * (size_t)(K + 1));
  for (int r = 0; r < M; ++r)
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      const int j = col[e];
      if (j < 0 || j >= K) return ORACLE_ERR_ARG;
      rowptr_t[j + 1]++;
    }
  for (int j = 0; j < K; ++j) rowptr_t[j + 1] += rowptr_t[j];
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(K > 0 ? K : 1));
  if (!fill) return ORACLE_ERR_ALLOC;
  memcpy(fill, rowptr_t, sizeof(int32_t) * (size_t)K);
  for (int r = 0; r < M; ++r)
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      const int j = col[e];
      const int p = fill[j]++;
      col_t[p] = r;
      perm[p] = e;
    }
  free(fill);
  return ORACLE_OK;
}

int oracle_version(void) { return 1; }
