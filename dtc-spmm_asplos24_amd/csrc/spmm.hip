// ME-TCF SpMM on gfx950: C = A * B with A in 16 x 8 TC blocks.
//
// Work decomposition (one wavefront = one work unit, 4 waves per workgroup):
//   unit = (run of TC blocks, column tile of 16*VEC*NSUB output columns).
//   The unbalanced schedule uses one whole window per unit
//   (DTCSpMM_kernel.cu:811-944 launches one CTA per window); the balanced
//   schedule cuts the TC-block list into fixed chunks, each unit walking the
//   windows that start in its chunk and flushing on every window change, with
//   windows longer than a chunk cut and merged atomically
//   (DTCSpMM_kernel.cu:4730-4901 walks 64-block chunks and flushes on a
//   window change).
//
// Inner loop: 4 consecutive TC blocks of the window form one K = 32 group,
// which is exactly the K of v_mfma_f32_16x16x32_{bf16,f16}.  Lane l works on
// block g = l / 16 of the group and MFMA column q = l % 16:
//   * B operand: the lane gathers the 8 rows SparseAToX[8*(bb+g) + j] of B,
//     VEC consecutive floats each at column tile + q*VEC (one 16-lane group
//     reads 16*VEC*4 contiguous bytes of a row);
//   * A operand: the lane needs A_g[q][0..7], read from a per-wave LDS tile
//     built by scattering the stored values at their tile ids;
//   * C: lane l holds rows 4*(l/16) + i, i < 4, of its MFMA column, i.e. the
//     output columns tile + q*VEC + v for the v-th MFMA.
// For FP32 the same group is fed as 8 steps of v_mfma_f32_16x16x4_f32 where
// step j pairs A_g[:, j] with B row j of block g.  The k order inside a
// group is a permutation shared by A and B, so every mapping sums the same
// products.
#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>

#include "dtc_internal.h"

// Workspace of dtc_spmm_balanced_noalloc, for callers that cannot allocate:
// a fixed device array, one copy per GPU, held until the process ends.  Host
// code never uses this name as an address (on the host it is a stand-in
// symbol); the device address comes from hipGetSymbolAddress, per device.
constexpr size_t kNoAllocScratchBytes = (size_t)8 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace dtc {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

constexpr int kBlkH = 16;
constexpr int kBlkW = 8;
constexpr int kGroup = 4;                        // TC blocks per K=32 group
constexpr int kTileFloats = kGroup * kBlkH * kBlkW;  // 512

// Per-wave LDS floats of a kernel instance: the A tiles of one group, or the
// 16-row output tile of one window while it is merged with atomics.
template <int VEC, int NSUB>
constexpr int wave_lds_floats() {
  return kBlkH * 16 * VEC * NSUB > kTileFloats ? kBlkH * 16 * VEC * NSUB : kTileFloats;
}
// ... followed by 4 words of wave state (the balanced chunk index for the
// deterministic merge), kept in LDS rather than in registers: the flush
// that reads it runs at the walk loop's register peak.
template <int VEC, int NSUB>
constexpr int wave_lds_words() {
  return wave_lds_floats<VEC, NSUB>() + 4;
}
constexpr int kWavesPerBlock = 4;

struct SpmmParams {
  const int32_t *rwo;    // unbalanced: row window offsets [W + 1]
  const int32_t *rowid;  // balanced: window of every TC block [T]
  const uint8_t *tile;
  const int32_t *off;
  const int32_t *atox;
  const float *vals;  // TC-block order, NULL = all ones
  const void *B;      // float32, or bf16 for the bf16-B schemes
  float *C;
  int32_t num_windows;
  int32_t num_blocks;
  int32_t M;
  int32_t b_rows;
  int32_t N;
  int32_t nnz;         // entries of tile / vals
  int32_t n_ct;        // column tiles
  int32_t num_items;   // windows (unbalanced) or chunks (balanced)
  int32_t chunk;       // balanced chunk length in TC blocks
  int32_t *queue;      // persistent work counter
  const int2 *runs;    // balanced: run [b, e) of every chunk, cut flags in the sign bits
  float *parts;        // balanced, deterministic merge: partial-tile slots (NULL: atomics)
  int32_t span;        // balanced cut span in TC blocks (a multiple of chunk)
};

// Kernels take the arrays as restrict-qualified arguments: the compiler may
// then read the wave-uniform ones (block offsets, window ids) through the
// scalar cache, whose waits do not queue behind the B gathers.
#define SPMM_KERNEL_ARGS                                                                   \
  const int32_t *__restrict__ a_rwo, const int32_t *__restrict__ a_rowid,                  \
      const uint8_t *__restrict__ a_tile, const int32_t *__restrict__ a_off,               \
      const int32_t *__restrict__ a_atox, const float *__restrict__ a_vals,                \
      const void *__restrict__ a_B, float *__restrict__ a_C, SpmmParams a_p
#define SPMM_PARAMS_FROM_ARGS                                                              \
  SpmmParams {                                                                             \
    a_rwo, a_rowid, a_tile, a_off, a_atox, a_vals, a_B, a_C, a_p.num_windows,              \
        a_p.num_blocks, a_p.M, a_p.b_rows, a_p.N, a_p.nnz, a_p.n_ct, a_p.num_items,        \
        a_p.chunk, a_p.queue, a_p.runs, a_p.parts, a_p.span                                \
  }

template <int VEC>
struct VecType {
  typedef float type __attribute__((ext_vector_type(VEC)));
};
template <>
struct VecType<1> {
  typedef float type;
};

template <int VEC>
__device__ __forceinline__ float comp(const typename VecType<VEC>::type &x, int v) {
  if constexpr (VEC == 1) return x;
  else return x[v];
}

template <int VEC>
__device__ __forceinline__ typename VecType<VEC>::type vzero() {
  typename VecType<VEC>::type z;
  if constexpr (VEC == 1) z = 0.f;
  else
#pragma unroll
    for (int v = 0; v < VEC; ++v) z[v] = 0.f;
  return z;
}

// Schemes for B stored as bf16 (values taken exactly): A enters as three
// bf16 limbs (A3) or as one (A1); accumulation is float32 either way.
constexpr int kBf16B_A3 = 16;
constexpr int kBf16B_A1 = 17;
template <int PREC>
constexpr bool is_bf16b() {
  return PREC == kBf16B_A3 || PREC == kBf16B_A1;
}

template <int VEC>
struct BfVec {
  typedef __bf16 type __attribute__((ext_vector_type(VEC)));
};

// Register type of one lane's slice of one gathered B row.
template <int PREC, int VEC>
using breg_t = typename std::conditional<is_bf16b<PREC>(), typename BfVec<VEC>::type,
                                         typename VecType<VEC>::type>::type;

__device__ __forceinline__ float bf16_to_f32(__bf16 x) { return (float)x; }

// Limb split of x into bf16 terms, x = l0 + l1 (+ l2) + residual.
__device__ __forceinline__ void split2(float x, __bf16 &h, __bf16 &l) {
  h = (__bf16)x;
  l = (__bf16)(x - bf16_to_f32(h));
}
__device__ __forceinline__ void split3(float x, __bf16 &h, __bf16 &m, __bf16 &l) {
  h = (__bf16)x;
  const float r1 = x - bf16_to_f32(h);
  m = (__bf16)r1;
  l = (__bf16)(r1 - bf16_to_f32(m));
}

#define MFMA_F32(a, b, c) __builtin_amdgcn_mfma_f32_16x16x4f32((a), (b), (c), 0, 0, 0)
#define MFMA_BF16(a, b, c) __builtin_amdgcn_mfma_f32_16x16x32_bf16((a), (b), (c), 0, 0, 0)
#define MFMA_F16(a, b, c) __builtin_amdgcn_mfma_f32_16x16x32_f16((a), (b), (c), 0, 0, 0)

// acc[s][v] += A_group * B_group for the columns of subtile s, component v.
template <int PREC, int VEC, int NSUB>
__device__ __forceinline__ void mfma_group(const float (&a)[8], const breg_t<PREC, VEC> (&breg)[NSUB][8],
                                           floatx4 (&acc)[NSUB][VEC], bool has_vals) {
  if constexpr (is_bf16b<PREC>()) {
    bf16x8 A0, A1, A2;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      __bf16 h, m, l;
      split3(a[j], h, m, l);
      A0[j] = h;
      A1[j] = m;
      A2[j] = l;
    }
#pragma unroll
    for (int s = 0; s < NSUB; ++s)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        bf16x8 Bv;
#pragma unroll
        for (int j = 0; j < 8; ++j) Bv[j] = breg[s][j][v];
        floatx4 c = acc[s][v];
        if (PREC == kBf16B_A3 && has_vals) {
          c = MFMA_BF16(A2, Bv, c);
          c = MFMA_BF16(A1, Bv, c);
        }
        acc[s][v] = MFMA_BF16(A0, Bv, c);
      }
  } else if constexpr (PREC == DTC_PRECISION_FP32) {
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int s = 0; s < NSUB; ++s)
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc[s][v] = MFMA_F32(a[j], comp<VEC>(breg[s][j], v), acc[s][v]);
  } else if constexpr (PREC == DTC_PRECISION_BF16) {
    bf16x8 A;
#pragma unroll
    for (int j = 0; j < 8; ++j) A[j] = (__bf16)a[j];
#pragma unroll
    for (int s = 0; s < NSUB; ++s)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        bf16x8 Bv;
#pragma unroll
        for (int j = 0; j < 8; ++j) Bv[j] = (__bf16)comp<VEC>(breg[s][j], v);
        acc[s][v] = MFMA_BF16(A, Bv, acc[s][v]);
      }
  } else if constexpr (PREC == DTC_PRECISION_FP16) {
    halfx8 A;
#pragma unroll
    for (int j = 0; j < 8; ++j) A[j] = (_Float16)a[j];
#pragma unroll
    for (int s = 0; s < NSUB; ++s)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        halfx8 Bv;
#pragma unroll
        for (int j = 0; j < 8; ++j) Bv[j] = (_Float16)comp<VEC>(breg[s][j], v);
        acc[s][v] = MFMA_F16(A, Bv, acc[s][v]);
      }
  } else if constexpr (PREC == DTC_PRECISION_BF16X3) {
    bf16x8 Ah, Al;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      __bf16 h, l;
      split2(a[j], h, l);
      Ah[j] = h;
      Al[j] = l;
    }
#pragma unroll
    for (int s = 0; s < NSUB; ++s)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        bf16x8 Bh, Bl;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          __bf16 h, l;
          split2(comp<VEC>(breg[s][j], v), h, l);
          Bh[j] = h;
          Bl[j] = l;
        }
        floatx4 c = acc[s][v];
        if (has_vals) c = MFMA_BF16(Al, Bh, c);
        c = MFMA_BF16(Ah, Bl, c);
        c = MFMA_BF16(Ah, Bh, c);
        acc[s][v] = c;
      }
  } else {  // DTC_PRECISION_BF16X6
    bf16x8 A0, A1, A2;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      __bf16 h, m, l;
      split3(a[j], h, m, l);
      A0[j] = h;
      A1[j] = m;
      A2[j] = l;
    }
#pragma unroll
    for (int s = 0; s < NSUB; ++s)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        bf16x8 B0, B1, B2;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          __bf16 h, m, l;
          split3(comp<VEC>(breg[s][j], v), h, m, l);
          B0[j] = h;
          B1[j] = m;
          B2[j] = l;
        }
        floatx4 c = acc[s][v];
        if (has_vals) {
          c = MFMA_BF16(A2, B0, c);
          c = MFMA_BF16(A1, B1, c);
        }
        c = MFMA_BF16(A0, B2, c);
        if (has_vals) c = MFMA_BF16(A1, B0, c);
        c = MFMA_BF16(A0, B1, c);
        c = MFMA_BF16(A0, B0, c);
        acc[s][v] = c;
      }
  }
}

// Metadata of one K=32 group (up to 4 TC blocks of one window), loaded one
// group ahead of its use so that its latency hides behind the previous
// group's B gathers.  Loading does not need the group's length: the offsets
// of all 4 candidate blocks are read and `resolve_meta` trims them once the
// length is known.
struct GroupMeta {
  int bb;                  // first block (wave-uniform)
  int f0, f1, f2, f3, f4;  // TCblock_offset[bb + k], raw
  int o0, o1, o2, o3, o4;  // element offsets of the group's blocks (resolved)
  int4 r0, r1;             // SparseAToX of block g (raw; masked where used)
  int nblk;                // blocks in the group (resolved, wave-uniform)
  int tile;                // tile id of element f0 + lane (raw)
  float val;               // its value (raw)
  int rid;                 // window of block bb + lane, -1 at or past the run's end
  int fe;                  // without WALK: TCblock_offset[min(bb + 4, b1)], the group's end
};

// Loads are unconditional and their results are only consumed one group
// later, so nothing in the current group waits on them.  With WALK the
// lane also reads the window id of block bb + lane: the 64-block look-ahead
// from which the group's length (it stops at the next window start) is
// found without another round trip.
template <bool WALK>
__device__ __forceinline__ void load_meta(const SpmmParams &p, int bb, int b1, bool has_vals,
                                          GroupMeta &m) {
  const int lane = lane_id();
  const int g = lane >> 4;
  const int T = p.num_blocks;
  m.bb = bb;
  m.f0 = p.off[bb];
  m.f1 = p.off[min(bb + 1, T)];
  m.f2 = p.off[min(bb + 2, T)];
  m.f3 = p.off[min(bb + 3, T)];
  m.f4 = p.off[min(bb + 4, T)];
  const int blk = min(bb + g, b1 - 1);
  const int4 *ap = reinterpret_cast<const int4 *>(p.atox + (int64_t)blk * kBlkW);
  m.r0 = ap[0];
  m.r1 = ap[1];
  // The first 64 stored elements of the group, one per lane (a group holds
  // 36 on average for R-MAT; more than 64 take the slow path in build_a).
  const int e0 = min(max(m.f0, 0), p.nnz);
  const int ec = e0 < p.nnz - lane ? e0 + lane : p.nnz - 1;
  m.tile = p.tile[ec];
  m.val = has_vals ? p.vals[ec] : 1.0f;
  if constexpr (WALK) {
    const int rb = bb + lane;
    m.rid = rb < b1 ? p.rowid[rb] : -1;
  } else {
    m.fe = p.off[min(bb + kGroup, b1)];
  }
}

// Group length and element offsets.  WALK: the group ends before the first
// block of the look-ahead whose window differs from its predecessor's (or
// at the run's end, where the look-ahead holds -1); else at b1.
template <bool WALK>
__device__ __forceinline__ void resolve_meta(const SpmmParams &p, int b1, GroupMeta &m) {
  if constexpr (!WALK) {
    // The group's end offset was loaded with the rest (its length is known
    // up front); offsets past it are clamped to it, which for non-decreasing
    // offsets is the end itself -- no index into the offsets by a run-time
    // count, which the compiler would place in scratch memory.
    m.nblk = min(kGroup, b1 - m.bb);
    const int o4 = min(m.fe, p.nnz);
    m.o0 = min(m.f0, o4);
    m.o1 = min(m.f1, o4);
    m.o2 = min(m.f2, o4);
    m.o3 = min(m.f3, o4);
    m.o4 = o4;
    return;
  }
  const int prev = __shfl_up(m.rid, 1, 64);
  const uint64_t starts = __ballot(lane_id() > 0 && m.rid != prev);
  const int nblk = starts ? min(kGroup, (int)__builtin_ctzll(starts)) : kGroup;
  // selects, not an indexed array (which would live in scratch memory)
  auto f = [&m](int k) {
    return k == 0 ? m.f0 : k == 1 ? m.f1 : k == 2 ? m.f2 : k == 3 ? m.f3 : m.f4;
  };
  const int o4 = min(f(nblk), p.nnz);
  m.o0 = min(m.f0, o4);
  m.o1 = min(f(min(1, nblk)), o4);
  m.o2 = min(f(min(2, nblk)), o4);
  m.o3 = min(f(min(3, nblk)), o4);
  m.o4 = o4;
  m.nblk = nblk;
}

// This lane's 8 B rows (-1 for lanes whose block lies past the run).
__device__ __forceinline__ void meta_rows(const GroupMeta &m, int (&rows)[8]) {
  const bool gv = (lane_id() >> 4) < m.nblk;
  rows[0] = gv ? m.r0.x : -1;
  rows[1] = gv ? m.r0.y : -1;
  rows[2] = gv ? m.r0.z : -1;
  rows[3] = gv ? m.r0.w : -1;
  rows[4] = gv ? m.r1.x : -1;
  rows[5] = gv ? m.r1.y : -1;
  rows[6] = gv ? m.r1.z : -1;
  rows[7] = gv ? m.r1.w : -1;
}

// B rows of the group into registers: loads from clamped addresses, then
// zero for slots without a row or columns past N (no branches around loads).
template <int PREC, int VEC, int NSUB>
__device__ __forceinline__ void gather_b(const SpmmParams &p, const int (&rows)[8], int colbase,
                                         breg_t<PREC, VEC> (&breg)[NSUB][8]) {
  typedef breg_t<PREC, VEC> vt;
  typedef typename std::conditional<is_bf16b<PREC>(), __bf16, float>::type et;
  const et *B = static_cast<const et *>(p.B);
  const int64_t N = p.N;
  vt zero;
  if constexpr (VEC == 1 && !is_bf16b<PREC>()) zero = 0.f;
  else
#pragma unroll
    for (int v = 0; v < VEC; ++v) zero[v] = (et)0.f;
#pragma unroll
  for (int s = 0; s < NSUB; ++s) {
    const int col = colbase + s * 16 * VEC;
    const bool colok = col < p.N;
    const int colc = colok ? col : 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int r = rows[j];
      const bool ok = colok && (unsigned)r < (unsigned)p.b_rows;
      const int rc = ok ? r : 0;
      const vt v = *reinterpret_cast<const vt *>(B + (int64_t)rc * N + colc);
      breg[s][j] = ok ? v : zero;
    }
  }
}

// Dense 16x8 A tiles of the group in this wave's LDS, then this lane's
// operand A_g[q][0..7].
__device__ __forceinline__ void build_a(const SpmmParams &p, const GroupMeta &m, bool has_vals,
                                        float *lds, float (&a)[8]) {
  const int lane = lane_id();
  const int g = lane >> 4;
  const int q = lane & 15;
  floatx4 *t4 = reinterpret_cast<floatx4 *>(lds);
  const floatx4 z4 = {0.f, 0.f, 0.f, 0.f};
  t4[lane * 2] = z4;
  t4[lane * 2 + 1] = z4;
  wave_lds_sync();
  // Stored elements are added into the tile (LDS float atomics): a CSR with
  // a repeated (row, column) pair puts several elements on one tile id, and
  // the product must count each of them, as the CSR product does.
  if (m.o0 + lane < m.o4) {
    const int e = m.o0 + lane;
    const int gg = (e >= m.o1) + (e >= m.o2) + (e >= m.o3);
    atomicAdd(&lds[gg * (kBlkH * kBlkW) + (m.tile & (kBlkH * kBlkW - 1))], m.val);
  }
  for (int e = m.o0 + 64 + lane; e < m.o4; e += 64) {
    const int gg = (e >= m.o1) + (e >= m.o2) + (e >= m.o3);
    const float v = has_vals ? p.vals[e] : 1.0f;
    atomicAdd(&lds[gg * (kBlkH * kBlkW) + (p.tile[e] & (kBlkH * kBlkW - 1))], v);
  }
  wave_lds_sync();
  const floatx4 lo = t4[(g * (kBlkH * kBlkW) + q * kBlkW) / 4];
  const floatx4 hi = t4[(g * (kBlkH * kBlkW) + q * kBlkW) / 4 + 1];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    a[j] = lo[j];
    a[j + 4] = hi[j];
  }
  wave_lds_sync();
}

// Accumulates the TC blocks [b0, b1) of one window into acc.  Software
// pipelined: group i's B gathers are issued, then group i+1's metadata, then
// group i's A tile is built and multiplied.
template <int PREC, int VEC, int NSUB>
__device__ __forceinline__ void run_blocks(const SpmmParams &p, int b0, int b1, int colbase,
                                           float *lds, floatx4 (&acc)[NSUB][VEC],
                                           bool has_vals) {
  if (b0 >= b1) return;
  GroupMeta cur;
  load_meta<false>(p, b0, b1, has_vals, cur);
  resolve_meta<false>(p, b1, cur);
  for (int bb = b0; bb < b1; bb += kGroup) {
    breg_t<PREC, VEC> breg[NSUB][8];
    int rows[8];
    meta_rows(cur, rows);
    gather_b<PREC, VEC, NSUB>(p, rows, colbase, breg);
    // The last group re-reads its own metadata instead of branching.
    const int bn = (bb + kGroup < b1) ? bb + kGroup : bb;
    GroupMeta nxt;
    load_meta<false>(p, bn, b1, has_vals, nxt);
    float a[8];
    build_a(p, cur, has_vals, lds, a);
    mfma_group<PREC, VEC, NSUB>(a, breg, acc, has_vals);
    cur = nxt;
    resolve_meta<false>(p, b1, cur);
  }
}

template <int VEC, int NSUB>
__device__ __forceinline__ void zero_acc(floatx4 (&acc)[NSUB][VEC]) {
#pragma unroll
  for (int s = 0; s < NSUB; ++s)
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[s][v] = floatx4{0.f, 0.f, 0.f, 0.f};
}

// Writes the 16 rows of window w held in acc.
template <int VEC, int NSUB>
__device__ __forceinline__ void store_window(const SpmmParams &p, int w, int colbase,
                                             const floatx4 (&acc)[NSUB][VEC]) {
  typedef typename VecType<VEC>::type vfloat;
  const int lane = lane_id();
  const int g = lane >> 4;
  const int64_t N = p.N;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t row = (int64_t)w * kBlkH + 4 * g + i;
    if (row < 0 || row >= p.M) continue;
#pragma unroll
    for (int s = 0; s < NSUB; ++s) {
      const int col = colbase + s * 16 * VEC;
      if (col >= p.N) continue;
      vfloat out;
      if constexpr (VEC == 1) out = acc[s][0][i];
      else
#pragma unroll
        for (int v = 0; v < VEC; ++v) out[v] = acc[s][v][i];
      __builtin_nontemporal_store(out, reinterpret_cast<vfloat *>(p.C + row * N + col));
    }
  }
}

// Zeroes this wave's column tile of windows [w0, w1) (empty windows: no TC
// block writes them, so any wave may, at any time of the launch).
template <int VEC, int NSUB>
__device__ __forceinline__ void zero_windows(const SpmmParams &p, int w0, int w1, int colbase) {
  typedef typename VecType<VEC>::type vfloat;
  const vfloat z = vzero<VEC>();
  const int g = lane_id() >> 4;
  const int64_t N = p.N;
  // Rolled: rare, and unrolled stores would hold their addresses in registers
  // across the whole SpMM loop.
#pragma unroll 1
  for (int r = max(w0, 0) * kBlkH; r < w1 * kBlkH; r += kBlkH) {
#pragma unroll 1
    for (int i = 0; i < 4; ++i) {
      const int64_t row = (int64_t)r + 4 * g + i;
      if (row >= p.M) continue;
#pragma unroll
      for (int s = 0; s < NSUB; ++s) {
        const int col = colbase + s * 16 * VEC;
        if (col < p.N) __builtin_nontemporal_store(z, reinterpret_cast<vfloat *>(p.C + row * N + col));
      }
    }
  }
}

// Adds the 16 rows of window w held in acc to C with atomics, through this
// wave's LDS: the tile is written in the accumulator layout and read back
// row-major, so that each atomic instruction covers 64 consecutive floats of
// a row (or two rows of a 32-column tile) instead of lanes 16 bytes apart
// spread over 4 rows -- a quarter of the cache lines per instruction.
// With the deterministic merge (p.parts set) the row-major tile is stored
// instead, whole, in a partial-tile slot: slot 2 c holds chunk c's share of
// its first window, slot 2 c + 1 of its last (`last_role`), each slot n_ct
// tiles of 16 x TW floats; k_balance_fixup sums them in chunk order.  The
// chunk index is the one run_unit left in the wave's LDS state word.
template <int VEC, int NSUB>
__device__ __forceinline__ void add_window_atomic(const SpmmParams &p, int w, int colbase,
                                                  float *lds, const floatx4 (&acc)[NSUB][VEC],
                                                  bool last_role) {
  constexpr int TW = 16 * VEC * NSUB;  // tile width in floats
  const int lane = lane_id();
  const int g = lane >> 4;
  const int q = lane & 15;
#pragma unroll
  for (int s = 0; s < NSUB; ++s)
#pragma unroll
    for (int v = 0; v < VEC; ++v)
#pragma unroll
      for (int i = 0; i < 4; ++i) lds[(4 * g + i) * TW + s * 16 * VEC + q * VEC + v] = acc[s][v][i];
  wave_lds_sync();
  const int col0 = colbase - q * VEC;  // first column of the tile
  const int64_t N = p.N;
  if (p.parts) {
    const int item = wave_uniform(reinterpret_cast<const int *>(lds)[wave_lds_floats<VEC, NSUB>()]);
    const int slot = 2 * item + (last_role ? 1 : 0);
    const int ct = wave_uniform(col0) / TW;
    float *dst = p.parts + ((int64_t)slot * p.n_ct + ct) * (kBlkH * TW);
#pragma unroll 1
    for (int k = 0; k < kBlkH * TW / 64; ++k) dst[k * 64 + lane] = lds[k * 64 + lane];
    wave_lds_sync();
    return;
  }
  // Rolled: at most two flushes per chunk take this path, and unrolling it
  // keeps every address live at once (twice the kernel's registers).
#pragma unroll 1
  for (int k = 0; k < kBlkH * TW / 64; ++k) {
    const int e = k * 64 + lane;
    const int r = e / TW;
    const int col = col0 + e % TW;
    const int64_t row = (int64_t)w * kBlkH + r;
    if (row >= 0 && row < p.M && col < p.N) unsafeAtomicAdd(p.C + row * N + col, lds[e]);
  }
  wave_lds_sync();
}

// Accumulates the TC blocks [b, e) of the balanced schedule, which may span
// many windows, as ONE software pipeline: groups never cross a window start
// (they are the groups a per-window loop would form), and when a group opens
// a new window the accumulators of the previous one are stored and cleared.
// Window `atomic_w0` (if >= 0) and `atomic_w1` (if >= 0) are merged with
// atomic adds (windows cut at chunk boundaries), every other window is
// stored whole.  Empty windows are zeroed by the wave that passes them: the
// gap before the run's first window (after the previous block's window), the
// gaps between consecutive windows, and -- for the run that ends at the last
// block -- the windows after it.  Window starts come from the look-ahead of
// window ids loaded with each group's metadata, so no window is searched for.
template <int PREC, int VEC, int NSUB>
__device__ __forceinline__ void walk_blocks(const SpmmParams &p, int b, int e, int colbase,
                                            float *lds, floatx4 (&acc)[NSUB][VEC],
                                            bool has_vals, int atomic_w0, int atomic_w1) {
  GroupMeta cur;
  load_meta<true>(p, b, e, has_vals, cur);
  const int prev_w = b > 0 ? wave_uniform(p.rowid[b - 1]) : -1;
  resolve_meta<true>(p, e, cur);
  int cw = wave_uniform(cur.rid);
  // pending empty windows [gap0, gap1): zeroed at one place in the loop
  int gap0 = prev_w + 1;
  int gap1 = min(cw, p.num_windows);
  zero_acc<VEC, NSUB>(acc);
  while (true) {
    breg_t<PREC, VEC> breg[NSUB][8];
    int rows[8];
    meta_rows(cur, rows);
    gather_b<PREC, VEC, NSUB>(p, rows, colbase, breg);
    // The previous window's stores go out after this group's gathers: loads
    // and stores retire through one in-order counter, so a store issued first
    // would hold up the wait for the gathers.
    const int w = wave_uniform(cur.rid);  // window of block cur.bb (lane 0)
    if (w != cw) {
      if (cw == atomic_w0 || cw == atomic_w1)
        add_window_atomic<VEC, NSUB>(p, cw, colbase, lds, acc, cw != atomic_w0);
      else store_window<VEC, NSUB>(p, cw, colbase, acc);
      zero_acc<VEC, NSUB>(acc);
      gap0 = cw + 1;
      gap1 = min(w, p.num_windows);
      cw = w;
    }
    if (gap0 < gap1) {
      zero_windows<VEC, NSUB>(p, gap0, gap1, colbase);
      gap0 = gap1;
    }
    const int bn = cur.bb + cur.nblk;
    const bool more = bn < e;
    // The last group re-reads its own metadata instead of branching.
    GroupMeta nxt;
    load_meta<true>(p, more ? bn : cur.bb, e, has_vals, nxt);
    float a[8];
    build_a(p, cur, has_vals, lds, a);
    mfma_group<PREC, VEC, NSUB>(a, breg, acc, has_vals);
    if (!more) break;
    cur = nxt;
    resolve_meta<true>(p, e, cur);
  }
  if (cw == atomic_w0 || cw == atomic_w1)
    add_window_atomic<VEC, NSUB>(p, cw, colbase, lds, acc, cw != atomic_w0);
  else store_window<VEC, NSUB>(p, cw, colbase, acc);
  if (e == p.num_blocks) zero_windows<VEC, NSUB>(p, cw + 1, p.num_windows, colbase);
}

// End of window w whose block b is known: one ballot over the next 64 block
// ids covers most windows; longer ones finish with a wave-parallel search.
__device__ __forceinline__ int window_end_from(const int32_t *rowid, int b, int T, int w) {
  const int idx = b + lane_id();
  const uint64_t m = __ballot(idx >= T || rowid[idx] != w);
  if (m) return b + (int)__builtin_ctzll(m);
  return wave_lower_bound<int32_t>(rowid, b + 64, T, w + 1);
}

// Start of window w whose block b is known (the mirror image).
__device__ __forceinline__ int window_start_from(const int32_t *rowid, int b, int w) {
  if (b == 0) return 0;
  const int idx = b - 1 - lane_id();  // b-1, b-2, ..., b-64
  const uint64_t m = __ballot(idx < 0 || rowid[idx] != w);
  if (m) return b - (int)__builtin_ctzll(m);
  return wave_lower_bound<int32_t>(rowid, 0, b - 64, w);
}

template <int PREC, int VEC, int NSUB, bool BALANCED>
__device__ __forceinline__ void run_unit(const SpmmParams &p, int item, int ct, float *lds) {
  const int lane = lane_id();
  const int colbase = ct * (16 * VEC * NSUB) + (lane & 15) * VEC;
  const bool has_vals = p.vals != nullptr;
  floatx4 acc[NSUB][VEC];
  if constexpr (!BALANCED) {
    const int w = item;
    // Offsets are clamped to the blocks actually passed in.
    const int b1 = max(0, min(p.rwo[w + 1], p.num_blocks));
    const int b0 = max(0, min(p.rwo[w], b1));
    zero_acc<VEC, NSUB>(acc);
    run_blocks<PREC, VEC, NSUB>(p, b0, b1, colbase, lds, acc, has_vals);
    store_window<VEC, NSUB>(p, w, colbase, acc);
  } else {
    // The chunk's run and the windows it merges atomically were worked out
    // by the preparation kernel (k_balance_prep).
    const int2 r = p.runs[item];
    const bool cut_first = r.x < 0;
    const bool cut_last = r.y < 0;
    const int b = cut_first ? ~r.x : r.x;
    const int e = cut_last ? ~r.y : r.y;
    if (b >= e) return;
    const int w_first = cut_first ? wave_uniform(p.rowid[b]) : -1;
    const int w_last = cut_last ? wave_uniform(p.rowid[e - 1]) : -1;
    if (p.parts && (cut_first || cut_last)) {
      reinterpret_cast<int *>(lds)[wave_lds_floats<VEC, NSUB>()] = item;
      wave_lds_sync();
    }
    walk_blocks<PREC, VEC, NSUB>(p, b, e, colbase, lds, acc, has_vals, w_first, w_last);
  }
}

// Waves per SIMD the compiler must leave room for (its register budget),
// per instance.  Left to itself it keeps the balanced kernels' accumulators
// in AGPRs and stops at 5 waves with 32-column tiles; asked for 7 it fits
// them in 70 VGPRs without scratch, and the SpMM runs ~5% faster (R-MAT
// scale 20: 4.23 -> 4.00 ms at N = 256, 0.60 -> 0.56 ms at N = 32).  With
// 64-column tiles, 5 waves instead of 4 measured no faster; the bf16x3 and
// bf16x6 splits would spill.  Those keep the default.
template <int PREC, int VEC, int NSUB, bool BALANCED>
constexpr int spmm_waves_per_eu() {
  if (!BALANCED || PREC == DTC_PRECISION_BF16X3 || PREC == DTC_PRECISION_BF16X6) return 1;
  return VEC * NSUB <= 2 ? 7 : 1;
}

// Workgroup i serves column tile i % n_ct.  Workgroups are dispatched to the
// 8 XCDs round robin, so when n_ct divides 8 every XCD (and its L2) only
// sees one column tile of B: 1/n_ct of each gathered row.
template <int PREC, int VEC, int NSUB, bool BALANCED>
__global__ __launch_bounds__(256) __attribute__((
    amdgpu_waves_per_eu(spmm_waves_per_eu<PREC, VEC, NSUB, BALANCED>()))) void spmm_kernel(SPMM_KERNEL_ARGS) {
  const SpmmParams p = SPMM_PARAMS_FROM_ARGS;
  __shared__ __attribute__((aligned(16))) float s_tile[kWavesPerBlock][wave_lds_words<VEC, NSUB>()];
  const int wave = wave_uniform(threadIdx.x >> 6);
  const int ct = blockIdx.x % p.n_ct;
  const int item = (blockIdx.x / p.n_ct) * kWavesPerBlock + wave;
  if (item >= p.num_items) return;
  run_unit<PREC, VEC, NSUB, BALANCED>(p, item, ct, s_tile[wave]);
}

// Persistent variant of the balanced schedule: every wave pulls units from a
// global counter until the list is exhausted.
template <int PREC, int VEC, int NSUB>
__global__ __launch_bounds__(256) __attribute__((
    amdgpu_waves_per_eu(spmm_waves_per_eu<PREC, VEC, NSUB, true>()))) void spmm_kernel_persistent(SPMM_KERNEL_ARGS) {
  const SpmmParams p = SPMM_PARAMS_FROM_ARGS;
  __shared__ __attribute__((aligned(16))) float s_tile[kWavesPerBlock][wave_lds_words<VEC, NSUB>()];
  const int wave = wave_uniform(threadIdx.x >> 6);
  const int lane = lane_id();
  const int total = p.num_items * p.n_ct;
  while (true) {
    int u = 0;
    if (lane == 0) u = atomicAdd(p.queue, 1);
    u = __shfl(u, 0, 64);
    u = wave_uniform(u);
    if (u >= total) break;
    run_unit<PREC, VEC, NSUB, true>(p, u / p.n_ct, u % p.n_ct, s_tile[wave]);
  }
}

// Zeroes the rows of window w (one wave).
__device__ __forceinline__ void zero_window_rows(float *C, int64_t w, int32_t M, int32_t N) {
  const int lane = lane_id();
  const int64_t r0 = w * kBlkH;
  const int64_t r1 = (r0 + kBlkH < (int64_t)M) ? r0 + kBlkH : (int64_t)M;
  if (r0 >= r1) return;
  float *base = C + r0 * (int64_t)N;
  const int64_t count = (r1 - r0) * (int64_t)N;
  if ((N & 3) == 0 && ((uintptr_t)base & 15) == 0) {
    floatx4 *b4 = reinterpret_cast<floatx4 *>(base);
    const floatx4 z = {0.f, 0.f, 0.f, 0.f};
    for (int64_t idx = lane; idx < count / 4; idx += 64) b4[idx] = z;
  } else {
    for (int64_t idx = lane; idx < count; idx += 64) base[idx] = 0.f;
  }
}

constexpr int kNearSpan = 512;  // chunks up to this many blocks use the bounded scans

// Start of window w (the window of block c) if it lies at most `span` blocks
// before c, else -1.  All span/64 loads per lane are issued before the first
// ballot: one memory round trip instead of a search's chain.
__device__ __forceinline__ int window_start_near(const int32_t *rowid, int c, int w, int span) {
  const int lane = lane_id();
  int v[kNearSpan / 64];
#pragma unroll
  for (int k = 0; k < kNearSpan / 64; ++k) {
    const int idx = c - 1 - lane - 64 * k;  // c-1, c-2, ...; -2 marks the array's start
    v[k] = (64 * k < span && idx >= c - span) ? (idx >= 0 ? rowid[idx] : -2) : w;
  }
#pragma unroll
  for (int k = 0; k < kNearSpan / 64; ++k) {
    const uint64_t m = __ballot(v[k] != w);
    if (m) return c - 64 * k - (int)__builtin_ctzll(m);
  }
  return -1;
}

// End of window w (the window of block c) if it lies at most `span` blocks
// after c, else -1 (same single round of loads).
__device__ __forceinline__ int window_end_near(const int32_t *rowid, int c, int T, int w, int span) {
  const int lane = lane_id();
  int v[kNearSpan / 64 + 1];  // blocks c .. c + span
#pragma unroll
  for (int k = 0; k < kNearSpan / 64 + 1; ++k) {
    const int idx = c + lane + 64 * k;  // c, c+1, ...; -2 marks the array's end
    v[k] = (64 * k <= span && idx <= c + span) ? (idx < T ? rowid[idx] : -2) : w;
  }
#pragma unroll
  for (int k = 0; k < kNearSpan / 64 + 1; ++k) {
    const uint64_t m = __ballot(v[k] != w);
    if (m) return c + 64 * k + (int)__builtin_ctzll(m);
  }
  return -1;
}

// Balanced-schedule preparation.  Ownership rule, for chunks of `chunk` TC
// blocks and a cut span `span` (a multiple of chunk): a window of at most
// `span` blocks belongs whole to the chunk holding its first block (and may
// run past that chunk's end); a longer window is cut at the multiples of
// `span` into pieces, each belonging to the chunk holding its first block,
// and the pieces are merged with atomic adds.  The multiples of `span` are
// chunk boundaries, so a chunk's run can only begin with a cut piece (one that
// starts at the chunk's first block) and only end with one (the first piece
// of a long window that starts inside the chunk, which then runs past the
// chunk's end); every window in between is short and whole.  A larger span
// cuts fewer windows into fewer pieces (fewer atomic tiles written) at the
// price of longer work units.  Wave v, one per chunk:
//   * for chunk v ([c0, c1) = blocks [v*chunk, (v+1)*chunk)), writes the
//     chunk's run [b, e) to runs[v]: b = c0 unless the window or piece
//     holding c0 began earlier (then b = its end), e = c1 unless a window or
//     piece that begins inside the chunk runs past c1 (then e = its end); ~b
//     marks a cut first piece, ~e a cut last piece (b >= e: nothing to do);
//     and zeroes the rows of the cut windows that start in the chunk, since
//     no chunk stores them whole;
//   * (wave 0) resets the persistent work counter.
// The window searches run here, once per chunk, on an idle memory system
// instead of at the start of every SpMM wave.  Empty windows are zeroed by
// the SpMM waves themselves (walk_blocks): nothing else writes them.
__global__ __launch_bounds__(256) void k_balance_prep(const int32_t *rowid, int32_t T,
                                                      int32_t M, int32_t N,
                                                      int32_t chunk, int32_t span,
                                                      int32_t num_items, float *C,
                                                      int32_t *queue, int2 *runs,
                                                      int32_t zero_cut) {
  if (queue && blockIdx.x == 0 && threadIdx.x == 0) *queue = 0;
  const int lane = lane_id();
  const int64_t v = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (v >= num_items) return;
  const int c0 = (int)v * chunk;
  const int c1 = min(c0 + chunk, T);
  // Window extents: exact for spans above kNearSpan (searches); for smaller
  // spans only within `span` blocks of the boundary, -1 meaning farther --
  // which already makes the window longer than the span.
  const bool near = span <= kNearSpan;
  auto start_of = [&](int c, int w) {
    return wave_uniform(near ? window_start_near(rowid, c, w, span) : window_start_from(rowid, c, w));
  };
  auto end_of = [&](int c, int w) {
    return wave_uniform(near ? window_end_near(rowid, c, T, w, span)
                             : max(c + 1, window_end_from(rowid, c, T, w)));
  };
  // first multiple of span after x (a chunk boundary; no multiple of span
  // lies strictly inside a chunk)
  auto next_cut = [&](int x) { return (int)min<int64_t>(((int64_t)x / span + 1) * span, T); };
  auto is_long = [&](int ws, int we) { return ws < 0 || we < 0 || we - ws > span; };
  const int w0 = wave_uniform(rowid[c0]);
  const int ws0 = start_of(c0, w0);
  const int we0 = end_of(c0, w0);
  const bool first_long = is_long(ws0, we0);
  bool cut_first = false;
  int b;
  if (!first_long) {
    b = ws0 < c0 ? we0 : c0;
  } else {
    // the piece holding c0 starts at c0 when c0 is a multiple of span or the
    // window's first block; it ends at the next multiple of span or the
    // window's end (an unknown end lies more than span away: past the cut)
    const int pe = (we0 >= 0 && we0 < next_cut(c0)) ? we0 : next_cut(c0);
    if (c0 % span == 0 || ws0 == c0) {
      b = c0;
      cut_first = true;
    } else {
      b = pe;
    }
  }
  int e = c1;
  bool cut_last = false;
  int wl = -1, wsl = -1;
  if (b < c1 && c1 < T) {
    const int r_in = wave_uniform(rowid[c1 - 1]);
    const int r_out = wave_uniform(rowid[c1]);
    if (r_in == r_out) {  // a window crosses c1
      if (r_in == w0) {
        // the run's first window or piece runs past c1 (b == c0 here): a
        // short window whole, or the same piece as the one at c0
        e = first_long ? ((we0 >= 0 && we0 < next_cut(c0)) ? we0 : next_cut(c0)) : we0;
      } else {
        wl = r_in;  // starts inside the chunk: within span of c1
        wsl = start_of(c1, r_in);
        const int wel = end_of(c1, r_in);
        if (!is_long(wsl, wel)) {
          e = wel;
        } else {
          e = (wel >= 0 && wel < next_cut(wsl)) ? wel : next_cut(wsl);
          cut_last = true;
        }
      }
    }
  }
  if (b >= c1) {
    b = e = c1;
    cut_first = cut_last = false;
  }
  if (lane == 0) runs[v] = make_int2(cut_first ? ~b : b, cut_last ? ~e : e);
  if (!zero_cut) return;  // deterministic merge: k_balance_fixup writes cut windows whole
  if (cut_first && ws0 == c0) zero_window_rows(C, w0, M, N);
  if (cut_last) zero_window_rows(C, wl, M, N);
}

// Deterministic merge of cut windows, after the SpMM kernel: wave (chunk c,
// column tile ct) takes the one cut window that starts in chunk c, if any
// (a cut window is longer than the span, so no chunk holds the start of two),
// sums its partial tiles in chunk order -- chunk c's slot for the window's
// role there, then slot 2 c' of every chunk c' holding a later piece (the
// chunks at the multiples of span inside the window), where it is the first
// piece -- and stores the window's rows.  Tiles are row-major 16 x tw floats
// (tw = the SpMM's column-tile width).
__global__ __launch_bounds__(256) void k_balance_fixup(const int32_t *__restrict__ rowid,
                                                       SpmmParams p, int32_t tw) {
  const int wave = wave_uniform(threadIdx.x >> 6);
  const int lane = lane_id();
  const int ct = blockIdx.x % p.n_ct;
  const int c = (blockIdx.x / p.n_ct) * kWavesPerBlock + wave;
  if (c >= p.num_items) return;
  const int2 r = p.runs[c];
  const bool cut_first = r.x < 0;
  const bool cut_last = r.y < 0;
  if (!cut_first && !cut_last) return;
  const int c0 = c * p.chunk;
  const int b = cut_first ? ~r.x : r.x;
  const int e = cut_last ? ~r.y : r.y;
  int w = -1, role = 0;
  if (cut_first) {  // the chunk's first window: cut, and starting at c0?
    const int wf = wave_uniform(rowid[b]);
    if (c0 == 0 || wave_uniform(rowid[c0 - 1]) != wf) w = wf;
  }
  if (w < 0 && cut_last) {  // the chunk's last window, starting inside the chunk?
    const int wl = wave_uniform(rowid[e - 1]);
    if (c0 == 0 || wave_uniform(rowid[c0 - 1]) != wl) {
      w = wl;
      role = 1;
    }
  }
  if (w < 0) return;
  const int we = wave_uniform(window_end_from(rowid, e - 1, p.num_blocks, w));
  // later pieces start at the multiples of span in (c0, we): chunks
  // first_later, first_later + step, ... (no multiple of span lies strictly
  // inside chunk c, so the first one after c0 is the first after the start)
  const int step = p.span / p.chunk;
  const int first_later = (int)(((int64_t)c0 / p.span + 1) * p.span / p.chunk);
  const int c_last = (we - 1) / p.chunk;
  const int64_t tile = (int64_t)kBlkH * tw;
  const int64_t slot_stride = (int64_t)p.n_ct * tile;
  const float *first = p.parts + ((int64_t)2 * c + role) * slot_stride + ct * tile;
  const float *later = p.parts + (int64_t)ct * tile;
  const int64_t N = p.N;
  for (int k = lane; k < kBlkH * tw; k += kWave) {
    float sum = first[k];
#pragma unroll 8
    for (int cc = first_later; cc <= c_last; cc += step) sum += later[(int64_t)2 * cc * slot_stride + k];
    const int64_t row = (int64_t)w * kBlkH + k / tw;
    const int col = ct * tw + k % tw;
    if (row < p.M && col < p.N) __builtin_nontemporal_store(sum, p.C + row * N + col);
  }
}

// ------------------------------------------------------------ host side

typedef void (*KernelFn)(const int32_t *, const int32_t *, const uint8_t *, const int32_t *,
                         const int32_t *, const float *, const void *, float *, SpmmParams);

struct Variant {
  int prec;  // DTC_PRECISION_* or a bf16-B scheme
  int vec;
  int nsub;
};

template <int PREC, int VEC, int NSUB>
KernelFn pick_schedule(int schedule) {
  switch (schedule) {
    case 0: return spmm_kernel<PREC, VEC, NSUB, false>;
    case 1: return spmm_kernel<PREC, VEC, NSUB, true>;
    default: return spmm_kernel_persistent<PREC, VEC, NSUB>;
  }
}

template <int PREC>
KernelFn pick_shape(int vec, int nsub, int schedule) {
  if (vec == 4 && nsub == 2) return pick_schedule<PREC, 4, 2>(schedule);
  if (vec == 4) return pick_schedule<PREC, 4, 1>(schedule);
  if (vec == 2) return pick_schedule<PREC, 2, 1>(schedule);
  return pick_schedule<PREC, 1, 1>(schedule);
}

template <int PREC>
KernelFn pick_shape_bf16b(int vec, int schedule) {
  if (vec == 8) return pick_schedule<PREC, 8, 1>(schedule);
  return pick_schedule<PREC, 4, 1>(schedule);
}

KernelFn pick_kernel(const Variant &v, int schedule) {
  switch (v.prec) {
    case DTC_PRECISION_FP32: return pick_shape<DTC_PRECISION_FP32>(v.vec, v.nsub, schedule);
    case DTC_PRECISION_BF16: return pick_shape<DTC_PRECISION_BF16>(v.vec, v.nsub, schedule);
    case DTC_PRECISION_FP16: return pick_shape<DTC_PRECISION_FP16>(v.vec, v.nsub, schedule);
    case DTC_PRECISION_BF16X3: return pick_shape<DTC_PRECISION_BF16X3>(v.vec, v.nsub, schedule);
    case kBf16B_A3: return pick_shape_bf16b<kBf16B_A3>(v.vec, schedule);
    case kBf16B_A1: return pick_shape_bf16b<kBf16B_A1>(v.vec, schedule);
    default: return pick_shape<DTC_PRECISION_BF16X6>(v.vec, v.nsub, schedule);
  }
}

// Kernel scheme and lane shape from the options, N and operand alignment.
// vec 0 for float32 B: 2 floats per lane (32-column tiles) in the balanced
// schedule, 4 in the unbalanced one -- the widths measured fastest for each
// across the selector's 2,000 training matrices -- or the widest narrower
// width N and the alignment allow.
dtc_status_t resolve_shape(const dtc_spmm_options_t *opt, int32_t n, const void *b, const float *c,
                           bool balanced, Variant *out) {
  const int prec = opt ? opt->precision : DTC_PRECISION_FP32;
  const int bdt = opt ? opt->b_dtype : DTC_B_FLOAT32;
  int vec = opt ? opt->vec : 0;
  int nsub = opt ? opt->subtiles : 0;
  if (bdt == DTC_B_BFLOAT16) {
    // bf16 B: 4 or 8 bf16 per lane (8 or 16 bytes), one subtile.
    if (vec != 0 && vec != 4 && vec != 8)
      return fail(DTC_ERROR_INVALID_VALUE, "vec must be 0, 4 or 8 for bfloat16 B");
    auto fits = [&](int v) {
      return n % v == 0 && ((uintptr_t)b % (2 * v)) == 0 && ((uintptr_t)c % (4 * v)) == 0;
    };
    // Balanced: 128-column tiles, faster than 64 on R-MAT scales 18 and 20 (0.44 vs
    // 0.49 ms and 1.89 vs 2.09 ms at chunk 64, N = 256); unbalanced: 64.
    if (vec == 0) vec = (balanced && fits(8)) ? 8 : 4;
    if (!fits(vec))
      return fail(DTC_ERROR_INVALID_VALUE,
                  "bfloat16 B needs N to be a multiple of 4 and 8-byte aligned B/C rows");
    out->prec = (prec == DTC_PRECISION_BF16) ? kBf16B_A1 : kBf16B_A3;
    out->vec = vec;
    out->nsub = 1;
    return DTC_SUCCESS;
  }
  if (bdt != DTC_B_FLOAT32) return fail(DTC_ERROR_INVALID_VALUE, "unknown b_dtype");
  if (vec != 0 && vec != 1 && vec != 2 && vec != 4)
    return fail(DTC_ERROR_INVALID_VALUE, "vec must be 0, 1, 2 or 4");
  if (nsub < 0 || nsub > 2) return fail(DTC_ERROR_INVALID_VALUE, "subtiles must be 0, 1 or 2");
  auto fits = [&](int v) {
    return n % v == 0 && ((uintptr_t)b % (4 * v)) == 0 && ((uintptr_t)c % (4 * v)) == 0;
  };
  if (vec == 0) {
    const int pref = balanced ? 2 : 4;
    vec = (pref >= 4 && fits(4)) ? 4 : (fits(2) ? 2 : 1);
  }
  if (!fits(vec))
    return fail(DTC_ERROR_INVALID_VALUE,
                "vec " + std::to_string(vec) + " needs N and the B/C base addresses to be multiples of it");
  if (nsub == 0) nsub = 1;  // one 16*vec column tile per wave measured fastest on R-MAT and SBM
  if (nsub == 2 && vec != 4) nsub = 1;
  out->prec = prec;
  out->vec = vec;
  out->nsub = nsub;
  return DTC_SUCCESS;
}

dtc_status_t check_common(const uint8_t *tile, const int32_t *off, const int32_t *atox,
                          int32_t num_blocks, int32_t nnz, int32_t num_nodes, const void *b,
                          int32_t b_rows, int32_t n, float *c, const dtc_spmm_options_t *opt) {
  if (num_nodes < 0 || b_rows < 0 || num_blocks < 0 || nnz < 0)
    return fail(DTC_ERROR_INVALID_VALUE, "sizes must be >= 0");
  if (n <= 0) return fail(DTC_ERROR_INVALID_VALUE, "N (feature width) must be >= 1");
  if (!off) return fail(DTC_ERROR_INVALID_VALUE, "TCblock_offset is NULL");
  if (num_nodes > 0 && !c) return fail(DTC_ERROR_INVALID_VALUE, "C is NULL");
  if (b_rows > 0 && !b) return fail(DTC_ERROR_INVALID_VALUE, "B is NULL");
  if (num_blocks > 0) {
    if (!tile || !atox)
      return fail(DTC_ERROR_INVALID_VALUE, "TCblocktile_id and sparse_AToX_idx are required");
    if (nnz < 1) return fail(DTC_ERROR_INVALID_VALUE, "TC blocks without stored elements");
    if ((uintptr_t)atox % 16 != 0)
      return fail(DTC_ERROR_INVALID_VALUE, "sparse_AToX_idx must be 16-byte aligned");
  }
  const int prec = opt ? opt->precision : DTC_PRECISION_FP32;
  if (prec < DTC_PRECISION_FP32 || prec > DTC_PRECISION_BF16X6)
    return fail(DTC_ERROR_INVALID_VALUE, "unknown precision");
  if ((int64_t)num_nodes * n > ((int64_t)1 << 40))
    return fail(DTC_ERROR_UNSUPPORTED, "output too large");
  return DTC_SUCCESS;
}

// Default balanced chunk: the largest power of two in [32, 512] that is at
// most T / 8192, i.e. about 8K chunks (x column tiles) of work units once
// the matrix is large enough.  On the selector's 400 held-out matrices this
// rule is 12% slower than the per-matrix best variant on average, a fixed
// 256 122% (small matrices leave most CUs idle with long chunks).
int default_chunk_blocks(int32_t num_blocks) {
  int c = 32;
  while (c < 512 && (int64_t)c * 2 * 8192 <= (int64_t)num_blocks) c *= 2;
  return c;
}

// Cut span of the balanced schedule: options->cut_span_blocks, a positive
// multiple of the chunk, or the chunk itself when 0.
dtc_status_t resolve_cut_span(const dtc_spmm_options_t *opt, int chunk, int *span) {
  const int s = opt ? opt->cut_span_blocks : 0;
  if (s < 0) return fail(DTC_ERROR_INVALID_VALUE, "cut_span_blocks must be >= 0");
  if (s > 0 && s % chunk != 0)
    return fail(DTC_ERROR_INVALID_VALUE, "cut_span_blocks must be a multiple of the chunk (" +
                                             std::to_string(chunk) + " blocks)");
  *span = s > 0 ? s : chunk;
  return DTC_SUCCESS;
}

// Balanced workspace: the persistent work counter, then one int2 run per
// chunk; with the deterministic merge, then two partial-tile slots per chunk
// of 16 rows x N floats (N rounded up to 128, a multiple of every
// column-tile width).  Without a chunk size, one run per block covers every
// chunk size.
constexpr size_t kQueueBytes = 256;
size_t balanced_workspace_bytes(int32_t num_blocks) {
  return kQueueBytes + sizeof(int2) * (size_t)std::max<int32_t>(num_blocks, 0);
}
int64_t num_chunks(int32_t num_blocks, int32_t chunk) {
  return ceil_div(std::max<int32_t>(num_blocks, 0), std::max<int32_t>(chunk, 1));
}
size_t balanced_parts_offset(int32_t num_blocks, int32_t chunk) {
  return align_up(kQueueBytes + sizeof(int2) * (size_t)num_chunks(num_blocks, chunk), 256);
}
size_t balanced_workspace_bytes_for(int32_t num_blocks, int32_t n, int32_t chunk, bool det) {
  if (!det) return kQueueBytes + sizeof(int2) * (size_t)num_chunks(num_blocks, chunk);
  const int64_t npad = ceil_div(std::max<int32_t>(n, 1), 128) * 128;
  return balanced_parts_offset(num_blocks, chunk) +
         (size_t)(2 * num_chunks(num_blocks, chunk)) * kBlkH * (size_t)npad * sizeof(float);
}

__global__ __launch_bounds__(256) void k_zero(float *c, int64_t count) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) c[i] = 0.f;
}

// C = 0 when B has no rows (every column of A is out of range).
dtc_status_t zero_output(float *c, int32_t m, int32_t n, hipStream_t stream) {
  const int64_t count = (int64_t)m * n;
  if (count == 0) return DTC_SUCCESS;
  const unsigned grid = (unsigned)std::min<int64_t>(65536, ceil_div(count, 256));
  k_zero<<<grid, 256, 0, stream>>>(c, count);
  DTC_LAUNCH_CHECK("zero_output");
  return DTC_SUCCESS;
}

}  // namespace
// TC-GNN's SpMM (DTCSpMM.cpp:97-115, DTCSpMM_kernel.cu:703-808 and
// :1566-1685), the comparison baseline without TC-block compaction.  Its
// format is ME-TCF's phase 1 only: blockPartition, edgeToColumn, edgeToRow.
// One unit per (row window, column tile); for every group of 4 blocks (32
// condensed columns) the wave rescans ALL of the window's edges, scattering
// those whose condensed column falls in the group into the LDS A tile and
// recording each condensed column's original column (TC-GNN's per-block
// sparse_AToX), then gathers B and runs the MFMAs as the ME-TCF kernels do.
// The rescans -- window edges times groups -- are what ME-TCF's compaction
// removes.
template <int VEC>
__global__ __launch_bounds__(256) void spmm_tcgnn_kernel(
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ col,
    const int32_t *__restrict__ e2c, const int32_t *__restrict__ e2r,
    const int32_t *__restrict__ bpart, const float *__restrict__ vals, int32_t W,
    SpmmParams p) {
  __shared__ __attribute__((aligned(16))) float s_tile[kWavesPerBlock][kTileFloats];
  __shared__ int32_t s_atox[kWavesPerBlock][kGroup * kBlkW];
  const int wave = wave_uniform(threadIdx.x >> 6);
  const int lane = lane_id();
  const int ct = blockIdx.x % p.n_ct;
  const int w = (blockIdx.x / p.n_ct) * kWavesPerBlock + wave;
  if (w >= W) return;
  const int colbase = ct * 16 * VEC + (lane & 15) * VEC;  // this lane's columns, as in run_unit
  const int64_t r0 = (int64_t)w * kBlkH;
  const int ws = wave_uniform(rowptr[min(r0, (int64_t)p.M)]);
  const int we = wave_uniform(rowptr[min(r0 + kBlkH, (int64_t)p.M)]);
  const int nb = wave_uniform(bpart[w]);
  float *tile = s_tile[wave];
  int32_t *atox = s_atox[wave];
  floatx4 acc[1][VEC];
  zero_acc<VEC, 1>(acc);
  const int g = lane >> 4;
  const int q = lane & 15;
  for (int k0 = 0; k0 < nb; k0 += kGroup) {
    floatx4 *t4 = reinterpret_cast<floatx4 *>(tile);
    const floatx4 z4 = {0.f, 0.f, 0.f, 0.f};
    t4[lane * 2] = z4;
    t4[lane * 2 + 1] = z4;
    if (lane < kGroup * kBlkW) atox[lane] = 0;  // padding slots gather row 0 times 0
    wave_lds_sync();
    for (int e = ws + lane; e < we; e += kWave) {
      const int c = e2c[e];
      const int gg = (c >> 3) - k0;
      if (gg >= 0 && gg < kGroup) {
        const int slot = c & (kBlkW - 1);
        atomicAdd(&tile[gg * (kBlkH * kBlkW) + (e2r[e] & (kBlkH - 1)) * kBlkW + slot],
                  vals ? vals[e] : 1.0f);  // repeated (row, column) pairs add up
        atox[gg * kBlkW + slot] = col[e];
      }
    }
    wave_lds_sync();
    float a[8];
    const floatx4 lo = t4[(g * (kBlkH * kBlkW) + q * kBlkW) / 4];
    const floatx4 hi = t4[(g * (kBlkH * kBlkW) + q * kBlkW) / 4 + 1];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a[j] = lo[j];
      a[j + 4] = hi[j];
    }
    int rows[8];
    const bool gv = g < nb - k0;
#pragma unroll
    for (int j = 0; j < 8; ++j) rows[j] = gv ? atox[g * kBlkW + j] : -1;
    wave_lds_sync();
    breg_t<DTC_PRECISION_FP32, VEC> breg[1][8];
    gather_b<DTC_PRECISION_FP32, VEC, 1>(p, rows, colbase, breg);
    mfma_group<DTC_PRECISION_FP32, VEC, 1>(a, breg, acc, true);
  }
  store_window<VEC, 1>(p, w, colbase, acc);
}

}  // namespace dtc

using namespace dtc;

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
dtc_status_t noalloc_scratch(void **out) {
  int dev = 0;
  DTC_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(DTC_ERROR_UNSUPPORTED, "device ordinal " + std::to_string(dev) + " beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(DTC_ERROR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return DTC_SUCCESS;
}

}  // namespace

extern "C" {

dtc_status_t dtc_spmm_tcgnn(const int32_t *node_pointer, const int32_t *edge_list,
                            const int32_t *block_partition, const int32_t *edge_to_column,
                            const int32_t *edge_to_row, const float *values, int32_t num_nodes,
                            int32_t nnz, const float *b, int32_t b_rows, int32_t n, float *c,
                            const dtc_spmm_options_t *options, dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  if (num_nodes < 0 || nnz < 0 || b_rows < 0)
    return fail(DTC_ERROR_INVALID_VALUE, "sizes must be >= 0");
  if (n <= 0) return fail(DTC_ERROR_INVALID_VALUE, "N (feature width) must be >= 1");
  if (options && (options->precision != DTC_PRECISION_FP32 || options->b_dtype != DTC_B_FLOAT32))
    return fail(DTC_ERROR_UNSUPPORTED, "the TC-GNN baseline runs float32 only");
  if (num_nodes == 0) return DTC_SUCCESS;
  if (!node_pointer || !block_partition || !c)
    return fail(DTC_ERROR_INVALID_VALUE, "node_pointer, block_partition and C are required");
  if (nnz > 0 && (!edge_list || !edge_to_column || !edge_to_row))
    return fail(DTC_ERROR_INVALID_VALUE, "edge arrays are required when nnz > 0");
  if (b_rows > 0 && !b) return fail(DTC_ERROR_INVALID_VALUE, "B is NULL");
  dtc_spmm_options_t o{};
  o.precision = DTC_PRECISION_FP32;
  o.vec = options ? options->vec : 0;
  o.subtiles = 1;
  Variant var;
  dtc_status_t st = resolve_shape(&o, n, b, c, false, &var);
  if (st != DTC_SUCCESS) return st;
  const int32_t W = (int32_t)ceil_div(num_nodes, kBlkH);
  if (b_rows == 0) return zero_output(c, num_nodes, n, stream);
  SpmmParams p{};
  p.B = b;
  p.C = c;
  p.M = num_nodes;
  p.b_rows = b_rows;
  p.N = n;
  p.n_ct = (int32_t)ceil_div(n, 16 * var.vec);
  const int64_t grid = (int64_t)p.n_ct * ceil_div(W, kWavesPerBlock);
  if (grid > INT32_MAX) return fail(DTC_ERROR_UNSUPPORTED, "grid too large");
  switch (var.vec) {
    case 4:
      spmm_tcgnn_kernel<4><<<(unsigned)grid, 256, 0, stream>>>(node_pointer, edge_list, edge_to_column,
                                                               edge_to_row, block_partition, values, W, p);
      break;
    case 2:
      spmm_tcgnn_kernel<2><<<(unsigned)grid, 256, 0, stream>>>(node_pointer, edge_list, edge_to_column,
                                                               edge_to_row, block_partition, values, W, p);
      break;
    default:
      spmm_tcgnn_kernel<1><<<(unsigned)grid, 256, 0, stream>>>(node_pointer, edge_list, edge_to_column,
                                                               edge_to_row, block_partition, values, W, p);
  }
  DTC_LAUNCH_CHECK("spmm_tcgnn");
  return DTC_SUCCESS;
}

dtc_status_t dtc_spmm_windows(const int32_t *row_window_offset, const uint8_t *tcblocktile_id,
                              const int32_t *tcblock_offset, const int32_t *sparse_atox,
                              const float *tcblock_values, int32_t num_windows,
                              int32_t num_blocks, int32_t nnz,
                              int32_t num_nodes, const void *b, int32_t b_rows, int32_t n,
                              float *c, const dtc_spmm_options_t *options,
                              dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  dtc_status_t st = check_common(tcblocktile_id, tcblock_offset, sparse_atox, num_blocks, nnz,
                                 num_nodes, b, b_rows, n, c, options);
  if (st != DTC_SUCCESS) return st;
  if (!row_window_offset) return fail(DTC_ERROR_INVALID_VALUE, "Rowwindow_offset is NULL");
  if (num_windows < 0 || num_blocks < 0 || nnz < 0)
    return fail(DTC_ERROR_INVALID_VALUE, "num_windows, num_blocks and nnz must be >= 0");
  if (num_blocks > 0 && (!tcblocktile_id || !sparse_atox))
    return fail(DTC_ERROR_INVALID_VALUE, "TCblocktile_id and sparse_AToX_idx are required");
  if ((int64_t)num_windows * kBlkH < num_nodes)
    return fail(DTC_ERROR_INVALID_VALUE,
                "Rowwindow_offset has " + std::to_string(num_windows) + " windows, fewer than num_nodes/16");
  Variant var;
  st = resolve_shape(options, n, b, c, false, &var);
  if (st != DTC_SUCCESS) return st;
  // Only windows that hold output rows are processed.
  const int32_t W = (int32_t)std::min<int64_t>(num_windows, ceil_div(num_nodes, kBlkH));
  if (W == 0) return DTC_SUCCESS;
  if (b_rows == 0) return zero_output(c, num_nodes, n, stream);
  SpmmParams p{};
  p.rwo = row_window_offset;
  p.tile = tcblocktile_id;
  p.off = tcblock_offset;
  p.atox = sparse_atox;
  p.vals = tcblock_values;
  p.B = b;
  p.C = c;
  p.num_windows = W;
  p.M = num_nodes;
  p.b_rows = b_rows;
  p.N = n;
  p.n_ct = (int32_t)ceil_div(n, 16 * var.vec * var.nsub);
  p.num_items = W;
  p.num_blocks = num_blocks;
  p.nnz = nnz;
  KernelFn fn = pick_kernel(var, 0);
  const int64_t grid = (int64_t)p.n_ct * ceil_div(W, kWavesPerBlock);
  if (grid > INT32_MAX) return fail(DTC_ERROR_UNSUPPORTED, "grid too large");
  hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(256), 0, stream, p.rwo, p.rowid, p.tile, p.off,
                     p.atox, p.vals, p.B, p.C, p);
  DTC_LAUNCH_CHECK("spmm_windows");
  return DTC_SUCCESS;
}

dtc_status_t dtc_spmm_balanced_workspace_size(int32_t num_blocks, size_t *bytes) {
  clear_error();
  if (!bytes) return fail(DTC_ERROR_INVALID_VALUE, "bytes is NULL");
  if (num_blocks < 0) return fail(DTC_ERROR_INVALID_VALUE, "num_blocks must be >= 0");
  *bytes = balanced_workspace_bytes(num_blocks);
  return DTC_SUCCESS;
}

dtc_status_t dtc_spmm_balanced_workspace_size_ex(int32_t num_blocks, int32_t n,
                                                 const dtc_spmm_options_t *options, size_t *bytes) {
  clear_error();
  if (!bytes) return fail(DTC_ERROR_INVALID_VALUE, "bytes is NULL");
  if (num_blocks < 0) return fail(DTC_ERROR_INVALID_VALUE, "num_blocks must be >= 0");
  if (n < 0) return fail(DTC_ERROR_INVALID_VALUE, "N must be >= 0");
  if (options && options->chunk_blocks < 0)
    return fail(DTC_ERROR_INVALID_VALUE, "chunk_blocks must be >= 0");
  const bool det = options && options->deterministic;
  const int chunk = (options && options->chunk_blocks > 0) ? options->chunk_blocks
                                                           : default_chunk_blocks(num_blocks);
  *bytes = balanced_workspace_bytes_for(num_blocks, n, chunk, det);
  return DTC_SUCCESS;
}

dtc_status_t dtc_spmm_balanced(const int32_t *tcblock_rowid, const uint8_t *tcblocktile_id,
                               const int32_t *tcblock_offset, const int32_t *sparse_atox,
                               const float *tcblock_values, int32_t num_blocks, int32_t nnz,
                               int32_t num_nodes, const void *b, int32_t b_rows, int32_t n,
                               float *c, const dtc_spmm_options_t *options, void *workspace,
                               size_t workspace_bytes, dtc_stream_t stream_) {
  clear_error();
  hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
  dtc_status_t st = check_common(tcblocktile_id, tcblock_offset, sparse_atox, num_blocks, nnz,
                                 num_nodes, b, b_rows, n, c, options);
  if (st != DTC_SUCCESS) return st;
  if (num_blocks > 0 && (!tcblock_rowid || !tcblocktile_id || !sparse_atox))
    return fail(DTC_ERROR_INVALID_VALUE, "TCblock_rowid, TCblocktile_id and sparse_AToX_idx are required");
  Variant var;
  st = resolve_shape(options, n, b, c, true, &var);
  if (st != DTC_SUCCESS) return st;
  const int chunk = (options && options->chunk_blocks > 0) ? options->chunk_blocks
                                                           : default_chunk_blocks(num_blocks);
  int span = chunk;
  st = resolve_cut_span(options, chunk, &span);
  if (st != DTC_SUCCESS) return st;
  const bool persistent = options && options->persistent;
  if (options && (options->deterministic < 0 || options->deterministic > 1))
    return fail(DTC_ERROR_INVALID_VALUE, "deterministic must be 0 or 1");
  const bool det = options && options->deterministic;
  const int32_t W = (int32_t)ceil_div(num_nodes, kBlkH);
  if (W == 0) return DTC_SUCCESS;
  const size_t need = balanced_workspace_bytes_for(num_blocks, n, chunk, det);
  if (num_blocks > 0 && (!workspace || workspace_bytes < need))
    return fail(DTC_ERROR_WORKSPACE,
                "balanced SpMM needs the workspace that dtc_spmm_balanced_workspace_size_ex reports (" +
                    std::to_string(need) + " bytes)");
  if (b_rows == 0 || num_blocks == 0) return zero_output(c, num_nodes, n, stream);
  SpmmParams p{};
  p.rowid = tcblock_rowid;
  p.tile = tcblocktile_id;
  p.off = tcblock_offset;
  p.atox = sparse_atox;
  p.vals = tcblock_values;
  p.B = b;
  p.C = c;
  p.num_windows = W;
  p.num_blocks = num_blocks;
  p.M = num_nodes;
  p.b_rows = b_rows;
  p.N = n;
  p.nnz = nnz;
  p.n_ct = (int32_t)ceil_div(n, 16 * var.vec * var.nsub);
  p.num_items = (int32_t)ceil_div(num_blocks, chunk);
  p.chunk = chunk;
  p.span = span;
  p.queue = persistent ? static_cast<int32_t *>(workspace) : nullptr;
  int2 *runs = reinterpret_cast<int2 *>(static_cast<char *>(workspace) + kQueueBytes);
  p.runs = runs;
  p.parts = det ? reinterpret_cast<float *>(static_cast<char *>(workspace) +
                                            balanced_parts_offset(num_blocks, chunk))
                : nullptr;
  const int64_t prep_waves = p.num_items;
  k_balance_prep<<<(unsigned)ceil_div(prep_waves, kWavesPerBlock), 256, 0, stream>>>(
      tcblock_rowid, num_blocks, num_nodes, n, chunk, span, p.num_items, c, p.queue, runs, det ? 0 : 1);
  DTC_LAUNCH_CHECK("balance_prep");
  if (!persistent) {
    KernelFn fn = pick_kernel(var, 1);
    const int64_t grid = (int64_t)p.n_ct * ceil_div(p.num_items, kWavesPerBlock);
    if (grid > INT32_MAX) return fail(DTC_ERROR_UNSUPPORTED, "grid too large");
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(256), 0, stream, p.rwo, p.rowid, p.tile, p.off,
                     p.atox, p.vals, p.B, p.C, p);
  } else {
    int dev = 0, cus = 0;
    DTC_HIP_CHECK(hipGetDevice(&dev));
    DTC_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t units = (int64_t)p.num_items * p.n_ct;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(ceil_div(units, kWavesPerBlock),
                                                                (int64_t)cus * 8));
    KernelFn fn = pick_kernel(var, 2);
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(256), 0, stream, p.rwo, p.rowid, p.tile, p.off,
                     p.atox, p.vals, p.B, p.C, p);
  }
  DTC_LAUNCH_CHECK("spmm_balanced");
  if (det) {
    const int64_t grid = (int64_t)p.n_ct * ceil_div(p.num_items, kWavesPerBlock);
    k_balance_fixup<<<(unsigned)grid, 256, 0, stream>>>(tcblock_rowid, p, 16 * var.vec * var.nsub);
    DTC_LAUNCH_CHECK("balance_fixup");
  }
  return DTC_SUCCESS;
}

// dtc_spmm_balanced with the library's fixed scratch array as its workspace:
// no hipMalloc/hipFree and no allocating call.  Everything the kernels read
// from the scratch is written earlier in the same call (chunk runs and the
// persistent counter by the preparation kernel, partial tiles by the SpMM
// kernel), so nothing needs clearing between calls.
dtc_status_t dtc_spmm_balanced_noalloc(const int32_t *tcblock_rowid, const uint8_t *tcblocktile_id,
                                       const int32_t *tcblock_offset, const int32_t *sparse_atox,
                                       const float *tcblock_values, int32_t num_blocks, int32_t nnz,
                                       int32_t num_nodes, const void *b, int32_t b_rows, int32_t n,
                                       float *c, const dtc_spmm_options_t *options,
                                       dtc_stream_t stream) {
  clear_error();
  if (num_blocks < 0 || n < 0) return fail(DTC_ERROR_INVALID_VALUE, "num_blocks and N must be >= 0");
  if (options && options->chunk_blocks < 0)
    return fail(DTC_ERROR_INVALID_VALUE, "chunk_blocks must be >= 0");
  const bool det = options && options->deterministic;
  const int chunk = (options && options->chunk_blocks > 0) ? options->chunk_blocks
                                                           : default_chunk_blocks(num_blocks);
  const size_t need = balanced_workspace_bytes_for(num_blocks, n, chunk, det);
  if (need > kNoAllocScratchBytes)
    return fail(DTC_ERROR_UNSUPPORTED,
                "this SpMM needs " + std::to_string(need) + " bytes of workspace, more than the " +
                    std::to_string(kNoAllocScratchBytes) +
                    " bytes of the allocation-free entry point; use dtc_spmm_balanced with a workspace");
  void *ws = nullptr;
  dtc_status_t st = noalloc_scratch(&ws);
  if (st != DTC_SUCCESS) return st;
  return dtc_spmm_balanced(tcblock_rowid, tcblocktile_id, tcblock_offset, sparse_atox, tcblock_values,
                           num_blocks, nnz, num_nodes, b, b_rows, n, c, options, ws, kNoAllocScratchBytes,
                           stream);
}

}  // extern "C"
