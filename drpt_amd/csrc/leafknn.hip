// Stage (c): leaf-local brute-force k-NN on gfx950.
//
// Replaces MDRPT::calculate_nns + MathOp::calculate_distance
// (cpp/src/drpt/algo/mdrpt.cpp:123-212, math_operations.cpp:385-396): every
// ordered pair of leafmates, per-row ordering, head kept.
//
// Exact distance (the ranking every path ends with):
//   D(x, y) = fma-chain over k = 0..d-1 of t_k = fl(x_k - y_k): acc = fma(t_k, t_k, acc)
// Results per row: the K nearest non-self leafmates ordered by (D, index).
//
// Paths:
//   FP32 (k_knn_exact)  D for all m^2 pairs, 64x64 tiles, 4x4 register
//         blocking (vector ALU), one 64-row block per work item.
//   BF16 (k_knn_lowp)   Gram X~ X~^T of the bf16-rounded rows on the MFMA cores
//         (v_mfma_f32_32x32x16_bf16, FP32 accumulate) gives approximate distances
//         a; a certified lower bound LB(a) <= D filters the leafmates, and
//         every distance that reaches a list is the exact D above:
//           tree mode  per-tree exact lists (bound pass + filter pass);
//           cand mode  the k' widened candidates of each row (first trees);
//           pool mode  filter against an external bound, survivors to pools.
//   k_row_merge         per row: distinct candidate ids, exact D, K best merged
//         with a known list (the graph build's first and last steps).
//   k_knn_fallback / k_row_exact  exact redo of single rows (list or pool
//         overflow, leaves too large to stage).
//
// Roofline: FP32 path vector-FP32 bound (3 flops per dim per pair); BF16 path
// MFMA bound, algorithmic flops 2*d*sum(m^2) per pass.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "internal.hpp"
#include "lowp_common.hpp"

namespace drpt {

namespace {

using namespace lowp;


constexpr int kThreads = 256;
constexpr int kRows = 64;
constexpr int kCols = 64;

__device__ __forceinline__ bool better(float d1, int i1, float d2, int i2) {
  return d1 < d2 || (d1 == d2 && i1 < i2);
}

// Register-resident sorted list of the KP best (distance, index) pairs.
template <int KP>
struct TopK {
  float d[KP];
  int i[KP];
  __device__ __forceinline__ void init() {
#pragma unroll
    for (int s = 0; s < KP; ++s) {
      d[s] = INFINITY;
      i[s] = INT_MAX;
    }
  }
  __device__ __forceinline__ void insert(float dv, int iv) {
    if (!better(dv, iv, d[KP - 1], i[KP - 1])) return;
    d[KP - 1] = dv;
    i[KP - 1] = iv;
#pragma unroll
    for (int s = KP - 1; s > 0; --s) {
      if (better(d[s], i[s], d[s - 1], i[s - 1])) {
        float td = d[s];
        d[s] = d[s - 1];
        d[s - 1] = td;
        int ti = i[s];
        i[s] = i[s - 1];
        i[s - 1] = ti;
      } else {
        break;
      }
    }
  }
  // insert unless the index is already held (equal indices carry equal distances)
  __device__ __forceinline__ void insert_unique(float dv, int iv) {
    if (!better(dv, iv, d[KP - 1], i[KP - 1])) return;
#pragma unroll
    for (int s = 0; s < KP; ++s)
      if (i[s] == iv) return;
    insert(dv, iv);
  }
  // write the first K entries (K <= KP) as (index or -1, distance)
  __device__ __forceinline__ void store(int K, int32_t* oi, float* od) const {
#pragma unroll
    for (int s = 0; s < KP; ++s) {
      if (s < K) {
        oi[s] = (i[s] == INT_MAX) ? -1 : i[s];
        od[s] = d[s];
      }
    }
  }
  __device__ __forceinline__ float dist_at(int k) const {
    float r = INFINITY;
#pragma unroll
    for (int s = 0; s < KP; ++s) r = (s == k) ? d[s] : r;
    return r;
  }
};

template <typename XT>
__device__ __forceinline__ float exact_dist(const XT* __restrict__ X, int64_t a, int64_t b, int d) {
  const XT* xa = X + a * d;
  const XT* xb = X + b * d;
  float acc = 0.0f;
  for (int k = 0; k < d; ++k) {
    float t = __fsub_rn((float)xa[k], (float)xb[k]);
    acc = fmaf(t, t, acc);
  }
  return acc;
}

// v[0..7] = X[id][k .. k+7] as float (zeros past d or for id < 0); vector
// loads when rows are 4-element aligned.
template <typename XT>
__device__ __forceinline__ void load8(const XT* __restrict__ X, int id, int d, int k, float* v) {
#pragma unroll
  for (int u = 0; u < 8; ++u) v[u] = 0.0f;
  if (id < 0) return;
  const XT* src = X + (int64_t)id * d;
  if ((d & 3) == 0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int kk = k + 4 * h;
      if (kk < d) {
        if (sizeof(XT) == 4) {
          const float4 f = *reinterpret_cast<const float4*>(src + kk);
          v[4 * h] = f.x;
          v[4 * h + 1] = f.y;
          v[4 * h + 2] = f.z;
          v[4 * h + 3] = f.w;
        } else {
          const uchar4 b = *reinterpret_cast<const uchar4*>(src + kk);
          v[4 * h] = (float)b.x;
          v[4 * h + 1] = (float)b.y;
          v[4 * h + 2] = (float)b.z;
          v[4 * h + 3] = (float)b.w;
        }
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k + u < d) v[u] = (float)src[k + u];
  }
}

__device__ __forceinline__ int next_item(unsigned* queue, int* s_item) {
  if (threadIdx.x == 0) *s_item = (int)atomicAdd(queue, 1u);
  __syncthreads();
  int item = *s_item;
  __syncthreads();
  return item;
}

// ---------------------------------------------------------------------------
// Wave-distributed sorted queues. For one row, lane l holds the l-th best
// (value, index) pair; better() is a strict total order. Candidates are
// offered one per lane: a ballot filters them against the queue's entry `kth`
// and each survivor is inserted with one DPP lane shift (wave_shr:1: lane l
// takes lane l-1's value). Uniform reads use v_readlane, so nothing goes
// through the LDS crossbar.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float lane_read(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ int lane_read(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float shift_up1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138,
                                                    0xf, 0xf, false));
}
__device__ __forceinline__ int shift_up1(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false);
}

__device__ __forceinline__ void wq_offer(float& qd, int& qi, float v, int vi, int kth, int lane) {
  float td = lane_read(qd, kth);
  int ti = lane_read(qi, kth);
  uint64_t mask = __ballot(better(v, vi, td, ti));
  while (mask) {
    const int src = __builtin_ctzll(mask);
    mask &= mask - 1;
    const float x = lane_read(v, src);
    const int xi = lane_read(vi, src);
    if (better(x, xi, td, ti)) {
      const float pd = shift_up1(qd);
      const int pi = shift_up1(qi);
      const bool here = better(x, xi, qd, qi);
      const bool shift = lane > 0 && better(x, xi, pd, pi);
      qd = here ? (shift ? pd : x) : qd;
      qi = here ? (shift ? pi : xi) : qi;
      td = lane_read(qd, kth);
      ti = lane_read(qi, kth);
    }
  }
}

// Lanes < cnt hold (dv, iv) pairs (unused lanes (+inf, INT_MAX)), distinct
// indices. Writes the K best in (value, index) order to oi/od (padded with
// -1/+inf) by rank counting; returns the K-th best value (+inf if fewer).
__device__ __forceinline__ float rank_store(float dv, int iv, int cnt, int K, int lane,
                                            int32_t* __restrict__ oi, float* __restrict__ od) {
  int rank = 0;
  for (int s = 0; s < cnt; ++s) rank += better(lane_read(dv, s), lane_read(iv, s), dv, iv) ? 1 : 0;
  const bool valid = lane < cnt && iv != INT_MAX;
  if (valid && rank < K) {
    oi[rank] = iv;
    od[rank] = dv;
  }
  const int nv = __popcll(__ballot(valid));
  if (lane >= nv && lane < K) {
    oi[lane] = -1;
    od[lane] = INFINITY;
  }
  const uint64_t m = __ballot(valid && rank == K - 1);
  return m ? lane_read(dv, __builtin_ctzll(m)) : INFINITY;
}

constexpr int kRowsPerWave = kRows / (kThreads / 64);  // 16

// ---------------------------------------------------------------------------
// FP32 exact path: D for every ordered pair of the 64-row block x leaf,
// 64x64 tiles with 4x4 register blocking, each wave keeps the queues of 16 rows.
// ---------------------------------------------------------------------------
template <typename XT>
__global__ void __launch_bounds__(kThreads)
k_knn_exact(const XT* __restrict__ X, int d, const int32_t* __restrict__ perm, int64_t n,
            const LeafItem* __restrict__ items, const int64_t* __restrict__ range, int K,
            int32_t* __restrict__ cand_idx, float* __restrict__ cand_dist,
            unsigned* __restrict__ queue) {
  __shared__ __align__(16) float As[32][kRows + 4];
  __shared__ __align__(16) float Bs[32][kCols + 4];
  __shared__ float Ds[kRows][kCols + 1];
  __shared__ int rowid[kRows];
  __shared__ int colid[kCols];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int wave = tid >> 6, lane = tid & 63;
  float qd[kRowsPerWave];
  int qi[kRowsPerWave];
  const int64_t begin = range[0], end = range[1];
  for (;;) {
    const int64_t item = begin + next_item(queue, &s_item);
    if (item >= end) break;
    const LeafItem it = items[item];
    const int32_t* pl = perm + (int64_t)it.tree * n + it.start;
    const int m = it.size;
    const int rows = min(kRows, it.rows);
    if (tid < kRows) rowid[tid] = tid < rows ? pl[it.row0 + tid] : -1;
#pragma unroll
    for (int rr = 0; rr < kRowsPerWave; ++rr) {
      qd[rr] = INFINITY;
      qi[rr] = INT_MAX;
    }
    for (int c0 = 0; c0 < m; c0 += kCols) {
      const int cols = min(kCols, m - c0);
      if (tid >= 64 && tid < 128) {
        int c = tid - 64;
        colid[c] = c < cols ? pl[c0 + c] : -1;
      }
      __syncthreads();
      float acc[4][4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0f;
      for (int k0 = 0; k0 < d; k0 += 32) {
        {
          const int r = tid >> 2, kb = (tid & 3) * 8;
          float va[8], vb[8];
          load8(X, rowid[r], d, k0 + kb, va);
          load8(X, colid[r], d, k0 + kb, vb);
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            As[kb + u][r] = va[u];
            Bs[kb + u][r] = vb[u];
          }
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < 32; ++kk) {
          const float4 a4 = *reinterpret_cast<const float4*>(&As[kk][ty * 4]);
          const float4 b4 = *reinterpret_cast<const float4*>(&Bs[kk][tx * 4]);
          const float av[4] = {a4.x, a4.y, a4.z, a4.w};
          const float bv[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
              const float t = __fsub_rn(av[a], bv[b]);
              acc[a][b] = fmaf(t, t, acc[a][b]);
            }
        }
        __syncthreads();
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) Ds[ty * 4 + a][tx * 4 + b] = acc[a][b];
      __syncthreads();
      {
        const int id = lane < cols ? colid[lane] : -1;
#pragma unroll
        for (int rr = 0; rr < kRowsPerWave; ++rr) {
          const int r = wave * kRowsPerWave + rr;
          if (r < rows) {
            const bool ok = id >= 0 && id != rowid[r];
            wq_offer(qd[rr], qi[rr], ok ? Ds[r][lane] : INFINITY, ok ? id : INT_MAX, K - 1, lane);
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int rr = 0; rr < kRowsPerWave; ++rr) {
      const int r = wave * kRowsPerWave + rr;
      if (r < rows && lane < K) {
        const int64_t o = ((int64_t)it.tree * n + rowid[r]) * K + lane;
        cand_idx[o] = qi[rr] == INT_MAX ? -1 : qi[rr];
        cand_dist[o] = qd[rr];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Low-precision path: BF16 Gram on the MFMA cores, exact FP32 distances for a
// filtered candidate set.
//
// One work item = one whole leaf (its point ids and |x~|^2 staged in LDS),
// processed in blocks of 64 query rows. For a block, the Gram tile against 64
// leafmates is computed with the leafmates as the MFMA's A operand and the
// query rows as B, so each lane owns one query row (l % 32) and holds 16
// leafmates of it per tile; four threads share a row, each seeing a fixed
// quarter of the columns. Approximate distance a = |x~_i|^2 + |x~_j|^2 - 2 G_ij.
//
// Lower bound. With e = |x~ - x| (rounded up), eta the FP32 error bound of a
// and c = (d + 8) 2^-24, every leafmate j satisfies
//   D(i, j) >= LB(a_ij) = (max(0, sqrt(max(0, a_ij - eta)) - e_i - e_max))^2 (1 - c)
// so for an upper bound B of the K-th exact distance that row i needs,
//   LB(a_ij) <= B  <=>  a_ij <= A_cut(B) = (sqrt(B / (1 - c)) + e_i + e_max)^2 + eta
// and every leafmate with D(i, j) <= B (in particular every member of the
// row's exact top K) passes the filter a_ij <= A_cut(B).
//
// Bound pass (own bound, used for per-tree lists): each thread keeps the
// ceil(k'/4) smallest a of its quarter as packed keys (a's FP32 bits with the
// low 11 bits replaced by the leaf position, compared as integers), so the row
// gets k' widened candidates; their exact D are computed and B = K-th smallest
// of them (an upper bound of the row's K-th exact distance within the leaf).
// Filter pass: leafmates with a <= A_cut(B) that are not already candidates
// are appended to the row's list (capacity kCap), their exact D computed, and
// the K best by (D, index) written out. A row whose list overflows is redone
// by k_knn_fallback. With an external bound (the K-th distance of a list merged
// over other trees) only the filter pass runs: the per-tree list then holds
// the K best of the leafmates that can still enter the merged top K, which
// leaves the merged result unchanged.
// ---------------------------------------------------------------------------
constexpr int kChunkBytes = 256;                  // bytes of a row per staged Gram chunk
constexpr int kTilePitchB = kChunkBytes + 16;     // LDS row pitch in bytes (16-byte aligned rows)
constexpr int kLeafStage = 2048;            // leaves above this go to the exact fallback
constexpr int kCap = 64;                    // candidate slots per query row
constexpr uint16_t kNoPos = 0xFFFF;

// Element types of the candidate Gram: bf16 (v_mfma_f32_32x32x16_bf16, 16 B
// operand per lane per k-step) and fp8 e4m3 with a power-of-two scale per row
// (v_mfma_f32_32x32x16_fp8_fp8, 8 B operand). Lane l supplies row l % 32 and
// elements 16 ks + 8 (l / 32) .. +7 of k-step ks.
struct ElBf16 {
  static constexpr int kBytes = 2;
  static constexpr int kOpBytes = 16;
  typedef uint4 op_t;
  __device__ static op_t lds(const unsigned char* p) { return *reinterpret_cast<const uint4*>(p); }
  __device__ static floatx16 mma(op_t a, op_t b, floatx16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  }
};
struct ElFp8 {
  static constexpr int kBytes = 1;
  static constexpr int kOpBytes = 8;
  typedef uint2 op_t;
  __device__ static op_t lds(const unsigned char* p) { return *reinterpret_cast<const uint2*>(p); }
  __device__ static floatx16 mma(op_t a, op_t b, floatx16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(__builtin_bit_cast(long, a),
                                                      __builtin_bit_cast(long, b), c, 0, 0, 0);
  }
};

// byte offset, within a row, of lane's operand for k-step ks
template <typename EL>
__device__ __forceinline__ int op_offset(int ks, int lane) {
  return ks * 2 * EL::kOpBytes + (lane >> 5) * EL::kOpBytes;
}

// 16 bytes of row id (rb bytes per row) from byte o; zeros past the row or for id < 0
__device__ __forceinline__ uint4 load16(const uint8_t* __restrict__ Xq, int id, int rb, int o) {
  uint4 v = make_uint4(0, 0, 0, 0);
  if (id < 0 || o >= rb) return v;
  const uint8_t* src = Xq + (int64_t)id * rb + o;
  if ((rb & 15) == 0) return *reinterpret_cast<const uint4*>(src);
  uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
  for (int u = 0; u < 16; ++u)
    if (o + u < rb) w[u >> 2] |= (uint32_t)src[u] << (8 * (u & 3));
  return make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ __forceinline__ float4 load4f(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ float4 load4f(const uint8_t* p) {
  const uchar4 b = *reinterpret_cast<const uchar4*>(p);
  return make_float4((float)b.x, (float)b.y, (float)b.z, (float)b.w);
}

// exact D (same fma chain as exact_dist), vector loads when rows are 4-aligned
template <typename XT>
__device__ __forceinline__ float exact_dist4(const XT* __restrict__ X, int64_t a, int64_t b, int d) {
  if ((d & 3) != 0) return exact_dist(X, a, b, d);
  const XT* xa = X + a * d;
  const XT* xb = X + b * d;
  float acc = 0.0f;
#pragma unroll 4
  for (int k = 0; k < d; k += 4) {
    const float4 u = load4f(xa + k), v = load4f(xb + k);
    float t = __fsub_rn(u.x, v.x);
    acc = fmaf(t, t, acc);
    t = __fsub_rn(u.y, v.y);
    acc = fmaf(t, t, acc);
    t = __fsub_rn(u.z, v.z);
    acc = fmaf(t, t, acc);
    t = __fsub_rn(u.w, v.w);
    acc = fmaf(t, t, acc);
  }
  return acc;
}


constexpr int kRing = 3;          // leafmate tile stages in flight (ring path)
constexpr int kRingRow = 256;     // bytes per leafmate row in a ring stage (d <= 128)

// LDS of the low-precision kernel, for NR query rows per block (4 threads per
// row). Tree mode (NR = 64): register-staged double buffer plus the
// candidates' exact distances. Candidate / pool modes (NR = 128): the tile
// ring, or for inputs the ring does not cover three 64-row tile buffers
// (leafmate double buffer; query rows staged in buffers 1-2).
template <bool kTree, int NR>
struct LowpShared;

#define DRPT_LOWP_COMMON(NR)                                                          \
  int lid[kLeafStage];              /* point id of every leaf position */           \
  float lnb[kLeafStage];            /* |x~|^2 of every leaf position */             \
  int8_t lexp[kLeafStage];          /* fp8: scale exponent of every leaf position */ \
  uint16_t cpos[NR][kCap + 2];      /* candidate leaf positions per query row */    \
  int ccnt[NR * 4];                 /* per row (tree) or per (row, quarter) (pool) */ \
  float cacut[NR];                                                                  \
  float red_n[NR / 16], red_e[NR / 16]; /* one per wave */                          \
  int ticket[2];

template <int NR>
struct LowpShared<true, NR> {
  // leafmate tiles (row bytes), double-buffered; for rows longer than
  // kChunkBytes buffer 1 holds the query rows' chunk and buffer 0 the
  // leafmates' (single-buffered)
  unsigned char Bs[2][kCols][kTilePitchB];
  DRPT_LOWP_COMMON(NR)
  float cdist[NR][kCap];            // exact distances of the candidates
  int cpre[NR + 1];
};

constexpr int kKnownLds = 16;  // pool mode: known ids per row staged in LDS (larger K: left to the row merge)
// row pitches in LDS one bank apart (33 / 17 dwords), so the rows a wave
// appends to or scans in the same step hit different banks
constexpr int kKnownPitch = kKnownLds + 1;

template <int NR>
struct LowpShared<false, NR> {
  union {
    unsigned char Bs[1 + NR / kCols][kCols][kTilePitchB];
    unsigned char ring[kRing][kCols * kRingRow];
  };
  DRPT_LOWP_COMMON(NR)
  int known[NR * kKnownPitch];      // pool mode: the block's known ids
  // pool mode: some list is half full after a tile of that parity (a wave may
  // set the next tile's flag while a slower one still reads this one)
  int flush_flag[2];
};
#undef DRPT_LOWP_COMMON

// Wave roles for NR query rows: wave w computes the 32 x 32 block of query
// rows qh .. qh + 31 and tile columns ch .. ch + 31 of each 64-column tile.
template <int NR>
struct WaveRole {
  static constexpr int kGroups = NR / 32;  // query groups of 32 rows
  int qh, ch;
  __device__ explicit WaveRole(int wave) : qh((wave % kGroups) * 32), ch((wave / kGroups) * 32) {}
};

// Query operand of the block (rows of at most kChunkBytes): lane l of wave w
// holds, for every k-step, its operand of query row qh + l % 32, i.e. the
// MFMA's B operand, for all passes (32 VGPRs for either element type).
template <typename EL>
struct QueryOp {
  typename EL::op_t v[kChunkBytes / (2 * EL::kOpBytes)];
};

// query rows staged in buffers 1.. (buffer 0 takes the first leafmate tile
// while other waves may still read their query operands)
template <typename EL, int NR, typename SH>
__device__ __forceinline__ void load_query(SH& S, const uint8_t* __restrict__ Xq, int rb,
                                           int row0, int rows, QueryOp<EL>& qo) {
  constexpr int NT = 4 * NR;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  unsigned char(*As)[kTilePitchB] = S.Bs[1];
#pragma unroll
  for (int s4 = 0; s4 < NR * 16 / NT; ++s4) {
    const int c = tid + NT * s4;
    const int r = c >> 4, part = c & 15;
    *reinterpret_cast<uint4*>(&As[r][part * 16]) =
        load16(Xq, r < rows ? S.lid[row0 + r] : -1, rb, part * 16);
  }
  __syncthreads();
  const WaveRole<NR> role(wave);
#pragma unroll
  for (int ks = 0; ks < kChunkBytes / (2 * EL::kOpBytes); ++ks)
    qo.v[ks] = EL::lds(&As[role.qh + (lane & 31)][op_offset<EL>(ks, lane)]);
  // buffers 1.. are next written after gram_pass's first barrier
}

// One pass of Gram tiles of the query block [row0, row0 + NR) against every
// leafmate tile; epi(c0, acc) sees each thread's 16 accumulators; hook(t) runs
// on every thread after tile t (and a barrier) when another tile follows. With the
// query operand in registers (rows of at most kChunkBytes) the leafmate tiles
// are double-buffered: one barrier per tile, the next tile's loads in flight
// during this tile's MFMAs and epilogue. rb = bytes per row, d = elements.
template <typename EL, int NR, bool kPipe, typename SH, typename Epi, typename Hook>
__device__ __forceinline__ void gram_pass(SH& S, const uint8_t* __restrict__ Xq, int d, int rb,
                                          int m, int row0, int rows, const QueryOp<EL>& qo,
                                          Epi&& epi, Hook&& hook) {
  constexpr int NT = 4 * NR;
  constexpr int kTileLoads = kCols * 16 / NT;  // 16-byte pieces of a tile per thread
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const WaveRole<NR> role(wave);
  const int qh = role.qh, ch = role.ch;
  constexpr int kStepBytes = 2 * EL::kOpBytes;  // bytes of a row per k-step
  if (rb <= kChunkBytes) {
    const int ksteps = (d + 15) / 16;
    uint4 rbv[kTileLoads];
#pragma unroll
    for (int s4 = 0; s4 < kTileLoads; ++s4) {
      const int c = tid + NT * s4;
      const int r = c >> 4, part = c & 15;
      *reinterpret_cast<uint4*>(&S.Bs[0][r][part * 16]) =
          load16(Xq, r < m ? S.lid[r] : -1, rb, part * 16);
    }
    __syncthreads();
    int b = 0;
    for (int c0 = 0; c0 < m; c0 += kCols, b ^= 1) {
      const bool more = c0 + kCols < m;
      if (more) {  // next leafmate tile in flight during this one
#pragma unroll
        for (int s4 = 0; s4 < kTileLoads; ++s4) {
          const int c = tid + NT * s4;
          const int r = c >> 4, part = c & 15;
          const int col = c0 + kCols + r;
          rbv[s4] = load16(Xq, col < m ? S.lid[col] : -1, rb, part * 16);
        }
      }
      floatx16 acc;
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
#pragma unroll
      for (int ks = 0; ks < kChunkBytes / kStepBytes; ++ks) {
        if (ks < ksteps) {
          const typename EL::op_t a = EL::lds(&S.Bs[b][ch + (lane & 31)][op_offset<EL>(ks, lane)]);
          acc = EL::mma(a, qo.v[ks], acc);
        }
      }
      epi(c0, acc);
      if (more) {
#pragma unroll
        for (int s4 = 0; s4 < kTileLoads; ++s4) {
          const int c = tid + NT * s4;
          *reinterpret_cast<uint4*>(&S.Bs[b ^ 1][c >> 4][(c & 15) * 16]) = rbv[s4];
        }
      }
      __syncthreads();
      if (more) hook(c0 / kCols);
    }
    return;
  }
  // longer rows: chunks of kChunkBytes, query chunk in buffers 1.., leafmates in buffer 0
  unsigned char(*As)[kTilePitchB] = S.Bs[1];
  unsigned char(*Bs)[kTilePitchB] = S.Bs[0];
  if constexpr (kPipe && NR == kCols) {
    // Two leafmate tiles per step: every wave multiplies its query operand
    // against columns ch.. of both tiles (one query read for two MFMAs), over
    // chunks of kPairChunk bytes (query and the two tiles fit the tile
    // buffers). The next chunk's pieces (this pair's next chunk, else the next
    // pair's first) are loaded into registers during this chunk's MFMAs and
    // the pair's epilogues.
    constexpr int kPC = 128;                              // bytes of a row per step
    constexpr int kPP = kPC + 16;                         // LDS row pitch
    constexpr int kQL = NR * (kPC / 16) / NT;             // query pieces per thread
    constexpr int kBL = 2 * kCols * (kPC / 16) / NT;      // leafmate pieces per thread
    static_assert((NR + 2 * kCols) * kPP <= (int)sizeof(S.Bs), "pair tiles fit the tile buffers");
    unsigned char* const Aq = &S.Bs[0][0][0];
    unsigned char* const Bl = Aq + NR * kPP;
    uint4 pa[kQL], pb[kBL];
    auto fetch = [&](int c0, int o0) {
#pragma unroll
      for (int s = 0; s < kQL; ++s) {
        const int c = tid + NT * s;
        const int r = c >> 3, part = c & 7;
        pa[s] = load16(Xq, r < rows ? S.lid[row0 + r] : -1, rb, o0 + part * 16);
      }
#pragma unroll
      for (int s = 0; s < kBL; ++s) {
        const int c = tid + NT * s;
        const int r = c >> 3, part = c & 7;
        const int col = c0 + r;
        pb[s] = load16(Xq, col < m ? S.lid[col] : -1, rb, o0 + part * 16);
      }
    };
    fetch(0, 0);
    for (int c0 = 0; c0 < m; c0 += 2 * kCols) {
      floatx16 acc0, acc1;
#pragma unroll
      for (int e = 0; e < 16; ++e) acc0[e] = acc1[e] = 0.0f;
      for (int o0 = 0; o0 < rb; o0 += kPC) {
        __syncthreads();  // every wave done with the previous chunk
#pragma unroll
        for (int s = 0; s < kQL; ++s) {
          const int c = tid + NT * s;
          *reinterpret_cast<uint4*>(Aq + (c >> 3) * kPP + (c & 7) * 16) = pa[s];
        }
#pragma unroll
        for (int s = 0; s < kBL; ++s) {
          const int c = tid + NT * s;
          *reinterpret_cast<uint4*>(Bl + (c >> 3) * kPP + (c & 7) * 16) = pb[s];
        }
        __syncthreads();
        if (o0 + kPC < rb) fetch(c0, o0 + kPC);
        else if (c0 + 2 * kCols < m) fetch(c0 + 2 * kCols, 0);
        const int ksteps = (min(kPC, rb - o0) + kStepBytes - 1) / kStepBytes;
        for (int ks = 0; ks < ksteps; ++ks) {
          const int o = op_offset<EL>(ks, lane);
          const typename EL::op_t bq = EL::lds(Aq + (qh + (lane & 31)) * kPP + o);
          const typename EL::op_t a0 = EL::lds(Bl + (ch + (lane & 31)) * kPP + o);
          const typename EL::op_t a1 = EL::lds(Bl + (kCols + ch + (lane & 31)) * kPP + o);
          acc0 = EL::mma(a0, bq, acc0);
          acc1 = EL::mma(a1, bq, acc1);
        }
      }
      epi(c0, acc0);
      if (c0 + kCols < m) epi(c0 + kCols, acc1);
      if (c0 + 2 * kCols < m) {
        __syncthreads();
        hook(c0 / kCols);
        hook(c0 / kCols + 1);
      }
    }
    __syncthreads();
    return;
  }
  for (int c0 = 0; c0 < m; c0 += kCols) {
    floatx16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
    for (int o0 = 0; o0 < rb; o0 += kChunkBytes) {
      __syncthreads();
      if constexpr (NR == kCols) {  // query and leafmate pieces loaded together
#pragma unroll
        for (int s4 = 0; s4 < kTileLoads; ++s4) {
          const int c = tid + NT * s4;
          const int r = c >> 4, part = c & 15;
          const int col = c0 + r;
          const uint4 va = load16(Xq, r < rows ? S.lid[row0 + r] : -1, rb, o0 + part * 16);
          const uint4 vb = load16(Xq, col < m ? S.lid[col] : -1, rb, o0 + part * 16);
          *reinterpret_cast<uint4*>(&As[r][part * 16]) = va;
          *reinterpret_cast<uint4*>(&Bs[r][part * 16]) = vb;
        }
      } else {
#pragma unroll
        for (int s4 = 0; s4 < NR * 16 / NT; ++s4) {
          const int c = tid + NT * s4;
          const int r = c >> 4, part = c & 15;
          *reinterpret_cast<uint4*>(&As[r][part * 16]) =
              load16(Xq, r < rows ? S.lid[row0 + r] : -1, rb, o0 + part * 16);
        }
#pragma unroll
        for (int s4 = 0; s4 < kTileLoads; ++s4) {
          const int c = tid + NT * s4;
          const int r = c >> 4, part = c & 15;
          const int col = c0 + r;
          *reinterpret_cast<uint4*>(&Bs[r][part * 16]) =
              load16(Xq, col < m ? S.lid[col] : -1, rb, o0 + part * 16);
        }
      }
      __syncthreads();
      // k-steps covering the chunk's elements (zeros past the row end)
      const int ksteps = (min(kChunkBytes, rb - o0) + kStepBytes - 1) / kStepBytes;
      for (int ks = 0; ks < ksteps; ++ks) {
        const typename EL::op_t a = EL::lds(&Bs[ch + (lane & 31)][op_offset<EL>(ks, lane)]);
        const typename EL::op_t bq = EL::lds(&As[qh + (lane & 31)][op_offset<EL>(ks, lane)]);
        acc = EL::mma(a, bq, acc);
      }
    }
    epi(c0, acc);
    if (c0 + kCols < m) {
      __syncthreads();
      hook(c0 / kCols);
    }
  }
  __syncthreads();
}



// Ring path of the Gram pass (rows of 16-byte multiples up to 256 bytes):
// leafmate tiles go straight from HBM into LDS (global_load_lds_dwordx4, 16 B
// per lane, lane l of an instruction writes 16 l bytes after its wave-uniform
// base), kRing stages, two tiles in flight during each tile's MFMAs. Stage
// layout: row r of the tile at r * 256 bytes, its 16-byte chunk c at physical
// chunk c ^ (r & 15) (bank spread for the MFMA operand reads); chunks past the
// row's bytes stay zero. Every wave issues exactly 16 / (NR / 16) load
// instructions per tile (each covers 4 whole rows, at least one valid chunk
// per row), so vmcnt <= that count means "all but the newest tile".
template <typename EL, int NR, typename Epi, typename Hook>
__device__ __forceinline__ void gram_pass_ring(LowpShared<false, NR>& S,
                                               const uint8_t* __restrict__ Xq, int d, int rb,
                                               int m, const QueryOp<EL>& qo, Epi&& epi,
                                               Hook&& hook) {
  constexpr int NT = 4 * NR, NW = NT / 64;
  constexpr int kIssues = 16 / NW;  // load instructions per wave per tile
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int ch = WaveRole<NR>(wave).ch;
  const int L = rb >> 4;           // 16-byte chunks per row
  const int ksteps = (d + 15) / 16;
  const int ntiles = (m + kCols - 1) / kCols;
  constexpr int kSteps = kChunkBytes / (2 * EL::kOpBytes);
  __syncthreads();  // query staging reads done (they share the ring's bytes)
  if (L < 16) {     // zero the chunks no load writes
    for (int j = tid; j < kRing * kCols * 16; j += NT) {
      const int st = j / (kCols * 16), slot = j % (kCols * 16);
      const int r = slot >> 4, c = (slot & 15) ^ (r & 15);
      if (c >= L) *reinterpret_cast<uint4*>(&S.ring[st][slot * 16]) = make_uint4(0, 0, 0, 0);
    }
  }
  auto issue = [&](int j) {
    unsigned char* base = S.ring[j % kRing];
#pragma unroll
    for (int i = 0; i < kIssues; ++i) {
      const int blk = i * NW + wave;  // 64 consecutive slots = rows 4 blk .. 4 blk + 3
      const int slot = blk * 64 + lane;
      const int r = slot >> 4, c = (slot & 15) ^ (r & 15);
      const int col = min(j * kCols + r, m - 1);  // rows past the leaf: any valid row
      const uint32_t dst = __builtin_amdgcn_readfirstlane(
          (uint32_t)(uintptr_t)(lds_ptr_t)(base + blk * 64 * 16));
      if (c < L) lds_dma16(Xq + (int64_t)S.lid[col] * rb + c * 16, dst);
    }
  };
  issue(0);
  if (ntiles > 1) issue(1);
  for (int t = 0; t < ntiles; ++t) {
    if (t + 1 < ntiles) wait_vmcnt<kIssues>(); else wait_vmcnt<0>();
    ring_barrier();  // tile t visible; every wave is done with tile t - 1's stage
    if (t + 2 < ntiles) issue(t + 2);
    if (t > 0) hook(t - 1);  // touches neither the ring nor the operand registers
    const unsigned char* st = S.ring[t % kRing];
    floatx16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
    const int r = ch + (lane & 31);
#pragma unroll
    for (int k4 = 0; k4 < kSteps; k4 += 4) {  // 4 operands read, then 4 MFMAs
      if (k4 >= ksteps) break;
      typename EL::op_t a[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int o = op_offset<EL>(k4 + u, lane);
        const int pc = (o >> 4) ^ (r & 15);
        a[u] = EL::lds(st + r * kRingRow + pc * 16 + (o & 15));
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (k4 + u < ksteps) acc = EL::mma(a[u], qo.v[k4 + u], acc);
    }
    epi(t * kCols, acc);
  }
  __syncthreads();  // ring reused by the next pass
}

// exact D of two pairs at once (independent chains, loads interleaved)
template <typename XT>
__device__ __forceinline__ void exact_dist4_x2(const XT* __restrict__ X, int64_t a0, int64_t b0,
                                               int64_t a1, int64_t b1, int d, float& D0,
                                               float& D1) {
  const XT* xa0 = X + a0 * d;
  const XT* xb0 = X + b0 * d;
  const XT* xa1 = X + a1 * d;
  const XT* xb1 = X + b1 * d;
  float acc0 = 0.0f, acc1 = 0.0f;
#pragma unroll 4
  for (int k = 0; k < d; k += 4) {
    const float4 u0 = load4f(xa0 + k), v0 = load4f(xb0 + k);
    const float4 u1 = load4f(xa1 + k), v1 = load4f(xb1 + k);
    float t = __fsub_rn(u0.x, v0.x);
    acc0 = fmaf(t, t, acc0);
    t = __fsub_rn(u1.x, v1.x);
    acc1 = fmaf(t, t, acc1);
    t = __fsub_rn(u0.y, v0.y);
    acc0 = fmaf(t, t, acc0);
    t = __fsub_rn(u1.y, v1.y);
    acc1 = fmaf(t, t, acc1);
    t = __fsub_rn(u0.z, v0.z);
    acc0 = fmaf(t, t, acc0);
    t = __fsub_rn(u1.z, v1.z);
    acc1 = fmaf(t, t, acc1);
    t = __fsub_rn(u0.w, v0.w);
    acc0 = fmaf(t, t, acc0);
    t = __fsub_rn(u1.w, v1.w);
    acc1 = fmaf(t, t, acc1);
  }
  D0 = acc0;
  D1 = acc1;
}

// Work distribution: the item range is cut into kQueues contiguous chunks and
// block b takes items of chunk b % kQueues from its own ticket counter. Blocks
// are dispatched round-robin over the 8 XCDs, so neighbouring items (the
// leaves of one subtree) share an XCD's L2.
constexpr int kQueues = 8;
constexpr int kSubCap = 16;  // pool mode: list slots per thread (4 threads per row)
// Kernel modes (kMode* in internal.hpp):
//   kModeTree  exact per-tree lists: bound pass, filter pass, exact D, top K
//   kModeCand  bound pass only: the row's k' widened candidates (point ids)
//   kModePool  filter pass against an external bound B (leaf-ordered), skipping
//              the row's known ids; survivors are appended to the row's pool

// Block shape: 4 threads per query row, 64 rows (256 threads, two blocks per
// CU). The passes also take 128 rows (512 threads, one block per CU: each
// staged leafmate tile feeds twice the MFMA work); measured on the 1M x 128,
// 1M x 960 and 10M x 96 workloads that was 0-8% slower, as one block's tile
// barriers then stall every wave of the CU instead of half of them.
template <int MODE>
constexpr int lowp_rows() { return 64; }

template <typename XT, typename EL, int T, int MODE, bool RING>
__global__ void __launch_bounds__(4 * lowp_rows<MODE>(), lowp_rows<MODE>() == 64 ? 2 : 1)
    k_knn_lowp(const KnnArgs A) {
  constexpr int NR = lowp_rows<MODE>();
  constexpr int NT = 4 * NR, NW = NT / 64;
  static_assert(MODE != kModeTree || NR == kRows, "tree mode keeps the 64-row layout");
  constexpr int KP = 4 * T;  // widened candidates per row in the bound pass
  constexpr bool kScaled = EL::kBytes == 1;  // fp8: per-row power-of-two scales
  __shared__ __align__(16) LowpShared<MODE == kModeTree, NR> S;
  const XT* __restrict__ X = static_cast<const XT*>(A.X);
  const uint8_t* __restrict__ Xq = static_cast<const uint8_t*>(A.Xq);
  const int d = A.d, K = A.K;
  const int rb = d * EL::kBytes;  // bytes per quantised row
  const int64_t n = A.n;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const bool a_resident = rb <= kChunkBytes;
  // this thread's query row within the block and quarter of the columns
  const WaveRole<NR> role(wave);
  const int q = role.qh + (lane & 31);
  const int sub = (role.ch / 32) * 2 + (lane >> 5);
  const int cq = role.ch + 4 * (lane >> 5);  // + 8 * (e / 4) + e % 4
  const int qx = blockIdx.x % kQueues;
  const int64_t chunk = (A.range[1] - A.range[0] + kQueues - 1) / kQueues;
  const int64_t begin = A.range[0] + qx * chunk;
  const int64_t end = min(A.range[1], begin + chunk);
  unsigned* myq = A.queue + qx;
  // the leaf-ordered norms and errors were not written: gather them from nbe
  const bool direct = A.lnb_gate != nullptr && *A.lnb_gate == 0;
  if (tid == 0) S.ticket[0] = (int)atomicAdd(myq, 1u);
  if constexpr (MODE == kModePool) {
    if (tid == 0) S.flush_flag[0] = S.flush_flag[1] = 0;
  }
  __syncthreads();
  int64_t item = begin + S.ticket[0];
  for (int parity = 1;; parity ^= 1) {
    if (item >= end) break;
    if (tid == 0) S.ticket[parity] = (int)atomicAdd(myq, 1u);  // next item, fetched early
    const LeafItem it = A.items[item];
    const int64_t lofs = (int64_t)it.tree * n + it.start;  // leaf offset in the per-tree arrays
    const int32_t* pl = A.perm + lofs;
    const int m = it.size;
    const int row_end = it.row0 + it.rows;
    auto lerr_at = [&](int j) -> float { return direct ? A.nbe[S.lid[j]].y : A.lerrT[lofs + j]; };
    if (m < A.size_lo || m > A.size_hi) {  // another launch's size bucket
      __syncthreads();
      item = begin + S.ticket[parity];
      continue;
    }
    if (m > kLeafStage) {  // too large to stage
      if constexpr (MODE == kModeTree) {  // exact path for these rows
        for (int r = it.row0 + tid; r < row_end; r += NT) {
          const int slot = atomicAdd(A.nfallback, 1);
          if (slot < A.fb_cap) {
            A.fallback[2 * slot] = (int32_t)item;
            A.fallback[2 * slot + 1] = r - it.row0;
          }
        }
      } else if constexpr (MODE == kModeCand) {  // no widened candidates
        for (int r = it.row0 + tid; r < row_end; r += NT)
          for (int s = 0; s < KP; ++s)
            A.out_key[((int64_t)(it.tree - A.tree0) * n + pl[r]) * KP + s] = ~0ull;
      } else {  // every row redone exactly over all trees
        for (int r = it.row0 + tid; r < row_end; r += NT) {
          const int id = pl[r];
          if (atomicExch(&A.ovf_flag[id], 1) == 0) A.ovf_list[atomicAdd(A.novf, 1)] = id;
        }
      }
      __syncthreads();
      item = begin + S.ticket[parity];
      continue;
    }
    // stage ids and norms of the leaf (leaf-ordered arrays: coalesced; or,
    // where they were not written, gathered from nbe); leaf maxima
    float mn = 0.0f, me_ = 0.0f;
    for (int j = tid; j < m; j += NT) {
      const int id = pl[j];
      const float2 ne = direct ? A.nbe[id] : make_float2(A.lnbT[lofs + j], A.lerrT[lofs + j]);
      S.lid[j] = id;
      S.lnb[j] = ne.x;
      if constexpr (kScaled) S.lexp[j] = A.lexpT[lofs + j];
      mn = fmaxf(mn, ne.x);
      me_ = fmaxf(me_, ne.y);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      mn = fmaxf(mn, __shfl_xor(mn, off));
      me_ = fmaxf(me_, __shfl_xor(me_, off));
    }
    if (lane == 0) {
      S.red_n[wave] = mn;
      S.red_e[wave] = me_;
    }
    __syncthreads();
    float nmax = S.red_n[0], emax = S.red_e[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) {
      nmax = fmaxf(nmax, S.red_n[w]);
      emax = fmaxf(emax, S.red_e[w]);
    }
    const double nm = sqrt((double)nmax * (1.0 + (d + 4.0) * 5.9604644775390625e-08));

    for (int row0 = it.row0; row0 < row_end; row0 += NR) {
      const int rows = min(NR, row_end - row0);
      const bool qv = q < rows;
      const int qpos = row0 + q;
      const float rn = qv ? S.lnb[qpos] : 0.0f;
      QueryOp<EL> qo;
      if (a_resident) load_query<EL, NR>(S, Xq, rb, row0, rows, qo);
      // fp8: the tile's Gram values scaled back by 2^(e_query + e_leafmate)
      // (exact: powers of two), so the epilogues see x~_i . x~_j as for bf16
      const int eq = kScaled && qv ? (int)S.lexp[qpos] : 0;
      auto scaled = [&](auto& epi) {
        return [&](int c0, floatx16 acc) {
          if constexpr (kScaled) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
              const uint32_t w = *reinterpret_cast<const uint32_t*>(&S.lexp[c0 + cq + 8 * g]);
#pragma unroll
              for (int u = 0; u < 4; ++u)
                acc[4 * g + u] = ldexpf(acc[4 * g + u], eq + (int)(int8_t)(w >> (8 * u)));
            }
          }
          epi(c0, acc);
        };
      };
      // the ring path is a separate instantiation (RING: rows of 16-byte
      // multiples up to 256 bytes, candidate / pool modes), so no
      // register-staged tile load is pending in its loop
      auto gram = [&](auto&& epi, auto&& hook) {
        auto e2 = scaled(epi);
        if constexpr (RING) gram_pass_ring<EL, NR>(S, Xq, d, rb, m, qo, e2, hook);
        else gram_pass<EL, NR, MODE != kModeTree>(S, Xq, d, rb, m, row0, rows, qo, e2, hook);
      };
      uint32_t key[T];
#pragma unroll
      for (int s = 0; s < T; ++s) key[s] = 0xFFFFFFFFu;
      if constexpr (MODE != kModePool) {
        // bound pass: per-thread T smallest packed keys
        gram([&](int c0, const floatx16& acc) {
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            const int cb = c0 + cq + 8 * g;
            const float4 cn4 = *reinterpret_cast<const float4*>(&S.lnb[cb]);
            const float cn[4] = {cn4.x, cn4.y, cn4.z, cn4.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const int c = cb + u;
              const float a = fmaf(-2.0f, acc[4 * g + u], __fadd_rn(rn, cn[u]));
              uint32_t kk = (c < m && c != qpos)
                                ? ((__float_as_uint(a > 0.0f ? a : 0.0f) & 0xFFFFF800u) | (uint32_t)c)
                                : 0xFFFFFFFFu;
              if (kk < key[T - 1]) {
#pragma unroll
                for (int s = 0; s < T; ++s) {
                  const uint32_t lo = min(kk, key[s]);
                  kk = max(kk, key[s]);
                  key[s] = lo;
                }
              }
            }
          }
        }, [](int) {});
      }
      if constexpr (MODE == kModeCand) {
        if (qv) {
          uint64_t* o = A.out_key + ((int64_t)(it.tree - A.tree0) * n + S.lid[qpos]) * KP + sub * T;
#pragma unroll
          for (int s = 0; s < T; ++s)
            o[s] = key[s] == 0xFFFFFFFFu
                       ? ~0ull
                       : ((uint64_t)(key[s] & 0xFFFFF800u) << 32) | (uint32_t)S.lid[key[s] & 0x7FFu];
        }
        __syncthreads();  // As / Bs reuse by the next block
        continue;
      }
      if constexpr (MODE == kModeTree) {
        if (qv) {
#pragma unroll
          for (int s = 0; s < T; ++s)
            S.cpos[q][sub * T + s] = key[s] == 0xFFFFFFFFu ? kNoPos : (uint16_t)(key[s] & 0x7FFu);
        }
        __syncthreads();
        const int npair = rows * KP;
        for (int p = tid; p < npair; p += 2 * NT) {
          const int p1 = p + NT;
          const int r0_ = p / KP, s0 = p - r0_ * KP;
          const uint16_t pos0 = S.cpos[r0_][s0];
          const bool two = p1 < npair;
          const int r1_ = two ? p1 / KP : r0_, s1 = two ? p1 - r1_ * KP : s0;
          const uint16_t pos1 = S.cpos[r1_][s1];
          const int a0 = S.lid[row0 + r0_], b0 = pos0 == kNoPos ? a0 : S.lid[pos0];
          const int a1 = S.lid[row0 + r1_], b1 = pos1 == kNoPos ? a1 : S.lid[pos1];
          float D0, D1;
          if ((d & 3) == 0) {
            exact_dist4_x2(X, a0, b0, a1, b1, d, D0, D1);
          } else {
            D0 = exact_dist(X, a0, b0, d);
            D1 = exact_dist(X, a1, b1, d);
          }
          S.cdist[r0_][s0] = pos0 == kNoPos ? INFINITY : D0;
          if (two) S.cdist[r1_][s1] = pos1 == kNoPos ? INFINITY : D1;
        }
        __syncthreads();
        // B = K-th smallest exact distance of the row's KP candidates
#pragma unroll 1
        for (int rr = 0; rr < kRowsPerWave; ++rr) {
          const int r = wave * kRowsPerWave + rr;
          if (r >= rows) break;
          const uint16_t pos = lane < KP ? S.cpos[r][lane] : kNoPos;
          const float dv = pos == kNoPos ? INFINITY : S.cdist[r][lane];
          const int iv = pos == kNoPos ? INT_MAX : S.lid[pos];
          int rank = 0;
          for (int s = 0; s < KP; ++s) rank += better(lane_read(dv, s), lane_read(iv, s), dv, iv) ? 1 : 0;
          const uint64_t hit = __ballot(iv != INT_MAX && rank == K - 1);
          const float B = hit ? lane_read(dv, __builtin_ctzll(hit)) : INFINITY;
          if (lane == 0) {
            S.cacut[r] = acut_of(B, S.lnb[row0 + r], lerr_at(row0 + r), nm, emax, d);
            S.ccnt[r] = KP;
          }
        }
      } else {  // kModePool: external bound, known ids of the block's rows
        if (tid < rows) {
          S.cacut[tid] = acut_of(A.boundT[lofs + row0 + tid], S.lnb[row0 + tid], lerr_at(row0 + tid),
                                 nm, emax, d);
          S.ccnt[tid] = 0;
        }
        if constexpr (MODE == kModePool) {
          if (K <= kKnownLds)  // read at the flush, after the Gram pass
            for (int j = tid; j < rows * K; j += NT) {
              const int r = j / K;
              S.known[r * kKnownPitch + (j - r * K)] = A.known[(int64_t)S.lid[row0 + r] * K + (j - r * K)];
            }
        }
      }
      __syncthreads();
      // filter pass: every leafmate that can enter the row's top K
      const float acut = qv ? S.cacut[q] : -INFINITY;
      int nsub = 0;  // kModePool: entries of this thread's list
      // kModePool flush: thread (row r = tid / 4, quarter tid % 4) drops the
      // row's known ids from the list of the thread that owned (r, quarter) in
      // the tiles, the 4 lanes of a row reserve pool slots with one atomic.
      // Runs at the end of the pass, and between tiles once some list is half
      // full: the epilogue of tile t sets flag[t & 1], the hook after tile t
      // reads and clears it. Every thread reads it after the barrier that
      // follows tile t, and the next write to it (tile t + 2) comes after a
      // barrier that follows the read, so every thread takes the same branch.
      auto flush = [&]() {
        if constexpr (MODE == kModePool) {
          S.ccnt[q * 4 + sub] = nsub;  // the owner's count, visible after the barrier
          __syncthreads();
          const int r = tid >> 2, qs = tid & 3;
          const int* kn = &S.known[r * kKnownPitch];
          const int cnt = r < rows ? S.ccnt[r * 4 + qs] : 0;
          uint16_t* lst = &S.cpos[r][qs * kSubCap];
          int keep = 0;
          if (cnt > kSubCap) {
            keep = A.pool_cap + 1;  // forces the row's overflow
          } else {
            if (K <= kKnownLds) {
              for (int j = 0; j < cnt; ++j) {
                const uint16_t c = lst[j];
                const int id = S.lid[c];
                bool known = false;
                for (int k = 0; k < K; ++k) known |= kn[k] == id;
                if (!known) lst[keep++] = c;
              }
            } else {
              keep = cnt;  // larger K: the row merge drops known ids (hash)
            }
          }
          int pre = keep;  // inclusive prefix over the row's 4 lanes
#pragma unroll
          for (int off = 1; off < 4; off <<= 1) {
            const int v = __shfl_up(pre, off, 4);
            if (qs >= off) pre += v;
          }
          const int total = __shfl(pre, 3, 4);
          int base = 0;
          if (qs == 0 && total > 0 && total <= A.pool_cap) base = atomicAdd(&A.pool_cnt[S.lid[row0 + r]], total);
          base = __shfl(base, 0, 4);
          if (total > 0) {
            const int id = S.lid[row0 + r];
            if (total <= A.pool_cap && base + total <= A.pool_cap) {
              const int o = base + pre - keep;
              for (int j = 0; j < keep; ++j)
                A.out_idx[(int64_t)id * A.pool_cap + o + j] = S.lid[lst[j]];
            } else if (qs == 0 && atomicExch(&A.ovf_flag[id], 1) == 0) {
              A.ovf_list[atomicAdd(A.novf, 1)] = id;
            }
          }
          nsub = 0;
        }
      };
      auto hook = [&](int tprev) {
        if constexpr (MODE == kModePool) {
          const int f = tprev & 1;
          if (S.flush_flag[f]) {
            flush();
            if (tid == 0) S.flush_flag[f] = 0;
            __syncthreads();
          }
        }
      };
      gram([&](int c0, const floatx16& acc) {
        uint32_t pm = 0;  // elements with a <= A_cut
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const float4 cn4 = *reinterpret_cast<const float4*>(&S.lnb[c0 + cq + 8 * g]);
          const float cn[4] = {cn4.x, cn4.y, cn4.z, cn4.w};
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const float a = fmaf(-2.0f, acc[4 * g + u], __fadd_rn(rn, cn[u]));
            pm |= (a <= acut ? 1u : 0u) << (4 * g + u);
          }
        }
        while (pm != 0) {
          const int e = __builtin_ctz(pm);
          pm &= pm - 1;
          const int c = c0 + cq + 8 * (e >> 2) + (e & 3);
          if (c >= m || c == qpos) continue;
          if constexpr (MODE == kModeTree) {
            bool dup = false;
#pragma unroll
            for (int s = 0; s < T; ++s)
              dup |= key[s] != 0xFFFFFFFFu && (key[s] & 0x7FFu) == (uint32_t)c;
            if (!dup) {
              const int slot = atomicAdd(&S.ccnt[q], 1);
              if (slot < kCap) S.cpos[q][slot] = (uint16_t)c;
            }
          } else if (nsub < kSubCap) {  // this thread's list (known ids dropped at the flush)
            S.cpos[q][sub * kSubCap + nsub] = (uint16_t)c;
            ++nsub;
          } else {  // list full within this tile: straight to the row's pool (the row
                    // merge drops known ids)
            const int id = S.lid[qpos];
            const int slot = atomicAdd(&A.pool_cnt[id], 1);
            if (slot < A.pool_cap) {
              A.out_idx[(int64_t)id * A.pool_cap + slot] = S.lid[c];
            } else if (atomicExch(&A.ovf_flag[id], 1) == 0) {
              A.ovf_list[atomicAdd(A.novf, 1)] = id;
            }
          }
        }
        if constexpr (MODE == kModePool) {
          // flushed at half capacity: a tile adds at most 16 entries per
          // thread but rarely more than a few; entries that still do not fit
          // go to the pool directly
          if (nsub > kSubCap / 2) S.flush_flag[(c0 / kCols) & 1] = 1;
        }
      }, hook);
      if constexpr (MODE == kModePool) {
        flush();
        if (tid == 0) S.flush_flag[0] = S.flush_flag[1] = 0;  // the last tile's, unread
        __syncthreads();  // lists and tiles reused by the next block
        continue;
      }
      if constexpr (MODE == kModeTree) {
        __syncthreads();
        // kModeTree: exact D of the appended candidates, flattened over the block's rows
        if (wave == 0) {
          int c = lane < rows ? min(S.ccnt[lane], kCap) - KP : 0;
  #pragma unroll
          for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(c, off);
            if (lane >= off) c += v;
          }
          S.cpre[lane + 1] = c;
          if (lane == 0) S.cpre[0] = 0;
        }
        __syncthreads();
        const int total = S.cpre[rows];
        for (int p = tid; p < total; p += 2 * NT) {
          int r_[2], s_[2];
  #pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int pp = min(p + h * NT, total - 1);
            int lo = 0, hi = rows;  // largest r with cpre[r] <= pp
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (S.cpre[mid] <= pp) lo = mid; else hi = mid;
            }
            r_[h] = lo;
            s_[h] = KP + pp - S.cpre[lo];
          }
          const int a0 = S.lid[row0 + r_[0]], b0 = S.lid[S.cpos[r_[0]][s_[0]]];
          const int a1 = S.lid[row0 + r_[1]], b1 = S.lid[S.cpos[r_[1]][s_[1]]];
          float D0, D1;
          if ((d & 3) == 0) {
            exact_dist4_x2(X, a0, b0, a1, b1, d, D0, D1);
          } else {
            D0 = exact_dist(X, a0, b0, d);
            D1 = exact_dist(X, a1, b1, d);
          }
          S.cdist[r_[0]][s_[0]] = D0;
          if (p + NT < total) S.cdist[r_[1]][s_[1]] = D1;
        }
        __syncthreads();
        // the K best of each row's list, or the exact fallback on overflow
  #pragma unroll 1
        for (int rr = 0; rr < kRowsPerWave; ++rr) {
          const int r = wave * kRowsPerWave + rr;
          if (r >= rows) break;
          const int cnt = S.ccnt[r];
          const int me = S.lid[row0 + r];
          if (cnt > kCap) {
            if (lane == 0) {
              const int slot = atomicAdd(A.nfallback, 1);
              if (slot < A.fb_cap) {
                A.fallback[2 * slot] = (int32_t)item;
                A.fallback[2 * slot + 1] = row0 + r - it.row0;
              }
            }
            continue;
          }
          const uint16_t pos = lane < cnt ? S.cpos[r][lane] : kNoPos;
          const float dv = pos == kNoPos ? INFINITY : S.cdist[r][lane];
          const int iv = pos == kNoPos ? INT_MAX : S.lid[pos];
          const int64_t o = ((int64_t)it.tree * n + me) * K;
          rank_store(dv, iv, cnt, K, lane, A.out_idx + o, A.out_dist + o);
        }

      }
      __syncthreads();
    }
    item = begin + S.ticket[parity];
  }
}

// Leaf-ordered copies of |x~|^2 and |x~ - x| (position j of tree t's leaf
// order); none when *gate == 0 (the passes that read them then gather nbe).
__global__ void k_leaf_gather(const int32_t* __restrict__ perm, int64_t total,
                              const float2* __restrict__ nbe, const int8_t* __restrict__ pexp,
                              float* __restrict__ lnbT, float* __restrict__ lerrT,
                              int8_t* __restrict__ lexpT, const int32_t* __restrict__ gate) {
  if (gate != nullptr && *gate == 0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int id = perm[i];
    const float2 v = nbe[id];  // one 8-byte gather per point
    lnbT[i] = v.x;
    lerrT[i] = v.y;
    if (pexp) lexpT[i] = pexp[id];
  }
}

// *lossy = 1 when some row's quantisation error is not 0
__global__ void k_lossy_flag(const float2* __restrict__ nbe, int64_t n, int32_t* __restrict__ lossy) {
  bool any = false;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    any = any || nbe[i].y != 0.0f;
  if (__ballot(any) != 0 && (threadIdx.x & 63) == 0) atomicOr(lossy, 1);
}

// leaf-ordered gather of one value per point: outT[i] = src[perm[i] * stride + offset];
// each thread loads four perm entries, then their four values, before storing
// (four independent gathers in flight instead of one dependent pair)
constexpr int kGatherUnroll = 4;
__global__ void __launch_bounds__(256)
k_leaf_gather_val(const int32_t* __restrict__ perm, int64_t total, const float* __restrict__ src,
                  int stride, int offset, float* __restrict__ outT) {
  const int64_t span = (int64_t)blockDim.x * kGatherUnroll;
  for (int64_t base = (int64_t)blockIdx.x * span; base < total; base += (int64_t)gridDim.x * span) {
    int32_t pi[kGatherUnroll];
    float v[kGatherUnroll];
#pragma unroll
    for (int u = 0; u < kGatherUnroll; ++u) {
      const int64_t i = base + u * blockDim.x + threadIdx.x;
      pi[u] = i < total ? perm[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < kGatherUnroll; ++u) {
      const int64_t i = base + u * blockDim.x + threadIdx.x;
      v[u] = i < total ? src[(int64_t)pi[u] * stride + offset] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kGatherUnroll; ++u) {
      const int64_t i = base + u * blockDim.x + threadIdx.x;
      if (i < total) outT[i] = v[u];
    }
  }
}

// ---------------------------------------------------------------------------
// Per-row merge (one wave per row). Candidate ids from up to 64*E slots
// (duplicates, -1 and the row itself allowed) minus the row's known ids are
// sorted and de-duplicated (bitonic network over the wave, E elements per
// lane), their exact D computed, sorted by (D, index) and merged with the
// known list; the K best are written. Keys (D bits << 32 | index) order like
// (D, index) because D >= 0.
// ---------------------------------------------------------------------------
// v of lane ^ j: quad DPP for j = 1, 2 (no LDS trip), a swizzle within each
// half wave for j = 4, 8, 16, a permute for larger j (j is a constant once the
// callers' loops are unrolled, so the switch folds away)
__device__ __forceinline__ uint32_t xor_lane(uint32_t v, int j) {
  switch (j) {
    case 1: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
    case 2: return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
    case 4: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (4 << 10));
    case 8: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (8 << 10));
    case 16: return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (16 << 10));
    default: return (uint32_t)__shfl_xor((int)v, j);
  }
}
__device__ __forceinline__ uint64_t xor_lane(uint64_t v, int j) {
  const uint32_t lo = xor_lane((uint32_t)v, j), hi = xor_lane((uint32_t)(v >> 32), j);
  return ((uint64_t)hi << 32) | lo;
}

// ascending sort of the 64*E values v[e] (element index e*64 + lane)
template <typename U, int E>
__device__ __forceinline__ void bitonic_wave(U (&v)[E], int lane) {
  constexpr int N = 64 * E;
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int e2 = e ^ (j >> 6);
          if (e2 > e) {
            const bool asc = (((e << 6) | lane) & k) == 0;
            const U a = v[e], b = v[e2];
            const bool sw = asc ? (a > b) : (a < b);
            v[e] = sw ? b : a;
            v[e2] = sw ? a : b;
          }
        }
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const U o = xor_lane(v[e], j);
          const bool asc = (((e << 6) | lane) & k) == 0;
          const bool lower = (lane & j) == 0;
          const U mn = v[e] < o ? v[e] : o, mx = v[e] < o ? o : v[e];
          v[e] = (asc == lower) ? mn : mx;
        }
      }
    }
  }
}

// ascending sort of the first nvalid of the 64 E values (the rest are the
// largest value): only the registers that hold valid values are sorted
template <typename U, int E>
__device__ __forceinline__ void bitonic_wave_prefix(U (&v)[E], int lane, int nvalid) {
  if (E == 1 || nvalid <= 64) {
    bitonic_wave<U, 1>(*reinterpret_cast<U(*)[1]>(&v[0]), lane);
  } else if (E == 2 || nvalid <= 128) {
    bitonic_wave<U, (E < 2 ? E : 2)>(*reinterpret_cast<U(*)[(E < 2 ? E : 2)]>(&v[0]), lane);
  } else if (E == 4 || nvalid <= 256) {
    bitonic_wave<U, (E < 4 ? E : 4)>(*reinterpret_cast<U(*)[(E < 4 ? E : 4)]>(&v[0]), lane);
  } else if (E == 8 || nvalid <= 512) {
    bitonic_wave<U, (E < 8 ? E : 8)>(*reinterpret_cast<U(*)[(E < 8 ? E : 8)]>(&v[0]), lane);
  } else {
    bitonic_wave<U, E>(v, lane);
  }
}

// exact D of (row, cand): the fma chain, 8 vector loads per batch (one
// 128-byte line of each row for FP32 input)
template <typename XT>
__device__ __forceinline__ float exact_dist8(const XT* __restrict__ xr, const XT* __restrict__ xc,
                                             int d) {
  float acc = 0.0f;
  if ((d & 3) == 0) {
#pragma unroll 8
    for (int k = 0; k < d; k += 4) {
      const float4 a = load4f(xr + k), b = load4f(xc + k);
      float t = __fsub_rn(a.x, b.x);
      acc = fmaf(t, t, acc);
      t = __fsub_rn(a.y, b.y);
      acc = fmaf(t, t, acc);
      t = __fsub_rn(a.z, b.z);
      acc = fmaf(t, t, acc);
      t = __fsub_rn(a.w, b.w);
      acc = fmaf(t, t, acc);
    }
  } else {
    for (int k = 0; k < d; ++k) {
      const float t = __fsub_rn((float)xr[k], (float)xc[k]);
      acc = fmaf(t, t, acc);
    }
  }
  return acc;
}

constexpr int kMergeWaves = 4;    // rows in flight per block (one per wave)
constexpr int kMergeRowLds = 1024; // elements of a row staged in LDS per wave (16 KB per block)

template <typename XT, int E, bool RL>
__global__ void __launch_bounds__(64 * kMergeWaves) k_row_merge(const RowMergeArgs A) {
  // per-wave open-addressing table for distinct ids: more slots than the K
  // known ids + 64 E insertions, so probing ends
  constexpr int kHashBits = E <= 4 ? 9 : E <= 8 ? 10 : 11;
  constexpr int kHashSlots = 1 << kHashBits;
  __shared__ uint32_t tab[kMergeWaves][kHashSlots];
  __shared__ uint64_t dense[kMergeWaves][64 * E];
  // RL: the row staged in LDS for its exact distances (a separate instantiation:
  // shorter rows keep the registers and LDS of the plain kernel)
  constexpr bool kRowLds = RL;
  __shared__ __align__(16) float qrow[kMergeWaves][kRowLds ? kMergeRowLds : 4];
  const XT* __restrict__ X = static_cast<const XT*>(A.X);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int K = A.K, d = A.d;
  uint32_t* tb = tab[w];
  uint64_t* dn = dense[w];
  for (int j = lane; j < kHashSlots; j += 64) tb[j] = 0xFFFFFFFFu;
  // rows in A.order (leaf order of one tree: neighbouring rows share most
  // candidates), one contiguous chunk per XCD (blocks go round-robin over 8)
  const int xcd = blockIdx.x & 7;
  const int64_t nblk = (gridDim.x - xcd + 7) >> 3;  // blocks on this XCD
  const int64_t nr = A.nrows_dev != nullptr ? (int64_t)*A.nrows_dev : A.n;
  const int64_t chunk = (nr + 7) >> 3;
  const int64_t lo = xcd * chunk, hi = min(nr, lo + chunk);
  const int nslots = A.S * A.ntsrc;
  for (int64_t i = lo + (int64_t)(blockIdx.x >> 3) * kMergeWaves + w; i < hi;
       i += nblk * kMergeWaves) {
    const int64_t row = A.order != nullptr ? (int64_t)A.order[i] : i;
    if (A.heavy != nullptr) {  // light pass (one source): rows beyond 64 E slots listed
      if (A.skip_flag == nullptr || A.skip_flag[row] == 0) {
        const int c = min(A.cnt_tree[row], A.S);
        if (c > 64 * E) {
          if (lane == 0) A.heavy[atomicAdd(A.nheavy, 1)] = (int32_t)row;
          continue;
        }
      }
    }
    if (A.skip_flag != nullptr && A.skip_flag[row] != 0) {
      // left to the exact row path; the known list passes through (a valid
      // bound for later passes)
      if (A.known_idx != nullptr && lane < K) {
        A.out_idx[row * K + lane] = A.known_idx[row * K + lane];
        A.out_dist[row * K + lane] = A.known_dist[row * K + lane];
      }
      if (A.out_bound != nullptr && lane == K - 1)
        A.out_bound[row] = A.known_idx != nullptr ? A.known_dist[row * K + lane] : INFINITY;
      continue;
    }
    // known list (sorted, -1 padded): lanes < K
    int kid = -1;
    float kd = INFINITY;
    if (A.known_idx != nullptr && lane < K) {
      kid = A.known_idx[row * K + lane];
      kd = A.known_dist[row * K + lane];
    }
    const int nknown = __popcll(__ballot(kid >= 0));
    // valid slots per source (lane t holds source t's count; ntsrc <= 64)
    const int ct = (A.cnt_tree != nullptr && lane < A.ntsrc)
                       ? min(A.cnt_tree[(int64_t)lane * A.n + row], A.S) : A.S;
    if (A.cnt_tree != nullptr && __ballot(lane < A.ntsrc && ct > 0) == 0) {
      // no candidates: the known list is the result
      if (lane < K) {
        A.out_idx[row * K + lane] = kid;
        A.out_dist[row * K + lane] = kd;
      }
      if (A.out_bound != nullptr && lane == K - 1) A.out_bound[row] = kd;
      continue;
    }
    // known ids first, then candidates: the first insertion of an id wins
    if (kid >= 0) {
      uint32_t h = ((uint32_t)kid * 2654435761u) >> (32 - kHashBits);
      while (true) {
        const uint32_t old = atomicCAS(&tb[h], 0xFFFFFFFFu, (uint32_t)kid);
        if (old == 0xFFFFFFFFu || old == (uint32_t)kid) break;
        h = (h + 1) & (kHashSlots - 1);
      }
    }
    int cnt_new = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = e * 64 + lane;
      const int t = min(j / A.S, 63), sl = j - t * A.S;
      const int cti = __shfl(ct, t);
      int id = -1;
      uint32_t ap = 0;  // approximate distance bits (src_key)
      if (j < nslots) {
        const bool ok = sl < cti;
        if (ok) {
          const int64_t at = t * A.src_tree_stride + row * A.S + sl;
          if (A.src_key != nullptr) {
            const uint64_t kv = A.src_key[at];
            if (kv != ~0ull) {
              id = (int)(uint32_t)kv;
              ap = (uint32_t)(kv >> 32);
            }
          } else {
            id = A.src[at];
          }
        }
      }
      bool fresh = false;
      if (id >= 0 && id != (int)row) {
        uint32_t h = ((uint32_t)id * 2654435761u) >> (32 - kHashBits);
        while (true) {
          const uint32_t old = atomicCAS(&tb[h], 0xFFFFFFFFu, (uint32_t)id);
          if (old == 0xFFFFFFFFu) { fresh = true; break; }
          if (old == (uint32_t)id) break;
          h = (h + 1) & (kHashSlots - 1);
        }
      }
      const uint64_t bal = __ballot(fresh);
      if (fresh) dn[cnt_new + __popcll(bal & ((1ull << lane) - 1))] = ((uint64_t)ap << 32) | (uint32_t)id;
      cnt_new += __popcll(bal);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // reset the table for the next row
    for (int j = lane; j < kHashSlots; j += 64) tb[j] = 0xFFFFFFFFu;
    // distinct new candidates in dense lanes; keep the rerank best approximate ones
    uint64_t key[E];
#pragma unroll
    for (int e = 0; e < E; ++e) key[e] = (e * 64 + lane < cnt_new) ? dn[e * 64 + lane] : ~0ull;
    int neval = cnt_new;
    if (A.rerank > 0 && cnt_new > A.rerank) {
      // approximate bits (low 11 bits zero) | dense index (< 64 E <= 1024): a
      // 32-bit sort; ties in the approximate bits go by dense index (any
      // choice among them is sound: the pooled passes re-find every leafmate
      // that can still enter the top K)
      uint32_t k32[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int j = e * 64 + lane;
        k32[e] = j < cnt_new ? (uint32_t)(key[e] >> 32) | (uint32_t)j : 0xFFFFFFFFu;
      }
      bitonic_wave_prefix<uint32_t, E>(k32, lane, cnt_new);
#pragma unroll
      for (int e = 0; e < E; ++e)
        key[e] = e * 64 + lane < A.rerank ? dn[k32[e] & 0x7FFu] : ~0ull;
      neval = A.rerank;
    }
    // exact D, keys (D bits << 32 | id)
    const XT* xr = X + (A.xpos != nullptr ? (int64_t)A.xpos[row] : row) * d;
    if (kRowLds) {
      // the row staged once in LDS as float: every lane reads it by broadcast
      // instead of loading it again beside its candidate (same fma chain)
      float* qf = qrow[w];
      for (int k = lane * 4; k < d; k += 256) *reinterpret_cast<float4*>(&qf[k]) = load4f(xr + k);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if (e * 64 + lane < neval) {
          const uint32_t id = (uint32_t)key[e];
          const XT* xc = X + (A.xpos != nullptr ? (int64_t)A.xpos[id] : (int64_t)id) * d;
          float acc = 0.0f;
#pragma unroll 8
          for (int k = 0; k < d; k += 4) {
            const float4 a = *reinterpret_cast<const float4*>(&qf[k]), b = load4f(xc + k);
            float t = __fsub_rn(a.x, b.x);
            acc = fmaf(t, t, acc);
            t = __fsub_rn(a.y, b.y);
            acc = fmaf(t, t, acc);
            t = __fsub_rn(a.z, b.z);
            acc = fmaf(t, t, acc);
            t = __fsub_rn(a.w, b.w);
            acc = fmaf(t, t, acc);
          }
          key[e] = ((uint64_t)__float_as_uint(acc) << 32) | id;
        } else {
          key[e] = ~0ull;
        }
      }
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if (e * 64 + lane < neval) {
          const uint32_t id = (uint32_t)key[e];
          const XT* xc = X + (A.xpos != nullptr ? (int64_t)A.xpos[id] : (int64_t)id) * d;
          key[e] = ((uint64_t)__float_as_uint(exact_dist8(xr, xc, d)) << 32) | id;
        } else {
          key[e] = ~0ull;
        }
      }
    }
    bitonic_wave_prefix<uint64_t, E>(key, lane, neval);
    // merge the K best new entries (lanes < K of key[0]) with the known list
    const uint64_t nk = lane < K ? key[0] : ~0ull;
    const uint64_t kk = kid >= 0 ? (((uint64_t)__float_as_uint(kd) << 32) | (uint32_t)kid) : ~0ull;
    int32_t* oi = A.out_idx + row * K;
    float* od = A.out_dist + row * K;
    if (K <= 32) {
      // the known list ascending in lanes 0-31 and the new K best descending
      // in lanes 32-63 form a bitonic sequence; its merge network sorts the
      // union (distinct ids: known ids are never fresh) and lanes < K hold
      // the K best
      const uint64_t nrev = __shfl(nk, 63 - lane);
      uint64_t v = lane < 32 ? kk : nrev;
#pragma unroll
      for (int j = 32; j > 0; j >>= 1) {
        const uint64_t o = xor_lane(v, j);
        const bool lower = (lane & j) == 0;
        v = lower ? (o < v ? o : v) : (o < v ? v : o);
      }
      if (lane < K) {
        oi[lane] = v == ~0ull ? -1 : (int32_t)(uint32_t)v;
        od[lane] = v == ~0ull ? INFINITY : __uint_as_float((uint32_t)(v >> 32));
      }
      if (A.out_bound != nullptr && lane == K - 1)
        A.out_bound[row] = v == ~0ull ? INFINITY : __uint_as_float((uint32_t)(v >> 32));
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      continue;
    }
    int rn = lane, rk = lane;  // rank of my new entry / of my known entry
    for (int s = 0; s < K; ++s) {
      const uint64_t o_new = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(nk >> 32), s) << 32) |
                             (uint32_t)__builtin_amdgcn_readlane((int)nk, s);
      const uint64_t o_kn = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(kk >> 32), s) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)kk, s);
      rn += o_kn < nk ? 1 : 0;
      rk += o_new < kk ? 1 : 0;
    }
    if (lane < K && nk != ~0ull && rn < K) {
      oi[rn] = (int32_t)(uint32_t)nk;
      od[rn] = __uint_as_float((uint32_t)(nk >> 32));
      if (A.out_bound != nullptr && rn == K - 1) A.out_bound[row] = __uint_as_float((uint32_t)(nk >> 32));
    }
    if (lane < K && kk != ~0ull && rk < K) {
      oi[rk] = kid;
      od[rk] = kd;
      if (A.out_bound != nullptr && rk == K - 1) A.out_bound[row] = kd;
    }
    const int nv = min(neval, K) + nknown;
    if (lane < K && lane >= nv) {
      oi[lane] = -1;
      od[lane] = INFINITY;
      if (A.out_bound != nullptr && lane == K - 1) A.out_bound[row] = INFINITY;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// Pool merge with two rows per wave (lanes 0-31 one row, 32-63 the next), for
// one source of ids (a pool) and K <= 32. Per row: its known list (sorted,
// -1 padded) is the running top K; the pool's ids minus known ones and
// duplicates (open-addressing hash per half wave) are listed in LDS; their
// exact D are taken 32 at a time, one per lane, each chunk sorted by (D, id)
// (32-lane bitonic network) and merged into the running list (half-cleaner
// against the reversed chunk, then 5 stages). Any pool count works, so no
// light / heavy split is needed, and the exact-distance phase fills the wave
// with both rows' candidates. Rows whose pools hold more than kM2Fresh
// entries are listed (heavy) for the full-width pass.
constexpr int kM2HashBits = 9;       // slots per half wave: more than its insertions (pool + K)
constexpr int kM2Fresh = 256;        // pool entries a row may hold here
constexpr int kM2RowLds = 256;       // elements of a row staged in LDS (two rows per wave)

// ascending sort of the 32 values of each half wave
__device__ __forceinline__ void sort32(uint64_t& v, int hl) {
#pragma unroll
  for (int k = 2; k <= 32; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint64_t o = xor_lane(v, j);
      const bool asc = (hl & k) == 0;
      const bool lower = (hl & j) == 0;
      const uint64_t mn = v < o ? v : o, mx = v < o ? o : v;
      v = (asc == lower) ? mn : mx;
    }
  }
}

// sorted half-wave lists r and c (ascending): r = the 32 smallest of both
__device__ __forceinline__ void merge32(uint64_t& r, uint64_t c, int hl) {
  const uint64_t rc = (uint64_t)(uint32_t)__shfl((int)(uint32_t)c, 31 - hl, 32) |
                      ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(c >> 32), 31 - hl, 32) << 32);
  r = r < rc ? r : rc;  // bitonic: the 32 smallest
#pragma unroll
  for (int j = 16; j > 0; j >>= 1) {
    const uint64_t o = xor_lane(r, j);
    const bool lower = (hl & j) == 0;
    const uint64_t mn = r < o ? r : o, mx = r < o ? o : r;
    r = lower ? mn : mx;
  }
}

template <typename XT, bool RL>
__global__ void __launch_bounds__(64 * kMergeWaves) k_row_merge2(const RowMergeArgs A) {
  constexpr int kSlots = 1 << kM2HashBits;
  __shared__ uint32_t tab[kMergeWaves][2][kSlots];
  __shared__ int32_t fresh[kMergeWaves][2][kM2Fresh];
  __shared__ __align__(16) float qrow[kMergeWaves][2][RL ? kM2RowLds : 4];
  const XT* __restrict__ X = static_cast<const XT*>(A.X);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int h = lane >> 5, hl = lane & 31;
  const int K = A.K, d = A.d;
  uint32_t* tb = tab[w][h];
  int32_t* fr = fresh[w][h];
  for (int j = hl; j < kSlots; j += 32) tb[j] = 0xFFFFFFFFu;
  const uint64_t hmask = h ? 0xFFFFFFFF00000000ull : 0x00000000FFFFFFFFull;
  // pairs of rows in A.order, one contiguous chunk per XCD (blocks go round-robin over 8)
  const int xcd = blockIdx.x & 7;
  const int64_t nblk = (gridDim.x - xcd + 7) >> 3;
  const int64_t nr = A.n;
  const int64_t npairs = (nr + 1) >> 1;
  const int64_t chunk = (npairs + 7) >> 3;
  const int64_t lo = xcd * chunk, hi = min(npairs, lo + chunk);
  const int64_t step = nblk * kMergeWaves;
  // the row of pair p (this half wave's), and the loads of a row that do not
  // depend on its pool count: its skip flag, count, known list entry and first
  // 32 pool slots; issued one pair (the row index two pairs) ahead
  auto row_of = [&](int64_t p) -> int64_t {
    const int64_t r = 2 * p + h;
    return p < hi && r < nr ? (A.order != nullptr ? (int64_t)A.order[r] : r) : -1;
  };
  struct Pre {
    int skip, cnt, kid, src0;
    float kdv;
  };
  auto load_pre = [&](int64_t r) -> Pre {
    Pre q{0, 0, -1, -1, INFINITY};
    if (r >= 0) {
      q.skip = A.skip_flag != nullptr ? A.skip_flag[r] : 0;
      q.cnt = A.cnt_tree[r];
      if (A.known_idx != nullptr && hl < K) {
        q.kid = A.known_idx[r * K + hl];
        q.kdv = A.known_dist[r * K + hl];
      }
      if (hl < A.S) q.src0 = A.src[r * A.S + hl];  // read before the count is known
    }
    return q;
  };
  const int64_t p0 = lo + (int64_t)(blockIdx.x >> 3) * kMergeWaves + w;
  int64_t row1 = row_of(p0);
  Pre pre1 = load_pre(row1);
  int64_t row2 = row_of(p0 + step);
  for (int64_t pi = p0; pi < hi; pi += step) {
    const int64_t ri = 2 * pi + h;
    const bool has = ri < nr;
    const int64_t row = has ? row1 : 0;
    const Pre cur = pre1;
    pre1 = load_pre(row2);
    row1 = row2;
    row2 = row_of(pi + 2 * step);
    const bool skip = has && cur.skip != 0;
    const int cnt0 = has && !skip ? min(cur.cnt, A.S) : 0;
    const bool heavy = cnt0 > kM2Fresh;  // to the full-width pass
    if (heavy && hl == 0) A.heavy[atomicAdd(A.nheavy, 1)] = (int32_t)row;
    const bool act = has && !skip && !heavy;
    // known list: the running top K (keys D bits << 32 | id, ascending, ~0 padded)
    int kid = -1;
    float kdv = INFINITY;
    if (has && A.known_idx != nullptr && hl < K) {
      kid = cur.kid;
      kdv = cur.kdv;
    }
    if (skip && hl < K) {  // left to the exact row path; the known list passes through
      A.out_idx[row * K + hl] = kid;
      A.out_dist[row * K + hl] = kdv;
    }
    if (skip && A.out_bound != nullptr && hl == K - 1) A.out_bound[row] = kdv;
    uint64_t run = (act && kid >= 0) ? (((uint64_t)__float_as_uint(kdv) << 32) | (uint32_t)kid) : ~0ull;
    if (act && kid >= 0) {
      uint32_t hh = ((uint32_t)kid * 2654435761u) >> (32 - kM2HashBits);
      while (true) {
        const uint32_t old = atomicCAS(&tb[hh], 0xFFFFFFFFu, (uint32_t)kid);
        if (old == 0xFFFFFFFFu || old == (uint32_t)kid) break;
        hh = (hh + 1) & (kSlots - 1);
      }
    }
    // distinct new ids of the pool
    const int cnt = act ? cnt0 : 0;
    const int cmax = max(cnt, __shfl_xor(cnt, 32));
    int nf = 0;
    for (int s0 = 0; s0 < cmax; s0 += 32) {
      const int sl = s0 + hl;
      const int id = sl < cnt ? (s0 == 0 ? cur.src0 : A.src[row * A.S + sl]) : -1;
      bool isnew = false;
      if (id >= 0 && id != (int)row) {
        uint32_t hh = ((uint32_t)id * 2654435761u) >> (32 - kM2HashBits);
        while (true) {
          const uint32_t old = atomicCAS(&tb[hh], 0xFFFFFFFFu, (uint32_t)id);
          if (old == 0xFFFFFFFFu) { isnew = true; break; }
          if (old == (uint32_t)id) break;
          hh = (hh + 1) & (kSlots - 1);
        }
      }
      const uint64_t bal = __ballot(isnew) & hmask;
      const uint64_t below = bal & ((1ull << lane) - 1);
      if (isnew && nf + __popcll(below) < kM2Fresh) fr[nf + __popcll(below)] = id;
      nf = min(kM2Fresh, nf + __popcll(bal));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int j = hl; j < kSlots; j += 32) tb[j] = 0xFFFFFFFFu;
    // exact D of the new ids, 32 at a time per row
    const int64_t qpos = A.xpos != nullptr ? (int64_t)A.xpos[row] : row;
    const XT* xr = X + qpos * d;
    // every value an integer in [0, 255]: the int8 copy x' = x - 128 and
    // D = n'_i + n'_j - 2 x'_i . x'_j in int32 (v_dot4: every partial sum an
    // exact integer, D < 2^24, so it equals the fma chain's value); else rows
    // exactly representable in bf16: the fma chain over the bf16 copy (the
    // same values in half the bytes)
    const bool u8c = RL && A.Xi8 != nullptr && (d & 15) == 0 && *A.notu8 == 0;
    const bool b16 = !u8c && RL && A.Xb != nullptr && (d & 7) == 0 && *A.lossy == 0;
    int ni = 0;  // u8c: n'_i
    if constexpr (RL) {
      float* qf = qrow[w][h];
      if (u8c) {  // the query row's int8 bytes
        if (act && hl < (d >> 4))
          reinterpret_cast<uint4*>(qf)[hl] = *reinterpret_cast<const uint4*>(A.Xi8 + row * d + hl * 16);
      } else if (act) {
        for (int k = hl * 4; k < d; k += 128) *reinterpret_cast<float4*>(&qf[k]) = load4f(xr + k);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (u8c && act) {
        const uint4* qa = reinterpret_cast<const uint4*>(qf);
        for (int c = 0; c < (d >> 4); ++c) {
          const uint4 a = qa[c];
          ni = __builtin_amdgcn_sdot4((int)a.x, (int)a.x, ni, false);
          ni = __builtin_amdgcn_sdot4((int)a.y, (int)a.y, ni, false);
          ni = __builtin_amdgcn_sdot4((int)a.z, (int)a.z, ni, false);
          ni = __builtin_amdgcn_sdot4((int)a.w, (int)a.w, ni, false);
        }
      }
    }
    const int nfmax = max(nf, __shfl_xor(nf, 32));
    for (int c0 = 0; c0 < nfmax; c0 += 32) {
      const int j = c0 + hl;
      uint64_t key = ~0ull;
      if (j < nf) {
        const int id = fr[j];
        const XT* xc = X + (A.xpos != nullptr ? (int64_t)A.xpos[id] : (int64_t)id) * d;
        float acc;
        if (u8c) {
          const uint4* qa = reinterpret_cast<const uint4*>(qrow[w][h]);
          const uint4* xb = reinterpret_cast<const uint4*>(A.Xi8 + (int64_t)id * d);
          int g = 0, nj = 0;
#pragma unroll 4
          for (int c = 0; c < (d >> 4); ++c) {
            const uint4 b = xb[c], a = qa[c];
            g = __builtin_amdgcn_sdot4((int)a.x, (int)b.x, g, false);
            g = __builtin_amdgcn_sdot4((int)a.y, (int)b.y, g, false);
            g = __builtin_amdgcn_sdot4((int)a.z, (int)b.z, g, false);
            g = __builtin_amdgcn_sdot4((int)a.w, (int)b.w, g, false);
            nj = __builtin_amdgcn_sdot4((int)b.x, (int)b.x, nj, false);
            nj = __builtin_amdgcn_sdot4((int)b.y, (int)b.y, nj, false);
            nj = __builtin_amdgcn_sdot4((int)b.z, (int)b.z, nj, false);
            nj = __builtin_amdgcn_sdot4((int)b.w, (int)b.w, nj, false);
          }
          acc = (float)(ni + nj - 2 * g);
        } else if (b16) {
          const float* qf = qrow[w][h];
          const uint16_t* xb = A.Xb + (int64_t)id * d;
          acc = 0.0f;
#pragma unroll 4
          for (int k = 0; k < d; k += 8) {
            const uint4 v = *reinterpret_cast<const uint4*>(xb + k);
            const float4 a0 = *reinterpret_cast<const float4*>(&qf[k]);
            const float4 a1 = *reinterpret_cast<const float4*>(&qf[k + 4]);
            const float b[8] = {__uint_as_float(v.x << 16), __uint_as_float(v.x & 0xFFFF0000u),
                                __uint_as_float(v.y << 16), __uint_as_float(v.y & 0xFFFF0000u),
                                __uint_as_float(v.z << 16), __uint_as_float(v.z & 0xFFFF0000u),
                                __uint_as_float(v.w << 16), __uint_as_float(v.w & 0xFFFF0000u)};
            const float a[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const float t = __fsub_rn(a[u], b[u]);
              acc = fmaf(t, t, acc);
            }
          }
        } else if constexpr (RL) {
          const float* qf = qrow[w][h];
          acc = 0.0f;
#pragma unroll 8
          for (int k = 0; k < d; k += 4) {
            const float4 a = *reinterpret_cast<const float4*>(&qf[k]), b = load4f(xc + k);
            float t = __fsub_rn(a.x, b.x);
            acc = fmaf(t, t, acc);
            t = __fsub_rn(a.y, b.y);
            acc = fmaf(t, t, acc);
            t = __fsub_rn(a.z, b.z);
            acc = fmaf(t, t, acc);
            t = __fsub_rn(a.w, b.w);
            acc = fmaf(t, t, acc);
          }
        } else {
          acc = exact_dist8(xr, xc, d);
        }
        key = ((uint64_t)__float_as_uint(acc) << 32) | (uint32_t)id;
      }
      sort32(key, hl);
      merge32(run, key, hl);
    }
    if (act && hl < K) {
      A.out_idx[row * K + hl] = run == ~0ull ? -1 : (int32_t)(uint32_t)run;
      A.out_dist[row * K + hl] = run == ~0ull ? INFINITY : __uint_as_float((uint32_t)(run >> 32));
    }
    if (act && A.out_bound != nullptr && hl == K - 1)
      A.out_bound[row] = run == ~0ull ? INFINITY : __uint_as_float((uint32_t)(run >> 32));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// First bounds from approximate distances alone (no row of X is read). Per
// row, the distinct ids among the candidate keys of the first trees; for each
// an upper bound of its exact distance D (the FP32 chain) from the key's
// approximate distance a (dropped low bits set: a_up >= a):
//   |x_i - x_j| <= |x~_i - x~_j| + e_i + e_j,  |x~_i - x~_j|^2 <= a + eta_ij,
//   eta_ij = (2d + 16) 2^-24 (n_i + n_j)^2 (the FP32 error of a, norms n
//   widened by (d + 4) 2^-24 as in the filter), D <= |x_i - x_j|^2 (1 + c),
//   c = (d + 8) 2^-24,
// so UB = (sqrt(a_up + eta_ij) + e_i + e_j)^2 (1 + c), evaluated in FP32 and
// widened by (1 + 64 2^-24) (the dozen roundings of the evaluation, sqrt
// within 2 ulp, stay below 16 2^-24 relative) plus 1e-30. The row's
// bound is the K-th smallest UB over distinct leafmates (+inf with fewer):
// at least K distinct leafmates lie within it, so it bounds the final K-th
// distance and every final neighbour passes the pooled filter. With more
// than 64 candidate slots the bound is the K-th smallest of the lanes' minima
// (one value per lane, each a distinct candidate): still at least K distinct
// leafmates within it, from a sort a quarter or half the size. (n_i^2, e_i)
// of each point come from nbe, the quantisation the keys were computed in.
template <int E>
__global__ void __launch_bounds__(64 * kMergeWaves)
k_row_bound(const uint64_t* __restrict__ src_key, int64_t tree_stride, int S, int ntsrc, int64_t n,
            int d, int K, const float2* __restrict__ nbe, float* __restrict__ out_bound) {
  constexpr int kHashBits = E <= 4 ? 9 : 10;
  constexpr int kHashSlots = 1 << kHashBits;
  __shared__ uint32_t tab[kMergeWaves][kHashSlots];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t* tb = tab[w];
  for (int j = lane; j < kHashSlots; j += 64) tb[j] = 0xFFFFFFFFu;
  const float u = 5.9604644775390625e-08f;  // 2^-24
  const float wn = 1.0f + (d + 4.0f) * u, wc = (1.0f + (d + 8.0f) * u) * (1.0f + 64.0f * u);
  const float ce = (2.0f * d + 16.0f) * u;
  const int nslots = S * ntsrc;
  const int64_t nwaves = (int64_t)gridDim.x * kMergeWaves;
  // a row's keys and own (|x~|^2, e), loaded one row ahead
  auto load_row = [&](int64_t r, uint64_t (&kv)[E], float2& mine) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = e * 64 + lane;
      const int t = j / S, sl = j - t * S;
      kv[e] = (r < n && j < nslots) ? src_key[t * tree_stride + r * S + sl] : ~0ull;
    }
    mine = r < n ? nbe[r] : make_float2(0.0f, 0.0f);
  };
  uint64_t kvn[E];
  float2 minen;
  load_row((int64_t)blockIdx.x * kMergeWaves + w, kvn, minen);
  for (int64_t row = (int64_t)blockIdx.x * kMergeWaves + w; row < n; row += nwaves) {
    uint64_t kvr[E];
#pragma unroll
    for (int e = 0; e < E; ++e) kvr[e] = kvn[e];
    const float2 mine = minen;
    load_row(row + nwaves, kvn, minen);
    const float ni = sqrtf(mine.x * wn);
    uint32_t ub[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint64_t kv = kvr[e];
      const int id = kv != ~0ull ? (int)(uint32_t)kv : -1;
      bool fresh = false;
      if (id >= 0 && id != (int)row) {
        uint32_t h = ((uint32_t)id * 2654435761u) >> (32 - kHashBits);
        while (true) {
          const uint32_t old = atomicCAS(&tb[h], 0xFFFFFFFFu, (uint32_t)id);
          if (old == 0xFFFFFFFFu) { fresh = true; break; }
          if (old == (uint32_t)id) break;
          h = (h + 1) & (kHashSlots - 1);
        }
      }
      ub[e] = 0xFFFFFFFFu;
      if (fresh) {
        const uint32_t abits = (uint32_t)(kv >> 32) | 0x7FFu;
        if (abits < 0x7F800000u) {
          const float2 o = nbe[id];
          const float nj = sqrtf(o.x * wn);
          const float eta = ce * (ni + nj) * (ni + nj);
          const float s = sqrtf(__uint_as_float(abits) + eta) + mine.y + o.y;
          const float b = s * s * wc + 1e-30f;
          ub[e] = __float_as_uint(b);  // non-negative: orders like its bits (+inf past the range)
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int j = lane; j < kHashSlots; j += 64) tb[j] = 0xFFFFFFFFu;
    uint32_t kth = 0xFFFFFFFFu;
    if (E == 1 || K > 64) {
      bitonic_wave<uint32_t, E>(ub, lane);
      // element K - 1 of the sorted 64 E values: register (K - 1) / 64, lane (K - 1) % 64
#pragma unroll
      for (int e = 0; e < E; ++e)
        if (e == (K - 1) / 64) kth = (uint32_t)__shfl((int)ub[e], (K - 1) % 64);
    } else {
      // the K-th smallest of the lanes' minima: K lanes hold a value at most
      // that large, each a distinct candidate, so it is a valid bound (at
      // least the K-th smallest UB, equal to it unless one lane holds two of
      // the K smallest); one 64-value sort instead of a 64 E-value one
      uint32_t m[1] = {ub[0]};
#pragma unroll
      for (int e = 1; e < E; ++e) m[0] = ub[e] < m[0] ? ub[e] : m[0];
      bitonic_wave<uint32_t, 1>(m, lane);
      kth = (uint32_t)__shfl((int)m[0], K - 1);
    }
    if (lane == 0) out_bound[row] = kth >= 0x7F800000u ? INFINITY : __uint_as_float(kth);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// Exact K best distinct leafmates over every tree, for rows whose pool
// overflowed or whose leaf is too large to stage. k_row_exact_tree: one block
// per (row, tree) pair, the row's exact KP best leafmates in that tree (each
// wave sorts its share by (D, id), one wave merges the waves' best);
// k_row_exact_merge: per row, the T sorted lists merged without duplicates.
constexpr int kFbThreadsRow = 256;
// largest per-tree list scratch of one batch of exact rows (bytes)
constexpr int64_t kRowExactScratch = (int64_t)512 << 20;
// one wave: exact keys (D bits << 32 | id, non-negative D orders like its
// bits, ties by id) of leafmates c = e * kFbThreadsRow + tid of pl[0, m),
// sorted over the wave; key[0] of lanes 0 .. 63 then holds the wave's 64
// smallest in order
template <typename XT, int E>
__device__ __forceinline__ void row_exact_wave(const XT* __restrict__ X, int d, const XT* xm, int me,
                                               const int32_t* __restrict__ pl, int m, int tid,
                                               int lane, uint64_t& k0) {
  uint64_t key[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int c = e * kFbThreadsRow + tid;
    key[e] = ~0ull;
    if (c < m) {
      const int id = pl[c];
      if (id != me)
        key[e] = ((uint64_t)__float_as_uint(exact_dist8(xm, X + (int64_t)id * d, d)) << 32) |
                 (uint32_t)id;
    }
  }
  bitonic_wave<uint64_t, E>(key, lane);
  k0 = key[0];
}

template <typename XT, int KP>
__global__ void __launch_bounds__(kFbThreadsRow)
k_row_exact_tree(const XT* __restrict__ X, int d, const int32_t* __restrict__ perm, int64_t n,
                 int T, const int32_t* __restrict__ invperm, const int32_t* __restrict__ rowitem,
                 const LeafItem* __restrict__ items, const int32_t* __restrict__ rows,
                 const int32_t* __restrict__ nrows, int64_t nrows_max,
                 int32_t* __restrict__ tmp_i, float* __restrict__ tmp_d) {
  constexpr int NW = kFbThreadsRow / 64;
  constexpr int kChunk = 8 * kFbThreadsRow;  // leafmates sorted per round
  static_assert((NW + 1) * KP <= 256, "the merge sorts four keys per lane");
  __shared__ uint64_t best[NW + 1][KP];      // per wave, then the running list
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t total = min((int64_t)*nrows, nrows_max) * T;
  for (int64_t p = blockIdx.x; p < total; p += gridDim.x) {
    const int64_t e = p / T;
    const int t = (int)(p - e * T);
    const int me = rows[e];
    const int pos = invperm[(int64_t)t * n + me];
    const LeafItem it = items[rowitem[(int64_t)t * n + pos]];
    const int32_t* pl = perm + (int64_t)t * n + it.start;
    const XT* xm = X + (int64_t)me * d;
    const int m = it.size;
    if (tid < KP) best[NW][tid] = ~0ull;
    // leafmates of one tree are distinct; per round each wave sorts its share
    // of kChunk of them, then wave 0 keeps the KP smallest of the waves' KP
    // smallest and the running list
    for (int c0 = 0; c0 < m; c0 += kChunk) {
      const int mc = min(m - c0, kChunk);
      uint64_t k0;
      if (mc <= 2 * kFbThreadsRow) row_exact_wave<XT, 2>(X, d, xm, me, pl + c0, mc, tid, lane, k0);
      else if (mc <= 4 * kFbThreadsRow) row_exact_wave<XT, 4>(X, d, xm, me, pl + c0, mc, tid, lane, k0);
      else row_exact_wave<XT, 8>(X, d, xm, me, pl + c0, mc, tid, lane, k0);
      if (lane < KP) best[w][lane] = k0;
      __syncthreads();
      if (w == 0) {
        uint64_t k4[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          const int j = h * 64 + lane;
          k4[h] = j < (NW + 1) * KP ? best[j / KP][j % KP] : ~0ull;
        }
        bitonic_wave<uint64_t, 4>(k4, lane);
        if (lane < KP) best[NW][lane] = k4[0];
      }
      __syncthreads();
    }
    if (tid < KP) {
      const uint64_t k = best[NW][tid];
      const bool ok = k != ~0ull;
      tmp_i[p * KP + tid] = ok ? (int32_t)(uint32_t)k : INT_MAX;
      tmp_d[p * KP + tid] = ok ? __uint_as_float((uint32_t)(k >> 32)) : INFINITY;
    }
    __syncthreads();
  }
}

template <int KP>
__global__ void k_row_exact_merge(int T, int K, const int32_t* __restrict__ rows,
                                  const int32_t* __restrict__ nrows, int64_t nrows_max,
                                  const int32_t* __restrict__ tmp_i, const float* __restrict__ tmp_d,
                                  int32_t* __restrict__ out_idx, float* __restrict__ out_dist) {
  const int64_t total = min((int64_t)*nrows, nrows_max);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    TopK<KP> top;
    top.init();
    for (int t = 0; t < T; ++t)
      for (int s = 0; s < KP; ++s) {
        const int64_t o = (e * T + t) * KP + s;
        const int id = tmp_i[o];
        if (id == INT_MAX) break;
        top.insert_unique(tmp_d[o], id);
      }
    const int me = rows[e];
    top.store(K, out_idx + (int64_t)me * K, out_dist + (int64_t)me * K);
  }
}

__global__ void k_invperm(const int32_t* __restrict__ perm, int64_t n, int T,
                          int32_t* __restrict__ inv) {
  const int64_t total = (int64_t)T * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = i / n;
    inv[t * n + perm[i]] = (int32_t)(i - t * n);
  }
}

// rowitem[tree * n + position] = the item covering that leaf position
__global__ void k_rowitem(const LeafItem* __restrict__ items, const int64_t* __restrict__ range,
                          int64_t n, int32_t* __restrict__ rowitem) {
  const int64_t b = range[0], e = range[1];
  for (int64_t i = b + blockIdx.x; i < e; i += gridDim.x) {
    const LeafItem it = items[i];
    const int r1 = it.row0 + it.rows;
    for (int r = it.row0 + threadIdx.x; r < r1; r += blockDim.x)
      rowitem[(int64_t)it.tree * n + it.start + r] = (int32_t)i;
  }
}

// Exact redo of single rows (candidate-list overflow, leaves too large to
// stage): one block per row, all leafmates.
constexpr int kFbThreads = 128;

template <typename XT, int KP>
__global__ void __launch_bounds__(kFbThreads)
k_knn_fallback(const XT* __restrict__ X, int d, const int32_t* __restrict__ perm, int64_t n,
               const LeafItem* __restrict__ items, const int32_t* __restrict__ fallback,
               const int32_t* __restrict__ nfallback, int64_t fb_cap, int K,
               int32_t* __restrict__ cand_idx, float* __restrict__ cand_dist) {
  __shared__ float ld[kFbThreads][KP];
  __shared__ int li[kFbThreads][KP];
  const int64_t total = min((int64_t)*nfallback, fb_cap);
  for (int64_t e = blockIdx.x; e < total; e += gridDim.x) {
    const LeafItem it = items[fallback[2 * e]];
    const int r = fallback[2 * e + 1];
    const int32_t* pl = perm + (int64_t)it.tree * n + it.start;
    const int me = pl[it.row0 + r];
    TopK<KP> top;
    top.init();
    for (int c = threadIdx.x; c < it.size; c += kFbThreads) {
      const int id = pl[c];
      if (id != me) top.insert(exact_dist(X, me, id, d), id);
    }
#pragma unroll
    for (int s = 0; s < KP; ++s) {
      ld[threadIdx.x][s] = top.d[s];
      li[threadIdx.x][s] = top.i[s];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int t = 1; t < kFbThreads; ++t)
        for (int s = 0; s < KP; ++s) {
          if (li[t][s] == INT_MAX) break;
          top.insert(ld[t][s], li[t][s]);
        }
      const int64_t o = ((int64_t)it.tree * n + me) * K;
      top.store(K, cand_idx + o, cand_dist + o);
    }
    __syncthreads();
  }
}

// bf16 copy (round to nearest even), |x~|^2 and |x~ - x| (rounded up) of one
// row (the whole wave). The squared norm is reduced in FP32, or in FP64
// (rounded once) for d >= 512, where d FP32 additions would loosen the filter
// bound.
template <typename XT>
__device__ __forceinline__ bool prep_bf16_row(const XT* __restrict__ X, int64_t row, int d, int lane,
                                              uint16_t* __restrict__ Xb, float2* __restrict__ nbe,
                                              uint8_t* __restrict__ Xu) {
  const bool wide = d >= 512;
  float s = 0.0f;
  double s64 = 0.0, e2 = 0.0;
  bool bad = false;  // Xu: some value of the lane's is not an integer in [0, 255]
  for (int k = lane; k < d; k += 64) {
    const float x = (float)X[row * d + k];
    uint32_t bits = __float_as_uint(x);
    bits += 0x7fffu + ((bits >> 16) & 1u);
    const uint16_t b = (uint16_t)(bits >> 16);
    Xb[row * d + k] = b;
    if (Xu != nullptr) {
      if (x >= 0.0f && x <= 255.0f && x == rintf(x)) Xu[row * d + k] = (uint8_t)x ^ 0x80u;
      else bad = true;
    }
    const float xb = __uint_as_float((uint32_t)b << 16);
    if (wide) s64 += (double)xb * (double)xb;
    else s = fmaf(xb, xb, s);
    const double df = (double)xb - (double)x;
    e2 += df * df;
  }
  if (wide) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s64 += __shfl_xor(s64, off);
  } else {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off);
  if (lane == 0) {
    float ef = (float)sqrt(e2);
    if (e2 > 0.0) ef = ef * (1.0f + 1.0e-6f) + 1.0e-30f;
    nbe[row] = make_float2(wide ? (float)s64 : s, ef);
  }
  return bad;
}

// One wave per row, the waves of the grid striding over the rows (a grid of
// a few thousand blocks instead of one block per four rows).
// Xu (optional): the int8 copy x - 128, written where the values are integers
// in [0, 255]; *notu8 = 1 when some value is not (the caller zeroes it).
template <typename XT>
__global__ void k_prep_bf16(const XT* __restrict__ X, int64_t n, int d, uint16_t* __restrict__ Xb,
                            float2* __restrict__ nbe, uint8_t* __restrict__ Xu,
                            int32_t* __restrict__ notu8) {
  const int lane = threadIdx.x & 63;
  const int64_t wstride = ((int64_t)gridDim.x * blockDim.x) >> 6;
  bool bad = false;
  for (int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; row < n;
       row += wstride)
    bad = prep_bf16_row(X, row, d, lane, Xb, nbe, Xu) || bad;
  if (Xu != nullptr && __ballot(bad) != 0 && lane == 0) atomicOr(notu8, 1);
}

// The same per-row results for float rows of d % 4 == 0 values read as float4
// chunks: G lanes per row (a power of two >= d / 4, at most 64), 64 / G rows
// per wave, the waves striding over groups of rows; each lane stores its four
// bf16 values as one 8-byte word and its four int8 values as one 4-byte word
// (where a value is not an integer in [0, 255] that word is never read: the
// flag is set).
template <int G>
__global__ void k_prep_bf16_v4(const float* __restrict__ X, int64_t n, int d,
                               uint16_t* __restrict__ Xb, float2* __restrict__ nbe,
                               uint8_t* __restrict__ Xu, int32_t* __restrict__ notu8) {
  constexpr int RPW = 64 / G;
  const int lane = threadIdx.x & 63;
  const int sub = lane & (G - 1), rw = lane / G;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int q = d >> 2;
  const bool wide = d >= 512;
  bool bad = false;
  for (int64_t r0 = wave * RPW; r0 < n; r0 += nwaves * RPW) {
    const int64_t row = r0 + rw;
    float s = 0.0f;
    double s64 = 0.0, e2 = 0.0;
    if (row < n) {
      const float4* xr = reinterpret_cast<const float4*>(X + row * d);
      uint2* br = reinterpret_cast<uint2*>(Xb + row * d);
      uint32_t* ur = Xu != nullptr ? reinterpret_cast<uint32_t*>(Xu + row * d) : nullptr;
      for (int k = sub; k < q; k += G) {
        const float4 x4 = xr[k];
        const float xs[4] = {x4.x, x4.y, x4.z, x4.w};
        uint32_t bw[4], u = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float x = xs[e];
          uint32_t bits = __float_as_uint(x);
          bits += 0x7fffu + ((bits >> 16) & 1u);
          bw[e] = bits >> 16;
          if (x >= 0.0f && x <= 255.0f && x == rintf(x)) u |= ((uint32_t)(uint8_t)x ^ 0x80u) << (8 * e);
          else bad = true;
          const float xb = __uint_as_float(bw[e] << 16);
          if (wide) s64 += (double)xb * (double)xb;
          else s = fmaf(xb, xb, s);
          const double df = (double)xb - (double)x;
          e2 += df * df;
        }
        br[k] = make_uint2(bw[0] | (bw[1] << 16), bw[2] | (bw[3] << 16));
        if (ur != nullptr) ur[k] = u;
      }
    }
    // sums over the G lanes of each row (xor offsets below G stay in the row)
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
      if (wide) s64 += __shfl_xor(s64, off);
      else s += __shfl_xor(s, off);
      e2 += __shfl_xor(e2, off);
    }
    if (sub == 0 && row < n) {
      float ef = (float)sqrt(e2);
      if (e2 > 0.0) ef = ef * (1.0f + 1.0e-6f) + 1.0e-30f;
      nbe[row] = make_float2(wide ? (float)s64 : s, ef);
    }
  }
  if (Xu != nullptr && __ballot(bad) != 0 && lane == 0) atomicOr(notu8, 1);
}

// FP8 quantisation, one wave per row: scale 2^e with e the smallest exponent
// for which max |x| / 2^e <= 448 (the e4m3 maximum, so no element saturates),
// clamped to [-120, 120]; q = e4m3(x / 2^e) rounded to nearest by the
// hardware conversion (the format the fp8 MFMA reads), x~ = 2^e q. The row's
// |x~|^2 and |x~ - x| (FP64, rounded up) are reduced as for bf16.
template <typename XT>
__global__ void k_prep_fp8(const XT* __restrict__ X, int64_t n, int d, uint8_t* __restrict__ Xq,
                           float2* __restrict__ nbe, int8_t* __restrict__ pexp) {
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= n) return;
  const XT* __restrict__ xr = X + row * d;
  float mx = 0.0f;
  for (int k = lane; k < d; k += 64) mx = fmaxf(mx, fabsf((float)xr[k]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  int e = 0;
  if (mx > 0.0f) {
    e = ilogbf(mx) - 8;  // mx / 2^e in [256, 512)
    if (ldexpf(mx, -e) > 448.0f) ++e;
    e = max(-120, min(120, e));
  }
  const bool wide = d >= 512;
  float s = 0.0f;
  double s64 = 0.0, e2 = 0.0;
  for (int k0 = 4 * lane; k0 < d; k0 += 256) {
    float x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = k0 + u < d ? (float)xr[k0 + u] : 0.0f;
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(x[0], -e), ldexpf(x[1], -e), 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(x[2], -e), ldexpf(x[3], -e), w, true);
    const float dq[4] = {__builtin_amdgcn_cvt_f32_fp8(w, 0), __builtin_amdgcn_cvt_f32_fp8(w, 1),
                         __builtin_amdgcn_cvt_f32_fp8(w, 2), __builtin_amdgcn_cvt_f32_fp8(w, 3)};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k0 + u >= d) break;
      const float xb = ldexpf(dq[u], e);
      if (wide) s64 += (double)xb * (double)xb;
      else s = fmaf(xb, xb, s);
      const double df = (double)xb - (double)x[u];
      e2 += df * df;
    }
    uint8_t* dst = Xq + row * d + k0;
    if ((d & 3) == 0) {
      *reinterpret_cast<uint32_t*>(dst) = (uint32_t)w;
    } else {
      for (int u = 0; u < 4 && k0 + u < d; ++u) dst[u] = (uint8_t)((uint32_t)w >> (8 * u));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off);
    s64 += __shfl_xor(s64, off);
    e2 += __shfl_xor(e2, off);
  }
  if (lane == 0) {
    float ef = (float)sqrt(e2);
    if (e2 > 0.0) ef = ef * (1.0f + 1.0e-6f) + 1.0e-30f;
    nbe[row] = make_float2(wide ? (float)s64 : s, ef);
    pexp[row] = (int8_t)e;
  }
}

template <typename T>
int persistent_grid(drpt_ctx* ctx, T kernel, int threads, unsigned* grid) {
  int per_cu = 0;
  DRPT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
  if (per_cu < 1) per_cu = 1;
  *grid = (unsigned)((int64_t)ctx->num_cus * per_cu);
  return DRPT_OK;
}

template <typename XT, typename EL, int T, int MODE, bool RING>
int launch_lowp_r(drpt_ctx* ctx, const KnnArgs& a) {
  unsigned grid;
  constexpr int NT = 4 * lowp_rows<MODE>();
  DRPT_CHECK(persistent_grid(ctx, k_knn_lowp<XT, EL, T, MODE, RING>, NT, &grid));
  hipLaunchKernelGGL((k_knn_lowp<XT, EL, T, MODE, RING>), dim3(grid), dim3(NT), 0, ctx->stream, a);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

template <typename XT, typename EL, int T, int MODE>
int launch_lowp_t(drpt_ctx* ctx, const KnnArgs& a) {
  if constexpr (MODE != kModeTree) {
    const int rb = a.d * EL::kBytes;
    if (a.ring && (rb & 15) == 0 && rb <= kRingRow)
      return launch_lowp_r<XT, EL, T, MODE, true>(ctx, a);
  }
  return launch_lowp_r<XT, EL, T, MODE, false>(ctx, a);
}

template <typename XT, typename EL, int T>
int launch_lowp_m(drpt_ctx* ctx, int mode, const KnnArgs& a) {
  if (mode == kModeTree) return launch_lowp_t<XT, EL, T, kModeTree>(ctx, a);
  if (mode == kModeCand) return launch_lowp_t<XT, EL, T, kModeCand>(ctx, a);
  return launch_lowp_t<XT, EL, T, kModePool>(ctx, a);
}

template <typename XT, typename EL>
int launch_lowp_x(drpt_ctx* ctx, int t, int mode, const KnnArgs& a) {
  if (t <= 3) return launch_lowp_m<XT, EL, 3>(ctx, mode, a);
  if (t <= 4) return launch_lowp_m<XT, EL, 4>(ctx, mode, a);
  if (t <= 6) return launch_lowp_m<XT, EL, 6>(ctx, mode, a);
  if (t <= 8) return launch_lowp_m<XT, EL, 8>(ctx, mode, a);
  if (t <= 12) return launch_lowp_m<XT, EL, 12>(ctx, mode, a);
  return launch_lowp_m<XT, EL, 16>(ctx, mode, a);
}

template <typename XT>
int launch_lowp_e(drpt_ctx* ctx, int precision, int t, int mode, const KnnArgs& a) {
  if (precision == DRPT_FP8) return launch_lowp_x<XT, ElFp8>(ctx, t, mode, a);
  return launch_lowp_x<XT, ElBf16>(ctx, t, mode, a);
}

// per-thread slots of the bound pass (4 threads per row)
int slots_of(int kp) {
  const int t = (kp + 3) / 4;
  return t <= 3 ? 3 : t <= 4 ? 4 : t <= 6 ? 6 : t <= 8 ? 8 : t <= 12 ? 12 : 16;
}

template <typename XT>
int dispatch_exact(drpt_ctx* ctx, const drpt_params* p, const XT* X, const int32_t* perm,
                   const LeafItem* items, const int64_t* range, int32_t* ci, float* cd,
                   unsigned* q) {
  unsigned grid;
  DRPT_CHECK(persistent_grid(ctx, k_knn_exact<XT>, kThreads, &grid));
  hipLaunchKernelGGL((k_knn_exact<XT>), dim3(grid), dim3(kThreads), 0, ctx->stream, X, p->d, perm,
                     p->n, items, range, p->nn, ci, cd, q);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

}  // namespace

int launch_prep_lowp(drpt_ctx* ctx, const void* X, int dtype, int64_t n, int d, int precision,
                     void* Xq, float2* nbe, int8_t* pexp, uint8_t* xi8, int32_t* notu8) {
  if (precision != DRPT_BF16 && precision != DRPT_FP8) {
    set_error("prep_lowp: precision not supported");
    return DRPT_ERR_UNSUPPORTED;
  }
  const int64_t threads = n * 64;
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  // k_prep_bf16 strides over the rows: a few resident rounds of blocks
  const unsigned sblocks = (unsigned)std::min<int64_t>(blocks, (int64_t)ctx->num_cus * 8);
  if (precision == DRPT_FP8) {
    if (dtype == DRPT_U8)
      hipLaunchKernelGGL(k_prep_fp8<uint8_t>, dim3(blocks), dim3(256), 0, ctx->stream,
                         (const uint8_t*)X, n, d, (uint8_t*)Xq, nbe, pexp);
    else
      hipLaunchKernelGGL(k_prep_fp8<float>, dim3(blocks), dim3(256), 0, ctx->stream,
                         (const float*)X, n, d, (uint8_t*)Xq, nbe, pexp);
  } else if (dtype == DRPT_U8)
    hipLaunchKernelGGL(k_prep_bf16<uint8_t>, dim3(sblocks), dim3(256), 0, ctx->stream,
                       (const uint8_t*)X, n, d, (uint16_t*)Xq, nbe, nullptr, nullptr);
  else if (d % 4 == 0 && ((uintptr_t)X & 15) == 0) {
    // float4 chunks: G lanes per row
    uint8_t* xu = notu8 ? xi8 : nullptr;
    const int q = d / 4;
    const int G = q <= 8 ? 8 : q <= 16 ? 16 : q <= 32 ? 32 : 64;
    const int64_t waves = (n * G + 63) / 64;
    const unsigned vblocks = (unsigned)std::min<int64_t>((waves + 3) / 4, (int64_t)ctx->num_cus * 8);
    auto* f = G == 8 ? k_prep_bf16_v4<8> : G == 16 ? k_prep_bf16_v4<16>
            : G == 32 ? k_prep_bf16_v4<32> : k_prep_bf16_v4<64>;
    hipLaunchKernelGGL(f, dim3(vblocks), dim3(256), 0, ctx->stream, (const float*)X, n, d,
                       (uint16_t*)Xq, nbe, xu, notu8);
  } else
    hipLaunchKernelGGL(k_prep_bf16<float>, dim3(sblocks), dim3(256), 0, ctx->stream,
                       (const float*)X, n, d, (uint16_t*)Xq, nbe, notu8 ? xi8 : nullptr, notu8);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_leaf_gather(drpt_ctx* ctx, const int32_t* perm, int64_t total, const float2* nbe,
                       const int8_t* pexp, float* lnbT, float* lerrT, int8_t* lexpT,
                       const int32_t* gate) {
  const int64_t blocks = std::min<int64_t>((total + 255) / 256, (int64_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(k_leaf_gather, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0,
                     ctx->stream, perm, total, nbe, pexp, lnbT, lerrT, lexpT, gate);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int lowp_slots(const drpt_params* p) { return 4 * slots_of(p->kprime); }

int launch_leaf_knn(drpt_ctx* ctx, const drpt_params* p, const void* X, const int32_t* perm,
                    const LeafItem* items, const int64_t* range, int32_t* cand_idx,
                    float* cand_dist) {
  if (p->nn < 1 || p->nn > 64) {
    set_error("leaf_knn: nn must be in [1, 64]");
    return DRPT_ERR_UNSUPPORTED;
  }
  unsigned* q = nullptr;
  DRPT_CHECK(ws(ctx, "knn.queue", kQueues, &q));
  DRPT_CHECK(fill_u32(ctx, q, 0u, kQueues));
  if (p->dtype == DRPT_U8)
    return dispatch_exact<uint8_t>(ctx, p, (const uint8_t*)X, perm, items, range, cand_idx,
                                   cand_dist, q);
  return dispatch_exact<float>(ctx, p, (const float*)X, perm, items, range, cand_idx, cand_dist, q);
}

int launch_lowp_knn(drpt_ctx* ctx, const drpt_params* p, int mode, KnnArgs a) {
  const int K = p->nn, kp = p->kprime;
  if (K < 1 || K > 64) {
    set_error("leaf_knn: nn must be in [1, 64]");
    return DRPT_ERR_UNSUPPORTED;
  }
  if (p->precision != DRPT_BF16 && p->precision != DRPT_FP8) {
    set_error("leaf_knn: precision not supported");
    return DRPT_ERR_UNSUPPORTED;
  }
  if (kp < K || kp > 64) {
    set_error("leaf_knn: kprime must be in [nn, 64] for BF16 / FP8");
    return DRPT_ERR_UNSUPPORTED;
  }
  if (a.precision != p->precision || (p->precision == DRPT_FP8 && !a.lexpT)) {
    set_error("leaf_knn: quantised inputs do not match the precision");
    return DRPT_ERR_ARG;
  }
  DRPT_CHECK(ws(ctx, "knn.queue", kQueues, &a.queue));
  DRPT_CHECK(fill_u32(ctx, a.queue, 0u, kQueues));
  const int t = slots_of(kp);
  if (p->dtype == DRPT_U8) return launch_lowp_e<uint8_t>(ctx, p->precision, t, mode, a);
  return launch_lowp_e<float>(ctx, p->precision, t, mode, a);
}

int launch_leaf_gather_val(drpt_ctx* ctx, const int32_t* perm, int64_t total, const float* src,
                           int stride, int offset, float* outT) {
  const int64_t blocks =
      std::min<int64_t>((total + 256 * kGatherUnroll - 1) / (256 * kGatherUnroll),
                        (int64_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(k_leaf_gather_val, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0,
                     ctx->stream, perm, total, src, stride, offset, outT);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

// the row staged in LDS for rows of 128 to kMergeRowLds elements (multiples
// of 4) at E <= 4: measured faster there (1M x 128, 1M x 960), slower at d = 96
template <typename XT, int E>
void launch_rm(drpt_ctx* ctx, unsigned grid, const RowMergeArgs& a) {
  if constexpr (E <= 4) {
    if ((a.d & 3) == 0 && a.d >= 128 && a.d <= kMergeRowLds) {
      hipLaunchKernelGGL((k_row_merge<XT, E, true>), dim3(grid), dim3(64 * kMergeWaves), 0, ctx->stream, a);
      return;
    }
  }
  hipLaunchKernelGGL((k_row_merge<XT, E, false>), dim3(grid), dim3(64 * kMergeWaves), 0, ctx->stream, a);
}

int launch_lossy_flag(drpt_ctx* ctx, const float2* nbe, int64_t n, int32_t* lossy) {
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1024));
  hipLaunchKernelGGL(k_lossy_flag, dim3(grid), dim3(256), 0, ctx->stream, nbe, n, lossy);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_row_bound(drpt_ctx* ctx, const uint64_t* src_key, int64_t tree_stride, int S, int ntsrc,
                     int64_t n, int d, int K, const float2* nbe, float* out_bound) {
  const int slots = S * ntsrc;
  if (K < 1 || K > 64 || slots > 256 || slots < 1) {
    set_error("row_bound: unsupported sizes (nn in [1, 64], at most 256 candidate slots)");
    return DRPT_ERR_UNSUPPORTED;
  }
  const unsigned grid = (unsigned)std::min<int64_t>((n + kMergeWaves - 1) / kMergeWaves,
                                                    (int64_t)ctx->num_cus * 16);
  const int E = slots <= 64 ? 1 : slots <= 128 ? 2 : 4;
  if (E < (K + 63) / 64) {
    set_error("row_bound: fewer candidate slots than nn");
    return DRPT_ERR_UNSUPPORTED;
  }
  if (E == 1)
    hipLaunchKernelGGL(k_row_bound<1>, dim3(grid), dim3(64 * kMergeWaves), 0, ctx->stream, src_key,
                       tree_stride, S, ntsrc, n, d, K, nbe, out_bound);
  else if (E == 2)
    hipLaunchKernelGGL(k_row_bound<2>, dim3(grid), dim3(64 * kMergeWaves), 0, ctx->stream, src_key,
                       tree_stride, S, ntsrc, n, d, K, nbe, out_bound);
  else
    hipLaunchKernelGGL(k_row_bound<4>, dim3(grid), dim3(64 * kMergeWaves), 0, ctx->stream, src_key,
                       tree_stride, S, ntsrc, n, d, K, nbe, out_bound);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_row_merge(drpt_ctx* ctx, int dtype, RowMergeArgs a) {
  const int slots = a.S * a.ntsrc;
  if (a.ntsrc == 1 && a.src != nullptr && a.src_key == nullptr && a.cnt_tree != nullptr &&
      a.K <= 32 && a.rerank == 0 && a.nrows_dev == nullptr && kM2Fresh + a.K < (1 << kM2HashBits) &&
      (a.S <= kM2Fresh || (a.heavy != nullptr && a.nheavy != nullptr))) {
    // pools: two rows per wave; rows of more than kM2Fresh entries listed and
    // merged by the full-width pass after it
    if (a.S > kM2Fresh) DRPT_CHECK(fill_u32(ctx, (uint32_t*)a.nheavy, 0u, 1));
    const int64_t want = std::min<int64_t>((a.n + 2 * kMergeWaves - 1) / (2 * kMergeWaves),
                                           (int64_t)ctx->num_cus * 16);
    const unsigned grid = (unsigned)(((std::max<int64_t>(want, 8) + 7) / 8) * 8);
    const bool rl = (a.d & 3) == 0 && a.d >= 128 && a.d <= kM2RowLds;
    if (dtype == DRPT_U8) {
      if (rl) hipLaunchKernelGGL((k_row_merge2<uint8_t, true>), dim3(grid), dim3(64 * kMergeWaves), 0, ctx->stream, a);
      else hipLaunchKernelGGL((k_row_merge2<uint8_t, false>), dim3(grid), dim3(64 * kMergeWaves), 0, ctx->stream, a);
    } else {
      if (rl) hipLaunchKernelGGL((k_row_merge2<float, true>), dim3(grid), dim3(64 * kMergeWaves), 0, ctx->stream, a);
      else hipLaunchKernelGGL((k_row_merge2<float, false>), dim3(grid), dim3(64 * kMergeWaves), 0, ctx->stream, a);
    }
    DRPT_HIP(hipGetLastError());
    if (a.S <= kM2Fresh) return DRPT_OK;
    a.order = a.heavy;
    a.nrows_dev = a.nheavy;
    a.heavy = nullptr;
    a.nheavy = nullptr;
  } else if (a.K < 1 || a.K > 64 || slots > 1024 || a.rerank > 64 * 4 || a.ntsrc > 64 ||
             (a.src_key != nullptr && slots > 256)) {
    set_error("row_merge: unsupported sizes (nn in [1, 64], at most 1024 candidate slots, "
              "256 with approximate keys)");
    return DRPT_ERR_UNSUPPORTED;
  }
  // a multiple of 8 blocks: every XCD chunk of rows has blocks (k_row_merge)
  const int64_t want = std::min<int64_t>((a.n + kMergeWaves - 1) / kMergeWaves,
                                         (int64_t)ctx->num_cus * 16);
  const unsigned grid = (unsigned)(((std::max<int64_t>(want, 8) + 7) / 8) * 8);
  const bool u8 = dtype == DRPT_U8;
  if (a.heavy != nullptr && a.nheavy != nullptr && a.ntsrc == 1 && a.cnt_tree != nullptr &&
      a.src != nullptr && a.rerank == 0 && slots > 256) {
    // light pass over every row (rows of at most 256 candidates merged, the
    // others listed), then the full pass over the listed rows only
    DRPT_CHECK(fill_u32(ctx, (uint32_t*)a.nheavy, 0u, 1));
    RowMergeArgs l = a;
    l.nrows_dev = nullptr;
    if (u8) launch_rm<uint8_t, 4>(ctx, grid, l);
    else launch_rm<float, 4>(ctx, grid, l);
    DRPT_HIP(hipGetLastError());
    a.order = a.heavy;
    a.nrows_dev = a.nheavy;
    a.heavy = nullptr;
    a.nheavy = nullptr;
  } else {
    a.heavy = nullptr;
    a.nheavy = nullptr;
  }
  const int E = slots <= 64 ? 1 : slots <= 128 ? 2 : slots <= 256 ? 4 : slots <= 512 ? 8 : 16;
  if (E == 16) {
    if (u8) launch_rm<uint8_t, 16>(ctx, grid, a);
    else launch_rm<float, 16>(ctx, grid, a);
  } else if (E == 8) {
    if (u8) launch_rm<uint8_t, 8>(ctx, grid, a);
    else launch_rm<float, 8>(ctx, grid, a);
  } else if (E == 1) {
    if (u8) launch_rm<uint8_t, 1>(ctx, grid, a);
    else launch_rm<float, 1>(ctx, grid, a);
  } else if (E == 2) {
    if (u8) launch_rm<uint8_t, 2>(ctx, grid, a);
    else launch_rm<float, 2>(ctx, grid, a);
  } else {
    if (u8) launch_rm<uint8_t, 4>(ctx, grid, a);
    else launch_rm<float, 4>(ctx, grid, a);
  }
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_row_exact(drpt_ctx* ctx, const drpt_params* p, const void* X, const int32_t* perm,
                     const LeafItem* items, const int64_t* range, const int32_t* ovf_list,
                     const int32_t* novf, int64_t nrows_max, int32_t* out_idx, float* out_dist) {
  const int64_t n = p->n;
  const int T = p->ntrees, K = p->nn;
  if (K > 32) {
    set_error("row_exact: nn > 32 is not supported");
    return DRPT_ERR_UNSUPPORTED;
  }
  int32_t *inv, *rowitem;
  DRPT_CHECK(ws(ctx, "knn.invperm", (size_t)T * n, &inv));
  DRPT_CHECK(ws(ctx, "knn.rowitem", (size_t)T * n, &rowitem));
  const unsigned g = (unsigned)std::min<int64_t>(((int64_t)T * n + 255) / 256, (int64_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(k_invperm, dim3(g), dim3(256), 0, ctx->stream, perm, n, T, inv);
  DRPT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_rowitem, dim3((unsigned)ctx->num_cus * 8), dim3(256), 0, ctx->stream, items,
                     range, n, rowitem);
  DRPT_HIP(hipGetLastError());
  // per-tree lists of at most kRowExactScratch bytes of rows at a time: the
  // rows [b0, b0 + batch) of ovf_list per pair of launches (nrows_max is the
  // host's count of listed rows, so every batch holds listed rows only)
  const int KPmax = K <= 16 ? 16 : 32;
  const int64_t per_row = (int64_t)T * KPmax * 8;
  int64_t batch = std::max<int64_t>(1, std::min<int64_t>(nrows_max, kRowExactScratch / per_row));
  if (ctx->row_exact_batch > 0) batch = std::min<int64_t>(batch, ctx->row_exact_batch);
  int32_t* ti;
  float* td;
  DRPT_CHECK(ws(ctx, "knn.rowx_i", (size_t)batch * T * KPmax, &ti));
  DRPT_CHECK(ws(ctx, "knn.rowx_d", (size_t)batch * T * KPmax, &td));
  for (int64_t b0 = 0; b0 < nrows_max; b0 += batch) {
    const int64_t nb = std::min(batch, nrows_max - b0);
    const int32_t* rows = ovf_list + b0;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nb * T, (int64_t)ctx->num_cus * 4));
    const unsigned mgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((nb + 255) / 256, 1024));
    auto go = [&](auto kp) -> int {
      constexpr int KP = decltype(kp)::value;
      if (p->dtype == DRPT_U8)
        hipLaunchKernelGGL((k_row_exact_tree<uint8_t, KP>), dim3(grid), dim3(kFbThreadsRow), 0,
                           ctx->stream, (const uint8_t*)X, p->d, perm, n, T, inv, rowitem, items,
                           rows, novf, nb, ti, td);
      else
        hipLaunchKernelGGL((k_row_exact_tree<float, KP>), dim3(grid), dim3(kFbThreadsRow), 0,
                           ctx->stream, (const float*)X, p->d, perm, n, T, inv, rowitem, items,
                           rows, novf, nb, ti, td);
      DRPT_HIP(hipGetLastError());
      hipLaunchKernelGGL((k_row_exact_merge<KP>), dim3(mgrid), dim3(256), 0, ctx->stream, T, K,
                         rows, novf, nb, ti, td, out_idx, out_dist);
      DRPT_HIP(hipGetLastError());
      return DRPT_OK;
    };
    if (K <= 16) DRPT_CHECK(go(std::integral_constant<int, 16>()));
    else DRPT_CHECK(go(std::integral_constant<int, 32>()));
  }
  return DRPT_OK;
}

int launch_fallback(drpt_ctx* ctx, const drpt_params* p, const void* X, const int32_t* perm,
                    const LeafItem* items, const int32_t* fallback, const int32_t* nfallback,
                    int64_t fallback_cap, int32_t* cand_idx, float* cand_dist) {
  const unsigned grid = (unsigned)(ctx->num_cus * 2);
  const int K = p->nn;
  if (K > 32) {
    set_error("fallback: nn > 32 is not supported");
    return DRPT_ERR_UNSUPPORTED;
  }
  if (p->dtype == DRPT_U8) {
    if (K <= 16)
      hipLaunchKernelGGL((k_knn_fallback<uint8_t, 16>), dim3(grid), dim3(kFbThreads), 0, ctx->stream,
                         (const uint8_t*)X, p->d, perm, p->n, items, fallback, nfallback,
                         fallback_cap, K, cand_idx, cand_dist);
    else
      hipLaunchKernelGGL((k_knn_fallback<uint8_t, 32>), dim3(grid), dim3(kFbThreads), 0, ctx->stream,
                         (const uint8_t*)X, p->d, perm, p->n, items, fallback, nfallback,
                         fallback_cap, K, cand_idx, cand_dist);
  } else {
    if (K <= 16)
      hipLaunchKernelGGL((k_knn_fallback<float, 16>), dim3(grid), dim3(kFbThreads), 0, ctx->stream,
                         (const float*)X, p->d, perm, p->n, items, fallback, nfallback,
                         fallback_cap, K, cand_idx, cand_dist);
    else
      hipLaunchKernelGGL((k_knn_fallback<float, 32>), dim3(grid), dim3(kFbThreads), 0, ctx->stream,
                         (const float*)X, p->d, perm, p->n, items, fallback, nfallback,
                         fallback_cap, K, cand_idx, cand_dist);
  }
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

}  // namespace drpt
