// Stage (c): leaf-local brute-force k-NN on gfx950.
//
// Replaces MDRPT::calculate_nns + MathOp::calculate_distance
// (cpp/src/drpt/algo/mdrpt.cpp:123-212, math_operations.cpp:385-396): every
// ordered pair of leafmates, per-row ordering, head kept.
//
// Work list: one item per (leaf, block of 64 rows); a persistent grid pulls
// items from an atomic queue (largest leaves first) so uneven leaves balance.
//
// Exact distance (the ranking every path ends with):
//   D(x, y) = fma-chain over k = 0..d-1 of t_k = fl(x_k - y_k): acc = fma(t_k, t_k, acc)
// Results per row: the K nearest non-self leafmates ordered by (D, index).
//
// Paths:
//   FP32   D for all m^2 pairs, 64x64 tiles, 4x4 register blocking (vector ALU).
//   BF16   Gram X~ X~^T of the bf16-rounded rows on the MFMA cores
//          (v_mfma_f32_32x32x16_bf16, FP32 accumulate), approximate distance
//          a = |x~_i|^2 + |x~_j|^2 - 2 G_ij, widened top-k' by (a, index),
//          exact D for the k' candidates, top-K by (D, index). A row is
//          certified when its K-th exact distance is below a lower bound of
//          every non-candidate's D (derived from the (k'+1)-th approximate
//          value, the bf16 rounding error norms and the FP32 accumulation
//          error); uncertified rows are redone exactly (k_knn_fallback), so
//          the lists equal the FP32 path's.
//
// Roofline: FP32 path vector-FP32 bound (3 flops per dim per pair); BF16 path
// MFMA bound, algorithmic flops 2*d*sum(m^2).
#include <climits>

#include "internal.hpp"

namespace drpt {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

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
    const int rows = min(kRows, m - it.row0);
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
// BF16 candidate path (MFMA) + exact re-rank + certification.
//
// One work item = 64 rows of one leaf. LDS holds, in turn: the bf16 row and
// column tiles and the 64x64 approximate-distance tile (Gram phase), then a
// chunk of dimensions of every point of the leaf as FP32 (re-rank). Each wave
// keeps the candidate queues of 16 rows. Candidates are leaf positions (ties
// broken by position: the certification below depends only on the (k'+1)-th
// approximate value, so the choice among tied candidates cannot change the
// result).
// ---------------------------------------------------------------------------
constexpr int kGramChunk = 128;                  // dims per Gram pass
constexpr int kTilePitch = kGramChunk + 8;       // bf16 elements per LDS row
constexpr int kBigBytes = 2 * kRows * kTilePitch * 2 + kRows * (kCols + 1) * 4;
constexpr int kMaxStagedLeaf = kBigBytes / 4;    // leaves above this go to the exact fallback

constexpr int kLeafStage = 2048;  // leaves above this go to the exact fallback

// 8 consecutive bf16 of row id from dim k (zeros past d, or for id < 0)
__device__ __forceinline__ uint4 load_piece(const uint16_t* __restrict__ Xb, int id, int d, int k) {
  uint4 v = make_uint4(0, 0, 0, 0);
  if (id < 0 || k >= d) return v;
  if ((d & 7) == 0) return *reinterpret_cast<const uint4*>(Xb + (int64_t)id * d + k);
  uint16_t t[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) t[u] = (k + u < d) ? Xb[(int64_t)id * d + k + u] : (uint16_t)0;
  return make_uint4(t[0] | (t[1] << 16), t[2] | (t[3] << 16), t[4] | (t[5] << 16), t[6] | (t[7] << 16));
}

template <typename XT>
__global__ void __launch_bounds__(kThreads, 2)
k_knn_bf16(const XT* __restrict__ X, const uint16_t* __restrict__ Xb,
           const float* __restrict__ nb, const float* __restrict__ err, int d,
           const int32_t* __restrict__ perm, int64_t n, const LeafItem* __restrict__ items,
           const int64_t* __restrict__ range, int K, int kp, int32_t* __restrict__ cand_idx,
           float* __restrict__ cand_dist, unsigned* __restrict__ queue,
           int32_t* __restrict__ fallback, int32_t* __restrict__ nfallback, int64_t fb_cap,
           const float* __restrict__ bound) {
  __shared__ __align__(16) unsigned char big[kBigBytes];
  uint16_t(*As)[kTilePitch] = reinterpret_cast<uint16_t(*)[kTilePitch]>(big);
  uint16_t(*Bs)[kTilePitch] = reinterpret_cast<uint16_t(*)[kTilePitch]>(big + kRows * kTilePitch * 2);
  float(*Ds)[kCols + 1] = reinterpret_cast<float(*)[kCols + 1]>(big + 2 * kRows * kTilePitch * 2);
  float* Ls = reinterpret_cast<float*>(big);
  __shared__ int lid[kLeafStage];     // point id of every leaf position
  __shared__ float lnb[kLeafStage];   // |x~|^2 of every leaf position
  __shared__ float red_n[kThreads / 64], red_e[kThreads / 64];
  __shared__ int s_ticket[2];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const bool a_resident = d <= kGramChunk;
  float qd[kRowsPerWave];
  int qi[kRowsPerWave];
  const int64_t begin = range[0], end = range[1];
  if (tid == 0) s_ticket[0] = (int)atomicAdd(queue, 1u);
  __syncthreads();
  int64_t item = begin + s_ticket[0];
  for (int parity = 1;; parity ^= 1) {
    if (item >= end) break;
    if (tid == 0) s_ticket[parity] = (int)atomicAdd(queue, 1u);  // next item, fetched early
    const LeafItem it = items[item];
    const int32_t* pl = perm + (int64_t)it.tree * n + it.start;
    const int m = it.size;
    const int rows = min(kRows, m - it.row0);
    if (m > kLeafStage) {  // too large to stage: exact path for these rows
      if (tid < rows) {
        const int slot = atomicAdd(nfallback, 1);
        if (slot < fb_cap) {
          fallback[2 * slot] = (int32_t)item;
          fallback[2 * slot + 1] = tid;
        }
      }
      __syncthreads();
      item = begin + s_ticket[parity];
      continue;
    }
    // stage ids and norms of the whole leaf; leaf maxima for the bound
    float mn = 0.0f, me_ = 0.0f;
    for (int j = tid; j < m; j += kThreads) {
      const int id = pl[j];
      const float v = nb[id];
      lid[j] = id;
      lnb[j] = v;
      mn = fmaxf(mn, v);
      me_ = fmaxf(me_, err[id]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      mn = fmaxf(mn, __shfl_xor(mn, off));
      me_ = fmaxf(me_, __shfl_xor(me_, off));
    }
    if (lane == 0) {
      red_n[wave] = mn;
      red_e[wave] = me_;
    }
    __syncthreads();
    float nmax = red_n[0], emax = red_e[0];
#pragma unroll
    for (int w = 1; w < kThreads / 64; ++w) {
      nmax = fmaxf(nmax, red_n[w]);
      emax = fmaxf(emax, red_e[w]);
    }
    // per-row inputs of the bound, one row per lane (rows wave*16 .. +15)
    float b_l = INFINITY, e_l = 0.0f, n_l = 0.0f;
    {
      const int r = wave * kRowsPerWave + lane;
      if (lane < kRowsPerWave && r < rows) {
        const int id = lid[it.row0 + r];
        n_l = lnb[it.row0 + r];
        e_l = err[id];
        if (bound != nullptr) b_l = bound[(int64_t)id * K + (K - 1)];
      }
    }
#pragma unroll
    for (int rr = 0; rr < kRowsPerWave; ++rr) {
      qd[rr] = INFINITY;
      qi[rr] = INT_MAX;
    }
    // first tile's operands in flight while the rest is set up
    uint4 ra[4], rb[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      const int c = tid + kThreads * s4;
      const int r = c >> 4, part = c & 15;
      ra[s4] = load_piece(Xb, r < rows ? lid[it.row0 + r] : -1, d, part * 8);
      rb[s4] = load_piece(Xb, r < m ? lid[r] : -1, d, part * 8);
    }
    // Pruning against a known K-th exact distance B(i) (from another tree's
    // list): a leafmate whose certified lower bound on D exceeds B(i) cannot
    // enter point i's final K, so it is skipped. LB(a) <= B  <=>  a <= A_cut
    // with A_cut = (sqrt(B/(1-c)) + e_i + e_max)^2 + eta (the bound used below).
    const double u = 5.9604644775390625e-08;  // 2^-24
    const double infl = 1.0 + (d + 4.0) * u;
    const double nm = sqrt((double)nmax * infl);
    float acut[kRowsPerWave];
#pragma unroll
    for (int rr = 0; rr < kRowsPerWave; ++rr) {
      acut[rr] = INFINITY;
      const double b = (double)lane_read(b_l, rr);
      if (b < INFINITY) {
        const double ni = sqrt((double)lane_read(n_l, rr) * infl);
        const double eta = (2.0 * d + 16.0) * u * (ni + nm) * (ni + nm);
        const double sb = sqrt(b / (1.0 - (d + 8.0) * u)) + (double)lane_read(e_l, rr) + (double)emax;
        acut[rr] = (float)((sb * sb + eta) * (1.0 + 1e-6) + 1e-30);
      }
    }
    for (int c0 = 0; c0 < m; c0 += kCols) {
      const int cols = min(kCols, m - c0);
      floatx16 acc;
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
      if (a_resident) {
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const int c = tid + kThreads * s4;
          const int r = c >> 4, part = c & 15;
          if (c0 == 0) *reinterpret_cast<uint4*>(&As[r][part * 8]) = ra[s4];
          *reinterpret_cast<uint4*>(&Bs[r][part * 8]) = rb[s4];
        }
        __syncthreads();
        if (c0 + kCols < m) {  // next column tile in flight during this one
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) {
            const int c = tid + kThreads * s4;
            const int r = c >> 4, part = c & 15;
            const int col = c0 + kCols + r;
            rb[s4] = load_piece(Xb, col < m ? lid[col] : -1, d, part * 8);
          }
        }
        const int ksteps = (d + 15) / 16;
        for (int ks = 0; ks < ksteps; ++ks) {
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wr + (lane & 31)][ks * 16 + 8 * (lane >> 5)]);
          const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wc + (lane & 31)][ks * 16 + 8 * (lane >> 5)]);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
      } else {
        for (int k0 = 0; k0 < d; k0 += kGramChunk) {
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) {
            const int c = tid + kThreads * s4;
            const int r = c >> 4, part = c & 15;
            const int col = c0 + r;
            const uint4 va = load_piece(Xb, r < rows ? lid[it.row0 + r] : -1, d, k0 + part * 8);
            const uint4 vb = load_piece(Xb, col < m ? lid[col] : -1, d, k0 + part * 8);
            *reinterpret_cast<uint4*>(&As[r][part * 8]) = va;
            *reinterpret_cast<uint4*>(&Bs[r][part * 8]) = vb;
          }
          __syncthreads();
          const int ksteps = min(kGramChunk, d - k0 + 15) / 16;
          for (int ks = 0; ks < ksteps; ++ks) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wr + (lane & 31)][ks * 16 + 8 * (lane >> 5)]);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wc + (lane & 31)][ks * 16 + 8 * (lane >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
          }
          __syncthreads();
        }
      }
      // 32x32 accumulator layout: element e of lane l is
      // row 8*(e/4) + 4*(l/32) + e%4, column l%32.
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int r = wr + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        const int c = wc + (lane & 31);
        const float rn = r < rows ? lnb[it.row0 + r] : 0.0f;
        const float cn = c < cols ? lnb[c0 + c] : 0.0f;
        Ds[r][c] = __fsub_rn(__fadd_rn(rn, cn), __fmul_rn(2.0f, acc[e]));
      }
      __syncthreads();
      {
        const int pos = c0 + lane;
#pragma unroll
        for (int rr = 0; rr < kRowsPerWave; ++rr) {
          const int r = wave * kRowsPerWave + rr;
          if (r < rows) {
            const float a = Ds[r][lane];
            const bool ok = lane < cols && pos != it.row0 + r && a <= acut[rr];
            wq_offer(qd[rr], qi[rr], ok ? a : INFINITY, ok ? pos : INT_MAX, kp, lane);
          }
        }
      }
      __syncthreads();
    }
    // candidates (lanes < k') and the (k'+1)-th approximate value per row
    int cpos[kRowsPerWave];
    float thr[kRowsPerWave], racc[kRowsPerWave];
#pragma unroll
    for (int rr = 0; rr < kRowsPerWave; ++rr) {
      thr[rr] = lane_read(qd[rr], kp);
      if (lane_read(qi[rr], kp) == INT_MAX) thr[rr] = INFINITY;
      cpos[rr] = (lane < kp && qi[rr] != INT_MAX) ? qi[rr] : -1;
      racc[rr] = 0.0f;
    }
    // exact re-rank: stage chunks of dims of every leaf point in LDS as FP32;
    // lane l of row r accumulates D(row, candidate l), dims ascending
    int RK = min(d, kBigBytes / 4 / m);
    if (RK >= 4) RK &= ~3;
    const bool vec = (RK & 3) == 0 && (d & 3) == 0;
    for (int k0 = 0; k0 < d; k0 += RK) {
      const int w = min(RK, d - k0);
      const bool vw = vec && (w & 3) == 0;
      for (int pos = tid; pos < m; pos += kThreads) {
        const XT* src = X + (int64_t)lid[pos] * d + k0;
        float* dst = Ls + pos * w;
        if (vw) {
          for (int q = 0; q < w; q += 4) {
            float4 v;
            if (sizeof(XT) == 4) {
              v = *reinterpret_cast<const float4*>(src + q);
            } else {
              const uchar4 b = *reinterpret_cast<const uchar4*>(src + q);
              v = make_float4((float)b.x, (float)b.y, (float)b.z, (float)b.w);
            }
            *reinterpret_cast<float4*>(dst + q) = v;
          }
        } else {
          for (int q = 0; q < w; ++q) dst[q] = (float)src[q];
        }
      }
      __syncthreads();
#pragma unroll
      for (int rr = 0; rr < kRowsPerWave; ++rr) {
        const int r = wave * kRowsPerWave + rr;
        if (r < rows && cpos[rr] >= 0) {
          const int rp = it.row0 + r;
          if (vw) {
            const float4* L4 = reinterpret_cast<const float4*>(Ls);
            const int w4 = w >> 2;
            for (int kq = 0; kq < w4; ++kq) {
              const float4 a = L4[rp * w4 + kq];
              const float4 b = L4[cpos[rr] * w4 + kq];
              float t = __fsub_rn(a.x, b.x);
              racc[rr] = fmaf(t, t, racc[rr]);
              t = __fsub_rn(a.y, b.y);
              racc[rr] = fmaf(t, t, racc[rr]);
              t = __fsub_rn(a.z, b.z);
              racc[rr] = fmaf(t, t, racc[rr]);
              t = __fsub_rn(a.w, b.w);
              racc[rr] = fmaf(t, t, racc[rr]);
            }
          } else {
            for (int kk = 0; kk < w; ++kk) {
              const float t = __fsub_rn(Ls[rp * w + kk], Ls[cpos[rr] * w + kk]);
              racc[rr] = fmaf(t, t, racc[rr]);
            }
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int rr = 0; rr < kRowsPerWave; ++rr) {
      const int r = wave * kRowsPerWave + rr;
      if (r < rows) {
        const float dv = cpos[rr] >= 0 ? racc[rr] : INFINITY;
        const int iv = cpos[rr] >= 0 ? lid[cpos[rr]] : INT_MAX;
        const int me = lid[it.row0 + r];
        const int64_t o = ((int64_t)it.tree * n + me) * K;
        const float dk = rank_store(dv, iv, kp, K, lane, cand_idx + o, cand_dist + o);
        if (lane == 0 && thr[rr] < INFINITY) {
          const double ni = sqrt((double)lnb[it.row0 + r] * infl);
          const double eta = (2.0 * d + 16.0) * u * (ni + nm) * (ni + nm);
          const double rt = sqrt(fmax(0.0, (double)thr[rr] - eta));
          const double rl = fmax(0.0, rt - ((double)err[me] + (double)emax));
          const double lb = rl * rl * (1.0 - (d + 8.0) * u);
          if (!((double)dk < lb)) {
            const int slot = atomicAdd(nfallback, 1);
            if (slot < fb_cap) {
              fallback[2 * slot] = (int32_t)item;
              fallback[2 * slot + 1] = r;
            }
          }
        }
      }
    }
    __syncthreads();
    item = begin + s_ticket[parity];
  }
}

// Exact redo of uncertified rows: one block per row, all leafmates.
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

// bf16 copy (round to nearest even), |x~|^2 (FP32) and |x~ - x| (rounded up).
template <typename XT>
__global__ void k_prep_bf16(const XT* __restrict__ X, int64_t n, int d, uint16_t* __restrict__ Xb,
                            float* __restrict__ nb, float* __restrict__ err) {
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (row >= n) return;
  float s = 0.0f;
  double e2 = 0.0;
  for (int k = lane; k < d; k += 64) {
    const float x = (float)X[row * d + k];
    uint32_t bits = __float_as_uint(x);
    bits += 0x7fffu + ((bits >> 16) & 1u);
    const uint16_t b = (uint16_t)(bits >> 16);
    Xb[row * d + k] = b;
    const float xb = __uint_as_float((uint32_t)b << 16);
    s = fmaf(xb, xb, s);
    const double df = (double)xb - (double)x;
    e2 += df * df;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off);
    e2 += __shfl_xor(e2, off);
  }
  if (lane == 0) {
    nb[row] = s;
    float ef = (float)sqrt(e2);
    if (e2 > 0.0) ef = ef * (1.0f + 1.0e-6f) + 1.0e-30f;
    err[row] = ef;
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

template <typename XT>
int dispatch(drpt_ctx* ctx, const drpt_params* p, const XT* X, const int32_t* perm,
             const LeafItem* items, const int64_t* range, const void* Xq, const float* nrm2,
             const float* err, int32_t* ci, float* cd, unsigned* q, int32_t* fb, int32_t* nfb,
             int64_t cap, const float* bound) {
  const int K = p->nn, d = p->d;
  const int64_t n = p->n;
  if (K < 1 || K > 64) {
    set_error("leaf_knn: nn must be in [1, 64]");
    return DRPT_ERR_UNSUPPORTED;
  }
  unsigned grid;
  if (p->precision == DRPT_FP32) {
    DRPT_CHECK(persistent_grid(ctx, k_knn_exact<XT>, kThreads, &grid));
    hipLaunchKernelGGL((k_knn_exact<XT>), dim3(grid), dim3(kThreads), 0, ctx->stream, X, d, perm, n,
                       items, range, K, ci, cd, q);
    DRPT_HIP(hipGetLastError());
    return DRPT_OK;
  }
  if (p->precision == DRPT_BF16) {
    const int kp = p->kprime;
    if (kp < K || kp > 63) {
      set_error("leaf_knn: kprime must be in [nn, 63] for BF16");
      return DRPT_ERR_UNSUPPORTED;
    }
    DRPT_CHECK(persistent_grid(ctx, k_knn_bf16<XT>, kThreads, &grid));
    hipLaunchKernelGGL((k_knn_bf16<XT>), dim3(grid), dim3(kThreads), 0, ctx->stream, X,
                       (const uint16_t*)Xq, nrm2, err, d, perm, n, items, range, K, kp, ci, cd, q,
                       fb, nfb, cap, bound);
    DRPT_HIP(hipGetLastError());
    return DRPT_OK;
  }
  set_error("leaf_knn: precision not supported");
  return DRPT_ERR_UNSUPPORTED;
}

}  // namespace

int launch_prep_lowp(drpt_ctx* ctx, const void* X, int dtype, int64_t n, int d, int precision,
                     void* Xq, float* nrm2, float* err, float* /*scale*/) {
  if (precision != DRPT_BF16) {
    set_error("prep_lowp: precision not supported");
    return DRPT_ERR_UNSUPPORTED;
  }
  const int64_t threads = n * 64;
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  if (dtype == DRPT_U8)
    hipLaunchKernelGGL(k_prep_bf16<uint8_t>, dim3(blocks), dim3(256), 0, ctx->stream,
                       (const uint8_t*)X, n, d, (uint16_t*)Xq, nrm2, err);
  else
    hipLaunchKernelGGL(k_prep_bf16<float>, dim3(blocks), dim3(256), 0, ctx->stream, (const float*)X,
                       n, d, (uint16_t*)Xq, nrm2, err);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_leaf_knn(drpt_ctx* ctx, const drpt_params* p, const void* X, const int32_t* perm,
                    const LeafItem* items, const int64_t* range, const void* Xq, const float* nrm2,
                    const float* err, int32_t* cand_idx, float* cand_dist, int32_t* fallback,
                    int32_t* nfallback, int64_t fallback_cap, const float* bound) {
  unsigned* q = nullptr;
  DRPT_CHECK(ws(ctx, "knn.queue", 1, &q));
  DRPT_CHECK(fill_u32(ctx, q, 0u, 1));
  if (p->dtype == DRPT_U8)
    return dispatch<uint8_t>(ctx, p, (const uint8_t*)X, perm, items, range, Xq, nrm2, err, cand_idx,
                             cand_dist, q, fallback, nfallback, fallback_cap, bound);
  return dispatch<float>(ctx, p, (const float*)X, perm, items, range, Xq, nrm2, err, cand_idx,
                         cand_dist, q, fallback, nfallback, fallback_cap, bound);
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
