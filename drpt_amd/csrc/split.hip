// Stage (b): per-level histogram-median splits of all trees at once.
//
// Replaces DRPTGlobal::grow_global_tree / grow_global_subtree
// (cpp/src/drpt/algo/drpt_global.cpp:153-352) and MathOp::distributed_mean,
// distributed_variance, distributed_median (math_operations.cpp:121-383).
//
// Layout: points never move. node[t][i] holds the heap index (root 1) of
// point i's current node in tree t; a point is at level l iff node>>l == 1.
// Level l of tree t reads the contiguous projection column P[t*D + l][.].
// Per-node state (counts, fixed-point sums, mean, 16-slot edge table,
// histogram, median) lives in small [T][2^l] arrays.
//
// Per level, all trees in one launch each:
//   k_sum1   count + sum of quantised values per node (root level only)
//   k_mu     split decision + mean                          (per node)
//   k_sum2   sum of quantised fl(v-mu)^2 per node            (per point)
//   k_edges  variance, sigma, edge table                     (per node)
//   k_hist   16-slot histogram (wave-ballot when a wave shares a node; the
//            tree's edge tables staged in LDS where they fit)
//   k_median cumulative search + interpolation               (per node)
//   k_part_sum1  v <= median -> 2h, else 2h+1, and the child's count and
//            sum of its next-level values (k_sum1 of level l + 1); k_part
//            alone at the last split level                  (per point)
//
// Arithmetic (bit-exact with the oracle's canonical statistics):
//   q(v) = rint(v * 2^-kq1) as int64; S = sum q (order-independent);
//   mu = fl32(S * 2^kq1 / n) (double); w = fl32(v - mu); w2 = fl32(w*w);
//   var = fl32(sum rint(w2 * 2^-kq2) * 2^kq2 / (n-1)); sigma = sqrtf(var);
//   edges/histogram/median exactly as math_operations.cpp:223-377 with
//   E[-1] = 0. No floating-point contraction anywhere in this file.
//
// Roofline: HBM-bound; algorithmic bytes 12 per point per tree per level
// (4 B value + 4 B node read + 4 B node write).
#include <algorithm>
#include <cmath>
#include <mutex>

#include "internal.hpp"

namespace drpt {

namespace {

constexpr int kThreads = 256;
// per-node LDS arrays of the LDS variants (bytes): sums 12 (sum1: count +
// int64 sum; sum2: 8), histograms 64 (16 slots), plus for the histogram the
// edge table (64) and split flag (4) when they fit as well; used while they
// fit kSumLdsMax / kHistLdsMax and the device's per-workgroup LDS
constexpr size_t kSumLdsMax = 96 * 1024;
constexpr size_t kHistLdsMax = 136 * 1024;

__device__ __forceinline__ void col_exponents(uint32_t cm, int bitlen, int& kq1, int& kq2) {
  int E = (int)((cm >> 23) & 0xffu);
  int e = (E == 0) ? -127 : E - 127;
  int shift = 63 - bitlen;
  kq1 = e + 1 - shift;
  kq2 = 2 * (e + 2) - shift;
}

__device__ __forceinline__ long long quant(float v, int kq) {
  return (long long)rint(ldexp((double)v, -kq));
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Points per thread whose node and value words are loaded before any is used
// (independent loads in flight; sums and counts do not depend on the order)
constexpr int kUnroll = 4;

__device__ __forceinline__ void load_points(const uint32_t* __restrict__ nt,
                                            const float* __restrict__ Pc, int64_t base,
                                            int64_t b1, uint32_t (&hu)[kUnroll],
                                            float (&pu)[kUnroll]) {
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) {
    const int64_t i = base + u * kThreads + threadIdx.x;
    hu[u] = i < b1 ? nt[i] : 0u;
    pu[u] = i < b1 ? Pc[i] : 0.0f;
  }
}

// Points [b0, b1) of tree blockIdx.y handled by this block.
__device__ __forceinline__ void block_range(int64_t n, int64_t ppb, int64_t& b0, int64_t& b1) {
  b0 = (int64_t)blockIdx.x * ppb;
  b1 = min(n, b0 + ppb);
}

template <bool kLds>
__global__ void __launch_bounds__(kThreads)
k_sum1(const float* __restrict__ P, const uint32_t* __restrict__ node, int64_t n, int D, int l,
       const uint32_t* __restrict__ colmax, int bitlen, uint32_t* __restrict__ cnt,
       unsigned long long* __restrict__ S1, int64_t ppb) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = blockIdx.y;
  const int NN = 1 << l;
  unsigned long long* lS = reinterpret_cast<unsigned long long*>(smem);
  uint32_t* lc = reinterpret_cast<uint32_t*>(lS + (kLds ? NN : 0));
  if (kLds) {
    for (int k = threadIdx.x; k < NN; k += kThreads) { lS[k] = 0; lc[k] = 0; }
    __syncthreads();
  }
  const int64_t col = (int64_t)t * D + l;
  int kq1, kq2;
  col_exponents(colmax[col], bitlen, kq1, kq2);
  const float* Pc = P + col * n;
  const uint32_t* nt = node + (int64_t)t * n;
  uint32_t* gc = cnt + (int64_t)t * NN;
  unsigned long long* gS = S1 + (int64_t)t * NN;
  int64_t b0, b1;
  block_range(n, ppb, b0, b1);
  const int lane = threadIdx.x & 63;
  for (int64_t base = b0; base < b1; base += kThreads * kUnroll) {
    uint32_t hu[kUnroll];
    float pu[kUnroll];
    load_points(nt, Pc, base, b1, hu, pu);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + u * kThreads + threadIdx.x;
      bool act = false;
      int nd = 0;
      long long q = 0;
      if (i < b1) {
        uint32_t h = hu[u];
        act = (h >> l) == 1u;
        nd = (int)(h - (uint32_t)NN);
        if (act) q = quant(pu[u], kq1);
      }
      unsigned long long am = __ballot(act);
      if (am == 0) continue;
      int leader = __ffsll((long long)am) - 1;
      int ndl = __shfl(nd, leader);
      unsigned long long same = __ballot(act && nd == ndl);
      if (same == am) {
        long long s = wave_sum_i64(q);
        if (lane == leader) {
          if (kLds) {
            atomicAdd(&lc[ndl], (uint32_t)__popcll(am));
            atomicAdd(&lS[ndl], (unsigned long long)s);
          } else {
            atomicAdd(&gc[ndl], (uint32_t)__popcll(am));
            atomicAdd(&gS[ndl], (unsigned long long)s);
          }
        }
      } else if (act) {
        if (kLds) {
          atomicAdd(&lc[nd], 1u);
          atomicAdd(&lS[nd], (unsigned long long)q);
        } else {
          atomicAdd(&gc[nd], 1u);
          atomicAdd(&gS[nd], (unsigned long long)q);
        }
      }
    }
  }
  if (kLds) {
    __syncthreads();
    for (int k = threadIdx.x; k < NN; k += kThreads) {
      if (lc[k]) {
        atomicAdd(&gc[k], lc[k]);
        atomicAdd(&gS[k], lS[k]);
      }
    }
  }
}

__global__ void k_mu(int T, int l, int D, int mode, int leaf_max, int bitlen,
                     const uint32_t* __restrict__ colmax, const uint32_t* __restrict__ cnt,
                     const unsigned long long* __restrict__ S1, float* __restrict__ mu,
                     uint32_t* __restrict__ split) {
  const int NN = 1 << l;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)T * NN) return;
  int t = (int)(idx / NN);
  int kq1, kq2;
  col_exponents(colmax[(int64_t)t * D + l], bitlen, kq1, kq2);
  uint32_t c = cnt[idx];
  split[idx] = (mode == DRPT_TREE_FIXED) ? 1u : (c > (uint32_t)leaf_max ? 1u : 0u);
  double s = ldexp((double)(long long)S1[idx], kq1);
  mu[idx] = (float)(s / (double)c);
}

template <bool kLds>
__global__ void __launch_bounds__(kThreads)
k_sum2(const float* __restrict__ P, const uint32_t* __restrict__ node, int64_t n, int D, int l,
       const uint32_t* __restrict__ colmax, int bitlen, const float* __restrict__ mu,
       const uint32_t* __restrict__ split, unsigned long long* __restrict__ S2, int64_t ppb) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = blockIdx.y;
  const int NN = 1 << l;
  unsigned long long* lS = reinterpret_cast<unsigned long long*>(smem);
  if (kLds) {
    for (int k = threadIdx.x; k < NN; k += kThreads) lS[k] = 0;
    __syncthreads();
  }
  const int64_t col = (int64_t)t * D + l;
  int kq1, kq2;
  col_exponents(colmax[col], bitlen, kq1, kq2);
  const float* Pc = P + col * n;
  const uint32_t* nt = node + (int64_t)t * n;
  const float* mt = mu + (int64_t)t * NN;
  const uint32_t* st = split + (int64_t)t * NN;
  unsigned long long* gS = S2 + (int64_t)t * NN;
  int64_t b0, b1;
  block_range(n, ppb, b0, b1);
  const int lane = threadIdx.x & 63;
  for (int64_t base = b0; base < b1; base += kThreads * kUnroll) {
    uint32_t hu[kUnroll];
    float pu[kUnroll];
    load_points(nt, Pc, base, b1, hu, pu);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + u * kThreads + threadIdx.x;
      bool act = false;
      int nd = 0;
      long long q = 0;
      if (i < b1) {
        uint32_t h = hu[u];
        act = (h >> l) == 1u;
        nd = (int)(h - (uint32_t)NN);
        if (act) act = st[nd] != 0u;
        if (act) {
          float w = __fsub_rn(pu[u], mt[nd]);
          float w2 = __fmul_rn(w, w);
          q = quant(w2, kq2);
        }
      }
      unsigned long long am = __ballot(act);
      if (am == 0) continue;
      int leader = __ffsll((long long)am) - 1;
      int ndl = __shfl(nd, leader);
      unsigned long long same = __ballot(act && nd == ndl);
      if (same == am) {
        long long s = wave_sum_i64(q);
        if (lane == leader) {
          if (kLds) atomicAdd(&lS[ndl], (unsigned long long)s);
          else atomicAdd(&gS[ndl], (unsigned long long)s);
        }
      } else if (act) {
        if (kLds) atomicAdd(&lS[nd], (unsigned long long)q);
        else atomicAdd(&gS[nd], (unsigned long long)q);
      }
    }
  }
  if (kLds) {
    __syncthreads();
    for (int k = threadIdx.x; k < NN; k += kThreads)
      if (lS[k]) atomicAdd(&gS[k], lS[k]);
  }
}

// Edge table: math_operations.cpp:250-288 (factor 1, no_of_bins 7).
__global__ void k_edges(int T, int l, int D, int bitlen, const uint32_t* __restrict__ colmax,
                        const uint32_t* __restrict__ cnt, const unsigned long long* __restrict__ S2,
                        const float* __restrict__ mu, const uint32_t* __restrict__ split,
                        float* __restrict__ E) {
  const int NN = 1 << l;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)T * NN) return;
  if (!split[idx]) return;
  int t = (int)(idx / NN);
  int kq1, kq2;
  col_exponents(colmax[(int64_t)t * D + l], bitlen, kq1, kq2);
  long long c = (long long)cnt[idx];
  double s2 = ldexp((double)(long long)S2[idx], kq2);
  float var = (float)(s2 / (double)(c - 1));
  float m = mu[idx];
  float sigma = sqrtf(var);
  float s1 = (float)((double)sigma / 30.0);
  float s2f = (float)((double)sigma / 6.0);
  float s3 = (float)((double)sigma / 2.0);
  float e[16];
#pragma unroll
  for (int j = 1; j <= 4; ++j) {
    float r = __fmul_rn((float)j, s1);
    e[6 + j] = __fadd_rn(m, r);
    e[7 - j] = __fsub_rn(m, r);
  }
#pragma unroll
  for (int j = 1; j <= 2; ++j) {
    float r = __fadd_rn(sigma, __fmul_rn((float)j, s2f));
    e[10 + j] = __fadd_rn(m, r);
    e[3 - j] = __fsub_rn(m, r);
  }
  {
    float r = __fadd_rn(__fmul_rn(2.0f, sigma), s3);
    e[13] = __fadd_rn(m, r);
    e[0] = __fsub_rn(m, r);
  }
  e[14] = 0.0f;
  e[15] = 0.0f;
  float4* dst = reinterpret_cast<float4*>(E + idx * 16);
  dst[0] = make_float4(e[0], e[1], e[2], e[3]);
  dst[1] = make_float4(e[4], e[5], e[6], e[7]);
  dst[2] = make_float4(e[8], e[9], e[10], e[11]);
  dst[3] = make_float4(e[12], e[13], e[14], e[15]);
}

// Slot of v: math_operations.cpp:292-308 -- the j in 1..15 with
// E[j-1] < v <= E[j], else 0. E[0..13] is non-decreasing (the steps sigma/30,
// sigma/6, sigma/2 are monotone and rounding keeps their order) and
// E[14] = E[15] = 0, so with c = #{j <= 13 : E[j] < v}: c in 1..13 is the slot,
// c = 0 gives 0, c = 14 gives 14 when v <= E[14] (= 0) and 0 otherwise (slot 15
// is never matched); NaN values or edges compare false and give 0, as there.
__device__ __forceinline__ int slot_of(float v, const float* e) {
  int c = 0;
#pragma unroll
  for (int j = 0; j < 14; ++j) c += e[j] < v ? 1 : 0;
  if (c == 14) c = v <= e[14] ? 14 : 0;
  return c;
}

template <bool kLds, bool kEL>
__global__ void __launch_bounds__(kThreads)
k_hist(const float* __restrict__ P, const uint32_t* __restrict__ node, int64_t n, int D, int l,
       const float* __restrict__ E, const uint32_t* __restrict__ split, uint32_t* __restrict__ H,
       int64_t ppb) {
  // kLds: histograms [NN][16] in LDS; kEL: also the tree's edge tables
  // [NN][16] and split flags [NN] (staged once per block, read per point)
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = blockIdx.y;
  const int NN = 1 << l;
  uint32_t* lH = reinterpret_cast<uint32_t*>(smem);
  float* lE = reinterpret_cast<float*>(lH + (kLds ? NN * 16 : 0));
  uint32_t* lS = reinterpret_cast<uint32_t*>(lE + (kEL ? NN * 16 : 0));
  const float* Eg = E + (int64_t)t * NN * 16;
  const uint32_t* Sg = split + (int64_t)t * NN;
  if (kLds) {
    for (int k = threadIdx.x; k < NN * 16; k += kThreads) lH[k] = 0;
    if (kEL) {
      for (int k = threadIdx.x; k < NN * 4; k += kThreads)
        reinterpret_cast<float4*>(lE)[k] = reinterpret_cast<const float4*>(Eg)[k];
      for (int k = threadIdx.x; k < NN; k += kThreads) lS[k] = Sg[k];
    }
    __syncthreads();
  }
  const int64_t col = (int64_t)t * D + l;
  const float* Pc = P + col * n;
  const uint32_t* nt = node + (int64_t)t * n;
  const float* Et = kEL ? lE : Eg;
  const uint32_t* st = kEL ? lS : Sg;
  uint32_t* gH = H + (int64_t)t * NN * 16;
  int64_t b0, b1;
  block_range(n, ppb, b0, b1);
  const int lane = threadIdx.x & 63;
  for (int64_t base = b0; base < b1; base += kThreads * kUnroll) {
    uint32_t hu[kUnroll];
    float pu[kUnroll];
    load_points(nt, Pc, base, b1, hu, pu);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + u * kThreads + threadIdx.x;
      bool act = false;
      int nd = 0, b = 0;
      if (i < b1) {
        uint32_t h = hu[u];
        act = (h >> l) == 1u;
        nd = (int)(h - (uint32_t)NN);
        if (act) act = st[nd] != 0u;
        if (act) {
          float e[16];
          const float4* src = reinterpret_cast<const float4*>(Et + (int64_t)nd * 16);
          float4 a0 = src[0], a1 = src[1], a2 = src[2], a3 = src[3];
          e[0] = a0.x; e[1] = a0.y; e[2] = a0.z; e[3] = a0.w;
          e[4] = a1.x; e[5] = a1.y; e[6] = a1.z; e[7] = a1.w;
          e[8] = a2.x; e[9] = a2.y; e[10] = a2.z; e[11] = a2.w;
          e[12] = a3.x; e[13] = a3.y; e[14] = a3.z; e[15] = a3.w;
          b = slot_of(pu[u], e);
        }
      }
      unsigned long long am = __ballot(act);
      if (am == 0) continue;
      int leader = __ffsll((long long)am) - 1;
      int ndl = __shfl(nd, leader);
      unsigned long long same = __ballot(act && nd == ndl);
      if (same == am) {
        // wave-ballot histogram: lane k < 16 adds the population of slot k
        uint32_t mine = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          uint32_t c = (uint32_t)__popcll(__ballot(act && b == k));
          mine = (lane == k) ? c : mine;
        }
        if (lane < 16 && mine) {
          if (kLds) atomicAdd(&lH[ndl * 16 + lane], mine);
          else atomicAdd(&gH[(int64_t)ndl * 16 + lane], mine);
        }
      } else if (act) {
        if (kLds) atomicAdd(&lH[nd * 16 + b], 1u);
        else atomicAdd(&gH[(int64_t)nd * 16 + b], 1u);
      }
    }
  }
  if (kLds) {
    __syncthreads();
    for (int k = threadIdx.x; k < NN * 16; k += kThreads)
      if (lH[k]) atomicAdd(&gH[k], lH[k]);
  }
}

// Cumulative search + interpolation: math_operations.cpp:333-377, E[-1] = 0.
__global__ void k_median(int T, int l, const uint32_t* __restrict__ cnt,
                         const uint32_t* __restrict__ H, const float* __restrict__ E,
                         const uint32_t* __restrict__ split, float* __restrict__ med) {
  const int NN = 1 << l;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)T * NN) return;
  if (!split[idx]) return;
  long long c = (long long)cnt[idx];
  const uint32_t* h = H + idx * 16;
  const float* e = E + idx * 16;
  float val = (float)((50 * c) / 100);
  float cf = 0.0f;
  int sel = -1;
  for (int k = 0; k < 16; ++k) {
    cf = (float)__dadd_rn((double)cf, (double)h[k]);
    if (cf >= val) {
      sel = k;
      break;
    }
  }
  if (sel < 0) sel = 15;
  double count = (double)h[sel];
  float lo = (sel == 0) ? 0.0f : e[sel - 1];
  float width = __fsub_rn(e[sel], lo);
  double num = __dsub_rn((double)(c / 2), __dsub_rn((double)cf, count));
  double frac = num / count;
  double m = __dadd_rn((double)lo, __dmul_rn(frac, (double)width));
  med[idx] = (float)m;
}

__global__ void __launch_bounds__(kThreads)
k_part(const float* __restrict__ P, uint32_t* __restrict__ node, int64_t n, int D, int l,
       const float* __restrict__ med, const uint32_t* __restrict__ split) {
  const int t = blockIdx.y;
  const int NN = 1 << l;
  const float* Pc = P + ((int64_t)t * D + l) * n;
  uint32_t* nt = node + (int64_t)t * n;
  const float* mt = med + (int64_t)t * NN;
  const uint32_t* st = split + (int64_t)t * NN;
  // block ranges of kUnroll * kThreads points; each thread loads its points'
  // node and value words before using any
  const int64_t span = (int64_t)kThreads * kUnroll;
  for (int64_t base = (int64_t)blockIdx.x * span; base < n; base += (int64_t)gridDim.x * span) {
    const int64_t b1 = min(n, base + span);
    uint32_t hu[kUnroll];
    float pu[kUnroll];
    load_points(nt, Pc, base, b1, hu, pu);
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + u * kThreads + threadIdx.x;
      const uint32_t h = hu[u];
      if (i >= b1 || (h >> l) != 1u) continue;
      const int nd = (int)(h - (uint32_t)NN);
      if (!st[nd]) continue;
      nt[i] = 2u * h + ((pu[u] <= mt[nd]) ? 0u : 1u);
    }
  }
}

// k_part of level l fused with k_sum1 of level l + 1: each point of a split
// node moves to its child (v <= median -> 2h, else 2h + 1) and, at once, adds
// its count and quantised level-(l+1) value to the child's sums (points of
// unsplit nodes are leaves and take no further part). kLds: the child sums
// in LDS ([2^(l+1)] int64 sums + counts).
template <bool kLds>
__global__ void __launch_bounds__(kThreads)
k_part_sum1(const float* __restrict__ P, uint32_t* __restrict__ node, int64_t n, int D, int l,
            const float* __restrict__ med, const uint32_t* __restrict__ split,
            const uint32_t* __restrict__ colmax, int bitlen, uint32_t* __restrict__ cnt,
            unsigned long long* __restrict__ S1, int64_t ppb) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = blockIdx.y;
  const int NN = 1 << l, NC = 2 << l;
  unsigned long long* lS = reinterpret_cast<unsigned long long*>(smem);
  uint32_t* lc = reinterpret_cast<uint32_t*>(lS + (kLds ? NC : 0));
  if (kLds) {
    for (int k = threadIdx.x; k < NC; k += kThreads) { lS[k] = 0; lc[k] = 0; }
    __syncthreads();
  }
  const float* Pc = P + ((int64_t)t * D + l) * n;
  const float* Pn = Pc + n;  // level l + 1 of the same tree
  int kq1, kq2;
  col_exponents(colmax[(int64_t)t * D + l + 1], bitlen, kq1, kq2);
  uint32_t* nt = node + (int64_t)t * n;
  const float* mt = med + (int64_t)t * NN;
  const uint32_t* st = split + (int64_t)t * NN;
  uint32_t* gc = cnt + (int64_t)t * NC;
  unsigned long long* gS = S1 + (int64_t)t * NC;
  int64_t b0, b1;
  block_range(n, ppb, b0, b1);
  const int lane = threadIdx.x & 63;
  for (int64_t base = b0; base < b1; base += kThreads * kUnroll) {
    uint32_t hu[kUnroll];
    float pu[kUnroll];
    load_points(nt, Pc, base, b1, hu, pu);
    float pn[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + u * kThreads + threadIdx.x;
      pn[u] = (i < b1 && (hu[u] >> l) == 1u) ? Pn[i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + u * kThreads + threadIdx.x;
      bool act = false;
      int nd = 0;
      long long q = 0;
      if (i < b1) {
        const uint32_t h = hu[u];
        if ((h >> l) == 1u) {
          const int np = (int)(h - (uint32_t)NN);
          if (st[np]) {
            const uint32_t hc = 2u * h + ((pu[u] <= mt[np]) ? 0u : 1u);
            nt[i] = hc;
            act = true;
            nd = (int)(hc - (uint32_t)NC);
            q = quant(pn[u], kq1);
          }
        }
      }
      unsigned long long am = __ballot(act);
      if (am == 0) continue;
      int leader = __ffsll((long long)am) - 1;
      int ndl = __shfl(nd, leader);
      unsigned long long same = __ballot(act && nd == ndl);
      if (same == am) {
        long long s = wave_sum_i64(q);
        if (lane == leader) {
          if (kLds) {
            atomicAdd(&lc[ndl], (uint32_t)__popcll(am));
            atomicAdd(&lS[ndl], (unsigned long long)s);
          } else {
            atomicAdd(&gc[ndl], (uint32_t)__popcll(am));
            atomicAdd(&gS[ndl], (unsigned long long)s);
          }
        }
      } else if (act) {
        if (kLds) {
          atomicAdd(&lc[nd], 1u);
          atomicAdd(&lS[nd], (unsigned long long)q);
        } else {
          atomicAdd(&gc[nd], 1u);
          atomicAdd(&gS[nd], (unsigned long long)q);
        }
      }
    }
  }
  if (kLds) {
    __syncthreads();
    for (int k = threadIdx.x; k < NC; k += kThreads) {
      if (lc[k]) {
        atomicAdd(&gc[k], lc[k]);
        atomicAdd(&gS[k], lS[k]);
      }
    }
  }
}

// up to four fills of 32-bit words in one launch (capture-safe)
__global__ void k_fill4(uint32_t* __restrict__ p0, size_t c0, uint32_t* __restrict__ p1, size_t c1,
                        uint32_t* __restrict__ p2, size_t c2, uint32_t* __restrict__ p3, size_t c3) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < c0 + c1 + c2 + c3; i += stride) {
    if (i < c0) p0[i] = 0u;
    else if (i < c0 + c1) p1[i - c0] = 0u;
    else if (i < c0 + c1 + c2) p2[i - c0 - c1] = 0u;
    else p3[i - c0 - c1 - c2] = 0u;
  }
}

int zero4(drpt_ctx* ctx, uint32_t* p0, size_t c0, uint32_t* p1, size_t c1, uint32_t* p2 = nullptr,
          size_t c2 = 0, uint32_t* p3 = nullptr, size_t c3 = 0) {
  const size_t total = c0 + c1 + c2 + c3;
  if (total == 0) return DRPT_OK;
  const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill4, dim3(blocks), dim3(256), 0, ctx->stream, p0, c0, p1, c1, p2, c2, p3,
                     c3);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int split_levels(drpt_ctx* ctx, const float* P, const uint32_t* colmax, int64_t n, int T, int D,
                 int mode, int leaf_max, int bitlen, uint32_t* node, uint32_t* cnt,
                 uint32_t* split, uint32_t* H, unsigned long long* S1, unsigned long long* S2,
                 float* mu, float* E, float* med);

int bitlen_of(int64_t n) {
  int b = 0;
  while (b < 63 && (n >> b) != 0) ++b;
  return b;
}

}  // namespace

// LDS sums and histograms beyond 64 KB of dynamic LDS where the device allows
// it (a function attribute, set once per context)
static void raise_split_lds(drpt_ctx* ctx) {
  if (!ctx->split_lds_big && ctx->max_lds > 64 * 1024) {
    const int bytes = (int)std::min<size_t>((size_t)ctx->max_lds, kSumLdsMax);
    const int hbytes = (int)std::min<size_t>((size_t)ctx->max_lds, kHistLdsMax);
    if (hipFuncSetAttribute((const void*)k_sum1<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            bytes) == hipSuccess &&
        hipFuncSetAttribute((const void*)k_sum2<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            bytes) == hipSuccess &&
        hipFuncSetAttribute((const void*)k_part_sum1<true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes) == hipSuccess &&
        hipFuncSetAttribute((const void*)k_hist<true, false>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, hbytes) == hipSuccess &&
        hipFuncSetAttribute((const void*)k_hist<true, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, hbytes) == hipSuccess)
      ctx->split_lds_big = true;
    else
      (void)hipGetLastError();
  }
}

static int check_tree_args(int64_t n, int T, int D) {
  if (D < 1 || D > 24 || T < 1 || n < 1) {
    set_error("grow_trees: depth must be in [1, 24], ntrees >= 1, n >= 1");
    return DRPT_ERR_ARG;
  }
  return DRPT_OK;
}

int grow_trees_impl(drpt_ctx* ctx, const float* P, const uint32_t* colmax, int64_t n, int T, int D,
                    int mode, int leaf_max, uint32_t* node) {
  DRPT_CHECK(check_tree_args(n, T, D));
  if (D == 1) return fill_u32(ctx, node, 1u, (size_t)T * n);
  const int64_t NMAX = (int64_t)1 << (D - 2);
  uint32_t *cnt, *split, *H;
  unsigned long long *S1, *S2;
  float *mu, *E, *med;
  DRPT_CHECK(ws(ctx, "split.cnt", T * NMAX, &cnt));
  DRPT_CHECK(ws(ctx, "split.flag", T * NMAX, &split));
  DRPT_CHECK(ws(ctx, "split.S1", T * NMAX, &S1));
  DRPT_CHECK(ws(ctx, "split.S2", T * NMAX, &S2));
  DRPT_CHECK(ws(ctx, "split.mu", T * NMAX, &mu));
  DRPT_CHECK(ws(ctx, "split.med", T * NMAX, &med));
  DRPT_CHECK(ws(ctx, "split.E", T * NMAX * 16, &E));
  DRPT_CHECK(ws(ctx, "split.H", T * NMAX * 16, &H));
  const int bitlen = bitlen_of(n);
  raise_split_lds(ctx);
  // the per-level launches (4 fills + 7 kernels per level) as one hipGraph,
  // replayed while sizes, parameters and buffers stay the same
  char key[512];
  snprintf(key, sizeof key, "%p %p %lld %d %d %d %d %p %p %p %p %p %p %p %p %p %p", (const void*)P,
           (const void*)colmax, (long long)n, T, D, mode, leaf_max, (void*)node, (void*)cnt,
           (void*)split, (void*)S1, (void*)S2, (void*)mu, (void*)med, (void*)E, (void*)H,
           (void*)ctx->stream);
  return run_graph(ctx, "split", key, [&]() -> int {
    return split_levels(ctx, P, colmax, n, T, D, mode, leaf_max, bitlen, node, cnt, split, H, S1, S2,
                        mu, E, med);
  });
}

namespace {

int split_levels(drpt_ctx* ctx, const float* P, const uint32_t* colmax, int64_t n, int T, int D,
                 int mode, int leaf_max, int bitlen, uint32_t* node, uint32_t* cnt,
                 uint32_t* split, uint32_t* H, unsigned long long* S1, unsigned long long* S2,
                 float* mu, float* E, float* med) {
  hipStream_t s = ctx->stream;
  const size_t lds_max = std::min<size_t>((size_t)ctx->max_lds,
                                          ctx->split_lds_big ? kSumLdsMax : 64 * 1024);
  const size_t hist_max = ctx->split_lds_big ? std::min<size_t>((size_t)ctx->max_lds, kHistLdsMax)
                                             : std::min<size_t>((size_t)ctx->max_lds, 64 * 1024);
  // points per block: with per-node LDS arrays, enough points to amortise their
  // flush; with global atomics, small blocks so every CU has work
  auto sum_ppb = [&](int64_t NN) -> int64_t {
    return (size_t)NN * 12 <= lds_max ? std::max<int64_t>(8192, NN * 8) : 8192;
  };
  DRPT_CHECK(fill_u32(ctx, node, 1u, (size_t)T * n));
  // sums of the root (later levels get theirs from the fused partition)
  DRPT_CHECK(zero4(ctx, cnt, (size_t)T, (uint32_t*)S1, 2 * (size_t)T));
  {
    const int64_t ppb = sum_ppb(1);
    const dim3 g((unsigned)((n + ppb - 1) / ppb), (unsigned)T);
    hipLaunchKernelGGL(k_sum1<true>, g, dim3(kThreads), 12, s, P, node, n, D, 0, colmax, bitlen, cnt,
                       S1, ppb);
    DRPT_HIP(hipGetLastError());
  }
  for (int l = 0; l <= D - 2; ++l) {
    const int64_t NN = (int64_t)1 << l;
    const int64_t TN = (int64_t)T * NN;
    DRPT_CHECK(zero4(ctx, (uint32_t*)S2, 2 * (size_t)TN, H, 16 * (size_t)TN));
    unsigned ngrid = (unsigned)((TN + 255) / 256);
    const bool lds_sum = (size_t)NN * 12 <= lds_max;
    const bool lds_hist = (size_t)NN * 64 <= hist_max;
    const bool lds_edges = lds_hist && (size_t)NN * (64 + 64 + 4) <= hist_max;
    const int64_t ppb = sum_ppb(NN);
    const int64_t ppb_hist = lds_hist ? std::max<int64_t>(8192, NN * 32) : 8192;
    const dim3 pgrid((unsigned)((n + ppb - 1) / ppb), (unsigned)T);
    const dim3 hgrid((unsigned)((n + ppb_hist - 1) / ppb_hist), (unsigned)T);
    hipLaunchKernelGGL(k_mu, dim3(ngrid), dim3(256), 0, s, T, l, D, mode, leaf_max, bitlen, colmax,
                       cnt, S1, mu, split);
    DRPT_HIP(hipGetLastError());
    if (lds_sum)
      hipLaunchKernelGGL(k_sum2<true>, pgrid, dim3(kThreads), NN * 8, s, P, node, n, D, l, colmax,
                         bitlen, mu, split, S2, ppb);
    else
      hipLaunchKernelGGL(k_sum2<false>, pgrid, dim3(kThreads), 0, s, P, node, n, D, l, colmax,
                         bitlen, mu, split, S2, ppb);
    DRPT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_edges, dim3(ngrid), dim3(256), 0, s, T, l, D, bitlen, colmax, cnt, S2, mu,
                       split, E);
    DRPT_HIP(hipGetLastError());
    if (lds_edges)
      hipLaunchKernelGGL((k_hist<true, true>), hgrid, dim3(kThreads), NN * (64 + 64 + 4), s, P,
                         node, n, D, l, E, split, H, ppb_hist);
    else if (lds_hist)
      hipLaunchKernelGGL((k_hist<true, false>), hgrid, dim3(kThreads), NN * 64, s, P, node, n, D, l,
                         E, split, H, ppb_hist);
    else
      hipLaunchKernelGGL((k_hist<false, false>), hgrid, dim3(kThreads), 0, s, P, node, n, D, l, E,
                         split, H, ppb_hist);
    DRPT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_median, dim3(ngrid), dim3(256), 0, s, T, l, cnt, H, E, split, med);
    DRPT_HIP(hipGetLastError());
    if (l < D - 2) {
      // partition fused with the children's sums (their counters cleared
      // first: the medians above were the last reads of this level's counts)
      const int64_t NC = 2 * NN;
      DRPT_CHECK(zero4(ctx, cnt, (size_t)T * NC, (uint32_t*)S1, 2 * (size_t)T * NC));
      const int64_t ppc = sum_ppb(NC);
      const dim3 cgrid((unsigned)((n + ppc - 1) / ppc), (unsigned)T);
      if ((size_t)NC * 12 <= lds_max)
        hipLaunchKernelGGL(k_part_sum1<true>, cgrid, dim3(kThreads), NC * 12, s, P, node, n, D, l,
                           med, split, colmax, bitlen, cnt, S1, ppc);
      else
        hipLaunchKernelGGL(k_part_sum1<false>, cgrid, dim3(kThreads), 0, s, P, node, n, D, l, med,
                           split, colmax, bitlen, cnt, S1, ppc);
      DRPT_HIP(hipGetLastError());
    } else {
      int64_t pbx = std::min<int64_t>((n + kThreads * 4 - 1) / (kThreads * 4), 4096);
      hipLaunchKernelGGL(k_part, dim3((unsigned)pbx, (unsigned)T), dim3(kThreads), 0, s, P, node,
                         n, D, l, med, split);
      DRPT_HIP(hipGetLastError());
    }
  }
  return DRPT_OK;
}

}  // namespace
}  // namespace drpt

// ---- allocation-free variant of drpt_grow_trees -----------------------------
// The per-node state of the splits (counts, fixed-point sums, mean, edge
// table, histogram, median: 160 B per node and tree at the deepest split
// level) lives in this file-scope device array instead of the context's
// workspace, for callers that cannot allocate device memory (code under
// stream capture, serving loops). Each GPU the library runs on holds its own
// copy until the process ends, and every call on that GPU uses it: calls that
// overlap on one GPU (several contexts or streams) race. Host code never uses
// the array's name as an address: drpt_create looks up the device address
// once per device (noalloc_setup) and the entry point passes that pointer.
constexpr size_t kNoallocScratchBytes = (size_t)32 << 20;
// trees x nodes of the deepest split level whose state fits the array
constexpr size_t kNoallocMaxNodes = 196608;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace drpt {

namespace {
constexpr int kNoallocDevices = 64;
void* g_scratch_ptr[kNoallocDevices];
std::mutex g_scratch_mu;

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

int noalloc_setup(int device) {
  if (device < 0 || device >= kNoallocDevices) {
    set_error("device index beyond the " + std::to_string(kNoallocDevices) +
              " devices the allocation-free entry points track");
    return DRPT_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return DRPT_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    (void)hipGetLastError();
    set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    return DRPT_ERR_HIP;
  }
  g_scratch_ptr[device] = ptr;
  return DRPT_OK;
}

}  // namespace drpt

extern "C" int drpt_grow_trees_noalloc(drpt_ctx* ctx, const float* P, const uint32_t* colmax,
                                       int64_t n, int ntrees, int depth, int tree_mode,
                                       int leaf_max, uint32_t* leaf_heap) {
  using namespace drpt;
  if (!ctx || !P || !colmax || !leaf_heap) {
    set_error("grow_trees_noalloc: null pointer");
    return DRPT_ERR_ARG;
  }
  if (depth <= 0 || (tree_mode == DRPT_TREE_FIXED && depth > std::log2((double)n))) {
    set_error("depth should be in range [1,....,log2(rows)]");
    return DRPT_ERR_DEPTH;
  }
  DRPT_CHECK(check_tree_args(n, ntrees, depth));
  DRPT_HIP(hipSetDevice(ctx->device));
  if (depth == 1) return fill_u32(ctx, leaf_heap, 1u, (size_t)ntrees * n);
  const size_t TN = (size_t)ntrees << (depth - 2);
  // [cnt u32 | flag u32 | S1 u64 | S2 u64 | mu f32 | med f32 | E f32 x16 | H u32 x16]
  const size_t off_flag = align256(TN * 4);
  const size_t off_S1 = off_flag + align256(TN * 4);
  const size_t off_S2 = off_S1 + align256(TN * 8);
  const size_t off_mu = off_S2 + align256(TN * 8);
  const size_t off_med = off_mu + align256(TN * 4);
  const size_t off_E = off_med + align256(TN * 4);
  const size_t off_H = off_E + align256(TN * 64);
  const size_t total = off_H + align256(TN * 64);
  if (TN > kNoallocMaxNodes || total > kNoallocScratchBytes) {
    set_error("grow_trees_noalloc: " + std::to_string(ntrees) + " trees of depth " +
              std::to_string(depth) + " need " + std::to_string(total) +
              " bytes of split state, more than the " + std::to_string(kNoallocScratchBytes) +
              " bytes this entry point holds (ntrees * 2^(depth-2) must be <= " +
              std::to_string(kNoallocMaxNodes) + "); use "
              "drpt_grow_trees");
    return DRPT_ERR_ARG;
  }
  unsigned char* base =
      ctx->device < kNoallocDevices ? (unsigned char*)g_scratch_ptr[ctx->device] : nullptr;
  if (!base) {
    set_error("grow_trees_noalloc: no scratch address for this device (context not set up)");
    return DRPT_ERR_HIP;
  }
  raise_split_lds(ctx);
  // launched eagerly on the context's stream (a caller may be capturing it);
  // split_levels clears every per-level array it accumulates into
  return split_levels(ctx, P, colmax, n, ntrees, depth, tree_mode, leaf_max, bitlen_of(n),
                      leaf_heap, (uint32_t*)base, (uint32_t*)(base + off_flag),
                      (uint32_t*)(base + off_H), (unsigned long long*)(base + off_S1),
                      (unsigned long long*)(base + off_S2), (float*)(base + off_mu),
                      (float*)(base + off_E), (float*)(base + off_med));
}
