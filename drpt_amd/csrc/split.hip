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
//   k_mu     split decision + mean; clears S2 and H           (per node)
//   k_sum2   sum of quantised fl(v-mu)^2 per node            (per point)
//   k_edges  variance, sigma, edge table; clears the next
//            level's counts and sums                         (per node)
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
#include <utility>
#include <cmath>
#include <mutex>

#include "internal.hpp"

namespace drpt {

namespace {

// Threads per block of the per-point kernels: 16 waves share one copy of a
// block's LDS arrays, so a CU that holds a single block (large arrays) still
// has 16 waves of loads in flight
constexpr int kThreads = 1024;
// per-node LDS arrays of the LDS variants (bytes): k_sum2 16 (int64 sum, mean,
// split flag), k_part_sum1 32 (two children's count + int64 sum, median,
// split flag), k_hist 68 (16 slots + pad) plus the 14 edges (60 with pad) and
// split flag (4) when they fit as well; used while they fit kSumLdsMax /
// kHistLdsMax and the device's per-workgroup LDS
constexpr size_t kSumLdsMax = 128 * 1024;
constexpr size_t kHistLdsMax = 144 * 1024;
// most node shards an LDS variant reads its points for before global atomics
constexpr int kMaxShard = 4;

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

// v + (v of the lane that DPP control kCtrl selects), in the rows of
// kRowMask (v + 0 in the others); the 64 bits move as two 32-bit halves
template <int kCtrl, int kRowMask>
__device__ __forceinline__ long long dpp_add_i64(long long v) {
  const int lo = (int)(uint32_t)(unsigned long long)v;
  const int hi = (int)(uint32_t)((unsigned long long)v >> 32);
  const uint32_t plo = (uint32_t)__builtin_amdgcn_update_dpp(0, lo, kCtrl, kRowMask, 0xF, false);
  const uint32_t phi = (uint32_t)__builtin_amdgcn_update_dpp(0, hi, kCtrl, kRowMask, 0xF, false);
  return (long long)((unsigned long long)v + (((unsigned long long)phi << 32) | plo));
}

// Sum over the wave's 64 lanes (modulo 2^64), the same value in every lane:
// quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror and row_mirror leave
// each row of 16 lanes with its row's sum; row_bcast:15 (rows 1, 3) and
// row_bcast:31 (rows 2, 3) carry the sums forward, so lane 63 holds the total.
__device__ __forceinline__ long long wave_sum_i64(long long v) {
  v = dpp_add_i64<0xB1, 0xF>(v);
  v = dpp_add_i64<0x4E, 0xF>(v);
  v = dpp_add_i64<0x141, 0xF>(v);
  v = dpp_add_i64<0x140, 0xF>(v);
  v = dpp_add_i64<0x142, 0xA>(v);
  v = dpp_add_i64<0x143, 0xC>(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(unsigned long long)v, 63);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((unsigned long long)v >> 32), 63);
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// Points per thread whose node and value words are loaded before any is used
// (independent loads in flight; sums and counts do not depend on the order)
constexpr int kUnroll = 4;

// Split flags: split[node] = 0 for a node that does not split, else
// kSplitBit | count (the count k_mu read, for k_edges and k_median, so the
// count array is free for the next level once k_mu has run).
constexpr uint32_t kSplitBit = 0x80000000u;

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

// Points [b0, b1) of tree blockIdx.y handled by point block pb.
__device__ __forceinline__ void block_range(int64_t n, int64_t ppb, int64_t& b0, int64_t& b1,
                                            int64_t pb = -1) {
  b0 = (pb < 0 ? (int64_t)blockIdx.x : pb) * ppb;
  b1 = min(n, b0 + ppb);
}

// Node shards of the LDS variants: the level's NN nodes in 2^shl ranges of
// NS = NN >> shl, block x = point block * 2^shl + shard; a shard's blocks
// read every point of their range and keep the points of their own nodes, so
// the per-node LDS arrays hold NS entries (levels whose arrays would not fit
// read their points 2^shl times instead of using global atomics). kSh false:
// one shard, with no shard arithmetic or range test in the kernel.
template <bool kSh>
struct Shard {
  int lo, NS;
  int64_t pb;
  __device__ __forceinline__ Shard(int l, int shl) {
    if (kSh) {
      NS = 1 << (l - shl);
      lo = (int)(blockIdx.x & ((1u << shl) - 1u)) << (l - shl);
      pb = (int64_t)(blockIdx.x >> shl);
    } else {
      NS = 1 << l;
      lo = 0;
      pb = blockIdx.x;
    }
  }
  // node index nd (relative to lo) of a level-l point is in this shard
  __device__ __forceinline__ bool has(int nd) const { return !kSh || (unsigned)nd < (unsigned)NS; }
};

// Accumulation of per-node int64 sums (and counts when kCnt) into S / C
// (LDS or global arrays). A batch whose active lanes all hold one node adds
// into per-lane registers while the wave stays on that node, with one wave
// reduction and one atomic per run of such batches; a batch with several
// nodes adds per lane. Every lane of the wave calls add() and flush().
template <bool kCnt>
struct NodeAcc {
  unsigned long long* S;
  uint32_t* C;
  int cur;        // node of the pending run (the same in every lane), -1: none
  long long acc;  // this lane's part of the run's sum
  uint32_t c;     // points of the run
  __device__ __forceinline__ NodeAcc(unsigned long long* S_, uint32_t* C_)
      : S(S_), C(C_), cur(-1), acc(0), c(0) {}
  __device__ __forceinline__ void flush() {
    if (cur >= 0) {
      const long long s = wave_sum_i64(acc);
      if ((threadIdx.x & 63) == 0) {
        atomicAdd(&S[cur], (unsigned long long)s);
        if (kCnt) atomicAdd(&C[cur], c);
      }
      cur = -1;
      acc = 0;
      c = 0;
    }
  }
  __device__ __forceinline__ void add(bool act, int nd, long long q) {
    const unsigned long long am = __ballot(act);
    if (am == 0) return;
    const int leader = __ffsll((long long)am) - 1;
    const int ndl = __builtin_amdgcn_readlane(nd, leader);
    if (__ballot(act && nd == ndl) == am) {
      if (ndl != cur) {
        flush();
        cur = ndl;
      }
      acc += q;  // 0 in inactive lanes
      c += (uint32_t)__popcll(am);
    } else if (act) {
      atomicAdd(&S[nd], (unsigned long long)q);
      if (kCnt) atomicAdd(&C[nd], 1u);
    }
  }
};

// Split decision and mean of one node from its count c and fixed-point sum s1
// (drpt_global.cpp:242-260 with the canonical sums).
__device__ __forceinline__ void node_mu(int mode, int leaf_max, int kq1, uint32_t c,
                                        unsigned long long s1, uint32_t& sp, float& mu) {
  const bool split = (mode == DRPT_TREE_FIXED) || c > (uint32_t)leaf_max;
  sp = split ? (kSplitBit | c) : 0u;
  const double s = ldexp((double)(long long)s1, kq1);
  mu = (float)(s / (double)c);
}

// Edge table of a split node (math_operations.cpp:250-288, factor 1,
// no_of_bins 7): variance from the fixed-point sum s2, then e[0..13] around
// the mean m; e[14] = e[15] = 0.
__device__ __forceinline__ void node_edges(uint32_t sp, int kq2, unsigned long long s2q, float m,
                                           float (&e)[16]) {
  const long long c = (long long)(sp & ~kSplitBit);
  const double s2 = ldexp((double)(long long)s2q, kq2);
  const float var = (float)(s2 / (double)(c - 1));
  const float sigma = sqrtf(var);
  const float s1 = (float)((double)sigma / 30.0);
  const float s2f = (float)((double)sigma / 6.0);
  const float s3 = (float)((double)sigma / 2.0);
#pragma unroll
  for (int j = 1; j <= 4; ++j) {
    const float r = __fmul_rn((float)j, s1);
    e[6 + j] = __fadd_rn(m, r);
    e[7 - j] = __fsub_rn(m, r);
  }
#pragma unroll
  for (int j = 1; j <= 2; ++j) {
    const float r = __fadd_rn(sigma, __fmul_rn((float)j, s2f));
    e[10 + j] = __fadd_rn(m, r);
    e[3 - j] = __fsub_rn(m, r);
  }
  {
    const float r = __fadd_rn(__fmul_rn(2.0f, sigma), s3);
    e[13] = __fadd_rn(m, r);
    e[0] = __fsub_rn(m, r);
  }
  e[14] = 0.0f;
  e[15] = 0.0f;
}

// Median of a split node: cumulative search over its 16-slot histogram h and
// interpolation in the edge table (edge(j) = E[j]), math_operations.cpp:333-377
// with E[-1] = 0.
template <class EdgeAt>
__device__ __forceinline__ float node_median(uint32_t sp, const uint32_t* h, EdgeAt edge) {
  const long long c = (long long)(sp & ~kSplitBit);
  const float val = (float)((50 * c) / 100);
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
  const double count = (double)h[sel];
  const float lo = (sel == 0) ? 0.0f : edge(sel - 1);
  const float width = __fsub_rn(edge(sel), lo);
  const double num = __dsub_rn((double)(c / 2), __dsub_rn((double)cf, count));
  const double frac = num / count;
  const double m = __dadd_rn((double)lo, __dmul_rn(frac, (double)width));
  return (float)m;
}

// Adds one batch of slots b (lanes where act) to the histograms Hb[nd * stride
// + slot]: when every active lane holds one node, lane k < 16 adds the
// population of slot k (wave ballots); otherwise one atomic per lane. Every
// lane of the wave calls it.
__device__ __forceinline__ void hist_add(bool act, int nd, int b, uint32_t* Hb, int stride,
                                         int lane) {
  const unsigned long long am = __ballot(act);
  if (am == 0) return;
  const int leader = __ffsll((long long)am) - 1;
  const int ndl = __builtin_amdgcn_readlane(nd, leader);
  if (__ballot(act && nd == ndl) == am) {
    uint32_t mine = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const uint32_t c = (uint32_t)__popcll(__ballot(act && b == k));
      mine = (lane == k) ? c : mine;
    }
    if (lane < 16 && mine) atomicAdd(&Hb[(int64_t)ndl * stride + lane], mine);
  } else if (act) {
    atomicAdd(&Hb[(int64_t)nd * stride + b], 1u);
  }
}

// Count and sum of quantised values of the root (level 0): every point of a
// tree is in node 1, so each wave adds one run.
__global__ void __launch_bounds__(kThreads)
k_sum1(const float* __restrict__ P, int64_t n, int D, const uint32_t* __restrict__ colmax,
       int bitlen, uint32_t* __restrict__ cnt, unsigned long long* __restrict__ S1, int64_t ppb) {
  const int t = blockIdx.y;
  const int64_t col = (int64_t)t * D;
  int kq1, kq2;
  col_exponents(colmax[col], bitlen, kq1, kq2);
  const float* Pc = P + col * n;
  __shared__ unsigned long long bS;
  __shared__ uint32_t bC;
  if (threadIdx.x == 0) {
    bS = 0;
    bC = 0;
  }
  __syncthreads();
  NodeAcc<true> acc(&bS, &bC);
  int64_t b0, b1;
  block_range(n, ppb, b0, b1);
  for (int64_t base = b0; base < b1; base += kThreads * kUnroll) {
    float pu[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + u * kThreads + threadIdx.x;
      pu[u] = i < b1 ? Pc[i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + u * kThreads + threadIdx.x;
      acc.add(i < b1, 0, i < b1 ? quant(pu[u], kq1) : 0);
    }
  }
  acc.flush();
  __syncthreads();
  if (threadIdx.x == 0 && bC) {
    atomicAdd(&cnt[t], bC);
    atomicAdd(&S1[t], bS);
  }
}

// Split decision and mean per node (drpt_global.cpp:242-260 with the
// canonical sums); clears this level's S2 and histogram entries for k_sum2
// and k_hist.
__global__ void k_mu(int T, int l, int D, int mode, int leaf_max, int bitlen,
                     const uint32_t* __restrict__ colmax, const uint32_t* __restrict__ cnt,
                     const unsigned long long* __restrict__ S1, float* __restrict__ mu,
                     uint32_t* __restrict__ split, unsigned long long* __restrict__ S2,
                     uint32_t* __restrict__ H) {
  const int NN = 1 << l;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)T * NN) return;
  int t = (int)(idx / NN);
  int kq1, kq2;
  col_exponents(colmax[(int64_t)t * D + l], bitlen, kq1, kq2);
  uint32_t sp;
  float m;
  node_mu(mode, leaf_max, kq1, cnt[idx], S1[idx], sp, m);
  split[idx] = sp;
  mu[idx] = m;
  S2[idx] = 0ull;
  uint4* h = reinterpret_cast<uint4*>(H + idx * 16);
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  h[0] = z;
  h[1] = z;
  h[2] = z;
  h[3] = z;
}

// Sum of quantised fl(v - mu)^2 per split node. kLds: the tree's sums, means
// and split flags in LDS ([NN] int64 + [NN] float + [NN] u32).
template <bool kLds, bool kSh>
__global__ void __launch_bounds__(kThreads)
k_sum2(const float* __restrict__ P, const uint32_t* __restrict__ node, int64_t n, int D, int l,
       const uint32_t* __restrict__ colmax, int bitlen, const float* __restrict__ mu,
       const uint32_t* __restrict__ split, unsigned long long* __restrict__ S2, int64_t ppb,
       int shl) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = blockIdx.y;
  const int NN = 1 << l;
  const Shard<kSh> sh(l, shl);
  const int NS = sh.NS;
  // this shard's nodes [lo, lo + NS): pointers shifted to them
  const float* mg = mu + (int64_t)t * NN + sh.lo;
  const uint32_t* sg = split + (int64_t)t * NN + sh.lo;
  unsigned long long* gS = S2 + (int64_t)t * NN + sh.lo;
  unsigned long long* lS = reinterpret_cast<unsigned long long*>(smem);
  float* lm = reinterpret_cast<float*>(lS + (kLds ? NS : 0));
  uint32_t* ls = reinterpret_cast<uint32_t*>(lm + (kLds ? NS : 0));
  if (kLds) {
    for (int k = threadIdx.x; k < NS; k += kThreads) {
      lS[k] = 0;
      lm[k] = mg[k];
      ls[k] = sg[k];
    }
    __syncthreads();
  }
  const float* mt = kLds ? lm : mg;
  const uint32_t* st = kLds ? ls : sg;
  const int64_t col = (int64_t)t * D + l;
  int kq1, kq2;
  col_exponents(colmax[col], bitlen, kq1, kq2);
  const float* Pc = P + col * n;
  const uint32_t* nt = node + (int64_t)t * n;
  NodeAcc<false> acc(kLds ? lS : gS, nullptr);
  int64_t b0, b1;
  block_range(n, ppb, b0, b1, sh.pb);
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
        nd = (int)(h - (uint32_t)NN) - sh.lo;
        act = act && sh.has(nd);
        if (act) act = st[nd] != 0u;
        if (act) {
          float w = __fsub_rn(pu[u], mt[nd]);
          float w2 = __fmul_rn(w, w);
          q = quant(w2, kq2);
        }
      }
      acc.add(act, nd, q);
    }
  }
  acc.flush();
  if (kLds) {
    __syncthreads();
    for (int k = threadIdx.x; k < NS; k += kThreads)
      if (lS[k]) atomicAdd(&gS[k], lS[k]);
  }
}

// Edge table: math_operations.cpp:250-288 (factor 1, no_of_bins 7). With
// zero_next, also clears the next level's count and sum entries of this
// node's two children (k_mu of this level was the last reader of both).
__global__ void k_edges(int T, int l, int D, int bitlen, const uint32_t* __restrict__ colmax,
                        const unsigned long long* __restrict__ S2,
                        const float* __restrict__ mu, const uint32_t* __restrict__ split,
                        float* __restrict__ E, int zero_next, uint32_t* __restrict__ cnt,
                        unsigned long long* __restrict__ S1) {
  const int NN = 1 << l;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)T * NN) return;
  if (zero_next) {
    // children of (t, nd) at level l + 1: t * 2NN + 2nd + {0, 1} = 2 idx + {0, 1}
    cnt[2 * idx] = 0u;
    cnt[2 * idx + 1] = 0u;
    S1[2 * idx] = 0ull;
    S1[2 * idx + 1] = 0ull;
  }
  const uint32_t sp = split[idx];
  if (!sp) return;
  int t = (int)(idx / NN);
  int kq1, kq2;
  col_exponents(colmax[(int64_t)t * D + l], bitlen, kq1, kq2);
  float e[16];
  node_edges(sp, kq2, S2[idx], mu[idx], e);
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

// Slot from the 14 edges alone (e[14] = 0 is a constant): the same count c,
// by binary search over the non-decreasing e[0..13] (4 dependent reads
// instead of 14; all-NaN edge tables still give 0).
__device__ __forceinline__ int slot_bsearch(float v, const float* e) {
  int c = e[7] < v ? 8 : 0;
  c += e[c + 3] < v ? 4 : 0;
  c += e[c + 1] < v ? 2 : 0;
  if (c < 14) c += e[c] < v ? 1 : 0;
  if (c == 14) c = v <= 0.0f ? 14 : 0;
  return c;
}

// LDS row strides (words) of the histogram and edge tables: odd, so lanes at
// random nodes spread over the 32 banks
constexpr int kHS = 17;
constexpr int kES = 15;

template <bool kLds, bool kEL, bool kSh>
__global__ void __launch_bounds__(kThreads)
k_hist(const float* __restrict__ P, const uint32_t* __restrict__ node, int64_t n, int D, int l,
       const float* __restrict__ E, const uint32_t* __restrict__ split, uint32_t* __restrict__ H,
       int64_t ppb, int shl) {
  // kLds: histograms [NS][kHS] in LDS; kEL: also the shard's edges [NS][kES]
  // and split flags [NS] (staged once per block, read per point)
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = blockIdx.y;
  const int NN = 1 << l;
  const Shard<kSh> sh(l, shl);
  const int NS = sh.NS;
  uint32_t* lH = reinterpret_cast<uint32_t*>(smem);
  float* lE = reinterpret_cast<float*>(lH + (kLds ? NS * kHS : 0));
  uint32_t* lS = reinterpret_cast<uint32_t*>(lE + (kEL ? NS * kES : 0));
  // this shard's nodes [lo, lo + NS): pointers shifted to them
  const float* Eg = E + ((int64_t)t * NN + sh.lo) * 16;
  const uint32_t* Sg = split + (int64_t)t * NN + sh.lo;
  if (kLds) {
    for (int k = threadIdx.x; k < NS * kHS; k += kThreads) lH[k] = 0;
    if (kEL) {
      for (int k = threadIdx.x; k < NS * 16; k += kThreads)
        if ((k & 15) < 14) lE[(k >> 4) * kES + (k & 15)] = Eg[k];
      for (int k = threadIdx.x; k < NS; k += kThreads) lS[k] = Sg[k];
    }
    __syncthreads();
  }
  const int64_t col = (int64_t)t * D + l;
  const float* Pc = P + col * n;
  const uint32_t* nt = node + (int64_t)t * n;
  const uint32_t* st = kEL ? lS : Sg;
  uint32_t* gH = H + ((int64_t)t * NN + sh.lo) * 16;
  int64_t b0, b1;
  block_range(n, ppb, b0, b1, sh.pb);
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
        nd = (int)(h - (uint32_t)NN) - sh.lo;
        act = act && sh.has(nd);
        if (act) act = st[nd] != 0u;
        if (act) {
          if (kEL) {
            b = slot_bsearch(pu[u], lE + nd * kES);
          } else {
            float e[16];
            const float4* src = reinterpret_cast<const float4*>(Eg + (int64_t)nd * 16);
            float4 a0 = src[0], a1 = src[1], a2 = src[2], a3 = src[3];
            e[0] = a0.x; e[1] = a0.y; e[2] = a0.z; e[3] = a0.w;
            e[4] = a1.x; e[5] = a1.y; e[6] = a1.z; e[7] = a1.w;
            e[8] = a2.x; e[9] = a2.y; e[10] = a2.z; e[11] = a2.w;
            e[12] = a3.x; e[13] = a3.y; e[14] = a3.z; e[15] = a3.w;
            b = slot_of(pu[u], e);
          }
        }
      }
      if (kLds) hist_add(act, nd, b, lH, kHS, lane);
      else hist_add(act, nd, b, gH, 16, lane);
    }
  }
  if (kLds) {
    __syncthreads();
    for (int k = threadIdx.x; k < NS * 16; k += kThreads) {
      const uint32_t v = lH[(k >> 4) * kHS + (k & 15)];
      if (v) atomicAdd(&gH[k], v);
    }
  }
}

// Cumulative search + interpolation: math_operations.cpp:333-377, E[-1] = 0.
__global__ void k_median(int T, int l, const uint32_t* __restrict__ H,
                         const float* __restrict__ E, const uint32_t* __restrict__ split,
                         float* __restrict__ med) {
  const int NN = 1 << l;
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)T * NN) return;
  const uint32_t sp = split[idx];
  if (!sp) return;
  const float* e = E + idx * 16;
  med[idx] = node_median(sp, H + idx * 16, [&](int j) { return e[j]; });
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
// ([2NN] int64 sums + [2NN] counts) and the tree's medians and split flags
// ([NN] float + [NN] u32) in LDS.
template <bool kLds, bool kSh>
__global__ void __launch_bounds__(kThreads)
k_part_sum1(const float* __restrict__ P, uint32_t* __restrict__ node, int64_t n, int D, int l,
            const float* __restrict__ med, const uint32_t* __restrict__ split,
            const uint32_t* __restrict__ colmax, int bitlen, uint32_t* __restrict__ cnt,
            unsigned long long* __restrict__ S1, int64_t ppb, int shl) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = blockIdx.y;
  const int NN = 1 << l, NC = 2 << l;
  const Shard<kSh> sh(l, shl);
  const int NS = sh.NS, NCS = 2 * NS;
  // this shard's parents [lo, lo + NS) and their children [2 lo, 2 lo + 2 NS)
  const float* mg = med + (int64_t)t * NN + sh.lo;
  const uint32_t* sg = split + (int64_t)t * NN + sh.lo;
  uint32_t* gc = cnt + (int64_t)t * NC + 2 * sh.lo;
  unsigned long long* gS = S1 + (int64_t)t * NC + 2 * sh.lo;
  unsigned long long* lS = reinterpret_cast<unsigned long long*>(smem);
  uint32_t* lc = reinterpret_cast<uint32_t*>(lS + (kLds ? NCS : 0));
  float* lm = reinterpret_cast<float*>(lc + (kLds ? NCS : 0));
  uint32_t* ls = reinterpret_cast<uint32_t*>(lm + (kLds ? NS : 0));
  if (kLds) {
    for (int k = threadIdx.x; k < NCS; k += kThreads) {
      lS[k] = 0;
      lc[k] = 0;
    }
    for (int k = threadIdx.x; k < NS; k += kThreads) {
      lm[k] = mg[k];
      ls[k] = sg[k];
    }
    __syncthreads();
  }
  const float* mt = kLds ? lm : mg;
  const uint32_t* st = kLds ? ls : sg;
  const float* Pc = P + ((int64_t)t * D + l) * n;
  const float* Pn = Pc + n;  // level l + 1 of the same tree
  int kq1, kq2;
  col_exponents(colmax[(int64_t)t * D + l + 1], bitlen, kq1, kq2);
  uint32_t* nt = node + (int64_t)t * n;
  NodeAcc<true> acc(kLds ? lS : gS, kLds ? lc : gc);
  int64_t b0, b1;
  block_range(n, ppb, b0, b1, sh.pb);
  for (int64_t base = b0; base < b1; base += kThreads * kUnroll) {
    uint32_t hu[kUnroll];
    float pu[kUnroll], pn[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + u * kThreads + threadIdx.x;
      hu[u] = i < b1 ? nt[i] : 0u;
      pu[u] = i < b1 ? Pc[i] : 0.0f;
      pn[u] = i < b1 ? Pn[i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int64_t i = base + u * kThreads + threadIdx.x;
      bool act = false;
      int nd = 0;
      long long q = 0;
      if (i < b1) {
        const uint32_t h = hu[u];
        const int np = (int)(h - (uint32_t)NN) - sh.lo;
        if ((h >> l) == 1u && sh.has(np)) {
          if (st[np]) {
            const uint32_t hc = 2u * h + ((pu[u] <= mt[np]) ? 0u : 1u);
            nt[i] = hc;
            act = true;
            nd = (int)(hc - (uint32_t)NC) - 2 * sh.lo;
            q = quant(pn[u], kq1);
          }
        }
      }
      acc.add(act, nd, q);
    }
  }
  acc.flush();
  if (kLds) {
    __syncthreads();
    for (int k = threadIdx.x; k < NCS; k += kThreads) {
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
    // every LDS variant, sharded or not
    const std::pair<const void*, int> fns[] = {
        {(const void*)k_sum2<true, false>, bytes},
        {(const void*)k_sum2<true, true>, bytes},
        {(const void*)k_part_sum1<true, false>, bytes},
        {(const void*)k_part_sum1<true, true>, bytes},
        {(const void*)k_hist<true, false, false>, hbytes},
        {(const void*)k_hist<true, false, true>, hbytes},
        {(const void*)k_hist<true, true, false>, hbytes}};
    bool ok = true;
    for (const auto& f : fns)
      ok = ok && hipFuncSetAttribute(f.first, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     f.second) == hipSuccess;
    if (ok)
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
  // points per block: with per-node LDS arrays, enough points to amortise
  // staging and flushing them; with global atomics, small blocks so every CU
  // has work
  auto grid_of = [&](int64_t ppb) { return dim3((unsigned)((n + ppb - 1) / ppb), (unsigned)T); };
  DRPT_CHECK(fill_u32(ctx, node, 1u, (size_t)T * n));
  // the root's count and sum (later levels get theirs from the fused partition)
  DRPT_CHECK(zero4(ctx, cnt, (size_t)T, (uint32_t*)S1, 2 * (size_t)T));
  hipLaunchKernelGGL(k_sum1, grid_of(16384), dim3(kThreads), 0, s, P, n, D, colmax, bitlen, cnt, S1,
                     (int64_t)16384);
  DRPT_HIP(hipGetLastError());
  for (int l = 0; l <= D - 2; ++l) {
    const int64_t NN = (int64_t)1 << l;
    const int64_t TN = (int64_t)T * NN;
    const unsigned ngrid = (unsigned)((TN + 255) / 256);
    const bool last = l == D - 2;
    // k_mu also clears this level's S2 and histograms
    hipLaunchKernelGGL(k_mu, dim3(ngrid), dim3(256), 0, s, T, l, D, mode, leaf_max, bitlen, colmax,
                       cnt, S1, mu, split, S2, H);
    DRPT_HIP(hipGetLastError());
    // LDS variants: the smallest node shard count 2^shl (up to kMaxShard)
    // whose per-shard arrays fit; beyond it, global atomics. Shards re-read
    // every point, which pays while most points are still active: at every
    // level of fixed-depth trees, and in leaf-max trees while an average node
    // holds more than leaf_max points (deeper, most points sit in leaves and
    // the few atomics of the active ones cost less than the extra reads).
    const bool busy = mode == DRPT_TREE_FIXED || ((int64_t)leaf_max << l) < n;
    auto shards = [&](size_t bytes_per_node, size_t cap) -> int {
      for (int shl = 0; (1 << shl) <= (busy ? kMaxShard : 1) && shl <= l; ++shl)
        if ((size_t)(NN >> shl) * bytes_per_node <= cap) return shl;
      return -1;
    };
    auto sgrid = [&](int64_t ppb, int shl) {
      return dim3((unsigned)(((n + ppb - 1) / ppb) << shl), (unsigned)T);
    };
    const int sh2 = shards(16, lds_max);
    if (sh2 >= 0) {
      const int64_t NS = NN >> sh2;
      const int64_t ppb = std::max<int64_t>(8192, NS * 8);
      auto* f = sh2 ? k_sum2<true, true> : k_sum2<true, false>;
      hipLaunchKernelGGL(f, sgrid(ppb, sh2), dim3(kThreads), (size_t)NS * 16, s, P, node, n, D, l,
                         colmax, bitlen, mu, split, S2, ppb, sh2);
    } else {
      hipLaunchKernelGGL((k_sum2<false, false>), grid_of(8192), dim3(kThreads), 0, s, P, node, n, D,
                         l, colmax, bitlen, mu, split, S2, (int64_t)8192, 0);
    }
    DRPT_HIP(hipGetLastError());
    // k_edges also clears the next level's counts and sums
    hipLaunchKernelGGL(k_edges, dim3(ngrid), dim3(256), 0, s, T, l, D, bitlen, colmax, S2, mu,
                       split, E, last ? 0 : 1, cnt, S1);
    DRPT_HIP(hipGetLastError());
    // histograms: edges staged as well where they fit unsharded; else the
    // histograms alone, sharded if need be
    const size_t hist_node = (size_t)kHS * 4, edge_node = (size_t)kES * 4 + 4;
    const int shh = shards(hist_node, hist_max);
    const bool lds_edges = shh == 0 && (size_t)NN * (hist_node + edge_node) <= hist_max;
    if (lds_edges) {
      const int64_t ppb = std::max<int64_t>(8192, NN * 32);
      hipLaunchKernelGGL((k_hist<true, true, false>), grid_of(ppb), dim3(kThreads),
                         (size_t)NN * (hist_node + edge_node), s, P, node, n, D, l, E, split, H,
                         ppb, 0);
    } else if (shh >= 0) {
      const int64_t NS = NN >> shh;
      const int64_t ppb = std::max<int64_t>(8192, NS * 32);
      auto* f = shh ? k_hist<true, false, true> : k_hist<true, false, false>;
      hipLaunchKernelGGL(f, sgrid(ppb, shh), dim3(kThreads), (size_t)NS * hist_node, s, P, node, n,
                         D, l, E, split, H, ppb, shh);
    } else {
      hipLaunchKernelGGL((k_hist<false, false, false>), grid_of(8192), dim3(kThreads), 0, s, P, node,
                         n, D, l, E, split, H, (int64_t)8192, 0);
    }
    DRPT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_median, dim3(ngrid), dim3(256), 0, s, T, l, H, E, split, med);
    DRPT_HIP(hipGetLastError());
    if (!last) {
      // partition fused with the children's counts and sums (per parent node:
      // two children's count + int64 sum, the median and the split flag)
      const int shp = shards(32, lds_max);
      if (shp >= 0) {
        const int64_t NS = NN >> shp;
        const int64_t ppc = std::max<int64_t>(8192, 2 * NS * 8);
        auto* f = shp ? k_part_sum1<true, true> : k_part_sum1<true, false>;
        hipLaunchKernelGGL(f, sgrid(ppc, shp), dim3(kThreads), (size_t)NS * 32, s, P, node, n, D,
                           l, med, split, colmax, bitlen, cnt, S1, ppc, shp);
      } else {
        hipLaunchKernelGGL((k_part_sum1<false, false>), grid_of(8192), dim3(kThreads), 0, s, P, node,
                           n, D, l, med, split, colmax, bitlen, cnt, S1, (int64_t)8192, 0);
      }
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
