// Stage (c): the pooled filter pass with each leaf resident in LDS.
//
// The graph build's certifying pass (MDRPT::calculate_nns,
// cpp/src/drpt/algo/mdrpt.cpp:123-212, restated as: every leafmate j of row i
// whose exact distance can be at most the row's bound B(i) goes to row i's
// pool) for leaves whose bf16 rows fit the LDS. The streaming variant
// (k_knn_lowp, pool mode) re-reads a leaf's rows once per 64-row query block;
// here a block stages the whole leaf once (global_load_lds_dwordx4, rows
// XOR-swizzled by 16-byte chunk for conflict-free MFMA operand reads) and its
// eight waves take (64 query rows x column range) tasks from an LDS counter,
// with no barrier until the leaf is done.
//
// Filter. The accumulator of the v_mfma_f32_32x32x16_bf16 tile starts at
// -|x~_j|^2 / 2 (column j, exact halving), so it ends at
//   acc_ij = G_ij - |x~_j|^2 / 2 = (|x~_i|^2 - a_ij) / 2,
// and a_ij <= A_cut(B_i) becomes acc_ij >= thr_i with the per-row threshold
//   thr_i = round_down((|x~_i|^2 - A_cut(B_i) - eta_i) / 2):
// per element the sign bit of fl(acc_ij - thr_i), appended to the lane's mask
// by one v_alignbit. The extra eta_i = (2d + 16) 2^-24 (n_i + nm)^2 (the
// bound A_cut already holds once) covers the different rounding of this form
// (the initialised accumulator, the stored norms and the subtraction, each
// within (d + 4) 2^-24 (n_i + nm)^2), so every leafmate the streaming pass
// keeps is kept here as well.
//
// Survivors are collected per lane (its query row and 16 columns of each
// tile) in a register list of 8 leaf positions and appended to the row's pool
// with one atomic per list; known ids are not dropped here (the row merge's
// hash drops them before any exact distance).
//
// Integer rows (kI8). When every value of X is an integer in [0, 255], the
// rows are staged as int8 x - 128 (distances do not change under the shift)
// and the Gram G' runs on v_mfma_i32_32x32x32_i8: exact in int32, twice the
// bf16 rate, 128 bytes per row of d <= 128, so two blocks share a CU and one
// block's leaf set-up overlaps the other's tasks. With n'_j = |x_j - 128|^2
// (exact, from the staged row) the accumulator starts at -floor(n'_j / 2) and
// the filter keeps acc_ij >= thr_i = floor((n'_i - B_i) / 2) in integers: any
// leafmate with D(i, j) = n'_i + n'_j - 2 G'_ij <= B_i passes (the test is
// the exact one up to the halving). Candidates start at -ceil(n'_j / 2), so
// that a = n'_i - 2 acc >= D, the bound their keys must give. The bf16 and
// int8 launches are both issued; a device flag (set by the quantisation
// kernel) lets exactly one of them run.
//
// Roofline: MFMA-bound, 2 d m^2 flops per leaf; HBM bytes per leaf m * 2d
// (rows; m * d for int8) + 16 m (ids, norms, errors, bounds).
#include <climits>
#include <type_traits>

#include "internal.hpp"
#include "lowp_common.hpp"

namespace drpt {

namespace {

using namespace lowp;

constexpr int kPlWaves = 8;
constexpr int kPlThreads = 64 * kPlWaves;
constexpr int kPlThreads8 = 256;  // int8 instances: 4 waves, two blocks per CU
constexpr int kPlPitch = 256;   // LDS bytes per leaf row (16 chunks of 16 bytes)
constexpr int kPlSteps = 8;     // k-steps of 16 elements: rows of at most 128 bf16
constexpr int kPlList = 8;      // pool entries a lane holds per query row
constexpr int kPlQueues = 8;    // ticket queues, one per XCD (blocks go round-robin)
constexpr int kPlMisc = 40;     // ints of bookkeeping after the leaf arrays

constexpr int kPlPitch8 = 128;  // int8 rows: 8 chunks of 16 bytes (d <= 128)
constexpr int kPlSteps8 = 4;    // k-steps of 32 int8 elements

typedef int intx4 __attribute__((ext_vector_type(4)));
typedef int intx16 __attribute__((ext_vector_type(16)));

// chunk c of LDS row r for row pitch P (chunks XOR-swizzled by the row)
template <int P>
__device__ __forceinline__ uint4 row_chunk(const unsigned char* rows, int r, int c) {
  return *reinterpret_cast<const uint4*>(rows + r * P + ((c ^ (r & (P / 16 - 1))) << 4));
}

// Per-lane list of 8 leaf positions in 4 registers (newest in w[0] bits 0-15).
struct PosList {
  uint32_t w[4];
  int cnt;
  __device__ __forceinline__ void reset() {
    w[0] = w[1] = w[2] = w[3] = 0u;
    cnt = 0;
  }
  __device__ __forceinline__ void push(uint32_t c) {
    w[3] = __builtin_amdgcn_alignbit(w[3], w[2], 16);
    w[2] = __builtin_amdgcn_alignbit(w[2], w[1], 16);
    w[1] = __builtin_amdgcn_alignbit(w[1], w[0], 16);
    w[0] = (w[0] << 16) | c;
    ++cnt;
  }
};

// Appends the list's entries to row id's pool (one atomic); a pool that
// cannot take them sends the row to the exact row path.
__device__ __forceinline__ void flush_list(const KnnArgs& A, const int32_t* lid, int id,
                                           PosList& L) {
  if (L.cnt <= 0) return;
  const int base = atomicAdd(&A.pool_cnt[id], L.cnt);
  if (base + L.cnt <= A.pool_cap) {
    int32_t* dst = A.out_idx + (int64_t)id * A.pool_cap + base;
#pragma unroll
    for (int j = 0; j < kPlList; ++j)
      if (j < L.cnt) dst[j] = lid[(L.w[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu];
  } else if (atomicExch(&A.ovf_flag[id], 1) == 0) {
    A.ovf_list[atomicAdd(A.novf, 1)] = id;
  }
  L.reset();
}

// Both query rows' lists: the two reservations issued before either is awaited.
__device__ __forceinline__ void flush_pair(const KnnArgs& A, const int32_t* lid, const int (&id)[2],
                                           PosList (&L)[2]) {
  int base[2];
#pragma unroll
  for (int g = 0; g < 2; ++g) base[g] = L[g].cnt > 0 ? atomicAdd(&A.pool_cnt[id[g]], L[g].cnt) : 0;
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    if (L[g].cnt <= 0) continue;
    if (base[g] + L[g].cnt <= A.pool_cap) {
      int32_t* dst = A.out_idx + (int64_t)id[g] * A.pool_cap + base[g];
#pragma unroll
      for (int j = 0; j < kPlList; ++j)
        if (j < L[g].cnt) dst[j] = lid[(L[g].w[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu];
    } else if (atomicExch(&A.ovf_flag[id[g]], 1) == 0) {
      A.ovf_list[atomicAdd(A.novf, 1)] = id[g];
    }
  }
}

// Filter of one lane's 16 tile elements against threshold thr: bit e (column
// 8 g + u of the lane's half for e = 4 g + u) set when acc[e] >= thr, from the
// sign of fl(acc[e] - thr) (one v_sub and one v_alignbit per element: the
// alignbit shifts the mask left and appends the sign bit). thr is never +0, so
// a zero difference is +0 and the sign test equals the comparison.
__device__ __forceinline__ uint32_t pass_bits(const floatx16& acc, float thr) {
  uint32_t fb = 0;
#pragma unroll
  for (int e = 15; e >= 0; --e)
    fb = __builtin_amdgcn_alignbit(fb, __float_as_uint(acc[e] - thr), 31);
  return ~fb & 0xFFFFu;
}
// integer accumulators: the sign of acc[e] - thr (|acc| < 2^23 and thr in
// [-2^30, 2^30], so the difference does not wrap)
__device__ __forceinline__ uint32_t pass_bits(const intx16& acc, int thr) {
  uint32_t fb = 0;
#pragma unroll
  for (int e = 15; e >= 0; --e) fb = __builtin_amdgcn_alignbit(fb, (uint32_t)(acc[e] - thr), 31);
  return ~fb & 0xFFFFu;
}

// sum of squares of the 16 int8 values of a chunk
__device__ __forceinline__ int sq16(uint4 v) {
  int s = __builtin_amdgcn_sdot4((int)v.x, (int)v.x, 0, false);
  s = __builtin_amdgcn_sdot4((int)v.y, (int)v.y, s, false);
  s = __builtin_amdgcn_sdot4((int)v.z, (int)v.z, s, false);
  return __builtin_amdgcn_sdot4((int)v.w, (int)v.w, s, false);
}

// LDS layout of k_pool_leaf for mmax leaf rows: the rows, then per row its id,
// -|x~|^2 / 2 and threshold (int8: the integer forms), then kPlMisc ints of
// bookkeeping.
__host__ __device__ constexpr size_t pool_leaf_row_bytes(int) { return kPlPitch + 12; }
__host__ __device__ constexpr size_t pool_leaf_row_bytes8() { return kPlPitch8 + 12; }

// misc (ints): next-item slots [8 s .. 8 s + 6] (a LeafItem and a valid flag),
// task counter [16], per-wave maxima (floats) [20 .. 35]
constexpr int kMiscTask = 16, kMiscRed = 20;

// tid 0 only: the next work item of this block's queue whose leaf size lies
// in [size_lo, size_hi] (items of other launches are passed over), or none,
// into slot s of misc
__device__ __forceinline__ void fetch_item(const KnnArgs& A, unsigned* myq, int64_t begin,
                                           int64_t end, int* misc, int s) {
  for (;;) {
    const int64_t item = begin + (int64_t)atomicAdd(myq, 1u);
    if (item >= end) {
      misc[8 * s + 6] = 0;
      return;
    }
    static_assert(sizeof(LeafItem) == 24, "LeafItem is three 8-byte words");
    const int2* src = reinterpret_cast<const int2*>(&A.items[item]);
    const int2 w0 = src[0], w1 = src[1], w2 = src[2];  // (tree, size), start, (row0, rows)
    if (w0.y >= A.size_lo && w0.y <= A.size_hi) {
      int2* dst = reinterpret_cast<int2*>(&misc[8 * s]);
      dst[0] = w0;
      dst[1] = w1;
      dst[2] = w2;
      misc[8 * s + 6] = 1;
      return;
    }
  }
}

// MODE kModePool: survivors of the certified filter to the rows' pools.
// MODE kModeCand: each row's k' = 4 TK widened candidates, the TK smallest
// approximate distances of each of its four column streams (two column ranges
// x two lane halves; of every 4 adjacent columns of a stream the nearest one),
// as keys (approx bits & ~0x7FF) << 32 | id.
//
// Per leaf: the next leaf's ids, norms, errors and bounds are loaded into
// registers, and the item after it is fetched by thread 0, while the leaf's
// tasks run; after them (one barrier) the next leaf is installed (ids and
// maxima in LDS, its rows by global_load_lds_dwordx4, thresholds).
//
// kI8: int8 rows (A.Xi8), KS k-steps of 32 elements, 256-thread blocks, two
// per CU (each holds one leaf, so one's set-up overlaps the other's tasks). With
// A.notu8 given, the bf16 instance returns at once when *notu8 == 0 and the
// int8 one when it is not.
template <int KS, int MODE, int TK, bool kI8>
__global__ void __launch_bounds__(kI8 ? kPlThreads8 : kPlThreads, kI8 ? 2 : 1) k_pool_leaf(const KnnArgs A, int mmax) {
  static_assert(MODE == kModePool || MODE == kModeCand, "pool or candidate mode");
  constexpr bool kPool = MODE == kModePool;
  if (A.notu8 != nullptr && ((*A.notu8 == 0) != kI8)) return;
  typedef typename std::conditional<kI8, intx16, floatx16>::type acc_t;
  typedef typename std::conditional<kI8, int, float>::type thr_t;
  constexpr int kPitch = kI8 ? kPlPitch8 : kPlPitch;
  extern __shared__ __align__(16) unsigned char smem[];
  unsigned char* rows = smem;
  int32_t* lid = reinterpret_cast<int32_t*>(smem + (size_t)mmax * kPitch);
  // per leaf row: the accumulator's start (-|x~|^2 / 2; int8: -floor or
  // -ceil(n' / 2)) and the filter threshold (pool) or squared norm (candidates)
  thr_t* chalf = reinterpret_cast<thr_t*>(lid + mmax);
  thr_t* thrs = chalf + mmax;
  int* misc = reinterpret_cast<int*>(thrs + mmax);
  float* red = reinterpret_cast<float*>(misc + kMiscRed);  // per wave: max |x~|^2 [w], max error [8 + w]
  const uint8_t* __restrict__ Xq = static_cast<const uint8_t*>(kI8 ? A.Xi8 : A.Xq);
  constexpr int d = kI8 ? 32 * KS : 16 * KS;
  constexpr int rb = 32 * KS;    // bytes per staged row
  constexpr int L = rb >> 4;     // 16-byte chunks per row
  constexpr int kRowsPerDma = 1024 / kPitch;  // rows one wave-wide DMA instruction covers
  // block: 8 waves (bf16, one block per CU) or 4 (int8, two per CU)
  constexpr int kThr = kI8 ? kPlThreads8 : kPlThreads, kWv = kThr / 64;
  constexpr int kRowsPerThread = kI8 ? 3 : 2;  // leaf rows whose metadata a thread loads
  const int64_t n = A.n;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int half = lane >> 5;
  const int qx = blockIdx.x % kPlQueues;
  const int64_t chunk = (A.range[1] - A.range[0] + kPlQueues - 1) / kPlQueues;
  const int64_t begin = A.range[0] + qx * chunk;
  const int64_t end = min(A.range[1], begin + chunk);
  unsigned* myq = A.queue + qx;
  // self bit of the diagonal tile: lane's query row (lane & 31) among its columns
  const int srel = (lane & 31) - 4 * half;
  const uint32_t selfbit =
      (srel >= 0 && (srel & 4) == 0) ? 1u << (((srel >> 3) << 2) | (srel & 3)) : 0u;
  const double u = 5.9604644775390625e-08;  // 2^-24

  // metadata of a leaf's rows j = tid + h * kThr, loaded into registers
  int rid[kRowsPerThread];
  float rnb[kRowsPerThread], rer[kRowsPerThread], rbd[kRowsPerThread];
  auto load_meta = [&](const LeafItem& it) {
    const int64_t lofs = (int64_t)it.tree * n + it.start;
#pragma unroll
    for (int h = 0; h < kRowsPerThread; ++h) {
      const int j = tid + h * kThr;
      rid[h] = 0;
      rnb[h] = rer[h] = rbd[h] = 0.0f;
      if (j < it.size) {
        rid[h] = A.perm[lofs + j];
        if constexpr (!kI8) {
          rnb[h] = A.lnbT[lofs + j];
          rer[h] = A.lerrT[lofs + j];
        }
        if constexpr (kPool) rbd[h] = A.boundT[lofs + j];
      }
    }
  };
  // ids, -|x~|^2 / 2 and the leaf's maxima; then its rows (slot s = 16 r +
  // physical chunk, logical chunk = physical ^ (r & 15); one instruction per
  // wave covers 64 slots, 4 rows) and, while they arrive, the thresholds
  // thr = round_down((|x~|^2 - A_cut(B) - eta) / 2)
  auto install = [&](const LeafItem& it) {
    const int m = it.size;
    if constexpr (kI8) {
      // ids; the rows; then from each staged row n' = |x - 128|^2 (exact),
      // the accumulator start and the threshold
#pragma unroll
      for (int h = 0; h < kRowsPerThread; ++h) {
        const int j = tid + h * kThr;
        if (j < m) lid[j] = rid[h];
      }
      if (tid == 0) misc[kMiscTask] = 0;
      __syncthreads();  // ids visible
      for (int blk = wave; blk * kRowsPerDma < m; blk += kWv) {
        const int slot = blk * 64 + lane;
        const int r = slot / (kPitch / 16), c = (slot % (kPitch / 16)) ^ (r & (kPitch / 16 - 1));
        const uint32_t dst = __builtin_amdgcn_readfirstlane(
            (uint32_t)(uintptr_t)(lds_ptr_t)(rows + blk * 64 * 16));
        if (r < m && c < L) lds_dma16(Xq + (int64_t)lid[r] * rb + c * 16, dst);
      }
      wait_vmcnt<0>();
      __syncthreads();  // rows visible
#pragma unroll
      for (int h = 0; h < kRowsPerThread; ++h) {
        const int j = tid + h * kThr;
        if (j < m) {
          int nq = 0;
#pragma unroll
          for (int c = 0; c < L; ++c) nq += sq16(row_chunk<kPitch>(rows, j, c));
          if constexpr (kPool) {
            const float B = rbd[h];
            int t = -(1 << 30);
            if (B < INFINITY) {
              const double v = floor(((double)nq - (double)B) * 0.5);
              t = (int)fmin(fmax(v, -1073741824.0), 1073741824.0);
            }
            chalf[j] = -(nq >> 1);
            thrs[j] = t;
          } else {
            chalf[j] = -((nq + 1) >> 1);
            thrs[j] = nq;
          }
        }
      }
      __syncthreads();  // starts and thresholds visible to every wave
    } else {
      float mn = 0.0f, me = 0.0f;
#pragma unroll
      for (int h = 0; h < kRowsPerThread; ++h) {
        const int j = tid + h * kThr;
        if (j < m) {
          lid[j] = rid[h];
          chalf[j] = -0.5f * rnb[h];
          mn = fmaxf(mn, rnb[h]);
          me = fmaxf(me, rer[h]);
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        mn = fmaxf(mn, __shfl_xor(mn, off));
        me = fmaxf(me, __shfl_xor(me, off));
      }
      if (lane == 0) {
        red[wave] = mn;
        red[8 + wave] = me;
      }
      if (tid == 0) misc[kMiscTask] = 0;
      __syncthreads();  // ids and maxima visible
      for (int blk = wave; blk * 4 < m; blk += kWv) {
        const int slot = blk * 64 + lane;
        const int r = slot >> 4, c = (slot & 15) ^ (r & 15);
        const uint32_t dst = __builtin_amdgcn_readfirstlane(
            (uint32_t)(uintptr_t)(lds_ptr_t)(rows + blk * 64 * 16));
        if (r < m && c < L) lds_dma16(Xq + (int64_t)lid[r] * rb + c * 16, dst);
      }
      float nmax = red[0], emax = red[8];
#pragma unroll
      for (int w = 1; w < kWv; ++w) {
        nmax = fmaxf(nmax, red[w]);
        emax = fmaxf(emax, red[8 + w]);
      }
      const double nm = sqrt((double)nmax * (1.0 + (d + 4.0) * u));
#pragma unroll
      for (int h = 0; h < kRowsPerThread; ++h) {
        const int j = tid + h * kThr;
        if (!kPool && j < m) thrs[j] = rnb[h];
        if (kPool && j < m) {
          const float acut = acut_of(rbd[h], rnb[h], rer[h], nm, emax, d);
          const double ni = sqrt((double)rnb[h] * (1.0 + (d + 4.0) * u));
          const double eta = (2.0 * d + 16.0) * u * (ni + nm) * (ni + nm);
          const float t = __double2float_rd(0.5 * ((double)rnb[h] - (double)acut - eta));
          thrs[j] = t == 0.0f ? -0.0f : t;  // a zero difference from it is then +0
        }
      }
      wait_vmcnt<0>();
      __syncthreads();  // rows and thresholds visible to every wave
    }
  };

  // items fetched one leaf ahead: the next item's metadata loads are in
  // flight during the current leaf's tasks
  if (tid == 0) {
    fetch_item(A, myq, begin, end, misc, 0);
    fetch_item(A, myq, begin, end, misc, 1);
  }
  __syncthreads();
  if (misc[6] == 0) return;
  LeafItem it = *reinterpret_cast<const LeafItem*>(&misc[0]);
  load_meta(it);
  install(it);
  for (int s = 0;; s ^= 1) {
    // slot s ^ 1: the next item; slot s (this leaf's, read) takes the one after
    // it, fetched by thread 0 while the tasks run (it takes tasks after that)
    const bool more = misc[8 * (s ^ 1) + 6] != 0;
    const LeafItem nx = *reinterpret_cast<const LeafItem*>(&misc[8 * (s ^ 1)]);
    if (more) {
      load_meta(nx);
      if (tid == 0) fetch_item(A, myq, begin, end, misc, s);
    }
    const int m = it.size;
    // tasks: 64 query rows of the item x a range of 32-column tiles; the split
    // minimises rounds x (tiles per task + 1) over the block's waves
    const int row0 = it.row0, row_end = it.row0 + it.rows;
    const int nq = (it.rows + 63) >> 6;
    const int ntile = (m + 31) >> 5;
    int nc = 1, best = INT_MAX;
    // candidates: one or two column ranges (each keeps TK keys per lane)
    const int ncmax = MODE == kModeCand ? 2 : 4 * kWv;
    for (int c = 1; c <= ntile && c <= ncmax; ++c) {
      const int tp = (ntile + c - 1) / c;
      const int cost = ((nq * c + kWv - 1) / kWv) * (tp + 1);
      if (cost < best) {
        best = cost;
        nc = c;
      }
    }
    const int tper = (ntile + nc - 1) / nc;
    nc = (ntile + tper - 1) / tper;  // (candidates: 2 stays 2, as ntile >= 2 there)
    const int ntask = nq * nc;
    // pool mode: the filter mask of the leaf's last tile (columns past m
    // cleared), for a branch-free epilogue
    uint32_t lastmask = 0u;
    if constexpr (kPool) {
      const int cbl = 32 * (ntile - 1) + 4 * half;
#pragma unroll
      for (int e = 0; e < 16; ++e) lastmask |= (cbl + 8 * (e >> 2) + (e & 3) < m ? 1u : 0u) << e;
    }
    for (;;) {
      int task = 0;
      if (lane == 0) task = atomicAdd(&misc[kMiscTask], 1);
      task = __builtin_amdgcn_readfirstlane(__shfl(task, 0));
      if (task >= ntask) break;
      const int qi = task / nc, cj = task - qi * nc;
      const int qb = row0 + 64 * qi;
      const int t0 = cj * tper, t1 = min(ntile, t0 + tper);
      uint4 qop[2][KS];
      thr_t thr[2];
      int id[2];
      PosList lst[2];
      uint32_t key[2][MODE == kModeCand ? TK : 1];  // candidates: ascending packed keys
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const int r = qb + 32 * g + (lane & 31);
        const bool valid = r < row_end;
        const int rr = valid ? r : row0;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) qop[g][ks] = row_chunk<kPitch>(rows, rr, 2 * ks + half);
        thr[g] = thrs[rr];
        id[g] = lid[rr];
        if (!valid) {  // and its bits are cleared below
          if constexpr (kI8) thr[g] = 1 << 30;
          else thr[g] = INFINITY;
        }
#pragma unroll
        for (int s2 = 0; s2 < (MODE == kModeCand ? TK : 1); ++s2) key[g][s2] = 0xFFFFFFFFu;
      }
      lst[0].reset();
      lst[1].reset();
      const bool valid0 = qb + (lane & 31) < row_end;
      const bool valid1 = qb + 32 + (lane & 31) < row_end;
      auto mma = [](uint4 a, uint4 b, const acc_t& c) -> acc_t {
        if constexpr (kI8)
          return __builtin_amdgcn_mfma_i32_32x32x32_i8(__builtin_bit_cast(intx4, a),
                                                       __builtin_bit_cast(intx4, b), c, 0, 0, 0);
        else
          return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                         __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
      };
      auto tile = [&](int t, acc_t& a0, acc_t& a1) {
        const int c0 = 32 * t;
        const int r = c0 + (lane & 31);
        uint4 av[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) av[ks] = row_chunk<kPitch>(rows, r, 2 * ks + half);
        acc_t ci;
        const int cb = c0 + 4 * half;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const uint4 v = *reinterpret_cast<const uint4*>(&chalf[cb + 8 * g]);
          ci[4 * g] = __builtin_bit_cast(thr_t, v.x);
          ci[4 * g + 1] = __builtin_bit_cast(thr_t, v.y);
          ci[4 * g + 2] = __builtin_bit_cast(thr_t, v.z);
          ci[4 * g + 3] = __builtin_bit_cast(thr_t, v.w);
        }
        a0 = mma(av[0], qop[0][0], ci);
        a1 = mma(av[0], qop[1][0], ci);
#pragma unroll
        for (int ks = 1; ks < KS; ++ks) {
          a0 = mma(av[ks], qop[0][ks], a0);
          a1 = mma(av[ks], qop[1][ks], a1);
        }
      };
      auto epilogue = [&](int t, const acc_t& a0, const acc_t& a1) {
        const int c0 = 32 * t;
        const int cb = c0 + 4 * half;
        uint32_t colmask = 0xFFFFu;
        if (c0 + 32 > m) {  // last tile: columns past the leaf
          colmask = 0u;
#pragma unroll
          for (int e = 0; e < 16; ++e)
            colmask |= (cb + 8 * (e >> 2) + (e & 3) < m ? 1u : 0u) << e;
        }
        if constexpr (MODE == kModeCand) {
          // key = approximate distance a = |x~_i|^2 - 2 acc (>= 0) with its low
          // 11 bits replaced by the leaf position; of each 4 adjacent columns
          // the smallest key enters the lane's TK smallest (branch-free
          // insertion: the candidates only seed the bounds, which any distinct
          // leafmates give)
#pragma unroll
          for (int g = 0; g < 2; ++g) {
            const acc_t& ac = g == 0 ? a0 : a1;
            uint32_t vm = (g == 0 ? valid0 : valid1) ? colmask : 0u;
            if (c0 == qb + 32 * g) vm &= ~selfbit;
#pragma unroll
            for (int q4 = 0; q4 < 4; ++q4) {
              uint32_t kk = 0xFFFFFFFFu;
#pragma unroll
              for (int u4 = 0; u4 < 4; ++u4) {
                const int e = 4 * q4 + u4;
                float a;
                if constexpr (kI8) a = (float)(thr[g] - 2 * ac[e]);  // >= D >= 0
                else a = fmaxf(fmaf(-2.0f, ac[e], thr[g]), 0.0f);
                const uint32_t ke = (vm >> e) & 1u
                                        ? (__float_as_uint(a) & 0xFFFFF800u) | (uint32_t)(cb + 8 * q4 + u4)
                                        : 0xFFFFFFFFu;
                kk = min(kk, ke);
              }
#pragma unroll
              for (int s2 = 0; s2 < TK; ++s2) {
                const uint32_t lo = min(kk, key[g][s2]);
                kk = max(kk, key[g][s2]);
                key[g][s2] = lo;
              }
            }
          }
          return;
        }
        uint32_t pm0 = valid0 ? pass_bits(a0, thr[0]) & colmask : 0u;
        uint32_t pm1 = valid1 ? pass_bits(a1, thr[1]) & colmask : 0u;
        if (c0 == qb) pm0 &= ~selfbit;
        if (c0 == qb + 32) pm1 &= ~selfbit;
        while (pm0 != 0u) {
          const int e = __builtin_ctz(pm0);
          pm0 &= pm0 - 1u;
          lst[0].push((uint32_t)(cb + 8 * (e >> 2) + (e & 3)));
          if (lst[0].cnt == kPlList) flush_list(A, lid, id[0], lst[0]);
        }
        while (pm1 != 0u) {
          const int e = __builtin_ctz(pm1);
          pm1 &= pm1 - 1u;
          lst[1].push((uint32_t)(cb + 8 * (e >> 2) + (e & 3)));
          if (lst[1].cnt == kPlList) flush_list(A, lid, id[1], lst[1]);
        }
      };
      // two accumulator sets: the next tile's MFMAs are in flight during this
      // tile's epilogue
      acc_t a0, a1, b0, b1;
      tile(t0, a0, a1);
      if constexpr (kPool) {
        const uint32_t vm0 = valid0 ? 0xFFFFu : 0u, vm1 = valid1 ? 0xFFFFu : 0u;
        // filter masks of tile t (branch-free, so that they share a basic
        // block with the next tile's MFMAs and the scheduler interleaves them)
        auto filter = [&](int t, const acc_t& x0, const acc_t& x1, uint32_t& pm0,
                          uint32_t& pm1) {
          const int c0 = 32 * t;
          const uint32_t cm = c0 + 32 > m ? lastmask : 0xFFFFu;
          pm0 = pass_bits(x0, thr[0]) & cm & vm0 & (c0 == qb ? ~selfbit : 0xFFFFFFFFu);
          pm1 = pass_bits(x1, thr[1]) & cm & vm1 & (c0 == qb + 32 ? ~selfbit : 0xFFFFFFFFu);
        };
        auto survivors = [&](int t, uint32_t pm0, uint32_t pm1) {
          const int pcb = 32 * t + 4 * half;
          while (pm0 != 0u) {
            const int e = __builtin_ctz(pm0);
            pm0 &= pm0 - 1u;
            lst[0].push((uint32_t)(pcb + 8 * (e >> 2) + (e & 3)));
            if (lst[0].cnt == kPlList) flush_list(A, lid, id[0], lst[0]);
          }
          while (pm1 != 0u) {
            const int e = __builtin_ctz(pm1);
            pm1 &= pm1 - 1u;
            lst[1].push((uint32_t)(pcb + 8 * (e >> 2) + (e & 3)));
            if (lst[1].cnt == kPlList) flush_list(A, lid, id[1], lst[1]);
          }
        };
        // tile pairs while two more tiles follow (the next tile's MFMAs and
        // this tile's filter in one basic block, so that they interleave),
        // then the last one or two
        int t = t0;
        for (; t + 2 < t1; t += 2) {  // a0/a1: tile t, in flight
          uint32_t p0, p1;
          tile(t + 1, b0, b1);
          filter(t, a0, a1, p0, p1);
          survivors(t, p0, p1);
          tile(t + 2, a0, a1);
          filter(t + 1, b0, b1, p0, p1);
          survivors(t + 1, p0, p1);
        }
        uint32_t p0, p1;
        if (t + 1 < t1) {
          tile(t + 1, b0, b1);
          filter(t, a0, a1, p0, p1);
          survivors(t, p0, p1);
          filter(t + 1, b0, b1, p0, p1);
          survivors(t + 1, p0, p1);
        } else {
          filter(t, a0, a1, p0, p1);
          survivors(t, p0, p1);
        }
      } else {
        for (int t = t0; t < t1; t += 2) {
          const bool has1 = t + 1 < t1;
          if (has1) tile(t + 1, b0, b1);
          epilogue(t, a0, a1);
          if (!has1) break;
          if (t + 2 < t1) tile(t + 2, a0, a1);
          epilogue(t + 1, b0, b1);
        }
      }
      if constexpr (kPool) {
        flush_pair(A, lid, id, lst);
      } else {
        // stream (column range cj, half) of a row: key slots (2 cj + half) TK
        // of its 4 TK; with one column range the other two streams are empty
#pragma unroll
        for (int g = 0; g < 2; ++g) {
          const int r = qb + 32 * g + (lane & 31);
          if (r < row_end) {
            uint64_t* o = A.out_key + ((int64_t)(it.tree - A.tree0) * n + id[g]) * (4 * TK);
#pragma unroll
            for (int s2 = 0; s2 < TK; ++s2) {
              const uint32_t k = key[g][s2];
              o[(2 * cj + half) * TK + s2] =
                  k == 0xFFFFFFFFu ? ~0ull
                                   : ((uint64_t)(k & 0xFFFFF800u) << 32) | (uint32_t)lid[k & 0x7FFu];
              if (nc == 1) o[(2 + half) * TK + s2] = ~0ull;
            }
          }
        }
      }
    }
    __syncthreads();  // every wave done with the leaf's tasks; slot s written
    if (!more) break;
    it = nx;
    install(it);
  }
}

template <int MODE, int TK>
const void* pool_leaf_fn(int ks) {
  switch (ks) {
    case 1: return (const void*)k_pool_leaf<1, MODE, TK, false>;
    case 2: return (const void*)k_pool_leaf<2, MODE, TK, false>;
    case 3: return (const void*)k_pool_leaf<3, MODE, TK, false>;
    case 4: return (const void*)k_pool_leaf<4, MODE, TK, false>;
    case 5: return (const void*)k_pool_leaf<5, MODE, TK, false>;
    case 6: return (const void*)k_pool_leaf<6, MODE, TK, false>;
    case 7: return (const void*)k_pool_leaf<7, MODE, TK, false>;
    default: return (const void*)k_pool_leaf<8, MODE, TK, false>;
  }
}

template <int MODE, int TK>
const void* pool_leaf_fn8(int ks) {
  switch (ks) {
    case 1: return (const void*)k_pool_leaf<1, MODE, TK, true>;
    case 2: return (const void*)k_pool_leaf<2, MODE, TK, true>;
    case 3: return (const void*)k_pool_leaf<3, MODE, TK, true>;
    default: return (const void*)k_pool_leaf<4, MODE, TK, true>;
  }
}

template <bool kI8>
const void* pool_leaf_instance(int mode, int tk, int ks) {
  if (mode == kModePool)
    return kI8 ? pool_leaf_fn8<kModePool, 1>(ks) : pool_leaf_fn<kModePool, 1>(ks);
  if (tk == 3) return kI8 ? pool_leaf_fn8<kModeCand, 3>(ks) : pool_leaf_fn<kModeCand, 3>(ks);
  if (tk == 4) return kI8 ? pool_leaf_fn8<kModeCand, 4>(ks) : pool_leaf_fn<kModeCand, 4>(ks);
  return kI8 ? pool_leaf_fn8<kModeCand, 6>(ks) : pool_leaf_fn<kModeCand, 6>(ks);
}

// one launch of instance fn over the items of b with sizes [b.size_lo,
// b.size_hi], mmax staged rows, `lds` bytes of dynamic LDS, `per_cu` blocks per CU
int launch_pool_leaf_one(drpt_ctx* ctx, const void* fn, KnnArgs b, int mmax, size_t lds, int per_cu,
                         int threads) {
  auto it = ctx->fn_lds.find(fn);
  if (it == ctx->fn_lds.end() || it->second < (int)lds) {
    DRPT_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    ctx->fn_lds[fn] = (int)lds;
  }
  unsigned* q;
  DRPT_CHECK(ws(ctx, "knn.queue_pl", kPlQueues, &q));
  DRPT_CHECK(fill_u32(ctx, q, 0u, kPlQueues));
  b.queue = q;
  void* args[] = {(void*)&b, (void*)&mmax};
  DRPT_HIP(hipLaunchKernel(fn, dim3((unsigned)(ctx->num_cus * per_cu)), dim3(threads), args, lds,
                           ctx->stream));
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

}  // namespace

// Largest leaf (rows, a multiple of 32: the tiles read whole 32-row groups)
// the LDS-resident pool pass stages at dimension d on this device, or 0 when
// it does not apply (BF16 rows of 32 to 256 bytes in 32-byte multiples).
int pool_leaf_rows(drpt_ctx* ctx, int d, int mode) {
  if (d < 16 || d > 16 * kPlSteps || (d & 15) != 0) return 0;
  const int64_t room = (int64_t)ctx->max_lds - kPlMisc * 4;
  const int64_t rows = (room / (int64_t)pool_leaf_row_bytes(mode)) & ~(int64_t)31;
  return rows >= 64 ? (int)std::min<int64_t>(rows, 2 * kPlThreads) : 0;
}

bool pool_leaf_int8(int d) { return d >= 32 && (d & 31) == 0 && d <= 32 * kPlSteps8; }

int launch_pool_leaf(drpt_ctx* ctx, int mode, int kp, const KnnArgs& a0) {
  if (mode != kModePool && mode != kModeCand) {
    set_error("internal: pool_leaf mode");
    return DRPT_ERR_ARG;
  }
  const int mmax = pool_leaf_rows(ctx, a0.d, mode);
  if (mmax <= 0 || a0.size_hi > mmax) {
    set_error("internal: pool_leaf launched for leaves it cannot stage");
    return DRPT_ERR_ARG;
  }
  int tk = 1;
  if (mode == kModeCand) {
    tk = kp / 4;
    if (kp != 4 * tk || (tk != 3 && tk != 4 && tk != 6)) {
      set_error("internal: pool_leaf candidate mode needs k' = 12, 16 or 24");
      return DRPT_ERR_ARG;
    }
  }
  // the int8 instances where the rows have an int8 copy: issued after the
  // bf16 one, the device flag picks which runs
  const bool i8 = a0.Xi8 != nullptr && a0.notu8 != nullptr && pool_leaf_int8(a0.d);
  KnnArgs a = a0;
  if (!i8) a.notu8 = nullptr;
  const size_t lds = (size_t)mmax * pool_leaf_row_bytes(mode) + kPlMisc * 4;
  DRPT_CHECK(launch_pool_leaf_one(ctx, pool_leaf_instance<false>(mode, tk, a0.d / 16), a, mmax, lds, 1,
                                  kPlThreads));
  if (!i8) return DRPT_OK;
  const void* fn8 = pool_leaf_instance<true>(mode, tk, a0.d / 32);
  // two blocks per CU for the leaves whose rows fit half the LDS, one for the rest
  const int m2 = (int)(((int64_t)ctx->max_lds / 2 - kPlMisc * 4) / (int64_t)pool_leaf_row_bytes8()) & ~31;
  if (m2 >= 64 && a0.size_lo <= m2) {
    KnnArgs b = a;
    b.size_hi = std::min(a0.size_hi, m2);
    DRPT_CHECK(launch_pool_leaf_one(ctx, fn8, b, m2, (size_t)m2 * pool_leaf_row_bytes8() + kPlMisc * 4, 2,
                                    kPlThreads8));
    if (a0.size_hi <= m2) return DRPT_OK;
    a.size_lo = m2 + 1;
  }
  if (a0.item_cap > 0) {
    // the few larger leaves: their items gathered first, so that the launch
    // does not walk (and skip) every other item
    LeafItem* big;
    int64_t* big_range;
    DRPT_CHECK(ws(ctx, "knn.items_big8", (size_t)a0.item_cap, &big));
    DRPT_CHECK(ws(ctx, "knn.range_big8", 2, &big_range));
    DRPT_CHECK(launch_select_items(ctx, a0.items, a0.range, a.size_lo, a.size_hi, a0.item_cap, big,
                                   big_range));
    a.items = big;
    a.range = big_range;
  }
  return launch_pool_leaf_one(ctx, fn8, a, mmax, (size_t)mmax * pool_leaf_row_bytes8() + kPlMisc * 4, 1,
                              kPlThreads8);
}

}  // namespace drpt
