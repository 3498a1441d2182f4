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

__device__ __forceinline__ int next_item(unsigned* queue, int* s_item) {
  if (threadIdx.x == 0) *s_item = (int)atomicAdd(queue, 1u);
  __syncthreads();
  int item = *s_item;
  __syncthreads();
  return item;
}

// ---------------------------------------------------------------------------
// FP32 exact path.
// ---------------------------------------------------------------------------
template <typename XT, int KP>
__global__ void __launch_bounds__(kThreads)
k_knn_exact(const XT* __restrict__ X, int d, const int32_t* __restrict__ perm, int64_t n,
            const LeafItem* __restrict__ items, int64_t nitems, int K,
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
  TopK<KP> top;
  for (;;) {
    const int64_t item = next_item(queue, &s_item);
    if (item >= nitems) break;
    const LeafItem it = items[item];
    const int32_t* pl = perm + (int64_t)it.tree * n + it.start;
    const int m = it.size;
    const int rows = min(kRows, m - it.row0);
    if (tid < kRows) rowid[tid] = tid < rows ? pl[it.row0 + tid] : -1;
    top.init();
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
          const int ida = rowid[r], idb = colid[r];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int k = k0 + kb + u;
            As[kb + u][r] = (ida >= 0 && k < d) ? (float)X[(int64_t)ida * d + k] : 0.0f;
            Bs[kb + u][r] = (idb >= 0 && k < d) ? (float)X[(int64_t)idb * d + k] : 0.0f;
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
      if (tid < rows) {
        const int me = rowid[tid];
        for (int c = 0; c < cols; ++c) {
          const int id = colid[c];
          if (id != me) top.insert(Ds[tid][c], id);
        }
      }
      __syncthreads();
    }
    if (tid < rows) {
      const int me = rowid[tid];
      const int64_t o = ((int64_t)it.tree * n + me) * K;
      top.store(K, cand_idx + o, cand_dist + o);
    }
  }
}

// ---------------------------------------------------------------------------
// BF16 candidate path (MFMA) + exact re-rank + certification.
// ---------------------------------------------------------------------------
template <typename XT, int KP>
__global__ void __launch_bounds__(kThreads)
k_knn_bf16(const XT* __restrict__ X, const uint16_t* __restrict__ Xb,
           const float* __restrict__ nb, const float* __restrict__ err, int d,
           const int32_t* __restrict__ perm, int64_t n, const LeafItem* __restrict__ items,
           int64_t nitems, int K, int kp, int32_t* __restrict__ cand_idx,
           float* __restrict__ cand_dist, unsigned* __restrict__ queue,
           int32_t* __restrict__ fallback, int32_t* __restrict__ nfallback, int64_t fb_cap) {
  __shared__ __align__(16) uint16_t As[kRows][72];
  __shared__ __align__(16) uint16_t Bs[kCols][72];
  __shared__ float Ds[kRows][kCols + 1];
  __shared__ int rowid[kRows];
  __shared__ int colid[kCols];
  __shared__ float rown[kRows];
  __shared__ float coln[kCols];
  __shared__ int cid[kRows][KP];
  __shared__ float cdx[kRows][KP];
  __shared__ float athr[kRows];
  __shared__ float red_n[kThreads / 64], red_e[kThreads / 64];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  TopK<KP> top;
  for (;;) {
    const int64_t item = next_item(queue, &s_item);
    if (item >= nitems) break;
    const LeafItem it = items[item];
    const int32_t* pl = perm + (int64_t)it.tree * n + it.start;
    const int m = it.size;
    const int rows = min(kRows, m - it.row0);
    // leaf maxima of |x~|^2 and |x~ - x| for the certification bound
    float mn = 0.0f, me_ = 0.0f;
    for (int j = tid; j < m; j += kThreads) {
      const int id = pl[j];
      mn = fmaxf(mn, nb[id]);
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
    if (tid < kRows) {
      const int id = tid < rows ? pl[it.row0 + tid] : -1;
      rowid[tid] = id;
      rown[tid] = id >= 0 ? nb[id] : 0.0f;
    }
    top.init();
    for (int c0 = 0; c0 < m; c0 += kCols) {
      const int cols = min(kCols, m - c0);
      if (tid >= 64 && tid < 128) {
        const int c = tid - 64;
        const int id = c < cols ? pl[c0 + c] : -1;
        colid[c] = id;
        coln[c] = id >= 0 ? nb[id] : 0.0f;
      }
      __syncthreads();
      floatx16 acc;
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
      for (int k0 = 0; k0 < d; k0 += 64) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int c = tid + kThreads * s;
          const int r = c >> 3, part = c & 7;
          const int k = k0 + part * 8;
          const int ida = rowid[r], idb = colid[r];
          uint4 va = make_uint4(0, 0, 0, 0), vb = make_uint4(0, 0, 0, 0);
          if ((d & 7) == 0 && k + 8 <= d) {  // 16-byte aligned rows only
            if (ida >= 0) va = *reinterpret_cast<const uint4*>(Xb + (int64_t)ida * d + k);
            if (idb >= 0) vb = *reinterpret_cast<const uint4*>(Xb + (int64_t)idb * d + k);
          } else if (k < d) {
            uint16_t ta[8], tb[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              ta[u] = (ida >= 0 && k + u < d) ? Xb[(int64_t)ida * d + k + u] : (uint16_t)0;
              tb[u] = (idb >= 0 && k + u < d) ? Xb[(int64_t)idb * d + k + u] : (uint16_t)0;
            }
            va = make_uint4(ta[0] | (ta[1] << 16), ta[2] | (ta[3] << 16), ta[4] | (ta[5] << 16),
                            ta[6] | (ta[7] << 16));
            vb = make_uint4(tb[0] | (tb[1] << 16), tb[2] | (tb[3] << 16), tb[4] | (tb[5] << 16),
                            tb[6] | (tb[7] << 16));
          }
          *reinterpret_cast<uint4*>(&As[r][part * 8]) = va;
          *reinterpret_cast<uint4*>(&Bs[r][part * 8]) = vb;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wr + (lane & 31)][ks * 16 + 8 * (lane >> 5)]);
          const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wc + (lane & 31)][ks * 16 + 8 * (lane >> 5)]);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
      }
      // 32x32 accumulator layout: element e of lane l is
      // row 8*(e/4) + 4*(l/32) + e%4, column l%32.
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int r = wr + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        const int c = wc + (lane & 31);
        const float s = __fadd_rn(rown[r], coln[c]);
        Ds[r][c] = __fsub_rn(s, __fmul_rn(2.0f, acc[e]));
      }
      __syncthreads();
      if (tid < rows) {
        const int me = rowid[tid];
        for (int c = 0; c < cols; ++c) {
          const int id = colid[c];
          if (id != me) top.insert(Ds[tid][c], id);
        }
      }
      __syncthreads();
    }
    // candidates and the (k'+1)-th approximate value
    if (tid < rows) {
      float thr = INFINITY;
#pragma unroll
      for (int s = 0; s < KP; ++s) {
        if (s < kp) cid[tid][s] = (top.i[s] == INT_MAX) ? -1 : top.i[s];
        if (s == kp) thr = (top.i[s] == INT_MAX) ? INFINITY : top.d[s];
      }
      athr[tid] = thr;
    }
    __syncthreads();
    // exact re-rank of the k' candidates of every row
    for (int pr = tid; pr < rows * kp; pr += kThreads) {
      const int r = pr / kp, c = pr - r * kp;
      const int id = cid[r][c];
      cdx[r][c] = id >= 0 ? exact_dist(X, rowid[r], id, d) : INFINITY;
    }
    __syncthreads();
    if (tid < rows) {
      top.init();
      for (int c = 0; c < kp; ++c) {
        const int id = cid[tid][c];
        if (id >= 0) top.insert(cdx[tid][c], id);
      }
      const int me = rowid[tid];
      const int64_t o = ((int64_t)it.tree * n + me) * K;
      top.store(K, cand_idx + o, cand_dist + o);
      const float thr = athr[tid];
      if (thr < INFINITY) {
        const double u = 5.9604644775390625e-08;  // 2^-24
        float nmax = red_n[0], emax = red_e[0];
#pragma unroll
        for (int w = 1; w < kThreads / 64; ++w) {
          nmax = fmaxf(nmax, red_n[w]);
          emax = fmaxf(emax, red_e[w]);
        }
        const double infl = 1.0 + (d + 4.0) * u;
        const double ni = sqrt((double)rown[tid] * infl);
        const double nm = sqrt((double)nmax * infl);
        const double eta = (2.0 * d + 16.0) * u * (ni + nm) * (ni + nm);
        const double rt = sqrt(fmax(0.0, (double)thr - eta));
        const double rl = fmax(0.0, rt - ((double)err[me] + (double)emax));
        const double lb = rl * rl * (1.0 - (d + 8.0) * u);
        const double dk = (double)top.dist_at(K - 1);
        if (!(dk < lb)) {
          const int slot = atomicAdd(nfallback, 1);
          if (slot < fb_cap) {
            fallback[2 * slot] = (int32_t)item;
            fallback[2 * slot + 1] = tid;
          }
        }
      }
    }
    __syncthreads();
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
int occupancy_grid(drpt_ctx* ctx, T kernel, int64_t nitems, unsigned* grid) {
  int per_cu = 0;
  DRPT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0));
  if (per_cu < 1) per_cu = 1;
  int64_t g = (int64_t)ctx->num_cus * per_cu;
  if (g > nitems) g = nitems;
  if (g < 1) g = 1;
  *grid = (unsigned)g;
  return DRPT_OK;
}

template <typename XT, int KP>
int run_exact(drpt_ctx* ctx, const XT* X, int d, const int32_t* perm, int64_t n,
              const LeafItem* items, int64_t nitems, int K, int32_t* ci, float* cd, unsigned* q) {
  unsigned grid;
  DRPT_CHECK(occupancy_grid(ctx, k_knn_exact<XT, KP>, nitems, &grid));
  hipLaunchKernelGGL((k_knn_exact<XT, KP>), dim3(grid), dim3(kThreads), 0, ctx->stream, X, d, perm,
                     n, items, nitems, K, ci, cd, q);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

template <typename XT, int KP>
int run_bf16(drpt_ctx* ctx, const XT* X, const uint16_t* Xb, const float* nb, const float* err,
             int d, const int32_t* perm, int64_t n, const LeafItem* items, int64_t nitems, int K,
             int kp, int32_t* ci, float* cd, unsigned* q, int32_t* fb, int32_t* nfb, int64_t cap) {
  unsigned grid;
  DRPT_CHECK(occupancy_grid(ctx, k_knn_bf16<XT, KP>, nitems, &grid));
  hipLaunchKernelGGL((k_knn_bf16<XT, KP>), dim3(grid), dim3(kThreads), 0, ctx->stream, X, Xb, nb,
                     err, d, perm, n, items, nitems, K, kp, ci, cd, q, fb, nfb, cap);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

template <typename XT>
int dispatch(drpt_ctx* ctx, const drpt_params* p, const XT* X, const int32_t* perm,
             const LeafItem* items, int64_t nitems, const void* Xq, const float* nrm2,
             const float* err, int32_t* ci, float* cd, unsigned* q, int32_t* fb, int32_t* nfb,
             int64_t cap) {
  const int K = p->nn, d = p->d;
  const int64_t n = p->n;
  if (p->precision == DRPT_FP32) {
    if (K <= 16) return run_exact<XT, 16>(ctx, X, d, perm, n, items, nitems, K, ci, cd, q);
    if (K <= 32) return run_exact<XT, 32>(ctx, X, d, perm, n, items, nitems, K, ci, cd, q);
    if (K <= 64) return run_exact<XT, 64>(ctx, X, d, perm, n, items, nitems, K, ci, cd, q);
    set_error("leaf_knn: nn > 64 is not supported");
    return DRPT_ERR_UNSUPPORTED;
  }
  if (p->precision == DRPT_BF16) {
    const int kp = p->kprime;
    const uint16_t* Xb = (const uint16_t*)Xq;
    if (kp + 1 <= 32)
      return run_bf16<XT, 32>(ctx, X, Xb, nrm2, err, d, perm, n, items, nitems, K, kp, ci, cd, q,
                              fb, nfb, cap);
    if (kp + 1 <= 64)
      return run_bf16<XT, 64>(ctx, X, Xb, nrm2, err, d, perm, n, items, nitems, K, kp, ci, cd, q,
                              fb, nfb, cap);
    set_error("leaf_knn: kprime > 63 is not supported for BF16");
    return DRPT_ERR_UNSUPPORTED;
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
                    const LeafItem* items, int64_t nitems, const void* Xq, const float* nrm2,
                    const float* err, const float* /*scale*/, int32_t* cand_idx, float* cand_dist,
                    int32_t* fallback, int32_t* nfallback, int64_t fallback_cap) {
  if (nitems <= 0) return DRPT_OK;
  unsigned* q = nullptr;
  DRPT_CHECK(ws(ctx, "knn.queue", 1, &q));
  DRPT_CHECK(fill_u32(ctx, q, 0u, 1));
  DRPT_CHECK(fill_u32(ctx, (uint32_t*)nfallback, 0u, 1));
  if (p->dtype == DRPT_U8)
    return dispatch<uint8_t>(ctx, p, (const uint8_t*)X, perm, items, nitems, Xq, nrm2, err,
                             cand_idx, cand_dist, q, fallback, nfallback, fallback_cap);
  return dispatch<float>(ctx, p, (const float*)X, perm, items, nitems, Xq, nrm2, err, cand_idx,
                         cand_dist, q, fallback, nfallback, fallback_cap);
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
