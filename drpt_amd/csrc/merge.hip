// Stage (d): cross-tree merge and dedup.
//
// Replaces the per-tree std::merge + adjacent std::unique of
// MDRPT::calculate_nns (cpp/src/drpt/algo/mdrpt.cpp:180-198) and the one-rank
// pass-through of communicate_nns (mdrpt.cpp:625-630). Every per-tree list is
// sorted by (D, index); a pair always carries the same D, so a candidate equal
// in index to a kept one is a duplicate and is dropped. One thread per point
// keeps a register list of the best KP >= K entries; each tree's list is read
// until its head can no longer enter the list.
//
// Roofline: HBM-bound; algorithmic bytes per point = 8*K*T read + 8*K written.
#include <climits>

#include "internal.hpp"

namespace drpt {

namespace {

__device__ __forceinline__ bool better(float d1, int i1, float d2, int i2) {
  return d1 < d2 || (d1 == d2 && i1 < i2);
}

template <int KP>
__global__ void __launch_bounds__(256)
k_merge(int64_t n, int T, int K, const int32_t* __restrict__ ci, const float* __restrict__ cd,
        int32_t* __restrict__ oi, float* __restrict__ od) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float d[KP];
  int ix[KP];
#pragma unroll
  for (int s = 0; s < KP; ++s) {
    d[s] = INFINITY;
    ix[s] = INT_MAX;
  }
  for (int t = 0; t < T; ++t) {
    const int64_t o = ((int64_t)t * n + p) * K;
    for (int k = 0; k < K; ++k) {
      const int iv = ci[o + k];
      if (iv < 0) break;
      const float dv = cd[o + k];
      if (!better(dv, iv, d[KP - 1], ix[KP - 1])) break;  // the rest of this list is worse
      bool dup = false;
#pragma unroll
      for (int s = 0; s < KP; ++s) dup |= (ix[s] == iv);
      if (dup) continue;
      d[KP - 1] = dv;
      ix[KP - 1] = iv;
#pragma unroll
      for (int s = KP - 1; s > 0; --s) {
        if (better(d[s], ix[s], d[s - 1], ix[s - 1])) {
          float td = d[s];
          d[s] = d[s - 1];
          d[s - 1] = td;
          int ti = ix[s];
          ix[s] = ix[s - 1];
          ix[s - 1] = ti;
        } else {
          break;
        }
      }
    }
  }
  const int64_t q = p * K;
#pragma unroll
  for (int s = 0; s < KP; ++s) {
    if (s < K) {
      oi[q + s] = ix[s] == INT_MAX ? -1 : ix[s];
      od[q + s] = d[s];
    }
  }
}

}  // namespace

int launch_merge(drpt_ctx* ctx, int64_t n, int T, int K, const int32_t* cand_idx,
                 const float* cand_dist, int32_t* out_idx, float* out_dist) {
  if (n <= 0) return DRPT_OK;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (K <= 16)
    hipLaunchKernelGGL(k_merge<16>, dim3(blocks), dim3(256), 0, ctx->stream, n, T, K, cand_idx,
                       cand_dist, out_idx, out_dist);
  else if (K <= 32)
    hipLaunchKernelGGL(k_merge<32>, dim3(blocks), dim3(256), 0, ctx->stream, n, T, K, cand_idx,
                       cand_dist, out_idx, out_dist);
  else if (K <= 64)
    hipLaunchKernelGGL(k_merge<64>, dim3(blocks), dim3(256), 0, ctx->stream, n, T, K, cand_idx,
                       cand_dist, out_idx, out_dist);
  else {
    set_error("merge: nn > 64 is not supported");
    return DRPT_ERR_UNSUPPORTED;
  }
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

}  // namespace drpt
