// Stage (a): sparse random projection P = X.R on gfx950.
//
// Replaces MathOp::multiply_mat (cpp/src/drpt/math/math_operations.cpp:15-46),
// a dense cblas_sgemm over a matrix R that is ~1/sqrt(d) dense. Here R is
// held as compressed columns (non-zeros of each column, ascending row), X rows
// are read once, coalesced, into an LDS tile, and every output is the
// double-precision sum of x_ij * r_jc over column c's non-zeros in ascending j,
// rounded once to float: a fixed order, so P is bit-reproducible (products of
// u8/f32 by f32 are exact in double, so FMA contraction cannot change it).
//
// Layout: X row-major [n][d] (u8 or f32); P column-major [L][n], so the tree
// split of level l of tree t reads one contiguous column; colmax[c] = max bits
// of |P[c][.]| (non-negative floats order like their bit patterns).
//
// Roofline: HBM-bound. Algorithmic bytes = n*d*sizeof(x) + 4*n*L.
#include <algorithm>

#include "internal.hpp"

namespace drpt {

namespace {

constexpr int kThreads = 256;

// bytes of the LDS header before the X tile (16-byte aligned): the column
// maxima, plus R's compressed columns when they are staged (RL)
__host__ __device__ inline size_t r_bytes(int L, int nnz, bool RL) {
  return ((size_t)L * 4 + (RL ? (size_t)(L + 1) * 4 + (size_t)nnz * 6 : 0) + 15) & ~(size_t)15;
}

// TR rows per tile; lane -> (row = lane % TR, column slot = lane / TR).
// Grid-stride over row tiles; the per-column max |P| is kept in LDS and
// flushed once per block. RL: R staged in LDS (else read through the caches:
// every block reads the same few KB).
template <typename XT, int TR, bool RL>
__global__ void __launch_bounds__(kThreads)
k_project(const XT* __restrict__ X, int64_t n, int d, int L, const int32_t* __restrict__ colptr,
          const uint16_t* __restrict__ rowidx, const float* __restrict__ w, float* __restrict__ P,
          uint32_t* __restrict__ colmax, int pitch) {
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS: column maxima [L], R's compressed columns (colptr [L+1], weights, row
  // indices) when RL, then the X tile
  const int nnz = colptr[L];
  uint32_t* lmax = reinterpret_cast<uint32_t*>(smem);
  const int32_t* lcp = colptr;
  const uint16_t* lri = rowidx;
  const float* lw = w;
  if constexpr (RL) {
    int32_t* scp = reinterpret_cast<int32_t*>(lmax + L);
    float* sw = reinterpret_cast<float*>(scp + L + 1);
    uint16_t* sri = reinterpret_cast<uint16_t*>(sw + nnz);
    for (int c = threadIdx.x; c <= L; c += kThreads) scp[c] = colptr[c];
    for (int k = threadIdx.x; k < nnz; k += kThreads) {
      sw[k] = w[k];
      sri[k] = rowidx[k];
    }
    lcp = scp;
    lw = sw;
    lri = sri;
  }
  auto wv = [&](int k) -> double { return (double)lw[k]; };
  XT* tile = reinterpret_cast<XT*>(smem + r_bytes(L, nnz, RL));
  for (int c = threadIdx.x; c < L; c += kThreads) lmax[c] = 0u;
  constexpr int kSlots = 64 / TR;  // columns a wave works on at once
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int r = lane % TR;
  const int slot = lane / TR;
  const int64_t ntiles = (n + TR - 1) / TR;
  for (int64_t tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
    const int64_t row0 = tl * TR;
    const int rows = (int)min((int64_t)TR, n - row0);
    __syncthreads();
    // Stage rows [row0, row0 + rows) into LDS (row pitch padded against bank conflicts).
    if (sizeof(XT) == 1 && (d & 3) == 0) {
      const uint32_t* src = reinterpret_cast<const uint32_t*>(X + row0 * d);
      uint32_t* dst = reinterpret_cast<uint32_t*>(tile);
      const int wpr = d >> 2;  // words per row
      const int wpitch = pitch >> 2;
      const int total = rows * wpr;
      for (int e = threadIdx.x; e < total; e += kThreads) {
        int rr = e / wpr, c = e - rr * wpr;
        dst[rr * wpitch + c] = src[e];
      }
    } else {
      const XT* src = X + row0 * d;
      const int total = rows * d;
      for (int e = threadIdx.x; e < total; e += kThreads) {
        int rr = e / d, c = e - rr * d;
        tile[rr * pitch + c] = src[e];
      }
    }
    __syncthreads();
    const bool valid_row = r < rows;
    for (int cb = wave * kSlots; cb < L; cb += 4 * kSlots) {
      const int c = cb + slot;
      float pv = 0.0f;
      if (c < L) {
        const int beg = lcp[c], end = lcp[c + 1];
        double acc = 0.0;
        const XT* xr = tile + r * pitch;
        int k = beg;
        // x * w is exact in double, so one fused multiply-add per term rounds
        // as the separate product and sum would
        for (; k + 4 <= end; k += 4) {  // independent loads, sums still in ascending k
          const double x0 = (double)xr[lri[k]], x1 = (double)xr[lri[k + 1]];
          const double x2 = (double)xr[lri[k + 2]], x3 = (double)xr[lri[k + 3]];
          acc = __fma_rn(x0, wv(k), acc);
          acc = __fma_rn(x1, wv(k + 1), acc);
          acc = __fma_rn(x2, wv(k + 2), acc);
          acc = __fma_rn(x3, wv(k + 3), acc);
        }
        for (; k < end; ++k) acc = __fma_rn((double)xr[lri[k]], wv(k), acc);
        pv = (float)acc;
        if (valid_row) P[(int64_t)c * n + row0 + r] = pv;
      }
      uint32_t bits = (valid_row && c < L) ? (__float_as_uint(pv) & 0x7fffffffu) : 0u;
      // reduce the max over the TR lanes sharing column c
      for (int off = TR / 2; off > 0; off >>= 1) bits = max(bits, (uint32_t)__shfl_xor((int)bits, off));
      if (r == 0 && c < L && bits) atomicMax(&lmax[c], bits);
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < L; c += kThreads)
    if (lmax[c]) atomicMax(&colmax[c], lmax[c]);
}

// Wave-uniform columns: the 64 lanes of a wave are the 64 rows of the tile, so
// each column's compressed entries (row index, weight as double) are read with
// scalar loads, 16 at a time, and every term is one LDS read of x plus one FMA;
// four columns run as independent FMA chains. The column maximum of |P| is a DPP wave
// reduction and one LDS atomic per column and tile.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  // quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror, row_bcast:15,
  // row_bcast:31: lane 63 ends with the wave's maximum
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xB1, 0xF, 0xF, false));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x4E, 0xF, 0xF, false));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x141, 0xF, 0xF, false));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x140, 0xF, 0xF, false));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x142, 0xA, 0xF, false));
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x143, 0xC, 0xF, false));
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

constexpr int kColsPerPass = 4;   // columns per group (independent FMA chains)
constexpr int kTermsPerBatch = 4; // terms of each column loaded per batch

// R for this kernel: column groups of kColsPerPass, their terms interleaved
// (term k of slot u of group g at gk[g] * 4 + 4 k + u), each group padded
// to a multiple of kTermsPerBatch terms with (row d, weight 0): row d of the
// tile is a zero slot, and fma(0, 0, acc) = acc exactly (acc starts at +0 and
// so is never -0). Slot u of group g computes column gcol[4 g + u] (-1: none);
// the columns are grouped in order of their non-zero counts, so a group's
// padding to its longest column stays small.
template <typename XT, int kStage>  // kStage: 16-byte chunks of the next X tile per thread
__global__ void __launch_bounds__(kThreads)
k_project_rows(const XT* __restrict__ X, int64_t n, int d, int L, const int32_t* __restrict__ gk,
               const int32_t* __restrict__ gcol, const int32_t* __restrict__ ridx,
               const double* __restrict__ wd, float* __restrict__ P,
               uint32_t* __restrict__ colmax, int pitch) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t* lmax = reinterpret_cast<uint32_t*>(smem);
  XT* tile = reinterpret_cast<XT*>(smem + (((size_t)L * 4 + 15) & ~(size_t)15));
  for (int c = threadIdx.x; c < L; c += kThreads) lmax[c] = 0u;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int ngroups = (L + kColsPerPass - 1) / kColsPerPass;
  const int64_t ntiles = (n + 63) / 64;
  // Rows of 16-byte multiples up to kStage * 256 bytes: the next tile of X is
  // loaded into registers (16-byte loads, all in flight at once) while the
  // current tile is computed, and written to LDS after it.
  const int rb = d * (int)sizeof(XT);
  const bool fast = (rb & 15) == 0 && rb <= kStage * 16 * kThreads / 64 && ((uintptr_t)X & 15) == 0;
  const int cpr = rb >> 4;  // 16-byte chunks per row
  uint4 pre[kStage];
  auto load_tile = [&](int64_t t) {
    const int64_t r0 = t * 64;
    const int nq = (int)min((int64_t)64, n - r0) * cpr;
    const uint4* src = reinterpret_cast<const uint4*>(X + r0 * d);
#pragma unroll
    for (int u = 0; u < kStage; ++u) {
      const int q = threadIdx.x + u * kThreads;
      pre[u] = q < nq ? src[q] : make_uint4(0u, 0u, 0u, 0u);
    }
  };
  if (fast && (int64_t)blockIdx.x < ntiles) load_tile(blockIdx.x);
  for (int64_t tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
    const int64_t row0 = tl * 64;
    const int rows = (int)min((int64_t)64, n - row0);
    __syncthreads();
    if (fast) {
      unsigned char* tb = reinterpret_cast<unsigned char*>(tile);
      const int pb = pitch * (int)sizeof(XT);  // row pitch in bytes (a multiple of 4)
#pragma unroll
      for (int u = 0; u < kStage; ++u) {
        const int q = threadIdx.x + u * kThreads;
        if (q < rows * cpr) {
          const int rr = q / cpr, c = q - rr * cpr;
          uint32_t* dst = reinterpret_cast<uint32_t*>(tb + rr * pb + c * 16);
          dst[0] = pre[u].x;
          dst[1] = pre[u].y;
          dst[2] = pre[u].z;
          dst[3] = pre[u].w;
        }
      }
    } else if (sizeof(XT) == 1 && (d & 3) == 0) {
      const uint32_t* src = reinterpret_cast<const uint32_t*>(X + row0 * d);
      uint32_t* dst = reinterpret_cast<uint32_t*>(tile);
      const int wpr = d >> 2, wpitch = pitch >> 2, total = rows * wpr;
      for (int e = threadIdx.x; e < total; e += kThreads) {
        const int rr = e / wpr, c = e - rr * wpr;
        dst[rr * wpitch + c] = src[e];
      }
    } else {
      const XT* src = X + row0 * d;
      const int total = rows * d;
      for (int e = threadIdx.x; e < total; e += kThreads) {
        const int rr = e / d, c = e - rr * d;
        tile[rr * pitch + c] = src[e];
      }
    }
    if (threadIdx.x < 64) tile[threadIdx.x * pitch + d] = (XT)0;  // the zero slot
    __syncthreads();
    if (fast && tl + gridDim.x < ntiles) load_tile(tl + gridDim.x);
    const XT* xr = tile + lane * pitch;
    const bool valid = lane < rows;
    for (int g = wave; g < ngroups; g += 4) {
      const int k0 = gk[g], k1 = gk[g + 1];
      double acc[kColsPerPass];
#pragma unroll
      for (int u = 0; u < kColsPerPass; ++u) acc[u] = 0.0;
      // x * w is exact in double, so one fused multiply-add per term rounds as
      // the separate product and sum would; each column's terms in ascending row
      for (int k = k0; k < k1; k += kTermsPerBatch) {
        int ri[kTermsPerBatch * kColsPerPass];
        double wv[kTermsPerBatch * kColsPerPass];
#pragma unroll
        for (int i = 0; i < kTermsPerBatch * kColsPerPass; ++i) {
          ri[i] = ridx[k * kColsPerPass + i];
          wv[i] = wd[k * kColsPerPass + i];
        }
        double xv[kTermsPerBatch * kColsPerPass];
#pragma unroll
        for (int i = 0; i < kTermsPerBatch * kColsPerPass; ++i) xv[i] = (double)xr[ri[i]];
#pragma unroll
        for (int t = 0; t < kTermsPerBatch; ++t)
#pragma unroll
          for (int u = 0; u < kColsPerPass; ++u)
            acc[u] = __fma_rn(xv[t * kColsPerPass + u], wv[t * kColsPerPass + u], acc[u]);
      }
#pragma unroll
      for (int u = 0; u < kColsPerPass; ++u) {
        const int c = gcol[g * kColsPerPass + u];
        if (c < 0) continue;
        const float pv = (float)acc[u];
        if (valid) P[(int64_t)c * n + row0 + lane] = pv;
        const uint32_t m = wave_max_u32(valid ? (__float_as_uint(pv) & 0x7fffffffu) : 0u);
        if (lane == 0 && m) atomicMax(&lmax[c], m);
      }
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < L; c += kThreads)
    if (lmax[c]) atomicMax(&colmax[c], lmax[c]);
}

constexpr size_t kLdsBudget = 64 * 1024;

// X tile row pitch in elements (odd word pitch: rows land on distinct banks)
template <typename XT>
int pitch_of(int d) {
  if (sizeof(XT) == 1) {
    int words = (d + 3) / 4;
    if ((words & 1) == 0) words += 1;
    return words * 4;
  }
  return d | 1;
}

// tile pitch of the wave-uniform kernel: one more element for the zero slot
template <typename XT>
int pitch_rows(int d) { return pitch_of<XT>(d + 1); }

template <typename XT, int TR, bool RL>
int run(drpt_ctx* ctx, const XT* X, int64_t n, int d, const int32_t* colptr, const uint16_t* rowidx,
        const float* w, int L, int nnz, float* P, uint32_t* colmax) {
  const int pitch = pitch_of<XT>(d);
  const size_t lds = r_bytes(L, nnz, RL) + (size_t)TR * pitch * sizeof(XT);
  if (lds > kLdsBudget) {
    set_error("project: LDS tile too large for this dimension / column count");
    return DRPT_ERR_UNSUPPORTED;
  }
  // as many blocks as are resident at once (LDS bounds them): every block
  // strides over the row tiles from the start, no late round of blocks
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_project<XT, TR, RL>, kThreads,
                                                   lds) != hipSuccess || per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 2;
  }
  int64_t blocks = std::min<int64_t>((n + TR - 1) / TR, (int64_t)ctx->num_cus * per_cu);
  hipLaunchKernelGGL((k_project<XT, TR, RL>), dim3((unsigned)blocks), dim3(kThreads), lds,
                     ctx->stream, X, n, d, L, colptr, rowidx, w, P, colmax, pitch);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

// the largest row tile that fits beside the header: R staged in LDS when
// some tile of at least 8 rows fits beside it, else R read from the caches
template <typename XT>
int dispatch(drpt_ctx* ctx, const XT* X, int64_t n, int d, const int32_t* colptr,
             const uint16_t* rowidx, const float* w, int L, int nnz, float* P, uint32_t* colmax) {
  const size_t row_bytes = (size_t)pitch_of<XT>(d) * sizeof(XT);
  for (int rl = 1; rl >= 0; --rl) {
    const size_t hdr = r_bytes(L, nnz, rl != 0);
    if (hdr >= kLdsBudget) continue;
    const size_t room = kLdsBudget - hdr;
#define DRPT_PROJ_TRY(TR)                                                                    \
  if ((TR) * row_bytes <= room)                                                              \
    return rl ? run<XT, TR, true>(ctx, X, n, d, colptr, rowidx, w, L, nnz, P, colmax)         \
              : run<XT, TR, false>(ctx, X, n, d, colptr, rowidx, w, L, nnz, P, colmax);
    DRPT_PROJ_TRY(64)
    DRPT_PROJ_TRY(32)
    DRPT_PROJ_TRY(16)
    DRPT_PROJ_TRY(8)
#undef DRPT_PROJ_TRY
  }
  set_error("project: dimension too large for the LDS tile");
  return DRPT_ERR_UNSUPPORTED;
}

}  // namespace

template <typename XT>
int run_rows(drpt_ctx* ctx, const XT* X, int64_t n, int d, const int32_t* gk, const int32_t* gcol,
             const int32_t* ridx, const double* wd, int L, float* P, uint32_t* colmax) {
  const int pitch = pitch_rows<XT>(d);
  const size_t lds = (((size_t)L * 4 + 15) & ~(size_t)15) + (size_t)64 * pitch * sizeof(XT);
  // register staging of the next tile: 8 chunks per thread for rows up to 512
  // bytes, else 16
  const bool small = (size_t)d * sizeof(XT) <= 8 * 16 * kThreads / 64;
  const void* f = small ? (const void*)k_project_rows<XT, 8> : (const void*)k_project_rows<XT, 16>;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kThreads, lds) != hipSuccess ||
      per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 2;
  }
  const int64_t blocks = std::min<int64_t>((n + 63) / 64, (int64_t)ctx->num_cus * per_cu);
  if (small)
    hipLaunchKernelGGL((k_project_rows<XT, 8>), dim3((unsigned)blocks), dim3(kThreads), lds,
                       ctx->stream, X, n, d, L, gk, gcol, ridx, wd, P, colmax, pitch);
  else
    hipLaunchKernelGGL((k_project_rows<XT, 16>), dim3((unsigned)blocks), dim3(kThreads), lds,
                       ctx->stream, X, n, d, L, gk, gcol, ridx, wd, P, colmax, pitch);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

template <typename XT>
bool rows_fit(int d, int L) {
  return (((size_t)L * 4 + 15) & ~(size_t)15) + (size_t)64 * pitch_rows<XT>(d) * sizeof(XT) <=
         kLdsBudget;
}

int launch_project(drpt_ctx* ctx, const void* X, int dtype, int64_t n, int d, const int32_t* colptr,
                   const uint16_t* rowidx, const float* w, const int32_t* gk, const int32_t* gcol,
                   const int32_t* ridx32, const double* wd, int L, int nnz, float* P,
                   uint32_t* colmax) {
  DRPT_CHECK(fill_u32(ctx, colmax, 0u, (size_t)L));
  if (n <= 0) return DRPT_OK;
  if (gk && gcol && ridx32 && wd) {  // 64-row tiles with wave-uniform columns where they fit
    if (dtype == DRPT_U8 && rows_fit<uint8_t>(d, L))
      return run_rows<uint8_t>(ctx, (const uint8_t*)X, n, d, gk, gcol, ridx32, wd, L, P, colmax);
    if (dtype != DRPT_U8 && rows_fit<float>(d, L))
      return run_rows<float>(ctx, (const float*)X, n, d, gk, gcol, ridx32, wd, L, P, colmax);
  }
  if (dtype == DRPT_U8)
    return dispatch<uint8_t>(ctx, (const uint8_t*)X, n, d, colptr, rowidx, w, L, nnz, P, colmax);
  return dispatch<float>(ctx, (const float*)X, n, d, colptr, rowidx, w, L, nnz, P, colmax);
}

}  // namespace drpt
