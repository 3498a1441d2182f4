// Leaf grouping and the reference local stage on the device.
//
// After the splits every point of tree t carries the heap index of its leaf.
// The leaf k-NN needs each leaf's points contiguous: counts per heap index,
// offsets (host prefix sum), then an atomic-cursor scatter. The reference's
// local stage (MDRPT::grow_local_trees, cpp/src/drpt/algo/mdrpt.cpp:321-400;
// DRPTLocal, drpt_local.cpp:81-180) addresses points by their position in the
// global leaf in ascending index order, so in that mode each global leaf is
// first sorted (LDS bitonic sort) and then permuted by the position table the
// host derives for the leaf's size (std::nth_element on equal keys).
//
// Roofline: HBM-bound, 4 B read + 4 B written per point per tree per pass.
#include <climits>

#include "internal.hpp"

namespace drpt {

namespace {

__global__ void k_count(const uint32_t* __restrict__ heap, int64_t n, int HS,
                        uint32_t* __restrict__ hcount) {
  const int t = blockIdx.y;
  const uint32_t* h = heap + (int64_t)t * n;
  uint32_t* hc = hcount + (int64_t)t * HS;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    atomicAdd(&hc[h[i]], 1u);
}

__global__ void k_scatter(const uint32_t* __restrict__ heap, int64_t n, int HS,
                          const int64_t* __restrict__ hoff, uint32_t* __restrict__ cursor,
                          int32_t* __restrict__ perm) {
  const int t = blockIdx.y;
  const uint32_t* h = heap + (int64_t)t * n;
  const int64_t* ho = hoff + (int64_t)t * HS;
  uint32_t* cu = cursor + (int64_t)t * HS;
  int32_t* pm = perm + (int64_t)t * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t hv = h[i];
    const uint32_t pos = atomicAdd(&cu[hv], 1u);
    pm[ho[hv] + pos] = (int32_t)i;
  }
}

// Ascending sort of each segment in LDS (bitonic network, size padded to a
// power of two with INT_MAX). One block per segment.
constexpr int kSortMax = 8192;

__global__ void __launch_bounds__(1024)
k_sort_segments(int32_t* __restrict__ perm, int64_t n, const int64_t* __restrict__ seg_start,
                const int32_t* __restrict__ seg_size, const int32_t* __restrict__ seg_tree) {
  __shared__ int32_t buf[kSortMax];
  const int s = blockIdx.x;
  const int m = seg_size[s];
  int32_t* p = perm + (int64_t)seg_tree[s] * n + seg_start[s];
  int P2 = 1;
  while (P2 < m) P2 <<= 1;
  for (int i = threadIdx.x; i < P2; i += blockDim.x) buf[i] = i < m ? p[i] : INT_MAX;
  __syncthreads();
  for (int k = 2; k <= P2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P2; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const bool up = (i & k) == 0;
          const int32_t a = buf[i], b = buf[ixj];
          if ((a > b) == up) {
            buf[i] = b;
            buf[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < m; i += blockDim.x) p[i] = buf[i];
}

// out[start + k] = in[start + table[k]] for each global leaf.
__global__ void k_apply_local(const int32_t* __restrict__ in, int32_t* __restrict__ out, int64_t n,
                              const int64_t* __restrict__ seg_start,
                              const int32_t* __restrict__ seg_tree,
                              const int64_t* __restrict__ seg_tbl,
                              const int32_t* __restrict__ seg_size,
                              const int32_t* __restrict__ tables) {
  const int s = blockIdx.x;
  const int64_t base = (int64_t)seg_tree[s] * n + seg_start[s];
  const int32_t* tb = tables + seg_tbl[s];
  const int m = seg_size[s];
  for (int k = threadIdx.x; k < m; k += blockDim.x) out[base + k] = in[base + tb[k]];
}

unsigned grid_for(int64_t n) {
  int64_t b = (n + 1023) / 1024;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

int launch_count_leaves(drpt_ctx* ctx, const uint32_t* heap, int64_t n, int T, int D,
                        uint32_t* hcount) {
  const int HS = 1 << D;
  DRPT_CHECK(fill_u32(ctx, hcount, 0u, (size_t)T * HS));
  hipLaunchKernelGGL(k_count, dim3(grid_for(n), T), dim3(256), 0, ctx->stream, heap, n, HS, hcount);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_scatter_leaves(drpt_ctx* ctx, const uint32_t* heap, int64_t n, int T, int D,
                          const int64_t* hoff, uint32_t* cursor, int32_t* perm) {
  const int HS = 1 << D;
  DRPT_CHECK(fill_u32(ctx, cursor, 0u, (size_t)T * HS));
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(n), T), dim3(256), 0, ctx->stream, heap, n, HS, hoff,
                     cursor, perm);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_sort_segments(drpt_ctx* ctx, int32_t* perm, int64_t n, const int64_t* seg_start,
                         const int32_t* seg_size, const int32_t* seg_tree, int nseg, int max_seg) {
  if (nseg <= 0) return DRPT_OK;
  if (max_seg > kSortMax) {
    set_error("reference local stage: a global leaf exceeds " + std::to_string(kSortMax) +
              " points (LDS sort capacity)");
    return DRPT_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(k_sort_segments, dim3(nseg), dim3(1024), 0, ctx->stream, perm, n, seg_start,
                     seg_size, seg_tree);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_apply_local(drpt_ctx* ctx, const int32_t* perm_in, int32_t* perm_out, int64_t n,
                       const int64_t* seg_start, const int32_t* seg_tree, const int64_t* seg_tbl,
                       const int32_t* seg_size, int nseg, const int32_t* tables) {
  if (nseg <= 0) return DRPT_OK;
  hipLaunchKernelGGL(k_apply_local, dim3(nseg), dim3(256), 0, ctx->stream, perm_in, perm_out, n,
                     seg_start, seg_tree, seg_tbl, seg_size, tables);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

}  // namespace drpt
