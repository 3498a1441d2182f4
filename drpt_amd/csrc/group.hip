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

// LDS-private variants: heap slots [blockIdx.z * CH, + CH) of tree blockIdx.y
// in LDS (4 B each; CH a power of two, the whole tree when it fits). A block
// counts its contiguous range of points in LDS and adds each touched slot's
// count to the global array with one atomic, instead of one global atomic per
// point; trees with more slots take several chunks, each reading the block's
// points again and skipping those of other chunks. The scatter counts again,
// reserves each touched slot's range of positions with one global atomic, and
// places its points through LDS cursors. Each thread loads kGroupLoads heap
// words before using any (loads in flight behind the LDS atomics).
constexpr int kGroupThreads = 1024;
constexpr int kGroupLoads = 8;

__device__ __forceinline__ void count_range(const uint32_t* __restrict__ h, int64_t b0, int64_t b1,
                                            uint32_t lo, int CH, uint32_t* lc) {
  for (int64_t base = b0; base < b1; base += (int64_t)kGroupThreads * kGroupLoads) {
    uint32_t hv[kGroupLoads];
#pragma unroll
    for (int u = 0; u < kGroupLoads; ++u) {
      const int64_t i = base + u * kGroupThreads + threadIdx.x;
      hv[u] = i < b1 ? h[i] - lo : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int u = 0; u < kGroupLoads; ++u)
      if (hv[u] < (uint32_t)CH) atomicAdd(&lc[hv[u]], 1u);
  }
}

__global__ void __launch_bounds__(kGroupThreads)
k_count_lds(const uint32_t* __restrict__ heap, int64_t n, int HS, int CH, int64_t ppb,
            uint32_t* __restrict__ hcount) {
  extern __shared__ uint32_t lc[];
  const int t = blockIdx.y;
  const uint32_t lo = (uint32_t)blockIdx.z * (uint32_t)CH;
  for (int k = threadIdx.x; k < CH; k += kGroupThreads) lc[k] = 0;
  __syncthreads();
  const int64_t b0 = (int64_t)blockIdx.x * ppb, b1 = min(n, b0 + ppb);
  count_range(heap + (int64_t)t * n, b0, b1, lo, CH, lc);
  __syncthreads();
  uint32_t* hc = hcount + (int64_t)t * HS + lo;
  for (int k = threadIdx.x; k < CH; k += kGroupThreads)
    if (lc[k]) atomicAdd(&hc[k], lc[k]);
}

__global__ void __launch_bounds__(kGroupThreads)
k_scatter_lds(const uint32_t* __restrict__ heap, int64_t n, int HS, int CH, int64_t ppb,
              const int64_t* __restrict__ hoff, uint32_t* __restrict__ cursor,
              int32_t* __restrict__ perm) {
  extern __shared__ uint32_t lc[];
  const int t = blockIdx.y;
  const uint32_t lo = (uint32_t)blockIdx.z * (uint32_t)CH;
  for (int k = threadIdx.x; k < CH; k += kGroupThreads) lc[k] = 0;
  __syncthreads();
  const uint32_t* h = heap + (int64_t)t * n;
  const int64_t b0 = (int64_t)blockIdx.x * ppb, b1 = min(n, b0 + ppb);
  count_range(h, b0, b1, lo, CH, lc);
  __syncthreads();
  const int64_t* ho = hoff + (int64_t)t * HS + lo;
  uint32_t* cu = cursor + (int64_t)t * HS + lo;
  for (int k = threadIdx.x; k < CH; k += kGroupThreads)
    if (lc[k]) lc[k] = (uint32_t)(ho[k] + atomicAdd(&cu[k], lc[k]));  // first position (< n)
  __syncthreads();
  int32_t* pm = perm + (int64_t)t * n;
  for (int64_t base = b0; base < b1; base += (int64_t)kGroupThreads * kGroupLoads) {
    uint32_t hv[kGroupLoads];
#pragma unroll
    for (int u = 0; u < kGroupLoads; ++u) {
      const int64_t i = base + u * kGroupThreads + threadIdx.x;
      hv[u] = i < b1 ? h[i] - lo : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int u = 0; u < kGroupLoads; ++u)
      if (hv[u] < (uint32_t)CH)
        pm[atomicAdd(&lc[hv[u]], 1u)] = (int32_t)(base + u * kGroupThreads + threadIdx.x);
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

// Real-key local tree level (-density >= 1: non-zero local projections):
// each segment of perm is sorted by (value, point index), value =
// LP[col(tree)][point], so its first ceil(m/2) positions hold the left child of
// drpt_local.cpp:121-161 (nth_element at m/2, left = [begin, end - m/2)). Keys
// are 64-bit: the float's order-preserving bits (-0 taken as +0, so equal
// values tie and the index decides) above the index. One block per segment,
// bitonic network in LDS (size padded to a power of two).
constexpr int kKeyedSortMax = 8192;

__device__ __forceinline__ uint32_t ordered_bits(float v) {
  uint32_t b = __float_as_uint(v);
  if (b == 0x80000000u) b = 0u;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void __launch_bounds__(1024)
k_sort_keyed(int32_t* __restrict__ perm, int64_t n, const float* __restrict__ LP, int ld, int level,
             const int64_t* __restrict__ seg_start, const int32_t* __restrict__ seg_size,
             const int32_t* __restrict__ seg_tree) {
  __shared__ unsigned long long buf[kKeyedSortMax];
  const int s = blockIdx.x;
  const int m = seg_size[s];
  const int t = seg_tree[s];
  int32_t* p = perm + (int64_t)t * n + seg_start[s];
  const float* v = LP + ((int64_t)t * ld + level) * n;
  int P2 = 1;
  while (P2 < m) P2 <<= 1;
  for (int i = threadIdx.x; i < P2; i += blockDim.x) {
    if (i < m) {
      const int32_t id = p[i];
      buf[i] = ((unsigned long long)ordered_bits(v[id]) << 32) | (uint32_t)id;
    } else {
      buf[i] = ~0ull;
    }
  }
  __syncthreads();
  for (int k = 2; k <= P2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P2; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const bool up = (i & k) == 0;
          const unsigned long long a = buf[i], b = buf[ixj];
          if ((a > b) == up) {
            buf[i] = b;
            buf[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < m; i += blockDim.x) p[i] = (int32_t)(uint32_t)buf[i];
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

// Block-wide exclusive scan of two int64 sequences (1024 threads).
__device__ __forceinline__ void block_scan2(long long& a, long long& b, long long& ta, long long& tb,
                                            long long* sa, long long* sb) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  long long ia = a, ib = b;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const long long xa = __shfl_up(ia, off), xb = __shfl_up(ib, off);
    if (lane >= off) {
      ia += xa;
      ib += xb;
    }
  }
  if (lane == 63) {
    sa[wave] = ia;
    sb[wave] = ib;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    long long ca = 0, cb = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      const long long va = sa[w], vb = sb[w];
      sa[w] = ca;
      sb[w] = cb;
      ca += va;
      cb += vb;
    }
    sa[16] = ca;
    sb[16] = cb;
  }
  __syncthreads();
  a = sa[wave] + ia - a;
  b = sb[wave] + ib - b;
  ta = sa[16];
  tb = sb[16];
  __syncthreads();
}

// Leaf offsets and work items in three passes over chunks of kScanChunk heap
// slots (one block per tree and chunk, then one block over the chunk totals),
// so the slots of all trees are scanned by T * HS / kScanChunk blocks at once.
constexpr int kScanChunk = 1024;

// Per (tree blockIdx.y, chunk blockIdx.x): cs[0] = points, cs[1] = work items
// (blocks of rpi rows), cs[2] = sum of m^2 over its leaves, cs[3] = leaves << 32
// | largest m, at cs = csum + 4 (t NC + j).
__global__ void __launch_bounds__(kScanChunk)
k_chunk_sums(const uint32_t* __restrict__ hcount, int HS, ItemPlan plan,
             int64_t* __restrict__ csum) {
  __shared__ long long sa[17], sb[17];
  __shared__ unsigned int smax;
  const int t = blockIdx.y, j = blockIdx.x, NC = gridDim.x;
  const int h = j * kScanChunk + threadIdx.x;
  if (threadIdx.x == 0) smax = 0;
  const long long c = h < HS ? (long long)hcount[(int64_t)t * HS + h] : 0;
  const int rpi = plan.rpi[size_bucket((int)min(c, (long long)INT_MAX))];
  long long a = c, b = (c + rpi - 1) / rpi, ta, tb;
  block_scan2(a, b, ta, tb, sa, sb);
  long long m2 = c * c, lv = c > 0 ? 1 : 0, tm2, tlv;
  block_scan2(m2, lv, tm2, tlv, sa, sb);
  unsigned int mx = (unsigned int)c;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = max(mx, (unsigned int)__shfl_xor((int)mx, off));
  if ((threadIdx.x & 63) == 0) atomicMax(&smax, mx);
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t* cs = csum + 4 * ((int64_t)t * NC + j);
    cs[0] = ta;
    cs[1] = tb;
    cs[2] = tm2;
    cs[3] = (int64_t)(((unsigned long long)tlv << 32) | smax);
  }
}

// One block over the T * NC chunk totals in tree-major order: cs[0] becomes
// the chunk's first position within its tree's perm (each tree's counts sum
// to n), cs[1] its first work item (all trees); base[t] = first item of tree
// t, base[T] = all items; range = {0, base[a], base[a], base[T], 0, base[T]};
// stats = {sum of m^2, largest m, leaves}.
__global__ void __launch_bounds__(1024)
k_chunk_bases(int64_t* __restrict__ csum, int T, int NC, int64_t n, int a_first,
              int64_t* __restrict__ base, int64_t* __restrict__ range,
              unsigned long long* __restrict__ stats) {
  __shared__ long long sa[17], sb[17];
  __shared__ unsigned int smax;
  if (threadIdx.x == 0) smax = 0;
  const int64_t total = (int64_t)T * NC;
  long long cp = 0, ci = 0, m2 = 0, lv = 0;
  unsigned int mx = 0;
  for (int64_t w = 0; w < total; w += blockDim.x) {
    const int64_t k = w + threadIdx.x;
    long long a = 0, b = 0, ta, tb;
    if (k < total) {
      const int64_t* cs = csum + 4 * k;
      a = cs[0];
      b = cs[1];
      m2 += cs[2];
      const unsigned long long pk = (unsigned long long)cs[3];
      lv += (long long)(pk >> 32);
      mx = max(mx, (unsigned int)pk);
    }
    block_scan2(a, b, ta, tb, sa, sb);
    if (k < total) {
      const int t = (int)(k / NC);
      csum[4 * k] = cp + a - (int64_t)t * n;
      csum[4 * k + 1] = ci + b;
      if (k % NC == 0) base[t] = ci + b;
    }
    cp += ta;
    ci += tb;
  }
  long long tm2, tlv;
  block_scan2(m2, lv, tm2, tlv, sa, sb);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = max(mx, (unsigned int)__shfl_xor((int)mx, off));
  if ((threadIdx.x & 63) == 0) atomicMax(&smax, mx);
  __syncthreads();  // base[] of every tree written (block scope) and smax complete
  if (threadIdx.x == 0) {
    base[T] = ci;
    const long long ba = a_first < T ? base[a_first] : ci;
    range[0] = 0;
    range[1] = ba;
    range[2] = ba;
    range[3] = ci;
    range[4] = 0;
    range[5] = ci;
    stats[0] = (unsigned long long)tm2;
    stats[1] = smax;
    stats[2] = (unsigned long long)tlv;
  }
}

// Per (tree, chunk): hoff[h] = first position of heap slot h's points in
// perm[t], and the work items (blocks of plan.rpi[bucket] rows) of every
// non-empty slot, in slot order from the chunk's first item.
__global__ void __launch_bounds__(kScanChunk)
k_chunk_emit(const uint32_t* __restrict__ hcount, int HS, ItemPlan plan,
             const int64_t* __restrict__ csum, int64_t* __restrict__ hoff,
             LeafItem* __restrict__ items) {
  __shared__ long long sa[17], sb[17];
  const int t = blockIdx.y, j = blockIdx.x, NC = gridDim.x;
  const int h = j * kScanChunk + threadIdx.x;
  const long long c = h < HS ? (long long)hcount[(int64_t)t * HS + h] : 0;
  const int rpi = plan.rpi[size_bucket((int)min(c, (long long)INT_MAX))];
  long long a = c, b = (c + rpi - 1) / rpi, ta, tb;
  block_scan2(a, b, ta, tb, sa, sb);
  const int64_t* cs = csum + 4 * ((int64_t)t * NC + j);
  const int64_t start = cs[0] + a;
  if (h < HS) hoff[(int64_t)t * HS + h] = start;
  if (c > 0) {
    LeafItem it;
    it.tree = t;
    it.size = (int32_t)c;
    it.start = start;
    const int64_t i0 = cs[1] + b;
    for (long long r0 = 0, k = 0; r0 < c; r0 += rpi, ++k) {
      it.row0 = (int32_t)r0;
      it.rows = (int32_t)min((long long)rpi, c - r0);
      items[i0 + k] = it;
    }
  }
}

__global__ void k_tree_range(const int64_t* __restrict__ base, int t0, int t1,
                             int64_t* __restrict__ range) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    range[0] = base[t0];
    range[1] = base[t1];
  }
}

// The items among [range[0], range[1]) whose leaf size lies in [lo, hi],
// copied to out (one reservation per wave, so their order is unspecified);
// out_range = {0, count} with its count word zeroed by the caller.
__global__ void __launch_bounds__(256)
k_select_items(const LeafItem* __restrict__ items, const int64_t* __restrict__ range, int lo,
               int hi, LeafItem* __restrict__ out, int64_t* __restrict__ out_range) {
  const int64_t b = range[0], e = range[1];
  const int lane = threadIdx.x & 63;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = b + (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); i0 < e;
       i0 += stride) {
    const int64_t i = i0 + lane;
    LeafItem it{};
    bool keep = false;
    if (i < e) {
      it = items[i];
      keep = it.size >= lo && it.size <= hi;
    }
    const uint64_t bal = __ballot(keep);
    if (bal == 0) continue;
    unsigned base = 0;
    if (lane == 0)
      base = atomicAdd(reinterpret_cast<unsigned*>(&out_range[1]), (unsigned)__popcll(bal));
    base = (unsigned)__shfl((int)base, 0);
    if (keep) out[base + __popcll(bal & ((1ull << lane) - 1))] = it;
  }
}

// Xp[i] = X[perm[i]], 16-, 4- or 1-byte words W per row; pos[perm[i]] = i
template <typename W>
__global__ void k_permute_rows(const W* __restrict__ X, int64_t words, const int32_t* __restrict__ perm,
                               int64_t n, W* __restrict__ Xp, int32_t* __restrict__ pos) {
  const int64_t total = n * words;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k / words, w = k - i * words;
    const int32_t src = perm[i];
    Xp[k] = X[(int64_t)src * words + w];
    if (w == 0) pos[src] = (int32_t)i;
  }
}

unsigned grid_for(int64_t n) {
  int64_t b = (n + 1023) / 1024;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

// heap slots per LDS chunk of the private variants (a power of two, at most
// HS, within the device's LDS; the function attribute raised beyond 64 KB; at
// most 8 chunks), else 0: global atomics
static int group_chunk(drpt_ctx* ctx, const void* f, int HS, int64_t n) {
  if (n < 16384) return 0;
  int ch = HS;
  while (ch > 4096 && (size_t)ch * 4 > (size_t)ctx->max_lds) ch >>= 1;
  if ((size_t)ch * 4 > (size_t)ctx->max_lds || HS / ch > 8) return 0;
  if ((size_t)ch * 4 > 64 * 1024 &&
      hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, ch * 4) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return ch;
}

// points per block of the private variants: about one block per CU over all
// trees and chunks, at least 4096 points each
static int64_t group_ppb(drpt_ctx* ctx, int64_t n, int T, int nch) {
  const int64_t bx = std::max<int64_t>(
      1, std::min<int64_t>((ctx->num_cus + T * nch - 1) / (T * nch), n / 4096));
  return (n + bx - 1) / bx;
}

int launch_count_leaves(drpt_ctx* ctx, const uint32_t* heap, int64_t n, int T, int D,
                        uint32_t* hcount) {
  const int HS = 1 << D;
  DRPT_CHECK(fill_u32(ctx, hcount, 0u, (size_t)T * HS));
  const int ch = group_chunk(ctx, (const void*)k_count_lds, HS, n);
  if (ch > 0) {
    const int nch = HS / ch;
    const int64_t ppb = group_ppb(ctx, n, T, nch);
    hipLaunchKernelGGL(k_count_lds, dim3((unsigned)((n + ppb - 1) / ppb), T, nch),
                       dim3(kGroupThreads), (size_t)ch * 4, ctx->stream, heap, n, HS, ch, ppb,
                       hcount);
  } else {
    hipLaunchKernelGGL(k_count, dim3(grid_for(n), T), dim3(256), 0, ctx->stream, heap, n, HS, hcount);
  }
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_scatter_leaves(drpt_ctx* ctx, const uint32_t* heap, int64_t n, int T, int D,
                          const int64_t* hoff, uint32_t* cursor, int32_t* perm) {
  const int HS = 1 << D;
  DRPT_CHECK(fill_u32(ctx, cursor, 0u, (size_t)T * HS));
  // one chunk only: with several, the LDS scatter measured slower than the
  // global cursors (10M x 96, 2^18 slots in 8 chunks: 21.0 against 12.4 ms)
  int ch = group_chunk(ctx, (const void*)k_scatter_lds, HS, n);
  if (ch > 0 && ch < HS) ch = 0;
  if (ch > 0) {
    const int nch = HS / ch;
    const int64_t ppb = group_ppb(ctx, n, T, nch);
    hipLaunchKernelGGL(k_scatter_lds, dim3((unsigned)((n + ppb - 1) / ppb), T, nch),
                       dim3(kGroupThreads), (size_t)ch * 4, ctx->stream, heap, n, HS, ch, ppb,
                       hoff, cursor, perm);
  } else {
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(n), T), dim3(256), 0, ctx->stream, heap, n, HS,
                       hoff, cursor, perm);
  }
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

int launch_sort_keyed(drpt_ctx* ctx, int32_t* perm, int64_t n, const float* LP, int ld, int level,
                      const int64_t* seg_start, const int32_t* seg_size, const int32_t* seg_tree,
                      int nseg, int max_seg) {
  if (nseg <= 0) return DRPT_OK;
  if (max_seg > kKeyedSortMax) {
    set_error("local random projection: a global leaf exceeds " + std::to_string(kKeyedSortMax) +
              " points (LDS sort capacity)");
    return DRPT_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(k_sort_keyed, dim3(nseg), dim3(1024), 0, ctx->stream, perm, n, LP, ld, level,
                     seg_start, seg_size, seg_tree);
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

namespace drpt {

int launch_device_items(drpt_ctx* ctx, const uint32_t* hcount, int64_t n, int T, int D,
                        const ItemPlan& plan, int first_trees, int64_t* hoff, int64_t* base,
                        int64_t* range, LeafItem* items, unsigned long long* stats) {
  const int HS = 1 << D;
  bool ok = first_trees >= 0 && first_trees <= T;
  for (int b = 0; b < kSizeBuckets; ++b) ok = ok && plan.rpi[b] >= 1;
  if (!ok) {
    set_error("internal: invalid work item layout");
    return DRPT_ERR_ARG;
  }
  const int NC = (HS + kScanChunk - 1) / kScanChunk;
  int64_t* csum;
  DRPT_CHECK(ws(ctx, "group.chunks", 4 * (size_t)T * NC, &csum));
  hipLaunchKernelGGL(k_chunk_sums, dim3(NC, T), dim3(kScanChunk), 0, ctx->stream, hcount, HS, plan,
                     csum);
  DRPT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_chunk_bases, dim3(1), dim3(1024), 0, ctx->stream, csum, T, NC, n,
                     first_trees, base, range, stats);
  DRPT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_chunk_emit, dim3(NC, T), dim3(kScanChunk), 0, ctx->stream, hcount, HS, plan,
                     csum, hoff, items);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_tree_range(drpt_ctx* ctx, const int64_t* base, int t0, int t1, int64_t* range) {
  if (t0 < 0 || t1 < t0) {
    set_error("tree_range: invalid tree interval");
    return DRPT_ERR_ARG;
  }
  hipLaunchKernelGGL(k_tree_range, dim3(1), dim3(64), 0, ctx->stream, base, t0, t1, range);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_permute_rows(drpt_ctx* ctx, const void* X, size_t rowbytes, const int32_t* perm,
                        int64_t n, void* Xp, int32_t* pos) {
  const unsigned grid = (unsigned)ctx->num_cus * 8;
  if (rowbytes % 16 == 0)
    hipLaunchKernelGGL(k_permute_rows<uint4>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint4*)X, (int64_t)(rowbytes / 16), perm, n, (uint4*)Xp, pos);
  else if (rowbytes % 4 == 0)
    hipLaunchKernelGGL(k_permute_rows<uint32_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint32_t*)X, (int64_t)(rowbytes / 4), perm, n, (uint32_t*)Xp, pos);
  else
    hipLaunchKernelGGL(k_permute_rows<uint8_t>, dim3(grid), dim3(256), 0, ctx->stream,
                       (const uint8_t*)X, (int64_t)rowbytes, perm, n, (uint8_t*)Xp, pos);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int launch_select_items(drpt_ctx* ctx, const LeafItem* items, const int64_t* range, int lo, int hi,
                        int64_t max_items, LeafItem* out, int64_t* out_range) {
  if (max_items >= ((int64_t)1 << 32)) {
    set_error("select_items: too many work items");
    return DRPT_ERR_ARG;
  }
  DRPT_CHECK(fill_u32(ctx, reinterpret_cast<uint32_t*>(out_range), 0u, 4));
  const unsigned grid = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>((max_items + 255) / 256, (int64_t)ctx->num_cus * 4));
  hipLaunchKernelGGL(k_select_items, dim3(grid), dim3(256), 0, ctx->stream, items, range, lo, hi,
                     out, out_range);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

}  // namespace drpt
