// Small device utilities.
#include <algorithm>

#include "internal.hpp"

namespace drpt {

namespace {
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

// acc += sum of p[0 .. count): per-thread sums, wave reductions, the block's
// waves summed in LDS and one atomic per block (a few hundred on the one
// address instead of one per wave)
__global__ void __launch_bounds__(256)
k_sum_u32(const uint32_t* __restrict__ p, size_t count, unsigned long long* __restrict__ acc) {
  __shared__ unsigned long long ws[4];
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  unsigned long long s = 0;
  for (; i < count; i += stride) s += p[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned lo = (unsigned)__shfl_xor((int)(unsigned)s, off);
    const unsigned hi = (unsigned)__shfl_xor((int)(unsigned)(s >> 32), off);
    s += ((unsigned long long)hi << 32) | lo;
  }
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = ws[0] + ws[1] + ws[2] + ws[3];
    if (t != 0) atomicAdd(acc, t);
  }
}

// Up to three fills of 32-bit words with their own values in one launch.
__global__ void k_fill3_u32(uint32_t* __restrict__ p0, uint32_t v0, size_t c0,
                            uint32_t* __restrict__ p1, uint32_t v1, size_t c1,
                            uint32_t* __restrict__ p2, uint32_t v2, size_t c2) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < c0 + c1 + c2; i += stride) {
    if (i < c0) p0[i] = v0;
    else if (i < c0 + c1) p1[i - c0] = v1;
    else p2[i - c0 - c1] = v2;
  }
}

// Up to four small device word arrays, one after another, into pinned host
// memory by plain stores from one block: no copy engine or staging buffer on
// the stream's path; the host reads dst after a later synchronisation.
__global__ void k_words_to_host(uint32_t* __restrict__ dst, const uint32_t* __restrict__ s0,
                                int c0, const uint32_t* __restrict__ s1, int c1,
                                const uint32_t* __restrict__ s2, int c2,
                                const uint32_t* __restrict__ s3, int c3) {
  for (int i = threadIdx.x; i < c0 + c1 + c2 + c3; i += blockDim.x) {
    uint32_t v;
    if (i < c0) v = s0[i];
    else if (i < c0 + c1) v = s1[i - c0];
    else if (i < c0 + c1 + c2) v = s2[i - c0 - c1];
    else v = s3[i - c0 - c1 - c2];
    dst[i] = v;
  }
}
}  // namespace

int fill3_u32(drpt_ctx* ctx, uint32_t* p0, uint32_t v0, size_t c0, uint32_t* p1, uint32_t v1,
              size_t c1, uint32_t* p2, uint32_t v2, size_t c2) {
  const size_t total = c0 + c1 + c2;
  if (total == 0) return DRPT_OK;
  const size_t blocks = std::min<size_t>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill3_u32, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, p0, v0, c0, p1,
                     v1, c1, p2, v2, c2);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int words_to_host(drpt_ctx* ctx, uint32_t* dst, const uint32_t* s0, int c0, const uint32_t* s1,
                  int c1, const uint32_t* s2, int c2, const uint32_t* s3, int c3) {
  const int total = c0 + c1 + c2 + c3;
  if (total <= 0) return DRPT_OK;
  hipLaunchKernelGGL(k_words_to_host, dim3(1), dim3(256), 0, ctx->stream, dst, s0, c0, s1, c1, s2,
                     c2, s3, c3);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int pinned_words(drpt_ctx* ctx, PinnedWords& pw, int count) {
  (void)ctx;
  if (pw.cap >= count) return DRPT_OK;
  if (pw.host) DRPT_HIP(hipHostFree(pw.host));
  pw.host = nullptr;
  pw.dev = nullptr;
  pw.cap = 0;
  // coherent: the kernel's stores reach host memory without a cache flush
  DRPT_HIP(hipHostMalloc((void**)&pw.host, 4 * (size_t)count,
                         hipHostMallocMapped | hipHostMallocCoherent));
  void* dev = nullptr;
  DRPT_HIP(hipHostGetDevicePointer(&dev, pw.host, 0));
  pw.dev = (uint32_t*)dev;
  pw.cap = count;
  return DRPT_OK;
}

// A fill kernel rather than hipMemsetAsync, so the sequence stays valid under
// stream capture into a hipGraph.
int fill_u32(drpt_ctx* ctx, uint32_t* p, uint32_t v, size_t count) {
  if (count == 0) return DRPT_OK;
  size_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, p, v, count);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int sum_u32(drpt_ctx* ctx, const uint32_t* p, size_t count, unsigned long long* acc) {
  if (count == 0) return DRPT_OK;
  size_t blocks = (count + 4095) / 4096;
  if (blocks > (size_t)ctx->num_cus * 2) blocks = (size_t)ctx->num_cus * 2;
  hipLaunchKernelGGL(k_sum_u32, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, p, count, acc);
  DRPT_HIP(hipGetLastError());
  return DRPT_OK;
}

int run_graph(drpt_ctx* ctx, const char* name, const std::string& key,
              const std::function<int()>& record) {
  if (!ctx->use_graphs) return record();
  drpt_ctx::Graph& g = ctx->graphs[name];
  if (g.exec == nullptr || g.key != key) {
    if (g.exec) {
      DRPT_HIP(hipStreamSynchronize(ctx->stream));  // a replay may still run
      DRPT_HIP(hipGraphExecDestroy(g.exec));
      g.exec = nullptr;
    }
    g.key.clear();
    DRPT_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = record();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != DRPT_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    DRPT_HIP(e);
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) g.exec = nullptr;
    DRPT_HIP(ei);
    g.key = key;
  }
  DRPT_HIP(hipGraphLaunch(g.exec, ctx->stream));
  return DRPT_OK;
}

}  // namespace drpt
