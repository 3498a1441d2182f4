// Small device utilities.
#include "internal.hpp"

namespace drpt {

namespace {
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

// acc += sum of p[0 .. count): per-thread sums, a wave reduction and one
// atomic per wave
__global__ void k_sum_u32(const uint32_t* __restrict__ p, size_t count,
                          unsigned long long* __restrict__ acc) {
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
  if ((threadIdx.x & 63) == 0 && s != 0) atomicAdd(acc, s);
}
}  // namespace

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
  size_t blocks = (count + 1023) / 1024;
  if (blocks > 1024) blocks = 1024;
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
