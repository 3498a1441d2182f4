// Small device utilities.
#include "internal.hpp"

namespace drpt {

namespace {
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
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

}  // namespace drpt
