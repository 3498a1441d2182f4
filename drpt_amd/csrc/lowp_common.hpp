// Device pieces shared by the low-precision leaf k-NN kernels (leafknn.hip,
// leafpool.hip): MFMA operand types, the certified filter bound, and the
// global-to-LDS load helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace drpt {
namespace lowp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Certified filter bound. For a row i with quantised squared norm nb_i,
// quantisation error e_i = |x~_i - x_i| and the leaf's largest squared norm and
// error (nm = its root, emax), every leafmate j with exact distance
// D(i, j) <= B has approximate distance a_ij = |x~_i|^2 + |x~_j|^2 - 2 G_ij
// (FP32, Gram of the quantised rows) at most
//   A_cut(B) = (sqrt(B / (1 - (d + 8) 2^-24)) + e_i + emax)^2 + eta,
//   eta = (2 d + 16) 2^-24 (n_i + nm)^2  (a bound of a's FP32 error).
// Returned rounded up to float (+inf when B is not finite). The absolute 1e-30 covers errors the relative terms do not: squared norms,
// Gram entries and FP8 rescalings that fall below the FP32 normal range
// (at most 2 d 2^-126 < 1e-33 for d < 2^16).
__device__ __forceinline__ float acut_of(float B, float nb_i, float e_i, double nm, float emax,
                                         int d) {
  if (!(B < INFINITY)) return INFINITY;
  const double u = 5.9604644775390625e-08;  // 2^-24
  const double ni = sqrt((double)nb_i * (1.0 + (d + 4.0) * u));
  const double eta = (2.0 * d + 16.0) * u * (ni + nm) * (ni + nm);
  const double sb = sqrt((double)B / (1.0 - (d + 8.0) * u)) + (double)e_i + (double)emax;
  return (float)((sb * sb + eta) * (1.0 + 1e-6) + 1e-30);
}

// vmcnt-only wait (expcnt and lgkmcnt left at their maxima), gfx9 encoding
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | (0xF << 8) | ((N >> 4) << 14));
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// 16 bytes per lane from global memory straight into LDS at lds_base + 16 *
// lane (global_load_lds_dwordx4, base in M0). Issued from asm so that the
// compiler does not wait for it before unrelated LDS reads; completion is
// awaited explicitly (wait_vmcnt). No register is written by the load; its
// order against the ring barrier (also volatile asm) is kept.
__device__ __forceinline__ void lds_dma16(const void* gptr, uint32_t lds_base) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off"
               :
               : "v"(gptr), "s"(lds_base)
               : "m0");
}

// Workgroup barrier for the ring: this wave's LDS operations complete, then
// s_barrier; global loads stay in flight (__syncthreads would wait for all of
// them). The memory clobber keeps the compiler from moving memory accesses
// across it.
__device__ __forceinline__ void ring_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

}  // namespace lowp
}  // namespace drpt
