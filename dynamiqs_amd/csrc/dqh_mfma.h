// Complex64 GEMM tiles on FP32 MFMA (v_mfma_f32_16x16x4_f32) shared by the dense
// kernels: a complex product is four real MFMAs per k-slice, FP32 accumulation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dqh {

typedef float f4 __attribute__((ext_vector_type(4)));


struct Acc {
  f4 re, im;
};

// Device buffers named in a descriptor are global memory: casting through
// address_space(1) lets the compiler use global_load/store with an SGPR base and a
// 32-bit VGPR offset instead of 64-bit flat addresses per access.
template <typename T>
__device__ __forceinline__ T *gp(uint64_t a) {
  return (T *)((__attribute__((address_space(1))) T *)a);
}

// x^e for the step-size controller via the hardware exp2/log2 (single precision is
// ample for a step factor and avoids the constant tables of the FP64 pow, which
// would otherwise stay resident in VGPRs across the whole integration loop).
__device__ __forceinline__ double ctrl_pow(double x, double e) {
  return (double)__builtin_amdgcn_exp2f((float)e * __builtin_amdgcn_logf((float)x));
}

__device__ __forceinline__ void acc_zero(Acc &a) {
  a.re = f4{0.f, 0.f, 0.f, 0.f};
  a.im = f4{0.f, 0.f, 0.f, 0.f};
}

// One complex k-slice: D += A * B with A = ar + i ai (lane's A element), B = b.
__device__ __forceinline__ void cmfma(Acc &a, float ar, float ai, float2 b) {
  a.re = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b.x, a.re, 0, 0, 0);
  a.re = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, b.y, a.re, 0, 0, 0);
  a.im = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b.y, a.im, 0, 0, 0);
  a.im = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b.x, a.im, 0, 0, 0);
}

// One complex k-slice with a conjugated B element: D += A * conj(m).
__device__ __forceinline__ void cmfma_cb(Acc &a, float ar, float ai, float mr, float mi) {
  a.re = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, mr, a.re, 0, 0, 0);
  a.re = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, mi, a.re, 0, 0, 0);
  a.im = __builtin_amdgcn_mfma_f32_16x16x4f32(-ar, mi, a.im, 0, 0, 0);
  a.im = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, mr, a.im, 0, 0, 0);
}

// acc += A[row0:row0+16, 0:K] @ B[0:K, col0:col0+16] for row-major complex64
// matrices (K multiple of 16). Lane (r = lane%16, g = lane/16) loads
// A[row0+r][k0+4g .. k0+4g+3] (32 contiguous bytes) and B[k0+4g+i][col0+r]; MFMA
// slice i then pairs A column k0+4g+i with B row k0+4g+i, which covers the 16
// k's of the block in the order {4g+i}.
__device__ __forceinline__ void tile_mac(Acc &acc, const float2 *__restrict__ A, int lda,
                                         const float2 *__restrict__ B, int ldb, int row0, int col0,
                                         int K, int lane) {
  const int r = lane & 15, g = lane >> 4;
  const float2 *Ap = A + (size_t)(row0 + r) * lda + 4 * g;
  const float2 *Bp = B + (size_t)(4 * g) * ldb + col0 + r;
#pragma unroll 2
  for (int k0 = 0; k0 < K; k0 += 16) {
    const float4 a01 = *reinterpret_cast<const float4 *>(Ap + k0);
    const float4 a23 = *reinterpret_cast<const float4 *>(Ap + k0 + 2);
    const float2 *bp = Bp + (size_t)k0 * ldb;
    const float2 b0 = bp[0];
    const float2 b1 = bp[ldb];
    const float2 b2 = bp[2 * ldb];
    const float2 b3 = bp[3 * ldb];
    cmfma(acc, a01.x, a01.y, b0);
    cmfma(acc, a01.z, a01.w, b1);
    cmfma(acc, a23.x, a23.y, b2);
    cmfma(acc, a23.z, a23.w, b3);
  }
}

// Store a 16x16 accumulator tile: lane (r, g) holds D[4g+v][r], v = 0..3.
__device__ __forceinline__ void tile_store(const Acc &acc, float2 *__restrict__ C, int ldc, int row0,
                                           int col0, int lane) {
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int v = 0; v < 4; ++v) C[(size_t)(row0 + 4 * g + v) * ldc + col0 + r] = make_float2(acc.re[v], acc.im[v]);
}

}  // namespace dqh
