// Expectation values of saved states (used by the DIA integrators).
#pragma once

#include <hip/hip_runtime.h>

namespace dqh {

// expects[inst][k][s] = Tr(E_k rho_s) = sum_ab (E_k^T)_ab (rho_s)_ab over the saved
// states, FP64 accumulation. grid: (nsave, ninst); block 256.
static __global__ __launch_bounds__(256) void k_expect_states(int64_t n_true, int64_t nsave, int64_t nE,
                                                const float2 *__restrict__ states, const float2 *__restrict__ EopsT,
                                                float2 *__restrict__ expects) {
  __shared__ double red[4][2];
  const int64_t s = blockIdx.x;
  const size_t inst = blockIdx.y;
  const size_t nn = (size_t)n_true * n_true;
  const float2 *R = states + (inst * nsave + s) * nn;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t k = 0; k < nE; ++k) {
    const float2 *Ek = EopsT + k * nn;
    double ar = 0.0, ai = 0.0;
    for (size_t e = threadIdx.x; e < nn; e += 256) {
      const float2 a = Ek[e], r = R[e];
      ar += (double)a.x * r.x - (double)a.y * r.y;
      ai += (double)a.x * r.y + (double)a.y * r.x;
    }
    for (int o = 32; o > 0; o >>= 1) {
      ar += __shfl_xor(ar, o, 64);
      ai += __shfl_xor(ai, o, 64);
    }
    __syncthreads();
    if (lane == 0) {
      red[wave][0] = ar;
      red[wave][1] = ai;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double sr = 0.0, si = 0.0;
      for (int w = 0; w < 4; ++w) {
        sr += red[w][0];
        si += red[w][1];
      }
      expects[(inst * nE + k) * nsave + s] = make_float2((float)sr, (float)si);
    }
  }
}

}  // namespace dqh
