// Quadrature of the adjoint gradient of a piecewise-constant control (gfx950).
//
// For J = Tr(O rho(T)) and H(t) = H0 + sum_j u_j(t) O_j with u_j constant on the
// segments [t_k, t_k+1), the gradient is
//   dJ/du_jk = int_{t_k}^{t_k+1} Re Tr(lambda(t) (-i [O_j, rho(t)])) dt,
// where lambda solves the adjoint Lindblad equation backward from lambda(T) = O.
// With P = rho lambda (so lambda rho = P^dag for Hermitian rho, lambda):
//   Re Tr(lambda (-i)[O, rho]) = Re( -i sum_ab O_ab (P - P^dag)_ba ).
// The integral is a Gauss-Legendre sum over the saved nodes of each segment. One
// workgroup per (instance, segment) accumulates its nodes in FP64; no atomics.
#include <hip/hip_runtime.h>

#include <string>

#include "dqh.h"
#include "dqh_common.h"

namespace dqh {
namespace adj {

constexpr int kThreads = 256;

// rho, lam: [ninst][nnodes][n][n] complex64; ops: [nterms][n][n] complex64
// weights: [nnodes] (already scaled by the segment length / 2)
// grad: [ninst][nterms][nseg] float64, nodes of one segment are contiguous
// seg_first: [nseg + 1] first node of each segment
//
// With P = rho lam, Tr(O (P - P^dag)) = Tr(O P) - conj(Tr(O^dag P)) and
// Tr(O P) = sum_cb (lam O)_cb rho_bc, Tr(O^dag P) = sum_cb (lam O^dag)_cb rho_bc.
// For n <= 64 the node's lam, rho and the operator sit in LDS (FP32 operands,
// FP64 accumulation); larger n read them from global memory.
constexpr int kMaxLds = 64;

__global__ __launch_bounds__(kThreads) void k_quad(int n, int nnodes, int nterms, int nseg,
                                                   const float2 *__restrict__ rho, const float2 *__restrict__ lam,
                                                   const float2 *__restrict__ ops, const double *__restrict__ weights,
                                                   const int *__restrict__ seg_first, double *__restrict__ grad) {
  __shared__ float2 sL[kMaxLds * (kMaxLds + 1)];
  __shared__ float2 sR[kMaxLds * (kMaxLds + 1)];
  __shared__ float2 sO[kMaxLds * (kMaxLds + 1)];
  __shared__ double red[kThreads / 64][2];
  const int seg = blockIdx.x;
  const size_t inst = blockIdx.y;
  const size_t nn = (size_t)n * n;
  const bool lds = n <= kMaxLds;
  const int ld = kMaxLds + 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = 0; j < nterms; ++j) {
    const float2 *Og = ops + (size_t)j * nn;
    if (lds) {
      __syncthreads();
      for (size_t e = threadIdx.x; e < nn; e += kThreads) sO[(e / n) * ld + e % n] = Og[e];
    }
    double acc = 0.0;
    for (int q = seg_first[seg]; q < seg_first[seg + 1]; ++q) {
      const float2 *R = rho + (inst * nnodes + q) * nn;
      const float2 *Lm = lam + (inst * nnodes + q) * nn;
      if (lds) {
        __syncthreads();
        for (size_t e = threadIdx.x; e < nn; e += kThreads) {
          sL[(e / n) * ld + e % n] = Lm[e];
          sR[(e / n) * ld + e % n] = R[e];
        }
        __syncthreads();
      }
      double tr = 0.0, ti = 0.0, ur = 0.0, ui = 0.0;  // Tr(O P), Tr(O^dag P)
      for (size_t e = threadIdx.x; e < nn; e += kThreads) {
        const int c = (int)(e / n), b = (int)(e % n);
        float wr = 0.f, wi = 0.f, vr = 0.f, vi = 0.f;  // (lam O)_cb, (lam O^dag)_cb
        for (int a = 0; a < n; ++a) {
          const float2 l = lds ? sL[c * ld + a] : Lm[(size_t)c * n + a];
          const float2 o = lds ? sO[a * ld + b] : Og[(size_t)a * n + b];
          const float2 od = lds ? sO[b * ld + a] : Og[(size_t)b * n + a];  // (O^dag)_ab = conj(O_ba)
          wr += l.x * o.x - l.y * o.y;
          wi += l.x * o.y + l.y * o.x;
          vr += l.x * od.x + l.y * od.y;
          vi += l.y * od.x - l.x * od.y;
        }
        const float2 r = lds ? sR[b * ld + c] : R[(size_t)b * n + c];
        tr += (double)wr * r.x - (double)wi * r.y;
        ti += (double)wr * r.y + (double)wi * r.x;
        ur += (double)vr * r.x - (double)vi * r.y;
        ui += (double)vr * r.y + (double)vi * r.x;
      }
      // T = Tr(O P) - conj(Tr(O^dag P));  Re(-i T) = Im(T) = ti + ui
      acc += weights[q] * (ti + ui);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __syncthreads();
    if (lane == 0) red[wave][0] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0;
      for (int w = 0; w < kThreads / 64; ++w) s += red[w][0];
      grad[(inst * nterms + j) * nseg + seg] = s;
    }
  }
}

}  // namespace adj
}  // namespace dqh

extern "C" int dqh_adjoint_pwc_quadrature(int64_t ninst, int64_t n, int64_t nnodes, int64_t nterms, int64_t nseg,
                                          const void *rho_nodes, const void *lam_nodes, const void *ops,
                                          const double *weights, const int32_t *seg_first, double *grad,
                                          void *stream) {
  if (ninst <= 0 || n <= 0 || nnodes <= 0 || nseg <= 0 || nterms <= 0 || nterms > 8 || !rho_nodes || !lam_nodes ||
      !ops || !weights || !seg_first || !grad)
    return dqh::fail(DQH_BAD_ARG, "dqh_adjoint_pwc_quadrature: bad arguments (1 <= nterms <= 8)");
  hipLaunchKernelGGL(dqh::adj::k_quad, dim3((unsigned)nseg, (unsigned)ninst), dim3(dqh::adj::kThreads), 0,
                     (hipStream_t)stream, (int)n, (int)nnodes, (int)nterms, (int)nseg,
                     reinterpret_cast<const float2 *>(rho_nodes), reinterpret_cast<const float2 *>(lam_nodes),
                     reinterpret_cast<const float2 *>(ops), weights, reinterpret_cast<const int *>(seg_first), grad);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess)
    return dqh::fail(DQH_HIP_ERROR, std::string("dqh_adjoint_pwc_quadrature: ") + hipGetErrorString(err));
  return DQH_OK;
}
