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
// weights: [nnodes] (already scaled by the segment length / 2); node_seg: [nnodes]
// grad: [ninst][nterms][nseg] float64, nodes of one segment are contiguous
// seg_first: [nseg + 1] first node of each segment
__global__ __launch_bounds__(kThreads) void k_quad(int n, int nnodes, int nterms, int nseg,
                                                   const float2 *__restrict__ rho, const float2 *__restrict__ lam,
                                                   const float2 *__restrict__ ops, const double *__restrict__ weights,
                                                   const int *__restrict__ seg_first, double *__restrict__ grad) {
  __shared__ double red[kThreads / 64][8];
  const int seg = blockIdx.x;
  const size_t inst = blockIdx.y;
  const size_t nn = (size_t)n * n;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = seg_first[seg]; q < seg_first[seg + 1]; ++q) {
    const float2 *R = rho + (inst * nnodes + q) * nn;
    const float2 *Lm = lam + (inst * nnodes + q) * nn;
    const double w = weights[q];
    // element (b, a) of P - P^dag, P = R Lm, paired with O[a][b]
    for (size_t e = threadIdx.x; e < nn; e += kThreads) {
      const int b = (int)(e / n), a = (int)(e % n);
      double pr = 0, pi = 0, qr = 0, qi = 0;  // P[b][a], P[a][b]
      for (int m = 0; m < n; ++m) {
        const float2 r1 = R[(size_t)b * n + m], l1 = Lm[(size_t)m * n + a];
        pr += (double)r1.x * l1.x - (double)r1.y * l1.y;
        pi += (double)r1.x * l1.y + (double)r1.y * l1.x;
        const float2 r2 = R[(size_t)a * n + m], l2 = Lm[(size_t)m * n + b];
        qr += (double)r2.x * l2.x - (double)r2.y * l2.y;
        qi += (double)r2.x * l2.y + (double)r2.y * l2.x;
      }
      // (P - P^dag)[b][a] = P[b][a] - conj(P[a][b])
      const double dr = pr - qr, di = pi + qi;
      for (int j = 0; j < nterms && j < 8; ++j) {
        const float2 o = ops[(size_t)j * nn + (size_t)a * n + b];
        // Re(-i * o * d) = Im(o * d)
        acc[j] += w * (o.x * di + o.y * dr);
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = 0; j < nterms && j < 8; ++j) {
    double v = acc[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave][j] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < nterms && j < 8; ++j) {
      double s = 0;
      for (int w = 0; w < kThreads / 64; ++w) s += red[w][j];
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
