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

// Same integrand for control operators stored by diagonals (O[a][a + p] =
// diags[u][a + p]): with the integrand Im(Tr(lam O rho) - Tr(lam rho O)),
//   Tr(lam O rho) = sum_u sum_{a,c} lam[c][a] O[a][a+p] rho[a+p][c],
//   Tr(lam rho O) = sum_u sum_{a,b} lam[b+p][a] rho[a][b] O[b][b+p],
// 2 d n^2 complex products per node instead of 2 n^3 (d diagonals). lam is
// staged in LDS (n <= 64), rho is read from global memory along its rows.
__global__ __launch_bounds__(kThreads) void k_quad_dia(int n, int nnodes, int nseg, dqh_dia_ops ops,
                                                       const float2 *__restrict__ rho, const float2 *__restrict__ lam,
                                                       const double *__restrict__ weights,
                                                       const int *__restrict__ seg_first, double *__restrict__ grad) {
  __shared__ float2 sL[kMaxLds * (kMaxLds + 1)];
  __shared__ double red[kThreads / 64];
  const int seg = blockIdx.x;
  const size_t inst = blockIdx.y;
  const size_t nn = (size_t)n * n;
  const int ld = kMaxLds + 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nterms = ops.nops;
  double acc[DQH_DIA_MAXOPS];
  for (int j = 0; j < DQH_DIA_MAXOPS; ++j) acc[j] = 0.0;
  for (int q = seg_first[seg]; q < seg_first[seg + 1]; ++q) {
    const float2 *R = rho + (inst * nnodes + q) * nn;
    const float2 *Lm = lam + (inst * nnodes + q) * nn;
    __syncthreads();
    for (size_t e = threadIdx.x; e < nn; e += kThreads) sL[(e / n) * ld + e % n] = Lm[e];
    __syncthreads();
    const double wq = weights[q];
    for (int j = 0; j < nterms; ++j) {
      const float2 *D = reinterpret_cast<const float2 *>(ops.diags[j]);
      double im = 0.0;
      for (size_t e = threadIdx.x; e < nn; e += kThreads) {
        const int a = (int)(e / n), c = (int)(e % n);  // rho[a][c] is contiguous in e
        const float2 r = R[e];
        float sr = 0.f, si = 0.f;
        for (int u = 0; u < ops.noff[j]; ++u) {
          const int p = ops.off[j][u];
          // Tr(lam O rho): rho[a][c] pairs with lam[c][a - p] O[a - p][a]
          const int a2 = a - p;
          if (a2 >= 0 && a2 < n) {
            const float2 o = D[(size_t)u * n + a];
            const float2 l = sL[c * ld + a2];
            const float lr = l.x * o.x - l.y * o.y, li = l.x * o.y + l.y * o.x;
            sr += lr * r.x - li * r.y;
            si += lr * r.y + li * r.x;
          }
          // Tr(lam rho O): rho[a][c] pairs with lam[c + p][a] O[c][c + p]
          const int c2 = c + p;
          if (c2 >= 0 && c2 < n) {
            const float2 o = D[(size_t)u * n + c2];
            const float2 l = sL[c2 * ld + a];
            const float lr = l.x * o.x - l.y * o.y, li = l.x * o.y + l.y * o.x;
            sr -= lr * r.x - li * r.y;
            si -= lr * r.y + li * r.x;
          }
        }
        im += (double)si;
      }
      acc[j] += wq * im;
    }
  }
  for (int j = 0; j < nterms; ++j) {
    double v = acc[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0;
      for (int w = 0; w < kThreads / 64; ++w) s += red[w];
      grad[(inst * nterms + j) * nseg + seg] = s;
    }
  }
}

// Gradient with respect to a jump-rate parameter: for L = sqrt(kappa) A,
// d/dkappa of the Lindbladian is the dissipator D[A], so
//   dJ/dkappa = int Tr(lam D[A] rho) dt,
//   Tr(lam D[A] rho) = sum_cb (A^dag lam A - (lam N + N lam) / 2)_cb rho_bc,  N = A^dag A.
// One workgroup per (instance, segment); lam, A, N and P = lam A sit in LDS
// (n <= 64), rho is read from global memory once per node; FP64 accumulation.
__global__ __launch_bounds__(kThreads) void k_rate_quad(int n, int nnodes, int nrates, int nseg,
                                                        const float2 *__restrict__ rho, const float2 *__restrict__ lam,
                                                        const float2 *__restrict__ Aops, const float2 *__restrict__ Nops,
                                                        const double *__restrict__ weights,
                                                        const int *__restrict__ seg_first, double *__restrict__ grad) {
  __shared__ float2 sL[kMaxLds * (kMaxLds + 1)];
  __shared__ float2 sA[kMaxLds * (kMaxLds + 1)];
  __shared__ float2 sN[kMaxLds * (kMaxLds + 1)];
  __shared__ float2 sP[kMaxLds * (kMaxLds + 1)];
  __shared__ double red[kThreads / 64];
  const int seg = blockIdx.x;
  const size_t inst = blockIdx.y;
  const size_t nn = (size_t)n * n;
  const int ld = kMaxLds + 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int r = 0; r < nrates; ++r) {
    __syncthreads();
    for (size_t e = threadIdx.x; e < nn; e += kThreads) {
      sA[(e / n) * ld + e % n] = Aops[(size_t)r * nn + e];
      sN[(e / n) * ld + e % n] = Nops[(size_t)r * nn + e];
    }
    double acc = 0.0;
    for (int q = seg_first[seg]; q < seg_first[seg + 1]; ++q) {
      const float2 *R = rho + (inst * nnodes + q) * nn;
      const float2 *Lm = lam + (inst * nnodes + q) * nn;
      __syncthreads();
      for (size_t e = threadIdx.x; e < nn; e += kThreads) sL[(e / n) * ld + e % n] = Lm[e];
      __syncthreads();
      for (size_t e = threadIdx.x; e < nn; e += kThreads) {  // P = lam A
        const int c = (int)(e / n), b = (int)(e % n);
        float pr = 0.f, pi = 0.f;
        for (int a = 0; a < n; ++a) {
          const float2 l = sL[c * ld + a], x = sA[a * ld + b];
          pr += l.x * x.x - l.y * x.y;
          pi += l.x * x.y + l.y * x.x;
        }
        sP[c * ld + b] = make_float2(pr, pi);
      }
      __syncthreads();
      double tr = 0.0;  // Re Tr(M rho), M = A^dag P - (lam N + N lam) / 2
      for (size_t e = threadIdx.x; e < nn; e += kThreads) {
        const int b = (int)(e / n), c = (int)(e % n);  // rho_bc is contiguous in e
        float mr = 0.f, mi = 0.f;
        for (int a = 0; a < n; ++a) {
          const float2 ad = sA[a * ld + c];  // (A^dag)_ca = conj(A_ac)
          const float2 p = sP[a * ld + b];
          mr += ad.x * p.x + ad.y * p.y;
          mi += ad.x * p.y - ad.y * p.x;
          const float2 l1 = sL[c * ld + a], n1 = sN[a * ld + b];  // (lam N)_cb
          const float2 n2 = sN[c * ld + a], l2 = sL[a * ld + b];  // (N lam)_cb
          mr -= 0.5f * ((l1.x * n1.x - l1.y * n1.y) + (n2.x * l2.x - n2.y * l2.y));
          mi -= 0.5f * ((l1.x * n1.y + l1.y * n1.x) + (n2.x * l2.y + n2.y * l2.x));
        }
        const float2 rv = R[e];
        tr += (double)mr * rv.x - (double)mi * rv.y;
      }
      acc += weights[q] * tr;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __syncthreads();
    if (lane == 0) red[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0;
      for (int w = 0; w < kThreads / 64; ++w) s += red[w];
      grad[(inst * nrates + r) * nseg + seg] = s;
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

extern "C" int dqh_adjoint_rate_quadrature(int64_t ninst, int64_t n, int64_t nnodes, int64_t nrates, int64_t nseg,
                                           const void *rho_nodes, const void *lam_nodes, const void *Aops,
                                           const void *Nops, const double *weights, const int32_t *seg_first,
                                           double *grad, void *stream) {
  if (ninst <= 0 || n <= 0 || n > dqh::adj::kMaxLds || nnodes <= 0 || nseg <= 0 || nrates <= 0 || nrates > 8 ||
      !rho_nodes || !lam_nodes || !Aops || !Nops || !weights || !seg_first || !grad)
    return dqh::fail(DQH_BAD_ARG, "dqh_adjoint_rate_quadrature: bad arguments (n <= 64, 1 <= nrates <= 8)");
  hipLaunchKernelGGL(dqh::adj::k_rate_quad, dim3((unsigned)nseg, (unsigned)ninst), dim3(dqh::adj::kThreads), 0,
                     (hipStream_t)stream, (int)n, (int)nnodes, (int)nrates, (int)nseg,
                     reinterpret_cast<const float2 *>(rho_nodes), reinterpret_cast<const float2 *>(lam_nodes),
                     reinterpret_cast<const float2 *>(Aops), reinterpret_cast<const float2 *>(Nops), weights,
                     reinterpret_cast<const int *>(seg_first), grad);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess)
    return dqh::fail(DQH_HIP_ERROR, std::string("dqh_adjoint_rate_quadrature: ") + hipGetErrorString(err));
  return DQH_OK;
}

extern "C" int dqh_adjoint_pwc_quadrature_dia(int64_t ninst, int64_t n, int64_t nnodes, int64_t nseg,
                                              const void *rho_nodes, const void *lam_nodes, const dqh_dia_ops *ops,
                                              const double *weights, const int32_t *seg_first, double *grad,
                                              void *stream) {
  if (ninst <= 0 || n <= 0 || n > dqh::adj::kMaxLds || nnodes <= 0 || nseg <= 0 || !ops || !rho_nodes || !lam_nodes ||
      !weights || !seg_first || !grad)
    return dqh::fail(DQH_BAD_ARG, "dqh_adjoint_pwc_quadrature_dia: bad arguments (n <= 64)");
  if (ops->nops < 1 || ops->nops > DQH_DIA_MAXOPS)
    return dqh::fail(DQH_BAD_ARG, "dqh_adjoint_pwc_quadrature_dia: 1 <= number of control operators <= 8");
  for (int k = 0; k < ops->nops; ++k) {
    if (ops->noff[k] < 1 || ops->noff[k] > DQH_DIA_MAXOFF || !ops->diags[k])
      return dqh::fail(DQH_BAD_ARG, "dqh_adjoint_pwc_quadrature_dia: bad number of diagonals or null diagonals");
    for (int u = 0; u < ops->noff[k]; ++u)
      if (ops->off[k][u] <= -n || ops->off[k][u] >= n)
        return dqh::fail(DQH_BAD_ARG, "dqh_adjoint_pwc_quadrature_dia: offset outside (-n, n)");
  }
  hipLaunchKernelGGL(dqh::adj::k_quad_dia, dim3((unsigned)nseg, (unsigned)ninst), dim3(dqh::adj::kThreads), 0,
                     (hipStream_t)stream, (int)n, (int)nnodes, (int)nseg, *ops,
                     reinterpret_cast<const float2 *>(rho_nodes), reinterpret_cast<const float2 *>(lam_nodes), weights,
                     reinterpret_cast<const int *>(seg_first), grad);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess)
    return dqh::fail(DQH_HIP_ERROR, std::string("dqh_adjoint_pwc_quadrature_dia: ") + hipGetErrorString(err));
  return DQH_OK;
}
