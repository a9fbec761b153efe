// Adaptive Rouchon 1-2 and 2-3 pairs for dense operators on gfx950, all in FP64:
// complex128 products on FP64 MFMA (v_mfma_f64_16x16x4_f64), the Cholesky factor,
// its triangular inverse and the Gauss-Jordan inverses in registers, FP64 state and
// error norm.
//
// One 512-thread workgroup owns one instance (np <= 64) for the whole solve. A
// trial step builds the no-jump propagators of the embedded pair from G(tau) at the
// stage times, applies the high and the low Kraus map to the same rho (each after
// its own Cholesky normalisation), and takes (rho_high - rho_low) / 2 as the error.
// Matrices live in the instance's work slots (global memory, L2-resident for one
// instance); products are 16x16 output tiles spread over the 8 waves.
//
// Register ownership for the factorisations: thread x owns the elements
// e = x + 512 m (m < 8, e < np^2) of an np x np matrix, i.e. at np = 64 column
// x % 64 of rows x / 64 + 8 m. A column or row needed by every thread at step k
// goes through LDS (double-buffered, one barrier per step).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "dqh.h"
#include "dqh_common.h"
#include "dqh_tsit5_tableau.h"

namespace dqh {
namespace ar {

constexpr int kWaves = 8;
constexpr int kThreads = kWaves * 64;
constexpr int kMaxNP = 64;
constexpr int kOwn = kMaxNP * kMaxNP / kThreads;  // elements per thread in the factorisations

typedef double d4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ T *gp(uint64_t a) {
  return (T *)((__attribute__((address_space(1))) T *)a);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// work slots (np^2 complex128 each)
enum Slot {
  K1, GB, T1, T2, U1, ULO, KK2, KK3, KK4, UA, UB, PP, Q2, INV, S, W, R, X0, X1, X2, Y, ST1, ST2, O1, O2, OLO, OHI,
  kSlots
};

struct Shared {
  double2 colb[2][kMaxNP];
  double2 rowb[2][kMaxNP];
  double2 rowc[2][kMaxNP];
  double dg[kMaxNP];
  double red[kWaves];
};

struct Ctx {
  int n, np, nL, nn;
  int normalize;
  const double2 *G0, *LdL, *L, *P, *O;
  double2 *w;
  double2 *rho;
  Shared *sh;
};

__device__ __forceinline__ double2 *slot(const Ctx &c, int s) { return c.w + (size_t)s * c.nn; }

// ------------------------------------------------------------------ reductions
__device__ __forceinline__ double block_sum(double v, Shared *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh->red[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) s += sh->red[w];
  return s;
}

// ------------------------------------------------------------------ products
// C = alpha op(A) op(B) (+ C with acc); op = conjugate transpose when TA / TB.
// Tile (i0, j0) per wave; lane (r = lane % 16, g = lane / 16) supplies op(A)[i0 + r][k]
// and op(B)[k][j0 + r] for k = k0 + g, and holds D[4v + g][r], v < 4, of the tile (the
// FP64 instruction interleaves the rows of its result, unlike the FP32 16x16x4 one,
// whose lane holds rows 4g..4g+3). C must not alias A or B. Ends with a barrier.
template <bool TA, bool TB>
__device__ __noinline__ void gemm(const Ctx &c, double2 *C, const double2 *A, const double2 *B, double alpha, bool acc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int np = c.np, NT = np >> 4;
  const int r = lane & 15, g = lane >> 4;
  for (int t = wave; t < NT * NT; t += kWaves) {
    const int i0 = (t / NT) * 16, j0 = (t % NT) * 16;
    d4 re = {0.0, 0.0, 0.0, 0.0}, im = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < np; k0 += 4) {
      const int k = k0 + g;
      double2 a = TA ? A[(size_t)k * np + i0 + r] : A[(size_t)(i0 + r) * np + k];
      double2 b = TB ? B[(size_t)(j0 + r) * np + k] : B[(size_t)k * np + j0 + r];
      if (TA) a.y = -a.y;
      if (TB) b.y = -b.y;
      re = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, re, 0, 0, 0);
      re = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, re, 0, 0, 0);
      im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, im, 0, 0, 0);
      im = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, im, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const size_t e = (size_t)(i0 + 4 * v + g) * np + j0 + r;
      double2 o = make_double2(alpha * re[v], alpha * im[v]);
      if (acc) {
        const double2 q = C[e];
        o.x += q.x;
        o.y += q.y;
      }
      C[e] = o;
    }
  }
  __syncthreads();
}

// C = alpha M X M^dag (+ C); T: scratch
__device__ void sand_n(const Ctx &c, double2 *C, const double2 *M, const double2 *X, double2 *T, double alpha,
                       bool acc) {
  gemm<false, false>(c, T, M, X, 1.0, false);
  gemm<false, true>(c, C, T, M, alpha, acc);
}

// C = alpha M^dag X M (+ C); T: scratch
__device__ void sand_c(const Ctx &c, double2 *C, const double2 *M, const double2 *X, double2 *T, double alpha,
                       bool acc) {
  gemm<true, false>(c, T, M, X, 1.0, false);
  gemm<false, false>(c, C, T, M, alpha, acc);
}

__device__ void zero(const Ctx &c, double2 *C) {
  for (int e = threadIdx.x; e < c.nn; e += kThreads) C[e] = make_double2(0.0, 0.0);
  __syncthreads();
}

// C = alpha sum_k L_k X L_k^dag (ADJ: alpha sum_k L_k^dag X L_k) (+ C); T: scratch
template <bool ADJ>
__device__ void diss(const Ctx &c, double2 *C, const double2 *X, double2 *T, double alpha, bool acc) {
  if (c.nL == 0) {
    if (!acc) zero(c, C);
    return;
  }
  for (int k = 0; k < c.nL; ++k) {
    const double2 *Lk = c.L + (size_t)k * c.nn;
    if (ADJ) {
      gemm<true, false>(c, T, Lk, X, 1.0, false);
      gemm<false, false>(c, C, T, Lk, alpha, acc || k > 0);
    } else {
      gemm<false, false>(c, T, Lk, X, 1.0, false);
      gemm<false, true>(c, C, T, Lk, alpha, acc || k > 0);
    }
  }
}

// ------------------------------------------------------------ factorisations
// W = T^-1 with S = T T^dag (T lower triangular, the Cholesky factor). Right-looking
// Cholesky on the owned elements (column k broadcast raw: T_rk = A_rk / sqrt(A_kk),
// A_rq -= A_rk conj(A_qk) / A_kk), then forward elimination of T W = I (row k of W
// and column k of T broadcast). Ends with a barrier.
__device__ __noinline__ void chol_inv(const Ctx &c, const double2 *Sm, double2 *Wm) {
  Shared *sh = c.sh;
  const int np = c.np, nn = c.nn;
  double2 a[kOwn], b[kOwn];
#pragma unroll
  for (int m = 0; m < kOwn; ++m) {
    const int e = threadIdx.x + kThreads * m;
    a[m] = e < nn ? Sm[e] : make_double2(0.0, 0.0);
  }
  for (int k = 0; k < np; ++k) {
    const int buf = k & 1;
#pragma unroll
    for (int m = 0; m < kOwn; ++m) {
      const int e = threadIdx.x + kThreads * m;
      if (e < nn && e % np == k && e / np >= k) sh->colb[buf][e / np] = a[m];
    }
    __syncthreads();
    const double dkk = sh->colb[buf][k].x;
    const double sq = sqrt(dkk), isq = 1.0 / sq, idk = 1.0 / dkk;
#pragma unroll
    for (int m = 0; m < kOwn; ++m) {
      const int e = threadIdx.x + kThreads * m;
      if (e >= nn) continue;
      const int r = e / np, q = e % np;
      if (q == k && r >= k) {
        a[m] = r == k ? make_double2(sq, 0.0) : make_double2(a[m].x * isq, a[m].y * isq);
      } else if (q > k && r >= q) {
        const double2 x = sh->colb[buf][r], y = sh->colb[buf][q];
        a[m].x -= (x.x * y.x + x.y * y.y) * idk;
        a[m].y -= (x.y * y.x - x.x * y.y) * idk;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < kOwn; ++m) {
    const int e = threadIdx.x + kThreads * m;
    const int r = e / np, q = e % np;
    b[m] = make_double2(e < nn && r == q ? 1.0 : 0.0, 0.0);
    if (e < nn && r == q) sh->dg[r] = a[m].x;
  }
  __syncthreads();
  for (int k = 0; k < np; ++k) {
    const int buf = k & 1;
    const double idg = 1.0 / sh->dg[k];
#pragma unroll
    for (int m = 0; m < kOwn; ++m) {
      const int e = threadIdx.x + kThreads * m;
      if (e >= nn) continue;
      const int r = e / np, q = e % np;
      if (r == k && q <= k) {
        b[m].x *= idg;
        b[m].y *= idg;
        sh->rowb[buf][q] = b[m];
      }
      if (q == k && r > k) sh->colb[buf][r] = a[m];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kOwn; ++m) {
      const int e = threadIdx.x + kThreads * m;
      if (e >= nn) continue;
      const int r = e / np, q = e % np;
      if (r > k && q <= k) {
        const double2 f = cmul(sh->colb[buf][r], sh->rowb[buf][q]);
        b[m].x -= f.x;
        b[m].y -= f.y;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < kOwn; ++m) {
    const int e = threadIdx.x + kThreads * m;
    if (e < nn) Wm[e] = b[m];
  }
  __syncthreads();
}

// Ainv = A^-1 by Gauss-Jordan elimination without pivoting (A is a no-jump
// propagator I + O(h G)); a vanishing pivot yields non-finite entries, which the
// error norm turns into a rejected step. Ends with a barrier.
__device__ __noinline__ void gj_inv(const Ctx &c, const double2 *A, double2 *Ainv) {
  Shared *sh = c.sh;
  const int np = c.np, nn = c.nn;
  double2 a[kOwn], b[kOwn];
#pragma unroll
  for (int m = 0; m < kOwn; ++m) {
    const int e = threadIdx.x + kThreads * m;
    a[m] = e < nn ? A[e] : make_double2(0.0, 0.0);
    b[m] = make_double2(e < nn && e / np == e % np ? 1.0 : 0.0, 0.0);
  }
  for (int k = 0; k < np; ++k) {
    const int buf = k & 1;
#pragma unroll
    for (int m = 0; m < kOwn; ++m) {
      const int e = threadIdx.x + kThreads * m;
      if (e >= nn) continue;
      const int r = e / np, q = e % np;
      if (q == k) sh->colb[buf][r] = a[m];
      if (r == k) {
        sh->rowb[buf][q] = a[m];
        sh->rowc[buf][q] = b[m];
      }
    }
    __syncthreads();
    const double2 p = sh->colb[buf][k];
    const double pn = 1.0 / (p.x * p.x + p.y * p.y);
    const double2 ip = make_double2(p.x * pn, -p.y * pn);
#pragma unroll
    for (int m = 0; m < kOwn; ++m) {
      const int e = threadIdx.x + kThreads * m;
      if (e >= nn) continue;
      const int r = e / np, q = e % np;
      const double2 ra = cmul(sh->rowb[buf][q], ip), rb = cmul(sh->rowc[buf][q], ip);
      if (r == k) {
        a[m] = ra;
        b[m] = rb;
      } else {
        const double2 f = sh->colb[buf][r];
        const double2 fa = cmul(f, ra), fb = cmul(f, rb);
        a[m].x -= fa.x;
        a[m].y -= fa.y;
        b[m].x -= fb.x;
        b[m].y -= fb.y;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < kOwn; ++m) {
    const int e = threadIdx.x + kThreads * m;
    if (e < nn) Ainv[e] = b[m];
  }
  __syncthreads();
}

// ---------------------------------------------------------------- generator
// f_j(tau): the cubic of the table interval that contains tau (clamped), Horner
__device__ __forceinline__ double2 mod_eval(const dqh_ar_inst &p, int j, double tau) {
  const int64_t nint = p.mod_nint;
  int64_t i = (int64_t)floor((tau - p.mod_t0) / p.mod_dt);
  i = i < 0 ? 0 : (i >= nint ? nint - 1 : i);
  const double x = tau - (p.mod_t0 + (double)i * p.mod_dt);
  const double2 *q = gp<const double2>(p.mod_coef) + ((size_t)j * nint + i) * 4;
  double re = q[0].x, im = q[0].y;
#pragma unroll
  for (int m = 1; m < 4; ++m) {
    re = re * x + q[m].x;
    im = im * x + q[m].y;
  }
  return make_double2(re, im);
}

// searchsorted(times, t, 'right') - 1; -1 before the first breakpoint
__device__ __forceinline__ int pwc_segment(const double *times, int nseg, double t) {
  int lo = 0, hi = nseg;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (times[mid] <= t)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

// Gout = G(tau) with the piecewise-constant terms of segment seg. Ends with a barrier.
__device__ __noinline__ void gen(const Ctx &c, const dqh_ar_inst &p, double2 *Gout, double tau, int seg) {
  const int npwc = (int)p.npwc, nmod = (int)p.nmod, nv = (int)p.nseg - 1;
  const double2 *vals = gp<const double2>(p.pwc_vals);
  double2 f[DQH_MAX_MOD];
  double2 cp[DQH_MAX_MOD];
#pragma unroll
  for (int j = 0; j < DQH_MAX_MOD; ++j) {
    f[j] = j < nmod ? mod_eval(p, j, tau) : make_double2(0.0, 0.0);
    cp[j] = (j < npwc && seg >= 0 && seg < nv) ? vals[(size_t)j * nv + seg] : make_double2(0.0, 0.0);
  }
  for (int e = threadIdx.x; e < c.nn; e += kThreads) {
    double2 g = c.G0[e];
#pragma unroll
    for (int j = 0; j < DQH_MAX_MOD; ++j) {
      if (j < npwc) {
        const double2 o = cmul(cp[j], c.P[(size_t)j * c.nn + e]);
        g.x += o.x;
        g.y += o.y;
      }
      if (j < nmod) {
        const double2 o = cmul(f[j], c.O[(size_t)j * c.nn + e]);
        g.x += o.x;
        g.y += o.y;
      }
    }
    Gout[e] = g;
  }
  __syncthreads();
}

// ------------------------------------------------------------------ the maps
// R = T^-dag rho T^-1 with S = T T^dag (cholesky_normalize, rouchon_integrator.py:552-582)
__device__ const double2 *normalised(const Ctx &c) {
  chol_inv(c, slot(c, S), slot(c, W));
  sand_c(c, slot(c, R), slot(c, W), c.rho, slot(c, T1), 1.0, false);
  return slot(c, R);
}

// KrausEuler (rouchon_integrator.py:226-269): OUT = M0 R M0^dag + h D(R),
// S = M0^dag M0 + h sum L^dag L
__device__ void euler_map(const Ctx &c, const double2 *M0, double h, double2 *OUT) {
  const double2 *Rp = c.rho;
  if (c.normalize) {
    double2 *Sm = slot(c, S);
    gemm<true, false>(c, Sm, M0, M0, 1.0, false);
    for (int e = threadIdx.x; e < c.nn; e += kThreads) {
      Sm[e].x += h * c.LdL[e].x;
      Sm[e].y += h * c.LdL[e].y;
    }
    __syncthreads();
    Rp = normalised(c);
  }
  sand_n(c, OUT, M0, Rp, slot(c, T1), 1.0, false);
  diss<false>(c, OUT, Rp, slot(c, T1), h, true);
}

// KrausHeun2 (c = (0, 1), a10 = 1, b = (1/2, 1/2); KrausMapRK :83-224) with U = U(1):
// S = U^dag U + h/2 D*(U^dag U) + h/2 (Z + h D*(Z)), Z = U^dag LdL U;
// OUT = A + h/2 B + h/2 D(A + h B), A = U R U^dag, B = U D(R) U^dag
__device__ void heun2_map(const Ctx &c, const double2 *U, double h, double2 *OUT) {
  double2 *t1 = slot(c, T1), *t2 = slot(c, T2), *o1 = slot(c, O1), *o2 = slot(c, O2);
  const double2 *Rp = c.rho;
  if (c.normalize) {
    double2 *Sm = slot(c, S);
    gemm<true, false>(c, o1, U, U, 1.0, false);
    for (int e = threadIdx.x; e < c.nn; e += kThreads) Sm[e] = o1[e];
    __syncthreads();
    diss<true>(c, Sm, o1, t1, 0.5 * h, true);
    sand_c(c, o2, U, c.LdL, t1, 1.0, false);
    for (int e = threadIdx.x; e < c.nn; e += kThreads) {
      Sm[e].x += 0.5 * h * o2[e].x;
      Sm[e].y += 0.5 * h * o2[e].y;
    }
    __syncthreads();
    diss<true>(c, Sm, o2, t1, 0.5 * h * h, true);
    Rp = normalised(c);
  }
  double2 *x0 = slot(c, X0), *st1 = slot(c, ST1);
  diss<false>(c, x0, Rp, t1, 1.0, false);
  sand_n(c, t2, U, Rp, t1, 1.0, false);
  sand_n(c, o1, U, x0, t1, 1.0, false);
  for (int e = threadIdx.x; e < c.nn; e += kThreads) {
    const double2 A = t2[e], B = o1[e];
    st1[e] = make_double2(A.x + h * B.x, A.y + h * B.y);
    OUT[e] = make_double2(A.x + 0.5 * h * B.x, A.y + 0.5 * h * B.y);
  }
  __syncthreads();
  diss<false>(c, OUT, st1, t1, 0.5 * h, true);
}

// KrausHeun3 (c = (0, 1/3, 2/3), a10 = 1/3, a21 = 2/3, b = (1/4, 0, 3/4)) with
// Ua = U(1/3), Ub = U(2/3), U1 = U(1), P = Ub Ua^-1, Q2 = U1 Ub^-1:
// S = U1^dag U1 + h/4 D*(U1^dag U1) + 3h/4 Ub^dag O Ub + h^2/2 Z + h^3/6 D*(Z) with
// O = D*(Q2^dag Q2), Z = Ua^dag D*(P^dag O P) Ua (the adjoint pass S_stage);
// stages s1 = Ua (R + h/3 D(R)) Ua^dag, s2 = Ub R Ub^dag + 2h/3 P D(s1) P^dag;
// OUT = U1 (R + h/4 D(R)) U1^dag + 3h/4 Q2 D(s2) Q2^dag
__device__ void heun3_map(const Ctx &c, double h, double2 *OUT) {
  double2 *t1 = slot(c, T1), *t2 = slot(c, T2), *o1 = slot(c, O1), *o2 = slot(c, O2);
  const double2 *u1 = slot(c, U1), *ua = slot(c, UA), *ub = slot(c, UB), *pp = slot(c, PP), *q2 = slot(c, Q2);
  const double2 *Rp = c.rho;
  if (c.normalize) {
    double2 *Sm = slot(c, S);
    gemm<true, false>(c, o1, u1, u1, 1.0, false);
    for (int e = threadIdx.x; e < c.nn; e += kThreads) Sm[e] = o1[e];
    __syncthreads();
    diss<true>(c, Sm, o1, t1, 0.25 * h, true);
    gemm<true, false>(c, t2, q2, q2, 1.0, false);
    diss<true>(c, o2, t2, t1, 1.0, false);
    sand_c(c, Sm, ub, o2, t1, 0.75 * h, true);
    sand_c(c, t2, pp, o2, t1, 1.0, false);
    diss<true>(c, o1, t2, t1, 1.0, false);
    sand_c(c, t2, ua, o1, t1, 1.0, false);
    for (int e = threadIdx.x; e < c.nn; e += kThreads) {
      Sm[e].x += 0.5 * h * h * t2[e].x;
      Sm[e].y += 0.5 * h * h * t2[e].y;
    }
    __syncthreads();
    diss<true>(c, Sm, t2, t1, h * h * h / 6.0, true);
    Rp = normalised(c);
  }
  double2 *x0 = slot(c, X0), *x1 = slot(c, X1), *x2 = slot(c, X2), *y = slot(c, Y);
  double2 *st1 = slot(c, ST1), *st2 = slot(c, ST2);
  diss<false>(c, x0, Rp, t1, 1.0, false);
  for (int e = threadIdx.x; e < c.nn; e += kThreads)
    y[e] = make_double2(Rp[e].x + h / 3.0 * x0[e].x, Rp[e].y + h / 3.0 * x0[e].y);
  __syncthreads();
  sand_n(c, st1, ua, y, t1, 1.0, false);
  diss<false>(c, x1, st1, t1, 1.0, false);
  sand_n(c, st2, ub, Rp, t1, 1.0, false);
  sand_n(c, st2, pp, x1, t1, 2.0 * h / 3.0, true);
  diss<false>(c, x2, st2, t1, 1.0, false);
  for (int e = threadIdx.x; e < c.nn; e += kThreads)
    y[e] = make_double2(Rp[e].x + 0.25 * h * x0[e].x, Rp[e].y + 0.25 * h * x0[e].y);
  __syncthreads();
  sand_n(c, OUT, u1, y, t1, 1.0, false);
  sand_n(c, OUT, q2, x2, t1, 0.75 * h, true);
}

// ------------------------------------------------------------- propagators
// Order 2 (explicit midpoint / Euler on dU/dt = G U): U1 = I + h Gm (I + h/2 G(t)),
// ULO = I + h Gm (KrausEuler's M0), Gm = G(t + h/2).
__device__ void props2(const Ctx &c, const dqh_ar_inst &p, double t, double h, int seg) {
  double2 *k1 = slot(c, K1), *gb = slot(c, GB), *t1 = slot(c, T1), *u1 = slot(c, U1), *ulo = slot(c, ULO);
  gen(c, p, k1, t, seg);
  gen(c, p, gb, t + 0.5 * h, seg);
  gemm<false, false>(c, t1, gb, k1, 1.0, false);
  for (int e = threadIdx.x; e < c.nn; e += kThreads) {
    const double d = (e / c.np == e % c.np) ? 1.0 : 0.0;
    const double2 g = gb[e], q = t1[e];
    ulo[e] = make_double2(d + h * g.x, h * g.y);
    u1[e] = make_double2(d + h * g.x + 0.5 * h * h * q.x, h * g.y + 0.5 * h * h * q.y);
  }
  __syncthreads();
}

// Order 3 (Bogacki-Shampine 3(2)): k1 = G(t), k2 = G(t + h/2)(I + h/2 k1),
// k3 = G(t + 3h/4)(I + 3h/4 k2), U1 = I + h (2/9 k1 + 1/3 k2 + 4/9 k3),
// k4 = G(t + h) U1, ULO = I + h (7/24 k1 + 1/4 k2 + 1/3 k3 + 1/8 k4); Ua, Ub on the
// cubic Hermite through (I, h k1) and (U1, h k4) at 1/3 and 2/3; P = Ub Ua^-1,
// Q2 = U1 Ub^-1.
__device__ void props3(const Ctx &c, const dqh_ar_inst &p, double t, double h, int seg) {
  double2 *k1 = slot(c, K1), *gb = slot(c, GB), *k2 = slot(c, KK2), *k3 = slot(c, KK3), *k4 = slot(c, KK4);
  double2 *u1 = slot(c, U1), *ulo = slot(c, ULO), *ua = slot(c, UA), *ub = slot(c, UB), *inv = slot(c, INV);
  gen(c, p, k1, t, seg);
  gen(c, p, gb, t + 0.5 * h, seg);
  gemm<false, false>(c, k2, gb, k1, 0.5 * h, false);
  for (int e = threadIdx.x; e < c.nn; e += kThreads) {
    k2[e].x += gb[e].x;
    k2[e].y += gb[e].y;
  }
  __syncthreads();
  gen(c, p, gb, t + 0.75 * h, seg);
  gemm<false, false>(c, k3, gb, k2, 0.75 * h, false);
  for (int e = threadIdx.x; e < c.nn; e += kThreads) {
    const double d = (e / c.np == e % c.np) ? 1.0 : 0.0;
    k3[e].x += gb[e].x;
    k3[e].y += gb[e].y;
    u1[e] = make_double2(d + h * (2.0 / 9.0 * k1[e].x + 1.0 / 3.0 * k2[e].x + 4.0 / 9.0 * k3[e].x),
                         h * (2.0 / 9.0 * k1[e].y + 1.0 / 3.0 * k2[e].y + 4.0 / 9.0 * k3[e].y));
  }
  __syncthreads();
  gen(c, p, gb, t + h, seg);
  gemm<false, false>(c, k4, gb, u1, 1.0, false);
  // Hermite basis at theta: h00 I + h10 (h k1) + h01 U1 + h11 (h k4)
  double hb[2][4];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const double th = (s + 1) / 3.0, t2 = th * th, t3 = t2 * th;
    hb[s][0] = 2 * t3 - 3 * t2 + 1;
    hb[s][1] = t3 - 2 * t2 + th;
    hb[s][2] = -2 * t3 + 3 * t2;
    hb[s][3] = t3 - t2;
  }
  for (int e = threadIdx.x; e < c.nn; e += kThreads) {
    const double d = (e / c.np == e % c.np) ? 1.0 : 0.0;
    const double2 a1 = k1[e], a2 = k2[e], a3 = k3[e], a4 = k4[e], uu = u1[e];
    ulo[e] = make_double2(d + h * (7.0 / 24.0 * a1.x + 0.25 * a2.x + 1.0 / 3.0 * a3.x + 0.125 * a4.x),
                          h * (7.0 / 24.0 * a1.y + 0.25 * a2.y + 1.0 / 3.0 * a3.y + 0.125 * a4.y));
    ua[e] = make_double2(hb[0][0] * d + hb[0][1] * h * a1.x + hb[0][2] * uu.x + hb[0][3] * h * a4.x,
                         hb[0][1] * h * a1.y + hb[0][2] * uu.y + hb[0][3] * h * a4.y);
    ub[e] = make_double2(hb[1][0] * d + hb[1][1] * h * a1.x + hb[1][2] * uu.x + hb[1][3] * h * a4.x,
                         hb[1][1] * h * a1.y + hb[1][2] * uu.y + hb[1][3] * h * a4.y);
  }
  __syncthreads();
  gj_inv(c, ua, inv);
  gemm<false, false>(c, slot(c, PP), ub, inv, 1.0, false);
  gj_inv(c, ub, inv);
  gemm<false, false>(c, slot(c, Q2), u1, inv, 1.0, false);
}

// ------------------------------------------------------------------ saving
__device__ __noinline__ void save(const Ctx &c, const dqh_ar_inst &p, int isave) {
  const bool save_states = (p.flags & DQH_FLAG_SAVE_STATES) != 0;
  double2 *st = gp<double2>(p.states);
  const double2 *ET = gp<const double2>(p.EopsT);
  double2 *ex = gp<double2>(p.expects);
  const int n = c.n, np = c.np;
  if (save_states) {
    for (int e = threadIdx.x; e < c.nn; e += kThreads) {
      const int i = e / np, j = e % np;
      if (i < n && j < n) st[((size_t)isave * n + i) * n + j] = c.rho[e];
    }
  }
  for (int q = 0; q < (int)p.nE; ++q) {
    double sr = 0.0, si = 0.0;
    for (int e = threadIdx.x; e < c.nn; e += kThreads) {
      const double2 v = cmul(ET[(size_t)q * c.nn + e], c.rho[e]);
      sr += v.x;
      si += v.y;
    }
    sr = block_sum(sr, c.sh);
    si = block_sum(si, c.sh);
    if (threadIdx.x == 0) ex[(size_t)q * p.ntsave + isave] = make_double2(sr, si);
  }
  __syncthreads();
}

// ------------------------------------------------------------------ kernel
__global__ __launch_bounds__(kThreads) void k_rouchon_adaptive(const dqh_ar_inst *__restrict__ insts) {
  __shared__ Shared sh;
  const dqh_ar_inst &p = insts[blockIdx.x];
  int64_t *stats = gp<int64_t>(p.stats);
  const int np = (int)p.np;
  if (np < 16 || np > kMaxNP || (np & 15) || p.n < 1 || p.n > np || (p.order != 2 && p.order != 3) ||
      p.nL < 0 || p.nE < 0 || p.ntsave < 1 || p.npwc < 0 || p.npwc > DQH_MAX_MOD || p.nmod < 0 ||
      p.nmod > DQH_MAX_MOD) {
    if (threadIdx.x == 0) stats[3] = DQH_UNSUPPORTED;
    return;
  }
  Ctx c;
  c.n = (int)p.n;
  c.np = np;
  c.nL = (int)p.nL;
  c.nn = np * np;
  c.normalize = (int)p.normalize;
  const double2 *ops = gp<const double2>(p.ops);
  c.G0 = ops;
  c.LdL = ops + c.nn;
  c.L = ops + 2 * (size_t)c.nn;
  c.P = c.L + (size_t)c.nL * c.nn;
  c.O = c.P + (size_t)p.npwc * c.nn;
  c.w = gp<double2>(p.work);
  c.rho = gp<double2>(p.rho);
  c.sh = &sh;
  const double *tsave = gp<const double>(p.tsave);
  const double *ptimes = gp<const double>(p.pwc_times);
  const int ntsave = (int)p.ntsave, nseg = (int)p.nseg;
  const int order = (int)p.order;
  const double rtol = p.rtol, atol = p.atol;
  const double inv_nn = 1.0 / ((double)c.n * c.n);
  const double t_end = tsave[ntsave - 1];

  double t = p.t0, dt = p.dt0;
  int isave = 0;
  int64_t nsteps = 0, nacc = 0, nrej = 0;
  int status = 0;
  while (isave < ntsave && tsave[isave] <= t) save(c, p, isave++);
  while (t < t_end) {
    if (nsteps >= p.max_steps) {
      status = DQH_MAX_STEPS;
      break;
    }
    const int seg = p.npwc > 0 ? pwc_segment(ptimes, nseg, t) : -1;
    double target = t_end;
    if (isave < ntsave && tsave[isave] < target) target = tsave[isave];
    if (p.npwc > 0 && seg + 1 < nseg && ptimes[seg + 1] > t && ptimes[seg + 1] < target) target = ptimes[seg + 1];
    double h = dt;
    double t_new = t + h;
    bool on_target = false;
    if (snaps_to(t, h, target)) {
      h = target - t;
      t_new = target;
      on_target = true;
    }
    if (!(h > 0.0) || !(h > 1e-14 * fmax(1.0, fabs(t)))) {
      status = DQH_NONFINITE;
      break;
    }
    double2 *ohi = slot(c, OHI), *olo = slot(c, OLO);
    if (order == 2) {
      props2(c, p, t, h, seg);
      euler_map(c, slot(c, ULO), h, olo);
      heun2_map(c, slot(c, U1), h, ohi);
    } else {
      props3(c, p, t, h, seg);
      heun3_map(c, h, ohi);
      heun2_map(c, slot(c, ULO), h, olo);
    }
    double ssum = 0.0;
    for (int e = threadIdx.x; e < c.nn; e += kThreads) {
      const int i = e / np, j = e % np;
      if (i >= c.n || j >= c.n) continue;
      const double2 y0 = c.rho[e], y1 = ohi[e], yl = olo[e];
      const double er = 0.5 * (y1.x - yl.x), ei = 0.5 * (y1.y - yl.y);
      const double sc = atol + rtol * fmax(sqrt(y0.x * y0.x + y0.y * y0.y), sqrt(y1.x * y1.x + y1.y * y1.y));
      ssum += (er * er + ei * ei) / (sc * sc);
    }
    const double norm = sqrt(block_sum(ssum, c.sh) * inv_nn);
    const bool finite = isfinite(norm);
    const bool keep = finite && norm < 1.0;
    double factor;
    if (!finite)
      factor = p.factormin;
    else if (norm == 0.0)
      factor = p.factormax;
    else
      factor = p.safety * pow(norm, -1.0 / order);
    factor = fmin(fmax(factor, keep ? 1.0 : p.factormin), p.factormax);
    ++nsteps;
    if (keep) {
      ++nacc;
      t = on_target ? target : t_new;
      for (int e = threadIdx.x; e < c.nn; e += kThreads) c.rho[e] = ohi[e];
      __syncthreads();
      while (isave < ntsave && tsave[isave] <= t) save(c, p, isave++);
    } else {
      ++nrej;
    }
    dt = factor * h;
  }
  while (isave < ntsave) save(c, p, isave++);
  if (threadIdx.x == 0) {
    stats[0] = nsteps;
    stats[1] = nacc;
    stats[2] = nrej;
    stats[3] = status;
  }
}

// C[b] = op(A[b]) op(B[b])
__global__ __launch_bounds__(kThreads) void k_zgemm(const double2 *A, const double2 *B, double2 *C, int np, int opA,
                                                    int opB) {
  Ctx c;
  c.np = np;
  c.nn = np * np;
  const size_t off = (size_t)blockIdx.x * c.nn;
  if (opA && opB)
    gemm<true, true>(c, C + off, A + off, B + off, 1.0, false);
  else if (opA)
    gemm<true, false>(c, C + off, A + off, B + off, 1.0, false);
  else if (opB)
    gemm<false, true>(c, C + off, A + off, B + off, 1.0, false);
  else
    gemm<false, false>(c, C + off, A + off, B + off, 1.0, false);
}

}  // namespace ar
}  // namespace dqh

using namespace dqh;

extern "C" int64_t dqh_sizeof_ar_inst(void) { return (int64_t)sizeof(dqh_ar_inst); }

extern "C" int64_t dqh_ar_work_elems(int64_t np) { return (int64_t)ar::kSlots * np * np; }

static int ar_launch_checked(const char *what) {
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(DQH_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(err));
  return DQH_OK;
}

extern "C" int dqh_mesolve_rouchon_adaptive_dense(const dqh_ar_inst *d_insts, int64_t ninst, void *stream) {
  if (!d_insts || ninst <= 0 || ninst > (1 << 30))
    return fail(DQH_BAD_ARG, "dqh_mesolve_rouchon_adaptive_dense: bad arguments");
  hipLaunchKernelGGL(ar::k_rouchon_adaptive, dim3((unsigned)ninst), dim3(ar::kThreads), 0, (hipStream_t)stream,
                     d_insts);
  return ar_launch_checked("dqh_mesolve_rouchon_adaptive_dense");
}

extern "C" int dqh_zgemm_dense(const void *A, const void *B, void *C, int64_t np, int64_t opA, int64_t opB,
                               int64_t nbatch, void *stream) {
  if (!A || !B || !C || nbatch <= 0 || nbatch > (1 << 30) || (opA != 0 && opA != 1) || (opB != 0 && opB != 1))
    return fail(DQH_BAD_ARG, "dqh_zgemm_dense: bad arguments");
  if (np < 16 || np > ar::kMaxNP || (np & 15)) return fail(DQH_BAD_SHAPE, "dqh_zgemm_dense: np must be 16, 32, 48 or 64");
  hipLaunchKernelGGL(ar::k_zgemm, dim3((unsigned)nbatch), dim3(ar::kThreads), 0, (hipStream_t)stream,
                     (const double2 *)A, (const double2 *)B, (double2 *)C, (int)np, (int)opA, (int)opB);
  return ar_launch_checked("dqh_zgemm_dense");
}
