// Dense master-equation kernels for gfx950: complex64 GEMMs on FP32 MFMA
// (v_mfma_f32_16x16x4_f32), FP64 state accumulation and FP64 reductions.
//
// One workgroup owns one instance for the whole solve: instances are
// independent, so a batch of B solves is a grid of B workgroups with no
// inter-workgroup synchronisation, and each instance keeps its own t, dt and
// accept/reject decision (per-instance adaptive stepping).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "dqh.h"
#include "dqh_common.h"
#include "dqh_mfma.h"
#include "dqh_tsit5_tableau.h"
#include "dqh_dopri8_tableau.h"

namespace dqh {

constexpr int kWaves = 8;               // waves per workgroup (2 per SIMD)
constexpr int kThreads = kWaves * 64;   // one workgroup = one instance
constexpr int kResMaxL = 3;             // resident path: up to 3 jump-operator terms

// ----------------------------------------------------------- reductions
template <int NW>
__device__ __forceinline__ double block_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) s += sh[w];
  return s;
}

// time-dependent jump operators (DQH_FLAG_LMOD): the per-stage copies of D and C
__device__ __forceinline__ const float2 *lmod_D(const dqh_me_inst &p) {
  return gp<const float2>(p.work) + (size_t)(10 + p.nL) * (size_t)(p.np * p.np);
}
__device__ __forceinline__ const float2 *lmod_C(const dqh_me_inst &p) {
  return lmod_D(p) + (size_t)p.nL * (size_t)(p.np * p.np);
}
// ----------------------------------------------------------- instance view
struct View {
  int n, np, nL, nE, ntsave, npwc, nseg, flags;
  size_t nn;
  const float2 *Hnh;  // current Hnh (constant or PWC-updated copy)
  const float2 *D, *C;
  float2 *kbase;  // 7 stage buffers; logical stages 0 and 6 swap slots after an accepted step
  int s0, s6;
  float2 *ys, *X, *tmp, *Hcur;
  double2 *y[2];
};

__device__ __forceinline__ void make_view(View &v, const dqh_me_inst &p) {
  v.n = (int)p.n;
  v.np = (int)p.np;
  v.nL = (int)p.nL;
  v.nE = (int)p.nE;
  v.ntsave = (int)p.ntsave;
  v.npwc = (int)p.npwc;
  v.nseg = (int)p.nseg;
  v.flags = (int)p.flags;
  v.nn = (size_t)v.np * v.np;
  float2 *w = gp<float2>(p.work);
  v.kbase = w;
  v.s0 = 0;
  v.s6 = 6;
  v.ys = w + 7 * v.nn;
  v.tmp = w + 8 * v.nn;
  v.X = w + 9 * v.nn;
  const bool tdep = v.npwc > 0 || p.nmod > 0;
  v.Hcur = tdep ? w + (9 + v.nL) * v.nn : nullptr;
  v.Hnh = tdep ? v.Hcur : gp<const float2>(p.Hnh);
  const bool lmod = (p.flags & DQH_FLAG_LMOD) != 0;
  v.D = lmod ? lmod_D(p) : gp<const float2>(p.Dops);
  v.C = lmod ? lmod_C(p) : gp<const float2>(p.Cops);
  double2 *y = gp<double2>(p.y);
  v.y[0] = y;
  v.y[1] = y + v.nn;
}

// out <- tmp + tmp^dag with tmp = [Hnh ys] + sum_k (D_k ys) C_k. ys is read from v.ys.
// Uses v.X (nL * nn) and v.tmp as scratch. Ends with a barrier.
template <int NW>
__device__ __forceinline__ void wg_rhs(const View &v, const float2 *D, const float2 *C, bool hasH, float2 *out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int np = v.np, NT = np >> 4;
  const int nt2 = NT * NT;
  // phase 1: X_k = D_k ys
  const int nt1 = v.nL * nt2;
  for (int t = wave; t < nt1; t += NW) {
    const int kk = t / nt2, ti = (t / NT) % NT, tj = t % NT;
    Acc acc;
    acc_zero(acc);
    tile_mac(acc, D + kk * v.nn, np, v.ys, np, ti * 16, tj * 16, np, lane);
    tile_store(acc, v.X + kk * v.nn, np, ti * 16, tj * 16, lane);
  }
  __syncthreads();
  // phase 2: tmp = Hnh ys + sum_k X_k C_k
  for (int t = wave; t < nt2; t += NW) {
    const int ti = t / NT, tj = t % NT;
    Acc acc;
    acc_zero(acc);
    if (hasH) tile_mac(acc, v.Hnh, np, v.ys, np, ti * 16, tj * 16, np, lane);
    for (int kk = 0; kk < v.nL; ++kk)
      tile_mac(acc, v.X + kk * v.nn, np, C + kk * v.nn, np, ti * 16, tj * 16, np, lane);
    tile_store(acc, v.tmp, np, ti * 16, tj * 16, lane);
  }
  __syncthreads();
  // phase 3: Hermitian symmetrisation
  for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
    const int i = (int)(e / np), j = (int)(e % np);
    const float2 a = v.tmp[e], b = v.tmp[(size_t)j * np + i];
    out[e] = make_float2(a.x + b.x, a.y - b.y);
  }
  __syncthreads();
}

__device__ __forceinline__ float2 *kp(const View &v, int j) {
  const int slot = j == 0 ? v.s0 : (j == 6 ? v.s6 : j);
  return v.kbase + (size_t)slot * v.nn;
}

__device__ __forceinline__ float2 to_f2(double2 a) { return make_float2((float)a.x, (float)a.y); }

// ys <- (float) y
template <int NW>
__device__ __forceinline__ void wg_round_state(const View &v, const double2 *y) {
  for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) v.ys[e] = to_f2(y[e]);
  __syncthreads();
}

// f_j(tau) of the smooth terms: the piecewise cubic of the interval that contains
// tau (clamped to the table), Horner in x = tau - t_i
__device__ __forceinline__ void mod_coeffs(const dqh_me_inst &p, double tau, double2 (&c)[DQH_MAX_MOD]) {
  const int nmod = (int)p.nmod;
  const int64_t nint = p.mod_nint;
  int64_t i = (int64_t)floor((tau - p.mod_t0) / p.mod_dt);
  i = i < 0 ? 0 : (i >= nint ? nint - 1 : i);
  const double x = tau - (p.mod_t0 + (double)i * p.mod_dt);
  const double2 *tab = gp<const double2>(p.mod_coef);
#pragma unroll
  for (int j = 0; j < DQH_MAX_MOD; ++j) {
    c[j] = make_double2(0.0, 0.0);
    if (j < nmod) {
      const double2 *q = tab + ((size_t)j * nint + i) * 4;
      double re = q[0].x, im = q[0].y;
#pragma unroll
      for (int m = 1; m < 4; ++m) {
        re = re * x + q[m].x;
        im = im * x + q[m].y;
      }
      c[j] = make_double2(re, im);
    }
  }
}

// f_j(tau) of table j of mod_coef (Horner on the interval that contains tau)
__device__ __forceinline__ double2 mod_eval(const dqh_me_inst &p, int j, double tau) {
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

// time-dependent jump operators (DQH_FLAG_LMOD): D_k <- f_k(tau) L_k and
// C_k <- conj(f_k(tau)) L_k^dag / 2 in work at (10 + nL) np^2, f_k = table nmod + k;
// with DQH_FLAG_LPWC, f_k also carries the factor of row npwc + k of pwc_vals on the
// segment seg of the step start (zero outside the breakpoints)
__device__ __forceinline__ void lmod_update(const dqh_me_inst &p, double tau, int seg, int nthreads) {
  const int nL = (int)p.nL, nmod = (int)p.nmod;
  const size_t nn = (size_t)(p.np * p.np);
  float2 *Dc = gp<float2>(p.work) + (size_t)(10 + nL) * nn;
  float2 *Cc = Dc + (size_t)nL * nn;
  const float2 *D0 = gp<const float2>(p.Dops), *C0 = gp<const float2>(p.Cops);
  const int nv = (int)p.nseg - 1;
  const bool lpwc = (p.flags & DQH_FLAG_LPWC) != 0;
  const bool lpair = (p.flags & DQH_FLAG_LPAIR) != 0;
  // coefficient q of D (q = 0) or of C (q = 1, DQH_FLAG_LPAIR): table nmod + q nL + k,
  // times row npwc + q nL + k of pwc_vals (DQH_FLAG_LPWC)
  auto coef = [&](int q, int k) {
    double2 f = mod_eval(p, nmod + q * nL + k, tau);
    if (lpwc) {
      const double *vals = gp<const double>(p.pwc_vals) + 2 * (size_t)(p.npwc + q * nL + k) * (size_t)(nv > 0 ? nv : 0);
      const double cr = (seg >= 0 && seg < nv) ? vals[2 * seg] : 0.0;
      const double ci = (seg >= 0 && seg < nv) ? vals[2 * seg + 1] : 0.0;
      f = make_double2(f.x * cr - f.y * ci, f.x * ci + f.y * cr);
    }
    return f;
  };
  for (int k = 0; k < nL; ++k) {
    const double2 f = coef(0, k);
    const double2 fc = lpair ? coef(1, k) : f;
    for (size_t e = threadIdx.x; e < nn; e += nthreads) {
      const float2 d = D0[k * nn + e], c = C0[k * nn + e];
      Dc[k * nn + e] = make_float2((float)(f.x * d.x - f.y * d.y), (float)(f.x * d.y + f.y * d.x));
      Cc[k * nn + e] = make_float2((float)(fc.x * c.x + fc.y * c.y), (float)(fc.x * c.y - fc.y * c.x));
    }
  }
}

// Hcur <- Hnh0 + sum_j c_j Hpwc_j for PWC segment seg (zero outside) + sum_j f_j(tau) Hmod_j
// (and the time-dependent jump operators at tau)
template <int NW>
__device__ __forceinline__ void wg_update_pwc(const View &v, const dqh_me_inst &p, int seg, double tau) {
  const float2 *H0 = gp<const float2>(p.Hnh);
  const float2 *Hp = gp<const float2>(p.Hpwc);
  const float2 *Hm = gp<const float2>(p.Hmod);
  const double *vals = gp<const double>(p.pwc_vals);
  const int nv = v.nseg - 1;
  const int nmod = (int)p.nmod;
  double2 fm[DQH_MAX_MOD];
  mod_coeffs(p, tau, fm);
  for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
    double re = H0[e].x, im = H0[e].y;
    if (seg >= 0 && seg < nv) {
      for (int j = 0; j < v.npwc; ++j) {
        const double cr = vals[2 * (j * nv + seg)], ci = vals[2 * (j * nv + seg) + 1];
        const float2 o = Hp[j * v.nn + e];
        re += cr * o.x - ci * o.y;
        im += cr * o.y + ci * o.x;
      }
    }
#pragma unroll
    for (int j = 0; j < DQH_MAX_MOD; ++j) {
      if (j < nmod) {
        const float2 o = Hm[j * v.nn + e];
        re += fm[j].x * o.x - fm[j].y * o.y;
        im += fm[j].x * o.y + fm[j].y * o.x;
      }
    }
    v.Hcur[e] = make_float2((float)re, (float)im);
  }
  if (p.flags & DQH_FLAG_LMOD) lmod_update(p, tau, seg, NW * 64);
  __syncthreads();
}

__device__ __forceinline__ int pwc_segment(const double *times, int nseg, double t) {
  // searchsorted(times, t, 'right') - 1; -1 before the first breakpoint
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

// Save the state interpolated as y + h * sum_j w_j k_j (or y itself when h == 0)
// into states[isave] (unpadded) and its expectation values.
template <int NW>
__device__ __forceinline__ void wg_save(const View &v, const dqh_me_inst &p, const double2 *y, double h, const double *w,
                        int nk, int isave, double *sh) {
  float2 *states = gp<float2>(p.states);
  const bool save_states = (v.flags & DQH_FLAG_SAVE_STATES) != 0;
  const float2 *ET = gp<const float2>(p.EopsT);
  double er[4] = {0, 0, 0, 0}, ei[4] = {0, 0, 0, 0};
  for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
    double re = y[e].x, im = y[e].y;
    for (int j = 0; j < nk; ++j) {
      const float2 kj = kp(v, j)[e];
      re += h * w[j] * kj.x;
      im += h * w[j] * kj.y;
    }
    const int i = (int)(e / v.np), jj = (int)(e % v.np);
    if (i < v.n && jj < v.n) {
      if (save_states) states[((size_t)isave * v.n + i) * v.n + jj] = make_float2((float)re, (float)im);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q < v.nE) {
          const float2 o = ET[q * v.nn + e];
          er[q] += o.x * re - o.y * im;
          ei[q] += o.x * im + o.y * re;
        }
      }
    }
  }
  float2 *ex = gp<float2>(p.expects);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q >= v.nE) continue;
    const double sr = block_sum<NW>(er[q], sh);
    const double si = block_sum<NW>(ei[q], sh);
    if (threadIdx.x == 0) ex[(size_t)q * v.ntsave + isave] = make_float2((float)sr, (float)si);
  }
  // more than 4 expectation operators: one extra pass per operator
  for (int q = 4; q < v.nE; ++q) {
    double a = 0, b = 0;
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const int i = (int)(e / v.np), jj = (int)(e % v.np);
      if (i >= v.n || jj >= v.n) continue;
      double re = y[e].x, im = y[e].y;
      for (int j = 0; j < nk; ++j) {
        const float2 kj = kp(v, j)[e];
        re += h * w[j] * kj.x;
        im += h * w[j] * kj.y;
      }
      const float2 o = ET[q * v.nn + e];
      a += o.x * re - o.y * im;
      b += o.x * im + o.y * re;
    }
    const double sr = block_sum<NW>(a, sh), si = block_sum<NW>(b, sh);
    if (threadIdx.x == 0) ex[(size_t)q * v.ntsave + isave] = make_float2((float)sr, (float)si);
  }
  __syncthreads();
}

template <int NW>
__device__ __forceinline__ void wg_write_final(const View &v, const dqh_me_inst &p, const double2 *y) {
  float2 *fin = gp<float2>(p.final_state);
  for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
    const int i = (int)(e / v.np), j = (int)(e % v.np);
    if (i < v.n && j < v.n) fin[(size_t)i * v.n + j] = to_f2(y[e]);
  }
}

// ------------------------------------------------------------ unit RHS kernel
template <int NW>
__device__ __forceinline__ void rhs_global(const dqh_me_inst &p) {
  View v;
  make_view(v, p);
  if (v.npwc > 0 || p.nmod > 0) {
    const int seg = v.npwc > 0 ? pwc_segment(gp<const double>(p.pwc_times), v.nseg, p.t0) : -1;
    wg_update_pwc<NW>(v, p, seg, p.t0);
  }
  wg_round_state<NW>(v, v.y[0]);
  wg_rhs<NW>(v, v.D, v.C, (v.flags & DQH_FLAG_HAS_HNH) != 0, kp(v, 0));
  float2 *fin = gp<float2>(p.final_state);
  for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
    const int i = (int)(e / v.np), j = (int)(e % v.np);
    if (i < v.n && j < v.n) fin[(size_t)i * v.n + j] = kp(v, 0)[e];
  }
}

// ----------------------------------------------------------- Tsit5 kernel
template <int NW>
__device__ __forceinline__ void tsit5_global(const dqh_me_inst &p, double *sh) {
  const RKTab &rk = rk_tab(p.flags);
  View v;
  make_view(v, p);
  const double *tsave = gp<const double>(p.tsave);
  const double *ptimes = gp<const double>(p.pwc_times);
  int64_t *stats = gp<int64_t>(p.stats);
  const double rtol = p.rtol, atol = p.atol;
  const double inv_nn = 1.0 / ((double)v.n * v.n);
  const double t_end = p.t1;
  const bool hasH = (v.flags & DQH_FLAG_HAS_HNH) != 0;

  double t = p.t0;
  int cur = 0;
  int isave = 0;
  int64_t nsteps = 0, nacc = 0, nrej = 0, nrhs = 0;
  int status = 0;
  int seg = -2;
  const double w1[1] = {0.0};

  // tsave points at t0 store the initial state
  while (isave < v.ntsave && tsave[isave] <= t) {
    wg_save<NW>(v, p, v.y[cur], 0.0, w1, 0, isave, sh);
    ++isave;
  }
  const bool smooth = p.nmod > 0;  // H changes within a step: updated at every stage time
  const bool lin = !smooth && !(p.flags & DQH_FLAG_LMOD);  // constant generator over each step
  if (v.npwc > 0 || smooth) {
    seg = v.npwc > 0 ? pwc_segment(ptimes, v.nseg, t) : -1;
    wg_update_pwc<NW>(v, p, seg, t);
  }
  // k0 = f(y0)
  wg_round_state<NW>(v, v.y[cur]);
  wg_rhs<NW>(v, v.D, v.C, hasH, kp(v, 0));
  ++nrhs;

  // initial step (Hairer-Norsett-Wanner II.4)
  double dt;
  {
    double s0 = 0, s1 = 0;
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const double2 y0 = v.y[cur][e];
      const double ay = sqrt(y0.x * y0.x + y0.y * y0.y);
      const double sc = atol + rtol * ay;
      const float2 f0 = kp(v, 0)[e];
      s0 += (y0.x * y0.x + y0.y * y0.y) / (sc * sc);
      s1 += ((double)f0.x * f0.x + (double)f0.y * f0.y) / (sc * sc);
    }
    const double d0 = sqrt(block_sum<NW>(s0, sh) * inv_nn);
    const double d1 = sqrt(block_sum<NW>(s1, sh) * inv_nn);
    const double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const double2 y0 = v.y[cur][e];
      const float2 f0 = kp(v, 0)[e];
      v.ys[e] = make_float2((float)(y0.x + h0 * f0.x), (float)(y0.y + h0 * f0.y));
    }
    __syncthreads();
    if (smooth) wg_update_pwc<NW>(v, p, seg, t + h0);
    wg_rhs<NW>(v, v.D, v.C, hasH, kp(v, 1));
    ++nrhs;
    double s2 = 0;
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const double2 y0 = v.y[cur][e];
      const double sc = atol + rtol * sqrt(y0.x * y0.x + y0.y * y0.y);
      const double dr = (double)kp(v, 1)[e].x - kp(v, 0)[e].x, di = (double)kp(v, 1)[e].y - kp(v, 0)[e].y;
      s2 += (dr * dr + di * di) / (sc * sc);
    }
    const double d2 = sqrt(block_sum<NW>(s2, sh) * inv_nn) / h0;
    const double md = fmax(d1, d2);
    const double h1 = md <= 1e-15 ? fmax(1e-6, h0 * 1e-3) : ctrl_pow(0.01 / md, kInitExp);
    dt = fmin(100.0 * h0, h1);
  }

  bool need_k0 = false;
  while (t < t_end) {
    if (nsteps >= p.max_steps) {
      status = DQH_MAX_STEPS;
      break;
    }
    if (v.npwc > 0) {
      const int s = pwc_segment(ptimes, v.nseg, t);
      if (s != seg) {
        seg = s;
        if (!smooth) wg_update_pwc<NW>(v, p, seg, t);
      }
    }
    if (need_k0) {
      wg_round_state<NW>(v, v.y[cur]);
      if (smooth) wg_update_pwc<NW>(v, p, seg, t);
      wg_rhs<NW>(v, v.D, v.C, hasH, kp(v, 0));
      ++nrhs;
      need_k0 = false;
    }
    // clip the step to the end time or the next breakpoint strictly after t
    double h = dt;
    double target = t_end;
    bool target_jump = false;
    if (v.npwc > 0) {
      const int m = pwc_segment(ptimes, v.nseg, t) + 1;
      if (m < v.nseg && ptimes[m] < t_end) {
        target = ptimes[m];
        target_jump = true;
      }
    }
    bool made_jump = false;
    double t_new = t + h;
    if (snaps_to(t, h, target)) {
      h = target - t;
      t_new = target;
      made_jump = target_jump;
    }
    if (!(h > 0.0) || !(h > 1e-14 * fmax(1.0, fabs(t)))) {
      status = DQH_NONFINITE;
      break;
    }
    const double2 *y = v.y[cur];
    double2 *ynew = v.y[1 - cur];
    // stages 1..6 (stage 6 input is the candidate solution). With a generator that
    // is constant over the step (lin), stages 1..5 evaluate the vector field on the
    // increment Delta_s = h sum_j a_sj k_j only and hold g_s = f(Delta_s), with
    // k_s = k_0 + g_s folded into the weight of k_0 wherever the stages combine.
    for (int s = 1; s < 7; ++s) {
      double c[6];
      for (int j = 0; j < 6; ++j) c[j] = (j < s) ? h * rk.A[s][j] : 0.0;
      if (lin) {
        double a0 = 0.0;
        for (int j = 0; j < s; ++j) a0 += rk.A[s][j];
        c[0] = h * a0;
      }
      const bool with_y = !lin || s == 6;
      for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
        double re = with_y ? y[e].x : 0.0, im = with_y ? y[e].y : 0.0;
        for (int j = 0; j < s; ++j) {
          const float2 kj = kp(v, j)[e];
          re += c[j] * kj.x;
          im += c[j] * kj.y;
        }
        if (s == 6) ynew[e] = make_double2(re, im);
        v.ys[e] = make_float2((float)re, (float)im);
      }
      __syncthreads();
      if (smooth) wg_update_pwc<NW>(v, p, seg, t + rk.C[s] * h);
      wg_rhs<NW>(v, v.D, v.C, hasH, kp(v, s));
      ++nrhs;
    }
    // error norm (FP64 accumulation)
    double ew[7];
    for (int j = 0; j < 7; ++j) ew[j] = rk.E[j];
    if (lin) fold_k0(ew);
    double s = 0;
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      double er = 0, ei = 0;
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        const float2 kj = kp(v, j)[e];
        er += ew[j] * kj.x;
        ei += ew[j] * kj.y;
      }
      er *= h;
      ei *= h;
      const double2 a = y[e], b = ynew[e];
      const double sc = atol + rtol * fmax(sqrt(a.x * a.x + a.y * a.y), sqrt(b.x * b.x + b.y * b.y));
      s += (er * er + ei * ei) / (sc * sc);
    }
    const double norm = sqrt(block_sum<NW>(s, sh) * inv_nn);
    const bool finite = isfinite(norm);
    const bool keep = finite && norm < 1.0;
    double factor;
    if (!finite)
      factor = p.factormin;
    else if (norm == 0.0)
      factor = p.factormax;
    else
      factor = p.safety * ctrl_pow(norm, -0.2);
    factor = fmin(fmax(factor, keep ? 1.0 : p.factormin), p.factormax);
    ++nsteps;
    if (keep) {
      if (threadIdx.x == 0 && nacc < p.step_log_cap) gp<double>(p.step_log)[nacc] = t_new;
      ++nacc;
      while (isave < v.ntsave && tsave[isave] <= t_new) {
        const double th = (tsave[isave] - t) / h;
        double w[7];
        for (int j = 0; j < 7; ++j) w[j] = th * (rk.R[j][0] + th * (rk.R[j][1] + th * (rk.R[j][2] + th * rk.R[j][3])));
        if (lin) fold_k0(w);
        wg_save<NW>(v, p, y, h, w, 7, isave, sh);
        ++isave;
      }
      t = t_new;
      cur = 1 - cur;
      const int ts = v.s0;
      v.s0 = v.s6;
      v.s6 = ts;
      need_k0 = made_jump;
    } else {
      ++nrej;
    }
    dt = factor * h;
  }
  // points not reached (only after a failure): repeat the last state
  {
    const double w0[1] = {0.0};
    while (isave < v.ntsave) {
      wg_save<NW>(v, p, v.y[cur], 0.0, w0, 0, isave, sh);
      ++isave;
    }
  }
  wg_write_final<NW>(v, p, v.y[cur]);
  if (!(v.flags & DQH_FLAG_SAVE_STATES)) {
    float2 *st = gp<float2>(p.states);
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const int i = (int)(e / v.np), j = (int)(e % v.np);
      if (i < v.n && j < v.n) st[(size_t)i * v.n + j] = to_f2(v.y[cur][e]);
    }
  }
  if (threadIdx.x == 0) {
    stats[0] = nsteps;
    stats[1] = nacc;
    stats[2] = nrej;
    stats[3] = status;
    stats[4] = nrhs;
  }
}

// ------------------------------------------------------------ Dopri8 body
// Stage slot j of the 13-stage pair: 0..6 in the Tsit5 stage buffers, 7..12 after
// the Hcur slot (work needs (16 + nL) np^2 elements; no DQH_FLAG_LMOD).
__device__ __forceinline__ float2 *k8(const View &v, float2 *w, int j) {
  return j < 7 ? v.kbase + (size_t)j * v.nn : w + (size_t)(10 + v.nL + j - 7) * v.nn;
}

// Adaptive Dormand-Prince 8(7) (DQH_FLAG_DOPRI8; the reference's dx.Dopri8() in
// diffrax_integrator.py:101-147): the same vector field, FP64 state and stage
// combination, RMS error norm and integral controller as the Tsit5 body, with the
// exponent 1/8 (error order 8) and 13 evaluations per step (no FSAL). Steps land on
// the save times (and on PWC breakpoints and t1), so every saved state is a step end
// of the 8th-order solution and no continuous extension is needed.
template <int NW>
__device__ __forceinline__ void rk8_global(const dqh_me_inst &p, double *sh) {
  const RK8Tab &rk = kRK8;
  View v;
  make_view(v, p);
  float2 *w = gp<float2>(p.work);
  const double *tsave = gp<const double>(p.tsave);
  const double *ptimes = gp<const double>(p.pwc_times);
  int64_t *stats = gp<int64_t>(p.stats);
  const double rtol = p.rtol, atol = p.atol;
  const double inv_nn = 1.0 / ((double)v.n * v.n);
  const double t_end = p.t1;
  const bool hasH = (v.flags & DQH_FLAG_HAS_HNH) != 0;
  const bool smooth = p.nmod > 0;

  double t = p.t0;
  int cur = 0;
  int isave = 0;
  int64_t nsteps = 0, nacc = 0, nrej = 0, nrhs = 0;
  int status = 0;
  int seg = -2;
  const double w1[1] = {0.0};

  while (isave < v.ntsave && tsave[isave] <= t) {
    wg_save<NW>(v, p, v.y[cur], 0.0, w1, 0, isave, sh);
    ++isave;
  }
  if (v.npwc > 0 || smooth) {
    seg = v.npwc > 0 ? pwc_segment(ptimes, v.nseg, t) : -1;
    wg_update_pwc<NW>(v, p, seg, t);
  }
  wg_round_state<NW>(v, v.y[cur]);
  wg_rhs<NW>(v, v.D, v.C, hasH, k8(v, w, 0));
  ++nrhs;

  // initial step (Hairer-Norsett-Wanner II.4), exponent 1/7 (the embedded order, as kInitExp)
  double dt;
  {
    double s0 = 0, s1 = 0;
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const double2 y0 = v.y[cur][e];
      const double sc = atol + rtol * sqrt(y0.x * y0.x + y0.y * y0.y);
      const float2 f0 = k8(v, w, 0)[e];
      s0 += (y0.x * y0.x + y0.y * y0.y) / (sc * sc);
      s1 += ((double)f0.x * f0.x + (double)f0.y * f0.y) / (sc * sc);
    }
    const double d0 = sqrt(block_sum<NW>(s0, sh) * inv_nn);
    const double d1 = sqrt(block_sum<NW>(s1, sh) * inv_nn);
    const double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const double2 y0 = v.y[cur][e];
      const float2 f0 = k8(v, w, 0)[e];
      v.ys[e] = make_float2((float)(y0.x + h0 * f0.x), (float)(y0.y + h0 * f0.y));
    }
    __syncthreads();
    if (smooth) wg_update_pwc<NW>(v, p, seg, t + h0);
    wg_rhs<NW>(v, v.D, v.C, hasH, k8(v, w, 1));
    ++nrhs;
    double s2 = 0;
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const double2 y0 = v.y[cur][e];
      const double sc = atol + rtol * sqrt(y0.x * y0.x + y0.y * y0.y);
      const double dr = (double)k8(v, w, 1)[e].x - k8(v, w, 0)[e].x, di = (double)k8(v, w, 1)[e].y - k8(v, w, 0)[e].y;
      s2 += (dr * dr + di * di) / (sc * sc);
    }
    const double d2 = sqrt(block_sum<NW>(s2, sh) * inv_nn) / h0;
    const double md = fmax(d1, d2);
    const double h1 = md <= 1e-15 ? fmax(1e-6, h0 * 1e-3) : ctrl_pow(0.01 / md, 1.0 / 7.0);
    dt = fmin(100.0 * h0, h1);
  }

  bool need_k0 = false;
  while (t < t_end) {
    if (nsteps >= p.max_steps) {
      status = DQH_MAX_STEPS;
      break;
    }
    if (v.npwc > 0) {
      const int sg = pwc_segment(ptimes, v.nseg, t);
      if (sg != seg) {
        seg = sg;
        if (!smooth) wg_update_pwc<NW>(v, p, seg, t);
      }
    }
    if (need_k0) {  // k_0 = f(y) at the start of every step (the pair has no FSAL stage)
      wg_round_state<NW>(v, v.y[cur]);
      if (smooth) wg_update_pwc<NW>(v, p, seg, t);
      wg_rhs<NW>(v, v.D, v.C, hasH, k8(v, w, 0));
      ++nrhs;
    }
    need_k0 = true;
    // clip the step to the next save time, breakpoint or the end time after t
    double target = t_end;
    if (isave < v.ntsave && tsave[isave] < target) target = tsave[isave];
    if (v.npwc > 0) {
      const int m = pwc_segment(ptimes, v.nseg, t) + 1;
      if (m < v.nseg && ptimes[m] < target) target = ptimes[m];
    }
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
    const double2 *y = v.y[cur];
    double2 *ynew = v.y[1 - cur];
    for (int st = 1; st < kRK8Stages; ++st) {
      for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
        double re = y[e].x, im = y[e].y;
        for (int j = 0; j < st; ++j) {
          const double c = rk.a[st][j];
          if (c != 0.0) {
            const float2 kj = k8(v, w, j)[e];
            re += h * c * kj.x;
            im += h * c * kj.y;
          }
        }
        v.ys[e] = make_float2((float)re, (float)im);
      }
      __syncthreads();
      if (smooth) wg_update_pwc<NW>(v, p, seg, t + rk.c[st] * h);
      wg_rhs<NW>(v, v.D, v.C, hasH, k8(v, w, st));
      ++nrhs;
    }
    // y1 (8th order) and the error against the 7th-order solution, in FP64
    double ssum = 0;
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const double2 a = y[e];
      double yr = a.x, yi = a.y, er = 0, ei = 0;
      for (int j = 0; j < kRK8Stages; ++j) {
        const float2 kj = k8(v, w, j)[e];
        yr += h * rk.b[j] * kj.x;
        yi += h * rk.b[j] * kj.y;
        er += rk.e[j] * kj.x;
        ei += rk.e[j] * kj.y;
      }
      er *= h;
      ei *= h;
      ynew[e] = make_double2(yr, yi);
      const double sc = atol + rtol * fmax(sqrt(a.x * a.x + a.y * a.y), sqrt(yr * yr + yi * yi));
      ssum += (er * er + ei * ei) / (sc * sc);
    }
    const double norm = sqrt(block_sum<NW>(ssum, sh) * inv_nn);
    const bool finite = isfinite(norm);
    const bool keep = finite && norm < 1.0;
    double factor;
    if (!finite)
      factor = p.factormin;
    else if (norm == 0.0)
      factor = p.factormax;
    else
      factor = p.safety * ctrl_pow(norm, -1.0 / 8.0);
    factor = fmin(fmax(factor, keep ? 1.0 : p.factormin), p.factormax);
    ++nsteps;
    if (keep) {
      if (threadIdx.x == 0 && nacc < p.step_log_cap) gp<double>(p.step_log)[nacc] = t_new;
      ++nacc;
      t = on_target ? target : t_new;
      cur = 1 - cur;
      while (isave < v.ntsave && tsave[isave] <= t) {
        wg_save<NW>(v, p, v.y[cur], 0.0, w1, 0, isave, sh);
        ++isave;
      }
    } else {
      ++nrej;
      need_k0 = false;  // k_0 = f(y) still holds
    }
    dt = factor * h;
  }
  while (isave < v.ntsave) {
    wg_save<NW>(v, p, v.y[cur], 0.0, w1, 0, isave, sh);
    ++isave;
  }
  wg_write_final<NW>(v, p, v.y[cur]);
  if (!(v.flags & DQH_FLAG_SAVE_STATES)) {
    float2 *stp = gp<float2>(p.states);
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const int i = (int)(e / v.np), j = (int)(e % v.np);
      if (i < v.n && j < v.n) stp[(size_t)i * v.n + j] = to_f2(v.y[cur][e]);
    }
  }
  if (threadIdx.x == 0) {
    stats[0] = nsteps;
    stats[1] = nacc;
    stats[2] = nrej;
    stats[3] = status;
    stats[4] = nrhs;
  }
}

// -------------------------------------------------------- Rouchon1 kernel
template <int NW>
__device__ __forceinline__ void rouchon1_global(const dqh_me_inst &p, double *sh) {
  View v;
  make_view(v, p);
  const double *tsave = gp<const double>(p.tsave);
  const double *step_t = gp<const double>(p.step_t);
  const int *step_set = gp<const int>(p.step_set);
  int64_t *stats = gp<int64_t>(p.stats);
  const int64_t nsteps = p.max_steps;
  const double w1[1] = {1.0};
  double2 *y = v.y[0];
  int isave = 0;
  while (isave < v.ntsave && tsave[isave] <= step_t[0]) {
    wg_save<NW>(v, p, y, 0.0, w1, 0, isave, sh);
    ++isave;
  }
  for (int64_t m = 0; m < nsteps; ++m) {
    const double t = step_t[m], tn = step_t[m + 1];
    const size_t so = (size_t)step_set[m] * v.nL * v.nn;
    wg_round_state<NW>(v, y);
    wg_rhs<NW>(v, v.D + so, v.C + so, false, kp(v, 0));
    // linear interpolation y + theta * increment at save points inside the step
    const double tol = 1e-12 * fmax(1.0, fabs(tn));
    while (isave < v.ntsave && tsave[isave] <= tn + tol) {
      const double th = fmin(fmax((tsave[isave] - t) / (tn - t), 0.0), 1.0);
      wg_save<NW>(v, p, y, th, w1, 1, isave, sh);
      ++isave;
    }
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const float2 d = kp(v, 0)[e];
      y[e].x += d.x;
      y[e].y += d.y;
    }
    __syncthreads();
  }
  while (isave < v.ntsave) {
    wg_save<NW>(v, p, y, 0.0, w1, 0, isave, sh);
    ++isave;
  }
  wg_write_final<NW>(v, p, y);
  if (!(v.flags & DQH_FLAG_SAVE_STATES)) {
    float2 *st = gp<float2>(p.states);
    for (size_t e = threadIdx.x; e < v.nn; e += NW * 64) {
      const int i = (int)(e / v.np), j = (int)(e % v.np);
      if (i < v.n && j < v.n) st[(size_t)i * v.n + j] = to_f2(y[e]);
    }
  }
  if (threadIdx.x == 0) {
    stats[0] = nsteps;
    stats[1] = nsteps;
    stats[2] = 0;
    stats[3] = 0;
    stats[4] = nsteps;
  }
}

}  // namespace dqh

#include "dqh_resident.h"

namespace dqh {

// One launch serves a heterogeneous batch: each workgroup picks the body that
// fits its instance (register/LDS-resident for np <= 64 and at most 3 jump
// operators, global-memory state otherwise).
// With DQH_FLAG_SIGNAL, publish the instance's completion: every thread makes its
// output stores visible at system scope, then thread 0 sets stats[7] = 1 with a
// system-scope release, so that a host polling stats[7] (outputs in pinned host
// memory) reads the finished outputs while the rest of the launch still runs.
// flags is the same for the whole workgroup, so every thread reaches the barrier.
__device__ __forceinline__ void signal_done(const dqh_me_inst &p) {
  if (!(p.flags & DQH_FLAG_SIGNAL)) return;
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(gp<int64_t>(p.stats) + 7, (int64_t)1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The body that fits one instance (called by every thread of the workgroup).
__device__ __forceinline__ void tsit5_dispatch(const dqh_me_inst &p, double *sh, float2 *lds, double2 *dg) {
  const int np = (int)p.np;
  const bool res = p.nL <= kResMaxL && !(p.flags & (DQH_FLAG_GLOBAL | DQH_FLAG_LMOD));
  if (p.hdiag != 0 || ((p.flags & DQH_FLAG_DOPRI8) && (p.flags & DQH_FLAG_LMOD))) {
    // the FP64 diagonal runs on k_tsit5_res<NP, true> only; Dopri8 takes no
    // modulated jump operators
    if (threadIdx.x == 0) gp<int64_t>(p.stats)[3] = DQH_UNSUPPORTED;
  } else if (p.flags & DQH_FLAG_DOPRI8)
    rk8_global<kWaves>(p, sh);
  else if (res && np == 16)
    tsit5_resident<16, kWaves, false>(p, sh, lds, dg);
  else if (res && np == 32)
    tsit5_resident<32, kWaves, false>(p, sh, lds, dg);
  else if (res && np == 48)
    tsit5_resident<48, kWaves, false>(p, sh, lds, dg);
  else if (res && np == 64)
    tsit5_resident<64, kWaves, false>(p, sh, lds, dg);
  else
    tsit5_global<kWaves>(p, sh);
}

__global__ __launch_bounds__(kThreads) void k_tsit5(const dqh_me_inst *__restrict__ insts) {
  __shared__ double sh[kWaves];
  __shared__ float2 lds[kResLdsFloat2];
  __shared__ double2 dg[64];
  const dqh_me_inst &p = insts[blockIdx.x];
  tsit5_dispatch(p, sh, lds, dg);
  signal_done(p);
}

// Operator preparation on the device (diffrax_integrator.py:319-323 builds
// Hnh = -iH - 1/2 sum L^dag L on every call): on entry Hnh holds -iH and Dops the
// jump operators L_k (zero-padded to np); on exit Hnh = -iH - 1/2 sum_k D_k^dag D_k,
// the products accumulated in FP64 from the complex64 D_k (the same D_k the jump
// term uses), and Cops = D_k^dag / 2. grid: (kPrepTiles, ninst); each workgroup
// strides over the np^2 elements of its instance (D_k[m][i] is the same for the
// threads of one row, D_k[m][j] coalesced).
constexpr int kPrepTiles = 16;

// Hnh <- Hnh - 1/2 sum_k D_k^dag D_k (FP64 products) and C_k = D_k^dag / 2 for the
// elements e0, e0 + stride, ... of one instance
__device__ __forceinline__ void prepare_elems(const dqh_me_inst &p, int e0, int stride) {
  const int np = (int)p.np, nL = (int)p.nL;
  const int nn = np * np;
  float2 *Hnh = gp<float2>(p.Hnh);
  const float2 *D = gp<const float2>(p.Dops);
  float2 *C = gp<float2>(p.Cops);
  for (int e = e0; e < nn; e += stride) {
    const int i = e / np, j = e - i * np;
    double ar = 0.0, ai = 0.0;
    for (int k = 0; k < nL; ++k) {
      const float2 *Dk = D + (size_t)k * nn;
      for (int m = 0; m < np; ++m) {
        const float2 a = Dk[(size_t)m * np + i], b = Dk[(size_t)m * np + j];
        // conj(a) b
        ar += (double)a.x * b.x + (double)a.y * b.y;
        ai += (double)a.x * b.y - (double)a.y * b.x;
      }
      const float2 t = Dk[(size_t)j * np + i];
      C[(size_t)k * nn + e] = make_float2(0.5f * t.x, -0.5f * t.y);
    }
    const float2 h = Hnh[e];
    Hnh[e] = make_float2((float)((double)h.x - 0.5 * ar), (float)((double)h.y - 0.5 * ai));
  }
}

__global__ __launch_bounds__(256) void k_prepare_ops(const dqh_me_inst *__restrict__ insts) {
  prepare_elems(insts[blockIdx.y], blockIdx.x * 256 + threadIdx.x, kPrepTiles * 256);
}

// Work-queue form of k_tsit5 for serving: a persistent grid whose workgroups take
// requests from a ring that the host keeps filling while the grid runs. A ring slot
// (kQueueSlotWords words) names a request's staged inputs and descriptor in pinned
// host memory and where they go in device memory: {src, dst, bytes, src_desc,
// dst_desc}. Thread 0 claims the next published slot (head < tail, compare-and-swap
// on head) or, with none published, sleeps and polls again; it gives up when the
// host sets stop, when no workgroup has claimed a request for idle_ticks of the wall
// clock, or once the workgroup is older than life_ticks (checked at every poll, so an
// idle workgroup leaves at once and a busy one after its request), so every
// workgroup exits and the launch ends regularly even under load. The workgroup then copies the inputs and the descriptor to the
// device itself, completes Hnh and C (as dqh_dense_prepare_ops), solves the request
// with the bodies of k_tsit5 and signals it; no host-side copy or launch is needed
// to start a request. busy_ticks accumulates the wall-clock ticks spent per request.
constexpr int kQueueSlotWords = DQH_QUEUE_SLOT_WORDS;

__global__ __launch_bounds__(kThreads) void k_tsit5_queue(dqh_queue q) {
  __shared__ double sh[kWaves];
  __shared__ float2 lds[kResLdsFloat2];
  __shared__ double2 dg[64];
  __shared__ uint64_t slot[kQueueSlotWords];
  int64_t *tail = gp<int64_t>(q.tail), *head = gp<int64_t>(q.head), *stop = gp<int64_t>(q.stop);
  int64_t *last_claim = head + 1;  // wall clock of the latest claim by any workgroup
  const uint64_t *ring = gp<const uint64_t>(q.ring);
  const uint64_t t_birth = wall_clock64();
  for (;;) {
    if (threadIdx.x == 0) {
      int64_t got = -1;
      const uint64_t t_idle = wall_clock64();
      for (;;) {
        if (wall_clock64() - t_birth > (uint64_t)q.life_ticks) break;
        const int64_t t = __hip_atomic_load(tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        int64_t h = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (h < t) {
          if (__hip_atomic_compare_exchange_strong(head, &h, h + 1, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)) {
            got = h;
            __hip_atomic_store(last_claim, (int64_t)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          continue;
        }
        if (__hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) break;
        // idle: no workgroup of the queue has claimed a request for idle_ticks (a
        // workgroup alone seeing no work under a steady load does not leave)
        const uint64_t last = (uint64_t)__hip_atomic_load(last_claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t now = wall_clock64();
        if (now - (last > t_idle ? last : t_idle) > (uint64_t)q.idle_ticks) break;
        __builtin_amdgcn_s_sleep(64);
      }
      const uint64_t *src = ring + (size_t)(got < 0 ? 0 : (got & (q.capacity - 1))) * kQueueSlotWords;
#pragma unroll
      for (int w = 0; w < kQueueSlotWords; ++w)
        slot[w] = got < 0 ? 0 : __hip_atomic_load(src + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    __syncthreads();
    if (slot[4] == 0) break;
    const uint64_t t0 = wall_clock64();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
    {
      // staged inputs (16-byte words) and the descriptor, host -> device
      const uint4 *src = gp<const uint4>(slot[0]);
      uint4 *dst = gp<uint4>(slot[1]);
      const int64_t nw = (int64_t)(slot[2] / 16);
      for (int64_t w = threadIdx.x; w < nw; w += kThreads) dst[w] = src[w];
      const uint64_t *sd = gp<const uint64_t>(slot[3]);
      uint64_t *dd = gp<uint64_t>(slot[4]);
      for (int w = threadIdx.x; w < (int)(sizeof(dqh_me_inst) / 8); w += kThreads) dd[w] = sd[w];
    }
    __threadfence();
    __syncthreads();
    // readfirstlane returns a signed int: each half goes through unsigned before
    // widening, or a low word >= 2^31 would sign-extend over the high word
    const uint32_t su_lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)slot[4]);
    const uint32_t su_hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(slot[4] >> 32));
    const uint64_t su = ((uint64_t)su_hi << 32) | (uint64_t)su_lo;
    const dqh_me_inst &p = *gp<const dqh_me_inst>(su);
    prepare_elems(p, threadIdx.x, kThreads);
    __threadfence();
    __syncthreads();
    tsit5_dispatch(p, sh, lds, dg);
    signal_done(p);
    if (threadIdx.x == 0)
      __hip_atomic_fetch_add(gp<int64_t>(q.busy_ticks), (int64_t)(wall_clock64() - t0), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();  // slot is rewritten for the next request
  }
}

// A batch whose instances all take the resident body for one NP: the kernel holds
// only that body, so its registers are allocated for it alone (the dispatcher's
// allocation covers every body it may call). An instance that does not fit is
// left untouched with status DQH_UNSUPPORTED.
// Waves per workgroup of the single-size kernel. 16 waves (128 VGPRs each) spill
// the np = 64 stages around every product, so every size runs 8.
template <int NP>
constexpr int res_waves() {
  return kWaves;
}

// HD: every instance carries its generator's diagonal in FP64 (hdiag).
template <int NP, bool HD>
__global__ __launch_bounds__(res_waves<NP>() * 64) void k_tsit5_res(const dqh_me_inst *__restrict__ insts) {
  constexpr int NW = res_waves<NP>();
  __shared__ double sh[NW];
  __shared__ float2 lds[kResLdsFloat2];
  __shared__ double2 dg[HD ? NP : 1];
  const dqh_me_inst &p = insts[blockIdx.x];
  if (p.np != NP || p.nL > kResMaxL || (p.flags & (DQH_FLAG_GLOBAL | DQH_FLAG_LMOD)) || (p.hdiag != 0) != HD) {
    if (threadIdx.x == 0) gp<int64_t>(p.stats)[3] = DQH_UNSUPPORTED;
  } else {
    tsit5_resident<NP, NW, HD>(p, sh, lds, dg);
  }
  signal_done(p);
}

__global__ __launch_bounds__(kThreads) void k_rhs(const dqh_me_inst *__restrict__ insts) {
  __shared__ float2 lds[kResLdsFloat2];
  const dqh_me_inst &p = insts[blockIdx.x];
  const int np = (int)p.np;
  if (p.nL <= kResMaxL && !(p.flags & (DQH_FLAG_GLOBAL | DQH_FLAG_LMOD))) {
    if (np == 16) return rhs_resident<16, kWaves>(p, lds);
    if (np == 32) return rhs_resident<32, kWaves>(p, lds);
    if (np == 48) return rhs_resident<48, kWaves>(p, lds);
    if (np == 64) return rhs_resident<64, kWaves>(p, lds);
  }
  rhs_global<kWaves>(p);
}

__global__ __launch_bounds__(kThreads) void k_rouchon1(const dqh_me_inst *__restrict__ insts) {
  __shared__ double sh[kWaves];
  const dqh_me_inst &p = insts[blockIdx.x];
  if (p.nmod > 0) {  // the Kraus sets are per step length and PWC segment only
    if (threadIdx.x == 0) gp<int64_t>(p.stats)[3] = DQH_UNSUPPORTED;
    return;
  }
  rouchon1_global<kWaves>(p, sh);
}

}  // namespace dqh

using namespace dqh;

extern "C" int64_t dqh_sizeof_me_inst(void) { return (int64_t)sizeof(dqh_me_inst); }

extern "C" int64_t dqh_me_work_elems(int64_t np, int64_t nL, int64_t npwc) {
  return (9 + nL + (npwc > 0 ? 1 : 0)) * np * np;
}

static int launch_checked(const char *what, hipError_t err) {
  if (err != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(err));
    return DQH_HIP_ERROR;
  }
  return DQH_OK;
}

extern "C" int dqh_dense_prepare_ops(const dqh_me_inst *d_insts, int64_t ninst, void *stream) {
  if (!d_insts || ninst <= 0 || ninst > 65535) return fail(DQH_BAD_ARG, "dqh_dense_prepare_ops: bad arguments");
  hipLaunchKernelGGL(k_prepare_ops, dim3(kPrepTiles, (unsigned)ninst), dim3(256), 0, (hipStream_t)stream, d_insts);
  return launch_checked("dqh_dense_prepare_ops", hipGetLastError());
}

extern "C" int dqh_lindblad_rhs_dense(const dqh_me_inst *d_insts, int64_t ninst, void *stream) {
  if (!d_insts || ninst <= 0 || ninst > (1 << 30)) return fail(DQH_BAD_ARG, "dqh_lindblad_rhs_dense: bad arguments");
  hipLaunchKernelGGL(k_rhs, dim3((unsigned)ninst), dim3(kThreads), 0, (hipStream_t)stream, d_insts);
  return launch_checked("dqh_lindblad_rhs_dense", hipGetLastError());
}

extern "C" int dqh_mesolve_tsit5_dense_queue(const dqh_queue *q, int64_t nworkgroups, void *stream) {
  if (!q || nworkgroups <= 0 || nworkgroups > 65535 || !q->ring || !q->tail || !q->head || !q->stop ||
      !q->busy_ticks || q->capacity <= 0 || (q->capacity & (q->capacity - 1)) || q->idle_ticks < 0 ||
      q->life_ticks < 0)
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dense_queue: bad arguments");
  hipLaunchKernelGGL(k_tsit5_queue, dim3((unsigned)nworkgroups), dim3(kThreads), 0, (hipStream_t)stream, *q);
  return launch_checked("dqh_mesolve_tsit5_dense_queue", hipGetLastError());
}

extern "C" int dqh_memcpy_h2d_async(void *dst, const void *src, int64_t bytes, void *stream) {
  if (!dst || !src || bytes < 0) return fail(DQH_BAD_ARG, "dqh_memcpy_h2d_async: bad arguments");
  if (bytes == 0) return DQH_OK;
  const hipError_t e = hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
  return e == hipSuccess ? DQH_OK : fail(DQH_HIP_ERROR, std::string("dqh_memcpy_h2d_async: ") + hipGetErrorString(e));
}

extern "C" int64_t dqh_wall_clock_khz(void) {
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess)
    return 0;
  return khz;
}

extern "C" int dqh_mesolve_tsit5_dense(const dqh_me_inst *d_insts, int64_t ninst, void *stream) {
  if (!d_insts || ninst <= 0 || ninst > (1 << 30)) return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dense: bad arguments");
  hipLaunchKernelGGL(k_tsit5, dim3((unsigned)ninst), dim3(kThreads), 0, (hipStream_t)stream, d_insts);
  return launch_checked("dqh_mesolve_tsit5_dense", hipGetLastError());
}

template <bool HD>
static int launch_uniform(const char *what, const dqh_me_inst *d_insts, int64_t ninst, int64_t np, void *stream) {
  if (!d_insts || ninst <= 0 || ninst > (1 << 30)) return fail(DQH_BAD_ARG, std::string(what) + ": bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)ninst);
  switch (np) {
    case 16: hipLaunchKernelGGL((k_tsit5_res<16, HD>), grid, dim3(res_waves<16>() * 64), 0, st, d_insts); break;
    case 32: hipLaunchKernelGGL((k_tsit5_res<32, HD>), grid, dim3(res_waves<32>() * 64), 0, st, d_insts); break;
    case 48: hipLaunchKernelGGL((k_tsit5_res<48, HD>), grid, dim3(res_waves<48>() * 64), 0, st, d_insts); break;
    case 64: hipLaunchKernelGGL((k_tsit5_res<64, HD>), grid, dim3(res_waves<64>() * 64), 0, st, d_insts); break;
    default: return fail(DQH_BAD_SHAPE, std::string(what) + ": np must be 16, 32, 48 or 64");
  }
  return launch_checked(what, hipGetLastError());
}

extern "C" int dqh_mesolve_tsit5_dense_uniform(const dqh_me_inst *d_insts, int64_t ninst, int64_t np, void *stream) {
  return launch_uniform<false>("dqh_mesolve_tsit5_dense_uniform", d_insts, ninst, np, stream);
}

extern "C" int dqh_mesolve_tsit5_dense_uniform_fp64diag(const dqh_me_inst *d_insts, int64_t ninst, int64_t np,
                                                        void *stream) {
  return launch_uniform<true>("dqh_mesolve_tsit5_dense_uniform_fp64diag", d_insts, ninst, np, stream);
}

extern "C" int dqh_mesolve_rouchon1_dense(const dqh_me_inst *d_insts, int64_t ninst, void *stream) {
  if (!d_insts || ninst <= 0 || ninst > (1 << 30)) return fail(DQH_BAD_ARG, "dqh_mesolve_rouchon1_dense: bad arguments");
  hipLaunchKernelGGL(k_rouchon1, dim3((unsigned)ninst), dim3(kThreads), 0, (hipStream_t)stream, d_insts);
  return launch_checked("dqh_mesolve_rouchon1_dense", hipGetLastError());
}
