// Schrodinger equation d psi / dt = -i H(t) psi (sesolve) with adaptive Tsit5 or
// Dopri5, for gfx950.
//
// One workgroup integrates one instance (its own adaptive step): one wave for
// np <= 64, four waves up to np = 1024. The FP64 state, the FP32 stage input and the
// seven FP32 stages live in LDS. The vector field is a GEMV with G = -i H read as
// G^T (stored transposed, so that consecutive rows are consecutive addresses) from
// global memory, where a batch that shares H finds it in L2. In the one-wave
// kernel each row is split over 64 / np lane groups (strided columns) and reduced
// with cross-lane shuffles, so that every lane works for np = 16 and 32; the
// four-wave kernel gives each thread whole rows. Piecewise-constant and modulated
// terms are separate products combined with their coefficients in FP64. The
// controller, error norm (RMS over the n entries), FSAL, step snapping, jump
// handling and dense-output saves are those of the mesolve kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "dqh.h"
#include "dqh_common.h"
#include "dqh_mfma.h"
#include "dqh_tsit5_tableau.h"

namespace dqh {
namespace se {

// workgroup size T (one wave or four)
template <int T>
struct SeCfg {
  static constexpr int kW = T / 64;
};

template <int T>
__device__ __forceinline__ double bsum(double v, double *red) {
  constexpr int kW = SeCfg<T>::kW;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kW; ++w) s += red[w];
  return s;
}

__device__ __forceinline__ int segment_of(const double *times, int nseg, double t) {
  int lo = 0, hi = nseg;  // searchsorted(times, t, 'right') - 1
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (times[mid] <= t)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

// coefficients of the time-dependent terms: PWC values of segment seg (zero outside)
// and the modulated f_j(tau) from their piecewise-cubic tables. The loops have
// compile-time bounds so that the coefficients stay in registers.
struct Coef {
  double2 c[8 + DQH_MAX_MOD];
  int npwc, nmod;
};

__device__ __forceinline__ void coefficients(const dqh_me_inst &p, int seg, double tau, Coef &cf) {
  cf.npwc = (int)p.npwc;
  cf.nmod = (int)p.nmod;
  const int nv = (int)p.nseg - 1;
  const double *vals = gp<const double>(p.pwc_vals);
  const bool inside = seg >= 0 && seg < nv;
#pragma unroll
  for (int j = 0; j < 8; ++j)
    cf.c[j] = (j < cf.npwc && inside) ? make_double2(vals[2 * (j * nv + seg)], vals[2 * (j * nv + seg) + 1])
                                      : make_double2(0.0, 0.0);
  if (cf.nmod > 0) {
    const int64_t nint = p.mod_nint;
    int64_t i = (int64_t)floor((tau - p.mod_t0) / p.mod_dt);
    i = i < 0 ? 0 : (i >= nint ? nint - 1 : i);
    const double x = tau - (p.mod_t0 + (double)i * p.mod_dt);
    const double2 *tab = gp<const double2>(p.mod_coef);
#pragma unroll
    for (int j = 0; j < DQH_MAX_MOD; ++j) {
      double re = 0.0, im = 0.0;
      if (j < cf.nmod) {
        const double2 *q = tab + ((size_t)j * nint + i) * 4;
        re = q[0].x;
        im = q[0].y;
        for (int m = 1; m < 4; ++m) {
          re = re * x + q[m].x;
          im = im * x + q[m].y;
        }
      }
      cf.c[8 + j] = make_double2(re, im);
    }
  }
}

// sum over j = j0, j0 + js, ... of M[i][j] ys[j] with M^T stored, so that
// consecutive rows (lanes) read consecutive elements (FP32 products and sums)
__device__ __forceinline__ float2 row_dot(const float2 *MT, int np, int i, const float2 *ys, int j0, int js) {
  float ar = 0.f, ai = 0.f;
  for (int j = j0; j < np; j += js) {
    const float2 g = MT[(size_t)j * np + i];
    const float2 v = ys[j];
    ar = fmaf(g.x, v.x, fmaf(-g.y, v.y, ar));
    ai = fmaf(g.x, v.y, fmaf(g.y, v.x, ai));
  }
  return make_float2(ar, ai);
}

// out = G(t) ys, G = -i H(t) = G0 + sum_k c_k G_k: one product per term, combined
// with the coefficients in FP64. G0^T is p.Hnh, the PWC terms' p.Hpwc and the
// modulated terms' p.Hmod. Lane group q of the one-wave kernel sums the columns
// j = q (mod 64 / np) of row i; the groups are reduced with shuffles.
template <int T>
__device__ __forceinline__ void field(const dqh_me_inst &p, const Coef &cf, const float2 *ys, float2 *out) {
  const int np = (int)p.np;
  const size_t nn = (size_t)np * np;
  const float2 *G0 = gp<const float2>(p.Hnh);
  const float2 *Gp = gp<const float2>(p.Hpwc);
  const float2 *Gm = gp<const float2>(p.Hmod);
  const int nq = T == 64 ? 64 / np : 1;  // np <= 64 in the one-wave kernel
  const int q = T == 64 ? (int)threadIdx.x / np : 0;
  for (int i0 = 0; i0 < np; i0 += T) {
    const int i = T == 64 ? (int)threadIdx.x - q * np : i0 + (int)threadIdx.x;
    const bool row = q < nq && i < np;
    double re = 0.0, im = 0.0;
    if (row) {
      const float2 a = row_dot(G0, np, i, ys, q, nq);
      re = a.x;
      im = a.y;
#pragma unroll
      for (int k = 0; k < 8 + DQH_MAX_MOD; ++k) {
        const bool on = k < 8 ? k < cf.npwc : k - 8 < cf.nmod;
        const double2 c = cf.c[k];
        if (on && (c.x != 0.0 || c.y != 0.0)) {
          const float2 r = row_dot(k < 8 ? Gp + (size_t)k * nn : Gm + (size_t)(k - 8) * nn, np, i, ys, q, nq);
          re += c.x * r.x - c.y * r.y;
          im += c.x * r.y + c.y * r.x;
        }
      }
    }
    if (T == 64) {
      for (int o = np; o < np * nq; o <<= 1) {
        re += __shfl_xor(re, o, 64);
        im += __shfl_xor(im, o, 64);
      }
    }
    if (row && q == 0) out[i] = make_float2((float)re, (float)im);
  }
  __syncthreads();
}

// save y + h sum_j w_j k_j (state and <psi|E|psi>) at save index isave
template <int T, int MAXNP>
__device__ void save_at(const dqh_me_inst &p, const double2 *y, float2 (*k)[MAXNP], double h, const double *w,
                        bool interp, int isave, double2 *v, double *red) {
  const int n = (int)p.n, np = (int)p.np, nE = (int)p.nE, ntsave = (int)p.ntsave;
  for (int i = threadIdx.x; i < np; i += T) {
    double re = y[i].x, im = y[i].y;
    if (interp)
      for (int j = 0; j < 7; ++j) {
        re += h * w[j] * k[j][i].x;
        im += h * w[j] * k[j][i].y;
      }
    v[i] = make_double2(re, im);
    if ((p.flags & DQH_FLAG_SAVE_STATES) && i < n)
      gp<float2>(p.states)[(size_t)isave * n + i] = make_float2((float)re, (float)im);
  }
  __syncthreads();
  const float2 *ET = gp<const float2>(p.EopsT);
  for (int qe = 0; qe < nE; ++qe) {
    // <v|E|v> = sum_i conj(v_i) sum_j E[i][j] v_j with E^T stored: E[i][j] = ET[j][i]
    double ar = 0.0, ai = 0.0;
    for (int i = threadIdx.x; i < np; i += T) {
      double sr = 0.0, si = 0.0;
      for (int j = 0; j < np; ++j) {
        const float2 e = ET[(size_t)qe * np * np + (size_t)j * np + i];
        sr += e.x * v[j].x - e.y * v[j].y;
        si += e.x * v[j].y + e.y * v[j].x;
      }
      ar += v[i].x * sr + v[i].y * si;
      ai += v[i].x * si - v[i].y * sr;
    }
    const double sr = bsum<T>(ar, red), si = bsum<T>(ai, red);
    if (threadIdx.x == 0) gp<float2>(p.expects)[(size_t)qe * ntsave + isave] = make_float2((float)sr, (float)si);
  }
  __syncthreads();
}

// MAXNP: the largest padded dimension the LDS arrays hold (64, 256 or 1024; the
// 1024 variant takes 112 KB of LDS, one workgroup per CU)
template <int T, int MAXNP>
__global__ __launch_bounds__(T) void k_sesolve(const dqh_me_inst *__restrict__ insts) {
  constexpr int kMaxNp = MAXNP;
  __shared__ double red[SeCfg<T>::kW];
  __shared__ double2 y[kMaxNp], y1[kMaxNp], vsave[kMaxNp];
  __shared__ float2 ys[kMaxNp];
  __shared__ float2 k[7][kMaxNp];
  const dqh_me_inst &p = insts[blockIdx.x];
  const RKTab &rk = rk_tab(p.flags);
  const int n = (int)p.n, np = (int)p.np, ntsave = (int)p.ntsave, nseg = (int)p.nseg;
  const bool pwc = p.npwc > 0, smooth = p.nmod > 0;
  const double *tsave = gp<const double>(p.tsave);
  const double *ptimes = gp<const double>(p.pwc_times);
  int64_t *stats = gp<int64_t>(p.stats);
  if (np > kMaxNp || n <= 0 || n > np || p.npwc > 8 || p.nmod > DQH_MAX_MOD || (p.npwc > 0 && nseg < 2) ||
      ((p.flags & DQH_FLAG_EULER) && (!p.step_t || p.max_steps < 0))) {
    if (threadIdx.x == 0) stats[3] = DQH_UNSUPPORTED;
    return;
  }
  const double rtol = p.rtol, atol = p.atol, inv_n = 1.0 / (double)n, t_end = p.t1;
  const double w0[7] = {0, 0, 0, 0, 0, 0, 0};
  const double2 *y0 = gp<const double2>(p.y);
  for (int i = threadIdx.x; i < np; i += T) {
    y[i] = i < n ? y0[i] : make_double2(0.0, 0.0);
    for (int j = 0; j < 7; ++j) k[j][i] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  double t = p.t0;
  int isave = 0, status = 0;
  int64_t nsteps = 0, nacc = 0, nrej = 0, nrhs = 0;
  while (isave < ntsave && tsave[isave] <= t) {
    save_at<T, MAXNP>(p, y, k, 0.0, w0, false, isave, vsave, red);
    ++isave;
  }
  if (p.flags & DQH_FLAG_EULER) {
    // fixed steps between the boundaries step_t[0..max_steps]: y <- y + h G(t) y with
    // the operators at the step start, linear interpolation to the save times
    const double *st = gp<const double>(p.step_t);
    Coef cf;
    double wl[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int64_t m = 0; m < p.max_steps; ++m) {
      const double ta = st[m], tb = st[m + 1], h = tb - ta;
      coefficients(p, pwc ? segment_of(ptimes, nseg, ta) : -1, ta, cf);
      for (int i = threadIdx.x; i < np; i += T) ys[i] = make_float2((float)y[i].x, (float)y[i].y);
      __syncthreads();
      field<T>(p, cf, ys, k[0]);
      ++nrhs;
      while (isave < ntsave && tsave[isave] <= tb + 1e-12 * fmax(1.0, fabs(tb))) {
        wl[0] = fmin(fmax((tsave[isave] - ta) / h, 0.0), 1.0);
        save_at<T, MAXNP>(p, y, k, h, wl, true, isave, vsave, red);
        ++isave;
      }
      for (int i = threadIdx.x; i < np; i += T)
        y[i] = make_double2(y[i].x + h * (double)k[0][i].x, y[i].y + h * (double)k[0][i].y);
      __syncthreads();
      ++nsteps;
      ++nacc;
    }
  } else {
    int seg = pwc ? segment_of(ptimes, nseg, t) : -1;
    Coef cf;
    coefficients(p, seg, t, cf);
    for (int i = threadIdx.x; i < np; i += T) ys[i] = make_float2((float)y[i].x, (float)y[i].y);
    __syncthreads();
    field<T>(p, cf, ys, k[0]);
    ++nrhs;
    // initial step (Hairer-Norsett-Wanner II.4), scale atol + rtol |y0|
    double dt;
    {
      double s0 = 0.0, s1 = 0.0;
      for (int i = threadIdx.x; i < n; i += T) {
        const double ay2 = y[i].x * y[i].x + y[i].y * y[i].y;
        const double sc = atol + rtol * sqrt(ay2);
        s0 += ay2 / (sc * sc);
        s1 += ((double)k[0][i].x * k[0][i].x + (double)k[0][i].y * k[0][i].y) / (sc * sc);
      }
      const double d0 = sqrt(bsum<T>(s0, red) * inv_n), d1 = sqrt(bsum<T>(s1, red) * inv_n);
      const double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
      for (int i = threadIdx.x; i < np; i += T)
        ys[i] = make_float2((float)(y[i].x + h0 * k[0][i].x), (float)(y[i].y + h0 * k[0][i].y));
      __syncthreads();
      if (smooth) coefficients(p, seg, t + h0, cf);
      field<T>(p, cf, ys, k[1]);
      ++nrhs;
      double s2 = 0.0;
      for (int i = threadIdx.x; i < n; i += T) {
        const double sc = atol + rtol * sqrt(y[i].x * y[i].x + y[i].y * y[i].y);
        const double dr = (double)k[1][i].x - k[0][i].x, di = (double)k[1][i].y - k[0][i].y;
        s2 += (dr * dr + di * di) / (sc * sc);
      }
      const double d2 = sqrt(bsum<T>(s2, red) * inv_n) / h0;
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
      if (pwc) {
        const int s = segment_of(ptimes, nseg, t);
        if (s != seg) seg = s;
      }
      if (need_k0) {
        coefficients(p, seg, t, cf);
        for (int i = threadIdx.x; i < np; i += T) ys[i] = make_float2((float)y[i].x, (float)y[i].y);
        __syncthreads();
        field<T>(p, cf, ys, k[0]);
        ++nrhs;
        need_k0 = false;
      }
      double h = dt, target = t_end;
      bool target_jump = false;
      if (pwc) {
        const int m = segment_of(ptimes, nseg, t) + 1;
        if (m < nseg && ptimes[m] < t_end) {
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
      if (!smooth) coefficients(p, seg, t, cf);
      for (int s = 1; s < 7; ++s) {
        for (int i = threadIdx.x; i < np; i += T) {
          double re = y[i].x, im = y[i].y;
          for (int j = 0; j < s; ++j) {
            re += h * rk.A[s][j] * k[j][i].x;
            im += h * rk.A[s][j] * k[j][i].y;
          }
          if (s == 6) y1[i] = make_double2(re, im);
          ys[i] = make_float2((float)re, (float)im);
        }
        __syncthreads();
        if (smooth) coefficients(p, seg, t + rk.C[s] * h, cf);
        field<T>(p, cf, ys, k[s]);
        ++nrhs;
      }
      double ssum = 0.0;
      for (int i = threadIdx.x; i < n; i += T) {
        double er = 0.0, ei = 0.0;
        for (int j = 0; j < 7; ++j) {
          er += rk.E[j] * k[j][i].x;
          ei += rk.E[j] * k[j][i].y;
        }
        er *= h;
        ei *= h;
        const double sc = atol + rtol * fmax(sqrt(y[i].x * y[i].x + y[i].y * y[i].y),
                                             sqrt(y1[i].x * y1[i].x + y1[i].y * y1[i].y));
        ssum += (er * er + ei * ei) / (sc * sc);
      }
      const double norm = sqrt(bsum<T>(ssum, red) * inv_n);
      const bool finite = isfinite(norm), keep = finite && norm < 1.0;
      double factor = !finite ? p.factormin : (norm == 0.0 ? p.factormax : p.safety * ctrl_pow(norm, -0.2));
      factor = fmin(fmax(factor, keep ? 1.0 : p.factormin), p.factormax);
      ++nsteps;
      if (keep) {
        if (threadIdx.x == 0 && nacc < p.step_log_cap) gp<double>(p.step_log)[nacc] = t_new;
        ++nacc;
        while (isave < ntsave && tsave[isave] <= t_new) {
          const double th = (tsave[isave] - t) / h;
          double w[7];
          for (int j = 0; j < 7; ++j) w[j] = th * (rk.R[j][0] + th * (rk.R[j][1] + th * (rk.R[j][2] + th * rk.R[j][3])));
          save_at<T, MAXNP>(p, y, k, h, w, true, isave, vsave, red);
          ++isave;
        }
        for (int i = threadIdx.x; i < np; i += T) {
          y[i] = y1[i];
          k[0][i] = k[6][i];
        }
        __syncthreads();
        t = t_new;
        need_k0 = made_jump;
      } else {
        ++nrej;
      }
      dt = factor * h;
    }
  }
  while (isave < ntsave) {
    save_at<T, MAXNP>(p, y, k, 0.0, w0, false, isave, vsave, red);
    ++isave;
  }
  for (int i = threadIdx.x; i < n; i += T) {
    const float2 v = make_float2((float)y[i].x, (float)y[i].y);
    gp<float2>(p.final_state)[i] = v;
    if (!(p.flags & DQH_FLAG_SAVE_STATES)) gp<float2>(p.states)[i] = v;
  }
  if (threadIdx.x == 0) {
    stats[0] = nsteps;
    stats[1] = nacc;
    stats[2] = nrej;
    stats[3] = status;
    stats[4] = nrhs;
  }
}

}  // namespace se
}  // namespace dqh

extern "C" int dqh_sesolve_tsit5(const dqh_me_inst *d_insts, int64_t ninst, int64_t np_max, void *stream) {
  if (!d_insts || ninst <= 0 || ninst > (1 << 30) || np_max <= 0 || np_max > 1024)
    return dqh::fail(DQH_BAD_ARG, "dqh_sesolve_tsit5: bad arguments");
  if (np_max <= 64)
    hipLaunchKernelGGL((dqh::se::k_sesolve<64, 64>), dim3((unsigned)ninst), dim3(64), 0, (hipStream_t)stream, d_insts);
  else if (np_max <= 256)
    hipLaunchKernelGGL((dqh::se::k_sesolve<256, 256>), dim3((unsigned)ninst), dim3(256), 0, (hipStream_t)stream,
                       d_insts);
  else
    hipLaunchKernelGGL((dqh::se::k_sesolve<256, 1024>), dim3((unsigned)ninst), dim3(256), 0, (hipStream_t)stream,
                       d_insts);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return dqh::fail(DQH_HIP_ERROR, std::string("dqh_sesolve_tsit5: ") + hipGetErrorString(err));
  return DQH_OK;
}
