// Grid-wide adaptive Tsit5 for dense master equations of large n (gfx950).
//
// The persistent kernel gives one workgroup (one CU) to each instance, which is
// the right shape for batches of small problems but leaves a single n = 256
// solve on one of 256 CUs. Here every stage of every instance is spread over the
// whole grid, in three launches:
//   k_xprod  X_k = L_k ys                          16 x 16 MFMA tiles, one per wave
//   k_tmp    tmp = Hnh ys + sum_k X_k (L_k^dag / 2) 16 x 16 MFMA tiles, one per wave
//   k_body   f = tmp + tmp^dag; k_s = f; ys <- the next stage input, in place (the
//            GEMMs that read ys are done): the increment h sum a_{s+1,j} k_j for
//            stages 1..5 (the generator is constant, so f(y + Delta) = k_0 + f(Delta)
//            and k_1..k_5 hold f(Delta_s) only), y + that increment for stage 6;
//            or on the last stage the FP64 partial sums of the scaled error
// and one commit launch per step (accept/reject, dense-output saves, y <- y1 in
// FP64, the next stage input), as in the DIA grid integrator: every workgroup
// reduces its instance's partial sums in the same order, control blocks alternate
// by step parity, and graph_steps steps are captured as a hipGraph and replayed
// until every instance is done. Instances are the dense descriptors of
// dqh_mesolve_tsit5_dense (same operators, scratch, outputs and controller,
// diffrax_integrator.py:59-81, 101-147), so the two entry points are
// interchangeable for constant generators.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "dqh.h"
#include "dqh_common.h"
#include "dqh_mfma.h"
#include "dqh_tsit5_tableau.h"

namespace dqh {
namespace dgrid {

struct Ctrl {
  double t, dt;
  int64_t nsteps, nacc, nrej, status, isave, rot;
};
static_assert(sizeof(Ctrl) == 64, "control block is 64 bytes");

struct G {
  int64_t ninst, np, npart;
  Ctrl *ctrl;     // [2][ninst]
  double *part;   // [ninst][2][npart]
  int64_t *flag;  // completion count
};

__device__ __forceinline__ bool is_done(const Ctrl &c, double t1) { return c.status != 0 || !(c.t < t1); }
__device__ __forceinline__ double step_h(const Ctrl &c, double t1) { return snaps_to(c.t, c.dt, t1) ? t1 - c.t : c.dt; }
__device__ __forceinline__ double step_end(const Ctrl &c, double t1) { return snaps_to(c.t, c.dt, t1) ? t1 : c.t + c.dt; }

// per-instance scratch (the descriptor's work buffer): k_0..k_6, ys, tmp, X_k
__device__ __forceinline__ float2 *kslot(const dqh_me_inst &p, int j, int64_t rot) {
  const size_t nn = (size_t)p.np * p.np;
  return gp<float2>(p.work) + (size_t)((j + rot) % 7) * nn;
}
__device__ __forceinline__ float2 *ysb(const dqh_me_inst &p) { return gp<float2>(p.work) + 7 * (size_t)p.np * p.np; }
__device__ __forceinline__ float2 *tmpb(const dqh_me_inst &p) { return gp<float2>(p.work) + 8 * (size_t)p.np * p.np; }
__device__ __forceinline__ float2 *xb(const dqh_me_inst &p) { return gp<float2>(p.work) + 9 * (size_t)p.np * p.np; }
__device__ __forceinline__ double *partial(const G &g, size_t inst, int which) {
  return g.part + (inst * 2 + which) * g.npart;
}

template <int NT>
__device__ __forceinline__ double block_sum(double x, double *sh) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = x;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) s += sh[w];
  return s;
}

template <int NT>
__device__ __forceinline__ double reduce_partials(const double *p, int64_t npart, double *sh) {
  double s = 0.0;
  for (int64_t b = threadIdx.x; b < npart; b += NT) s += p[b];
  return block_sum<NT>(s, sh);
}

__device__ __forceinline__ double sq(double2 v) { return v.x * v.x + v.y * v.y; }

// control block `cur` of this instance, or of block 0 when cur < 0 (initial step)
__device__ __forceinline__ Ctrl ctrl_of(const G &g, int cur) {
  return g.ctrl[(size_t)(cur < 0 ? 0 : cur) * g.ninst + blockIdx.y];
}

// ----------------------------------------------------------- vector field
// X_k = L_k ys for every jump operator; grid (ceil(maxL * T^2 / 4), ninst), 4 waves
__global__ __launch_bounds__(256) void k_xprod(const dqh_me_inst *__restrict__ insts, G g, int cur) {
  const dqh_me_inst &p = insts[blockIdx.y];
  if (cur >= 0 && is_done(ctrl_of(g, cur), p.t1)) return;
  const int np = (int)p.np, T = np >> 4;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= (int)p.nL * T * T) return;  // wave-uniform: no barrier follows
  const int lane = threadIdx.x & 63;
  const size_t nn = (size_t)np * np;
  const int kk = t / (T * T), ti = (t / T) % T, tj = t % T;
  Acc acc;
  acc_zero(acc);
  tile_mac(acc, gp<const float2>(p.Dops) + kk * nn, np, ysb(p), np, ti * 16, tj * 16, np, lane);
  tile_store(acc, xb(p) + kk * nn, np, ti * 16, tj * 16, lane);
}

// tmp = Hnh ys + sum_k X_k C_k; grid (ceil(T^2 / 4), ninst), 4 waves
__global__ __launch_bounds__(256) void k_tmp(const dqh_me_inst *__restrict__ insts, G g, int cur) {
  const dqh_me_inst &p = insts[blockIdx.y];
  if (cur >= 0 && is_done(ctrl_of(g, cur), p.t1)) return;
  const int np = (int)p.np, T = np >> 4;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= T * T) return;
  const int lane = threadIdx.x & 63;
  const size_t nn = (size_t)np * np;
  const int ti = t / T, tj = t % T;
  Acc acc;
  acc_zero(acc);
  if (p.flags & DQH_FLAG_HAS_HNH) tile_mac(acc, gp<const float2>(p.Hnh), np, ysb(p), np, ti * 16, tj * 16, np, lane);
  const float2 *X = xb(p), *C = gp<const float2>(p.Cops);
  for (int kk = 0; kk < (int)p.nL; ++kk) tile_mac(acc, X + kk * nn, np, C + kk * nn, np, ti * 16, tj * 16, np, lane);
  tile_store(acc, tmpb(p), np, ti * 16, tj * 16, lane);
}

__device__ __forceinline__ float2 field_at(const float2 *tmp, int np, int i, int j) {
  const float2 a = tmp[(size_t)i * np + j], b = tmp[(size_t)j * np + i];
  return make_float2(a.x + b.x, a.y - b.y);
}

// ------------------------------------------------------------------ one step
// stage s = 1..6 after k_xprod / k_tmp on ys; grid (ceil(np^2 / 256), ninst)
__global__ __launch_bounds__(256) void k_body(const dqh_me_inst *__restrict__ insts, G g, int s, int cur) {
  __shared__ double sh[4];
  const dqh_me_inst &p = insts[blockIdx.y];
  const RKTab &rk = rk_tab(p.flags);
  const Ctrl c = ctrl_of(g, cur);
  if (is_done(c, p.t1)) return;
  const double h = step_h(c, p.t1);
  const int np = (int)p.np, n = (int)p.n;
  const size_t nn = (size_t)np * np;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double acc = 0.0;
  if (e < (int64_t)nn) {
    const int i = (int)e / np, j = (int)e - ((int)e / np) * np;  // e < np^2 < 2^31
    const float2 f = field_at(tmpb(p), np, i, j);
    kslot(p, s, c.rot)[e] = f;
    // increment form (constant generators): stages 1..5 hold g_s = f(Delta_s),
    // k_s = k_0 + g_s, folded into the weight of k_0 (stage 6 is f(y1))
    const float2 k0 = kslot(p, 0, c.rot)[e];
    if (s < 6) {
      double a0 = 0.0;
      for (int q = 0; q <= s; ++q) a0 += rk.A[s + 1][q];
      double re = h * a0 * k0.x + h * rk.A[s + 1][s] * f.x, im = h * a0 * k0.y + h * rk.A[s + 1][s] * f.y;
      if (s + 1 == 6) {
        const double2 yv = gp<const double2>(p.y)[e];
        re += yv.x;
        im += yv.y;
      }
      for (int q = 1; q < s; ++q) {
        const float2 kq = kslot(p, q, c.rot)[e];
        re += h * rk.A[s + 1][q] * kq.x;
        im += h * rk.A[s + 1][q] * kq.y;
      }
      ysb(p)[e] = make_float2((float)re, (float)im);
    } else if (i < n && j < n) {
      const double2 yv = gp<const double2>(p.y)[e];
      double ew[7], bw[7];
      for (int q = 0; q < 7; ++q) ew[q] = rk.E[q], bw[q] = rk.A[6][q];
      fold_k0(ew);
      fold_k0(bw);
      double er = ew[6] * f.x + ew[0] * k0.x, ei = ew[6] * f.y + ew[0] * k0.y;
      double yr = yv.x + h * bw[0] * k0.x, yi = yv.y + h * bw[0] * k0.y;
      for (int q = 1; q < 6; ++q) {
        const float2 kq = kslot(p, q, c.rot)[e];
        er += ew[q] * kq.x;
        ei += ew[q] * kq.y;
        yr += h * bw[q] * kq.x;
        yi += h * bw[q] * kq.y;
      }
      er *= h;
      ei *= h;
      const double sc = p.atol + p.rtol * fmax(sqrt(sq(yv)), sqrt(yr * yr + yi * yi));
      acc += (er * er + ei * ei) / (sc * sc);
    }
  }
  if (s == 6) {
    const double tot = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) partial(g, blockIdx.y, 0)[blockIdx.x] = tot;
  }
}

// accept/reject, saves, y update, next stage input; grid (ceil(np^2 / 256), ninst)
__global__ __launch_bounds__(256) void k_commit(const dqh_me_inst *__restrict__ insts, G g, int cur) {
  __shared__ double sh[4];
  const size_t inst = blockIdx.y;
  const dqh_me_inst &p = insts[inst];
  const RKTab &rk = rk_tab(p.flags);
  const Ctrl c = ctrl_of(g, cur);
  Ctrl *nx = g.ctrl + (size_t)(1 - cur) * g.ninst + inst;
  const bool writer = blockIdx.x == 0 && threadIdx.x == 0;
  if (is_done(c, p.t1)) {
    if (writer) *nx = c;
    return;
  }
  const double h = step_h(c, p.t1);
  const double t_new = step_end(c, p.t1);
  const double inv_nn = 1.0 / ((double)p.n * (double)p.n);
  const double norm = sqrt(reduce_partials<256>(partial(g, inst, 0), g.npart, sh) * inv_nn);
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
  const double *tsave = gp<const double>(p.tsave);
  Ctrl o = c;
  o.nsteps += 1;
  int64_t is_end = c.isave;
  if (keep) {
    if (writer && c.nacc < p.step_log_cap) gp<double>(p.step_log)[c.nacc] = t_new;
    o.nacc += 1;
    o.t = t_new;
    o.rot = (c.rot + 6) % 7;  // FSAL: k_6 of this step is k_0 of the next
    while (is_end < p.ntsave && tsave[is_end] <= t_new) ++is_end;
    o.isave = is_end;
  } else {
    o.nrej += 1;
  }
  o.dt = factor * h;
  if (o.t < p.t1) {
    if (o.nsteps >= p.max_steps) {
      o.status = DQH_MAX_STEPS;
    } else {
      const double hn = step_h(o, p.t1);
      if (!(hn > 0.0) || !(hn > 1e-14 * fmax(1.0, fabs(o.t)))) o.status = DQH_NONFINITE;
    }
  }
  const int np = (int)p.np, n = (int)p.n;
  const size_t nn = (size_t)np * np;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < (int64_t)nn) {
    const int i = (int)e / np, j = (int)e - ((int)e / np) * np;  // e < np^2 < 2^31
    double2 *yp = gp<double2>(p.y) + e;
    const double2 yv = *yp;
    double2 y1 = yv;
    float2 k0n = kslot(p, 0, c.rot)[e];
    if (keep) {
      float2 kk[7];
#pragma unroll
      for (int q = 0; q < 7; ++q) kk[q] = kslot(p, q, c.rot)[e];
      if (p.flags & DQH_FLAG_SAVE_STATES) {
        float2 *st = gp<float2>(p.states);
        for (int64_t is = c.isave; is < is_end; ++is) {
          const double th = (tsave[is] - c.t) / h;
          double w[7];
#pragma unroll
          for (int q = 0; q < 7; ++q) w[q] = th * (rk.R[q][0] + th * (rk.R[q][1] + th * (rk.R[q][2] + th * rk.R[q][3])));
          fold_k0(w);
          double re = yv.x, im = yv.y;
#pragma unroll
          for (int q = 0; q < 7; ++q) {
            re += h * w[q] * kk[q].x;
            im += h * w[q] * kk[q].y;
          }
          if (i < n && j < n) st[(size_t)is * n * n + (size_t)i * n + j] = make_float2((float)re, (float)im);
        }
      }
      double bw[7];
#pragma unroll
      for (int q = 0; q < 7; ++q) bw[q] = rk.A[6][q];
      fold_k0(bw);
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        y1.x += h * bw[q] * kk[q].x;
        y1.y += h * bw[q] * kk[q].y;
      }
      *yp = y1;
      k0n = kk[6];
    }
    // stage 1 input of the next attempt: the increment hn a_10 k_0
    if (!is_done(o, p.t1)) {
      const double hn = step_h(o, p.t1);
      ysb(p)[e] = make_float2((float)(hn * rk.A[1][0] * k0n.x), (float)(hn * rk.A[1][0] * k0n.y));
    }
  }
  if (writer) *nx = o;
}

// number of instances that are done, after an even number of steps (parity 0)
__global__ __launch_bounds__(256) void k_count(const dqh_me_inst *__restrict__ insts, G g) {
  __shared__ double sh[4];
  double cnt = 0.0;
  for (int64_t b = threadIdx.x; b < g.ninst; b += 256) cnt += is_done(g.ctrl[b], insts[b].t1) ? 1.0 : 0.0;
  const double tot = block_sum<256>(cnt, sh);
  if (threadIdx.x == 0) *g.flag = (int64_t)(tot + 0.5);
}

// ------------------------------------------------------------- initial step
__global__ __launch_bounds__(256) void k_init(const dqh_me_inst *__restrict__ insts, G g) {
  const size_t inst = blockIdx.y;
  const dqh_me_inst &p = insts[inst];
  const double *tsave = gp<const double>(p.tsave);
  int64_t is0 = 0;
  while (is0 < p.ntsave && tsave[is0] <= p.t0) ++is0;
  const int np = (int)p.np, n = (int)p.n;
  const size_t nn = (size_t)np * np;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < (int64_t)nn) {
    const double2 yv = gp<const double2>(p.y)[e];
    const float2 f = make_float2((float)yv.x, (float)yv.y);
    ysb(p)[e] = f;
    const int i = (int)e / np, j = (int)e - ((int)e / np) * np;  // e < np^2 < 2^31
    if ((p.flags & DQH_FLAG_SAVE_STATES) && i < n && j < n) {
      float2 *st = gp<float2>(p.states);
      for (int64_t is = 0; is < is0; ++is) st[(size_t)is * n * n + (size_t)i * n + j] = f;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Ctrl c;
    c.t = p.t0;
    c.dt = 0.0;
    c.nsteps = c.nacc = c.nrej = 0;
    c.status = p.max_steps <= 0 && p.t0 < p.t1 ? DQH_MAX_STEPS : 0;
    c.isave = is0;
    c.rot = 0;
    g.ctrl[inst] = c;
    g.ctrl[g.ninst + inst] = c;
  }
}

// pass 0: k_0 = f(y0) and the partials of d0, d1; pass 1: partials of d2 for
// f(y0 + h0 k_0) (Hairer-Norsett-Wanner II.4, as the persistent kernel)
__global__ __launch_bounds__(256) void k_init_body(const dqh_me_inst *__restrict__ insts, G g, int pass) {
  __shared__ double sh[4];
  const size_t inst = blockIdx.y;
  const dqh_me_inst &p = insts[inst];
  if (is_done(g.ctrl[inst], p.t1)) return;
  const int np = (int)p.np, n = (int)p.n;
  const size_t nn = (size_t)np * np;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double a0 = 0.0, a1 = 0.0;
  if (e < (int64_t)nn) {
    const int i = (int)e / np, j = (int)e - ((int)e / np) * np;  // e < np^2 < 2^31
    const float2 f = field_at(tmpb(p), np, i, j);
    const double2 yv = gp<const double2>(p.y)[e];
    const double sc = p.atol + p.rtol * sqrt(sq(yv));
    float2 *k0 = kslot(p, 0, 0);
    if (pass == 0) {
      k0[e] = f;
      if (i < n && j < n) {
        a0 = sq(yv) / (sc * sc);
        a1 = ((double)f.x * f.x + (double)f.y * f.y) / (sc * sc);
      }
    } else if (i < n && j < n) {
      const float2 kq = k0[e];
      const double dr = (double)f.x - kq.x, di = (double)f.y - kq.y;
      a0 = (dr * dr + di * di) / (sc * sc);
    }
  }
  const double t0s = block_sum<256>(a0, sh);
  const double t1s = block_sum<256>(a1, sh);
  if (threadIdx.x == 0) {
    partial(g, inst, 0)[blockIdx.x] = t0s;
    if (pass == 0) partial(g, inst, 1)[blockIdx.x] = t1s;
  }
}

__global__ __launch_bounds__(256) void k_init_h0(const dqh_me_inst *__restrict__ insts, G g) {
  __shared__ double sh[4];
  const size_t inst = blockIdx.y;
  const dqh_me_inst &p = insts[inst];
  const Ctrl c = g.ctrl[inst];
  if (is_done(c, p.t1)) return;
  const double inv_nn = 1.0 / ((double)p.n * (double)p.n);
  const double d0 = sqrt(reduce_partials<256>(partial(g, inst, 0), g.npart, sh) * inv_nn);
  const double d1 = sqrt(reduce_partials<256>(partial(g, inst, 1), g.npart, sh) * inv_nn);
  const double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
  const size_t nn = (size_t)p.np * p.np;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < (int64_t)nn) {
    const double2 yv = gp<const double2>(p.y)[e];
    const float2 kq = kslot(p, 0, 0)[e];
    ysb(p)[e] = make_float2((float)(yv.x + h0 * kq.x), (float)(yv.y + h0 * kq.y));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Ctrl o = c;
    o.dt = h0;
    g.ctrl[g.ninst + inst] = o;
  }
}

__global__ __launch_bounds__(256) void k_init_dt(const dqh_me_inst *__restrict__ insts, G g) {
  __shared__ double sh[4];
  const size_t inst = blockIdx.y;
  const dqh_me_inst &p = insts[inst];
  const RKTab &rk = rk_tab(p.flags);
  const Ctrl c = g.ctrl[g.ninst + inst];
  if (is_done(c, p.t1)) return;
  const double inv_nn = 1.0 / ((double)p.n * (double)p.n);
  const double h0 = c.dt;
  const double d1 = sqrt(reduce_partials<256>(partial(g, inst, 1), g.npart, sh) * inv_nn);
  const double d2 = sqrt(reduce_partials<256>(partial(g, inst, 0), g.npart, sh) * inv_nn) / h0;
  const double md = fmax(d1, d2);
  const double h1 = md <= 1e-15 ? fmax(1e-6, h0 * 1e-3) : ctrl_pow(0.01 / md, kInitExp);
  Ctrl o = c;
  o.dt = fmin(100.0 * h0, h1);
  const double hn = step_h(o, p.t1);
  if (!(hn > 0.0) || !(hn > 1e-14 * fmax(1.0, fabs(o.t)))) o.status = DQH_NONFINITE;
  const size_t nn = (size_t)p.np * p.np;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < (int64_t)nn && o.status == 0) {  // stage 1 input (increment form)
    const float2 kq = kslot(p, 0, 0)[e];
    ysb(p)[e] = make_float2((float)(hn * rk.A[1][0] * kq.x), (float)(hn * rk.A[1][0] * kq.y));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) g.ctrl[inst] = o;
}

// saves not reached keep the last state; final state; stats
__global__ __launch_bounds__(256) void k_finish(const dqh_me_inst *__restrict__ insts, G g) {
  const size_t inst = blockIdx.y;
  const dqh_me_inst &p = insts[inst];
  const Ctrl c = g.ctrl[inst];
  const int np = (int)p.np, n = (int)p.n;
  const size_t nn = (size_t)np * np;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e < (int64_t)nn) {
    const int i = (int)e / np, j = (int)e - ((int)e / np) * np;  // e < np^2 < 2^31
    if (i < n && j < n) {
      const double2 yv = gp<const double2>(p.y)[e];
      const float2 f = make_float2((float)yv.x, (float)yv.y);
      gp<float2>(p.final_state)[(size_t)i * n + j] = f;
      float2 *st = gp<float2>(p.states);
      if (p.flags & DQH_FLAG_SAVE_STATES) {
        for (int64_t is = c.isave; is < p.ntsave; ++is) st[(size_t)is * n * n + (size_t)i * n + j] = f;
      } else {
        st[(size_t)i * n + j] = f;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < 8) {
    int64_t *st = gp<int64_t>(p.stats);
    const int64_t vals[8] = {c.nsteps, c.nacc, c.nrej, c.status, 2 + 6 * c.nsteps, 0, 0, 0};
    st[threadIdx.x] = vals[threadIdx.x];
  }
}

// expects[k][s] = Tr(E_k rho_s) from the saved states (E_k^T padded to np);
// grid (ntsave, ninst)
__global__ __launch_bounds__(256) void k_expect(const dqh_me_inst *__restrict__ insts) {
  __shared__ double sh[4];
  const dqh_me_inst &p = insts[blockIdx.y];
  const int64_t s = blockIdx.x;
  if (p.nE <= 0 || s >= p.ntsave) return;
  const int np = (int)p.np, n = (int)p.n;
  const float2 *R = gp<const float2>(p.states) + (size_t)s * n * n;
  for (int64_t q = 0; q < p.nE; ++q) {
    const float2 *ET = gp<const float2>(p.EopsT) + (size_t)q * np * np;
    double ar = 0.0, ai = 0.0;
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += 256) {
      const int i = (int)(e / n), j = (int)(e % n);
      const float2 a = ET[(size_t)i * np + j], r = R[e];
      ar += (double)a.x * r.x - (double)a.y * r.y;
      ai += (double)a.x * r.y + (double)a.y * r.x;
    }
    const double sr = block_sum<256>(ar, sh);
    const double si = block_sum<256>(ai, sh);
    if (threadIdx.x == 0) gp<float2>(p.expects)[(size_t)q * p.ntsave + s] = make_float2((float)sr, (float)si);
  }
}

struct Layout {
  size_t ctrl, part, flag, total;
};

inline size_t align256(size_t x) { return (x + 255) / 256 * 256; }

inline Layout layout(int64_t np, int64_t ninst) {
  const size_t npart = ((size_t)np * np + 255) / 256;
  Layout L;
  L.ctrl = 0;
  L.part = align256(2 * (size_t)ninst * sizeof(Ctrl));
  L.flag = L.part + align256((size_t)ninst * 2 * npart * sizeof(double));
  L.total = L.flag + 256;
  return L;
}

}  // namespace dgrid
}  // namespace dqh

using namespace dqh;

extern "C" int64_t dqh_tsit5_grid_scratch_bytes(int64_t np, int64_t ninst) {
  if (np <= 0 || ninst <= 0) return -1;
  return (int64_t)dgrid::layout(np, ninst).total;
}

namespace {

struct Launch {
  dim3 gx, gt, gflat;
};

void launch_field(const dqh_me_inst *d, const dgrid::G &g, const Launch &L, int cur, hipStream_t s) {
  hipLaunchKernelGGL(dgrid::k_xprod, L.gx, dim3(256), 0, s, d, g, cur);
  hipLaunchKernelGGL(dgrid::k_tmp, L.gt, dim3(256), 0, s, d, g, cur);
}

void launch_step(const dqh_me_inst *d, const dgrid::G &g, const Launch &L, int cur, hipStream_t s) {
  for (int sg = 1; sg <= 6; ++sg) {
    launch_field(d, g, L, cur, s);
    hipLaunchKernelGGL(dgrid::k_body, L.gflat, dim3(256), 0, s, d, g, sg, cur);
  }
  hipLaunchKernelGGL(dgrid::k_commit, L.gflat, dim3(256), 0, s, d, g, cur);
}

}  // namespace

extern "C" int dqh_mesolve_tsit5_dense_grid(const dqh_me_inst *d_insts, const dqh_me_inst *h_insts, int64_t ninst,
                                            void *scratch, int64_t scratch_bytes, int64_t graph_steps,
                                            int64_t *host_flags, void *stream) {
  if (!d_insts || !h_insts || ninst <= 0 || ninst > 65535 || !scratch || !host_flags)
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dense_grid: bad arguments");
  if (graph_steps < 0 || (graph_steps & 1))
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dense_grid: graph_steps must be 0 or a positive even number");
  const int64_t np = h_insts[0].np;
  int64_t maxL = 0, max_steps = 0;
  for (int64_t b = 0; b < ninst; ++b) {
    const dqh_me_inst &p = h_insts[b];
    if (p.np != np || p.np <= 0 || p.np % 16 != 0 || p.np > 1024 || p.n <= 0 || p.n > p.np)
      return fail(DQH_BAD_SHAPE, "dqh_mesolve_tsit5_dense_grid: every instance needs the same np (multiple of 16)");
    if (p.npwc != 0 || p.nmod != 0 || p.nL < 0 || p.nL > 8 || p.ntsave < 1 || !p.work || !p.y || !p.tsave || !p.stats ||
        !p.states || !p.final_state || (p.nL > 0 && (!p.Dops || !p.Cops)))
      return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dense_grid: constant generators only (npwc == nmod == 0), nL <= 8");
    if (p.flags & DQH_FLAG_DOPRI8)
      return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dense_grid: Dopri8 instances run on dqh_mesolve_tsit5_dense");
    if (p.hdiag != 0)
      return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dense_grid: descriptors with hdiag (FP64 diagonal) run on "
                               "dqh_mesolve_tsit5_dense_uniform_fp64diag only");
    if (p.nE > 0 && (!(p.flags & DQH_FLAG_SAVE_STATES) || !p.EopsT || !p.expects))
      return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dense_grid: expectation values need saved states");
    if (!(p.t1 >= p.t0) || !(p.rtol + p.atol > 0))
      return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dense_grid: need t1 >= t0 and positive tolerances");
    maxL = std::max(maxL, p.nL);
    max_steps = std::max(max_steps, p.max_steps);
  }
  const dgrid::Layout Ly = dgrid::layout(np, ninst);
  if (scratch_bytes < (int64_t)Ly.total)
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dense_grid: scratch smaller than dqh_tsit5_grid_scratch_bytes");
  char *w = reinterpret_cast<char *>(scratch);
  dgrid::G g;
  g.ninst = ninst;
  g.np = np;
  g.ctrl = reinterpret_cast<dgrid::Ctrl *>(w + Ly.ctrl);
  g.part = reinterpret_cast<double *>(w + Ly.part);
  g.flag = reinterpret_cast<int64_t *>(w + Ly.flag);
  const int64_t T = np / 16;
  Launch L;
  L.gx = dim3((unsigned)std::max<int64_t>(1, (maxL * T * T + 3) / 4), (unsigned)ninst);
  L.gt = dim3((unsigned)((T * T + 3) / 4), (unsigned)ninst);
  L.gflat = dim3((unsigned)((np * np + 255) / 256), (unsigned)ninst);
  g.npart = L.gflat.x;
  hipStream_t st = (hipStream_t)stream;

  hipLaunchKernelGGL(dgrid::k_init, L.gflat, dim3(256), 0, st, d_insts, g);
  launch_field(d_insts, g, L, -1, st);
  hipLaunchKernelGGL(dgrid::k_init_body, L.gflat, dim3(256), 0, st, d_insts, g, 0);
  hipLaunchKernelGGL(dgrid::k_init_h0, L.gflat, dim3(256), 0, st, d_insts, g);
  launch_field(d_insts, g, L, -1, st);
  hipLaunchKernelGGL(dgrid::k_init_body, L.gflat, dim3(256), 0, st, d_insts, g, 1);
  hipLaunchKernelGGL(dgrid::k_init_dt, L.gflat, dim3(256), 0, st, d_insts, g);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess)
    return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_tsit5_dense_grid init: ") + hipGetErrorString(err));

  const int64_t G = graph_steps > 0 ? graph_steps : 2;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cs = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  auto cleanup = [&]() {
    for (int q = 0; q < 2; ++q)
      if (ev[q]) (void)hipEventDestroy(ev[q]);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cs) (void)hipStreamDestroy(cs);
  };
  if (graph_steps > 0) {
    if ((err = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) == hipSuccess &&
        (err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) == hipSuccess) {
      for (int64_t q = 0; q < G; ++q) launch_step(d_insts, g, L, (int)(q & 1), cs);
      hipLaunchKernelGGL(dgrid::k_count, dim3(1), dim3(256), 0, cs, d_insts, g);
      err = hipStreamEndCapture(cs, &graph);
      if (err == hipSuccess) err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    }
    if (err != hipSuccess) {
      cleanup();
      return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_tsit5_dense_grid graph: ") + hipGetErrorString(err));
    }
  }
  for (int q = 0; q < 2 && err == hipSuccess; ++q) err = hipEventCreateWithFlags(&ev[q], hipEventDisableTiming);
  volatile int64_t *hf = host_flags;
  hf[0] = hf[1] = -1;
  const int64_t max_iters = max_steps / G + 3;  // every instance stops within its max_steps
  bool finished = false;
  for (int64_t it = 0; it < max_iters && err == hipSuccess; ++it) {
    if (exec) {
      err = hipGraphLaunch(exec, st);
    } else {
      for (int64_t q = 0; q < G; ++q) launch_step(d_insts, g, L, (int)(q & 1), st);
      hipLaunchKernelGGL(dgrid::k_count, dim3(1), dim3(256), 0, st, d_insts, g);
      err = hipGetLastError();
    }
    if (err == hipSuccess)
      err = hipMemcpyAsync(const_cast<int64_t *>(hf + (it & 1)), g.flag, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (err == hipSuccess) err = hipEventRecord(ev[it & 1], st);
    if (err == hipSuccess && it > 0) {
      err = hipEventSynchronize(ev[(it - 1) & 1]);
      if (err == hipSuccess && hf[(it - 1) & 1] >= ninst) {
        finished = true;
        break;
      }
    }
  }
  if (err == hipSuccess && !finished) {
    err = hipStreamSynchronize(st);
    int64_t last = -1;
    if (err == hipSuccess) err = hipMemcpy(&last, g.flag, sizeof(int64_t), hipMemcpyDeviceToHost);
    finished = err == hipSuccess && last >= ninst;
  }
  cleanup();
  if (err != hipSuccess) return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_tsit5_dense_grid: ") + hipGetErrorString(err));
  hipLaunchKernelGGL(dgrid::k_finish, L.gflat, dim3(256), 0, st, d_insts, g);
  int64_t maxsave = 0;
  for (int64_t b = 0; b < ninst; ++b) maxsave = std::max(maxsave, h_insts[b].nE > 0 ? h_insts[b].ntsave : 0);
  if (maxsave > 0) hipLaunchKernelGGL(dgrid::k_expect, dim3((unsigned)maxsave, (unsigned)ninst), dim3(256), 0, st, d_insts);
  err = hipGetLastError();
  if (err != hipSuccess) return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_tsit5_dense_grid: ") + hipGetErrorString(err));
  if (!finished) return fail(DQH_HIP_ERROR, "dqh_mesolve_tsit5_dense_grid: instances still running after max_steps steps");
  return DQH_OK;
}
