// Adaptive Tsit5 for sparse-DIA master equations (gfx950).
//
// The large-n counterpart of the persistent dense kernel: a state of n = 1024
// does not fit one workgroup, so every Tsit5 stage is one grid-wide launch over
// all instances (64 x 64 output tiles per workgroup, the DIA Lindblad stencil
// of dqh_dia_tile.h) and the step controller lives on the device:
//   stage s = 1..6  k_s = f(ys_s) at the tile, then ys_{s+1} = y + h sum a_{s+1,j} k_j
//                   at the same elements (only k_j of the element itself is
//                   needed, so no second sweep); stage 6 also accumulates the
//                   scaled error of y1 = y + h sum b_j k_j in FP64 per workgroup;
//   commit          every workgroup reduces its instance's partial sums in the same
//                   order (identical accept/reject and dt everywhere, no atomics),
//                   writes dense-output saves, y <- y1 in FP64 on accept, and forms
//                   ys_1 = y + h' a_10 k_0 of the next attempt; workgroup 0 writes
//                   the instance's control block for the next step.
// Control blocks are double-buffered by step parity (every kernel of a step reads
// block `cur`, the commit writes block `1 - cur`), so a fixed sequence of launches
// is valid for any step and is captured once as a hipGraph of graph_steps steps
// and replayed. Instances keep their own t, dt, counters and save cursor; an
// instance that is done turns every later launch into an early return. The host
// loop reads a completion count after each replay; max_steps bounds the number of
// replays. Semantics follow diffrax_integrator.py:59-81, 101-147 (Tsit5 + PID
// controller + SaveAt dense output) as the dense kernel implements them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "dqh.h"
#include "dqh_common.h"
#include "dqh_dia_tile.h"
#include "dqh_expect.h"
#include "dqh_tsit5_tableau.h"

namespace dqh {
namespace diats {

struct Ctrl {
  double t, dt;
  int64_t nsteps, nacc, nrej, status, isave, rot;  // rot: slot of k_0 in the 7 stage buffers
};
static_assert(sizeof(Ctrl) == 64, "control block is 64 bytes");

struct V {
  int64_t n, n_true, ninst, nsave, max_steps, npart, flags;
  double rtol, atol, safety, fmin_, fmax_, t0, t1, inv_nn;
  float2 *k;      // [7][ninst][n][n]
  float2 *ys;     // [2][ninst][n][n]
  double *part;   // [ninst][2][npart]
  Ctrl *ctrl;     // [2][ninst]
  int64_t *flag;  // completion count
  double2 *y;     // [ninst][n][n]
  const double *tsave;
  float2 *states;
  int64_t *stats;
  // time-dependent Hamiltonian terms (npwc + nmod > 0): hcur[inst] holds the
  // diagonals of Hnh at the current stage time on the noff0 offsets of operator 0
  int64_t npwc, nseg, nmod, noff0, mod_nint;
  double mod_t0, mod_dt;
  const double *ptimes;
  const double *pvals;
  const double2 *mcoef;
  const float2 *H0;
  const float2 *Hterms;
  float2 *hcur;
};

__device__ __forceinline__ bool is_done(const Ctrl &c, double t1) { return c.status != 0 || !(c.t < t1); }
// step of the attempt described by c: the proposed dt, or exactly the rest of the
// interval when it would end within the snapping tolerance of t1
// Ctrl::rot holds the slot of k_0 (bits 0-2) and kJump: the last accepted step ended on
// a PWC breakpoint, so k_0 of the next step is evaluated anew (k_stage with s = 0)
constexpr int64_t kJump = 8;

// searchsorted(ptimes, t, 'right') - 1: the PWC segment of t (-1 before the first breakpoint)
__device__ __forceinline__ int seg_of(const V &v, double t) {
  int lo = 0, hi = (int)v.nseg;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (v.ptimes[mid] <= t)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}
// the end time, or the first breakpoint strictly after t when it comes before it
__device__ __forceinline__ double target_of(const V &v, double t) {
  if (v.npwc > 0) {
    const int m = seg_of(v, t) + 1;
    if (m < v.nseg && v.ptimes[m] < v.t1) return v.ptimes[m];
  }
  return v.t1;
}
__device__ __forceinline__ double step_h(const V &v, const Ctrl &c) {
  const double tg = target_of(v, c.t);
  return snaps_to(c.t, c.dt, tg) ? tg - c.t : c.dt;
}
__device__ __forceinline__ double step_end(const V &v, const Ctrl &c) {
  const double tg = target_of(v, c.t);
  return snaps_to(c.t, c.dt, tg) ? tg : c.t + c.dt;
}
// the attempt ends on a breakpoint before t1
__device__ __forceinline__ bool ends_on_jump(const V &v, const Ctrl &c) {
  const double tg = target_of(v, c.t);
  return tg < v.t1 && snaps_to(c.t, c.dt, tg);
}

__device__ __forceinline__ float2 *kslot(const V &v, int j, int64_t rot, size_t inst) {
  const size_t nn = (size_t)v.n * v.n;
  return v.k + ((size_t)((j + (rot & 7)) % 7) * v.ninst + inst) * nn;
}
__device__ __forceinline__ float2 *ysbuf(const V &v, int which, size_t inst) {
  const size_t nn = (size_t)v.n * v.n;
  return v.ys + ((size_t)which * v.ninst + inst) * nn;
}
__device__ __forceinline__ double *partial(const V &v, size_t inst, int which) {
  return v.part + (inst * 2 + which) * v.npart;
}

// FP64 sum over the workgroup; every thread returns the same value.
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

// Sum of one partial array of an instance, the same order in every workgroup.
template <int NT>
__device__ __forceinline__ double reduce_partials(const double *p, int64_t npart, double *sh) {
  double s = 0.0;
  for (int64_t b = threadIdx.x; b < npart; b += NT) s += p[b];
  return block_sum<NT>(s, sh);
}

__device__ __forceinline__ double sq(double2 v) { return v.x * v.x + v.y * v.y; }

// Visit the output elements of this workgroup: the tile (HALO > 0, 512 threads,
// the stage input staged in LDS with that halo) or one element per thread
// (HALO == 0, 256 threads, neighbours read through the caches). fn(i, j, f)
// receives the vector field f at (i, j).
template <int HALO, class Fn>
__device__ __forceinline__ void sweep(const V &v, const dqh_dia_ops &ops, const float2 *in, float2 *sx, Fn fn,
                                      const float2 *D0 = nullptr) {
  const int64_t n = v.n;
  if constexpr (HALO > 0) {
    int i0, j0;
    diat::tile_of(n, i0, j0);
    diat::load_halo<HALO>(sx, in, nullptr, n, i0, j0);
    __syncthreads();
    const int jj = threadIdx.x & (diat::kTJ - 1);
#pragma unroll 1
    for (int m = 0; m < diat::kRowsPerThread; ++m) {
      const int ii = (threadIdx.x >> 6) + (diat::kThreads / diat::kTJ) * m;
      const int i = i0 + ii, j = j0 + jj;
      fn(i, j, diat::stencil(diat::LdsX<HALO>::at(sx, ii, jj), ops, false, n, i, j, D0));
    }
  } else {
    int i, j;
    if (diat::row_elem(n, i, j)) {
      const diat::GlobalX X{in, nullptr, n, i, j};
      fn(i, j, diat::stencil(X, ops, false, n, i, j, D0));
    }
  }
}

template <int HALO>
constexpr int nthreads() {
  return HALO > 0 ? diat::kThreads : 256;
}

// ------------------------------------------------------------------ one step
template <int HALO>
__global__ __launch_bounds__(nthreads<HALO>()) void k_stage(V v, dqh_dia_ops ops, int s, int cur) {
  const RKTab &rk = rk_tab(v.flags);
  constexpr int NT = nthreads<HALO>();
  __shared__ double sh[NT / 64];
  float2 *sx = nullptr;
  if constexpr (HALO > 0) {
    __shared__ float2 lds[diat::Tile<HALO>::NX];
    sx = lds;
  }
  const size_t inst = blockIdx.y;
  const Ctrl c = v.ctrl[(size_t)cur * v.ninst + inst];
  if (is_done(c, v.t1)) return;
  // s = 0: k_0 = f(y) after a step that ended on a breakpoint (input: y in ys buffer 1)
  if (s == 0 && !(c.rot & kJump)) return;
  const double h = step_h(v, c);
  const int64_t n = v.n;
  const size_t nn = (size_t)n * n;
  const double2 *y = v.y + inst * nn;
  float2 *ks = kslot(v, s, c.rot, inst);
  float2 *nxt = ysbuf(v, s & 1, inst);
  double acc = 0.0;
  sweep<HALO>(v, ops, ysbuf(v, (s + 1) & 1, inst), sx, [&](int i, int j, float2 f) {
    const size_t e = (size_t)i * n + j;
    ks[e] = f;
    const double2 yv = y[e];
    if (s < 6) {
      double re = yv.x + h * rk.A[s + 1][s] * f.x, im = yv.y + h * rk.A[s + 1][s] * f.y;
      for (int q = 0; q < s; ++q) {
        const float2 kq = kslot(v, q, c.rot, inst)[e];
        re += h * rk.A[s + 1][q] * kq.x;
        im += h * rk.A[s + 1][q] * kq.y;
      }
      nxt[e] = make_float2((float)re, (float)im);
    } else if (i < v.n_true && j < v.n_true) {
      double er = rk.E[6] * f.x, ei = rk.E[6] * f.y, yr = yv.x, yi = yv.y;
      for (int q = 0; q < 6; ++q) {
        const float2 kq = kslot(v, q, c.rot, inst)[e];
        er += rk.E[q] * kq.x;
        ei += rk.E[q] * kq.y;
        yr += h * rk.A[6][q] * kq.x;
        yi += h * rk.A[6][q] * kq.y;
      }
      er *= h;
      ei *= h;
      const double sc = v.atol + v.rtol * fmax(sqrt(sq(yv)), sqrt(yr * yr + yi * yi));
      acc += (er * er + ei * ei) / (sc * sc);
    }
  }, v.hcur ? v.hcur + inst * (size_t)v.noff0 * n : nullptr);
  if (s == 6) {
    const double tot = block_sum<NT>(acc, sh);
    if (threadIdx.x == 0) partial(v, inst, 0)[blockIdx.x] = tot;
  }
}

// accept/reject, saves, y update, next stage input; grid (ceil(n^2/256), ninst)
__global__ __launch_bounds__(256) void k_commit(V v, int cur) {
  const RKTab &rk = rk_tab(v.flags);
  __shared__ double sh[4];
  const size_t inst = blockIdx.y;
  const Ctrl c = v.ctrl[(size_t)cur * v.ninst + inst];
  Ctrl *nx = v.ctrl + (size_t)(1 - cur) * v.ninst + inst;
  const bool writer = blockIdx.x == 0 && threadIdx.x == 0;
  if (is_done(c, v.t1)) {
    if (writer) *nx = c;
    return;
  }
  const double h = step_h(v, c);
  const double t_new = step_end(v, c);
  const bool jump = ends_on_jump(v, c);
  const double norm = sqrt(reduce_partials<256>(partial(v, inst, 0), v.npart, sh) * v.inv_nn);
  const bool finite = isfinite(norm);
  const bool keep = finite && norm < 1.0;
  double factor;
  if (!finite)
    factor = v.fmin_;
  else if (norm == 0.0)
    factor = v.fmax_;
  else
    factor = v.safety * pow(norm, -0.2);
  factor = fmin(fmax(factor, keep ? 1.0 : v.fmin_), v.fmax_);
  Ctrl o = c;
  o.nsteps += 1;
  int64_t is_end = c.isave;
  if (keep) {
    o.nacc += 1;
    o.t = t_new;
    o.rot = ((c.rot & 7) + 6) % 7 + (jump ? kJump : 0);  // FSAL: k_6 of this step is k_0 of the next
    while (is_end < v.nsave && v.tsave[is_end] <= t_new) ++is_end;
    o.isave = is_end;
  } else {
    o.nrej += 1;
    o.rot = c.rot & 7;
  }
  o.dt = factor * h;
  if (o.t < v.t1) {
    if (o.nsteps >= v.max_steps) {
      o.status = DQH_MAX_STEPS;
    } else {
      const double hn = step_h(v, o);
      if (!(hn > 0.0) || !(hn > 1e-14 * fmax(1.0, fabs(o.t)))) o.status = DQH_NONFINITE;
    }
  }
  const int64_t n = v.n;
  const size_t nn = (size_t)n * n;
  int i, j;
  if (diat::row_elem(n, i, j)) {
    const int64_t e = (int64_t)i * n + j;
    double2 *yp = v.y + inst * nn + e;
    const double2 yv = *yp;
    double2 y1 = yv;
    float2 k0n = kslot(v, 0, c.rot, inst)[e];
    if (keep) {
      float2 kk[7];
#pragma unroll
      for (int q = 0; q < 7; ++q) kk[q] = kslot(v, q, c.rot, inst)[e];
      const int64_t ntn = v.n_true * v.n_true;
      for (int64_t is = c.isave; is < is_end; ++is) {
        const double th = (v.tsave[is] - c.t) / h;
        double re = yv.x, im = yv.y;
#pragma unroll
        for (int q = 0; q < 7; ++q) {
          const double w = th * (rk.R[q][0] + th * (rk.R[q][1] + th * (rk.R[q][2] + th * rk.R[q][3])));
          re += h * w * kk[q].x;
          im += h * w * kk[q].y;
        }
        if (i < v.n_true && j < v.n_true)
          v.states[(inst * v.nsave + is) * ntn + (size_t)i * v.n_true + j] = make_float2((float)re, (float)im);
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        y1.x += h * rk.A[6][q] * kk[q].x;
        y1.y += h * rk.A[6][q] * kk[q].y;
      }
      *yp = y1;
      k0n = kk[6];
      // on a breakpoint, the next step evaluates k_0 = f(y1) with the new segment's H
      if (jump) ysbuf(v, 1, inst)[e] = make_float2((float)y1.x, (float)y1.y);
    }
    if (!is_done(o, v.t1)) {
      const double hn = step_h(v, o);
      ysbuf(v, 0, inst)[e] = make_float2((float)(y1.x + hn * rk.A[1][0] * k0n.x), (float)(y1.y + hn * rk.A[1][0] * k0n.y));
    }
  }
  if (writer) *nx = o;
}

// number of instances that are done, after an even number of steps (parity 0)
__global__ __launch_bounds__(256) void k_count(V v) {
  __shared__ double sh[4];
  double cnt = 0.0;
  for (int64_t b = threadIdx.x; b < v.ninst; b += 256) cnt += is_done(v.ctrl[b], v.t1) ? 1.0 : 0.0;
  const double tot = block_sum<256>(cnt, sh);
  if (threadIdx.x == 0) *v.flag = (int64_t)(tot + 0.5);
}

// ------------------------------------------------------------- initial step
// Control block, stage input ys_0 = y0 and the saves at tsave <= t0.
__global__ __launch_bounds__(256) void k_init(V v) {
  const size_t inst = blockIdx.y;
  int64_t is0 = 0;
  while (is0 < v.nsave && v.tsave[is0] <= v.t0) ++is0;
  const int64_t n = v.n;
  const size_t nn = (size_t)n * n;
  int i, j;
  if (diat::row_elem(n, i, j)) {
    const int64_t e = (int64_t)i * n + j;
    const double2 yv = v.y[inst * nn + e];
    const float2 f = make_float2((float)yv.x, (float)yv.y);
    ysbuf(v, 0, inst)[e] = f;
    const int64_t ntn = v.n_true * v.n_true;
    if (i < v.n_true && j < v.n_true)
      for (int64_t is = 0; is < is0; ++is) v.states[(inst * v.nsave + is) * ntn + (size_t)i * v.n_true + j] = f;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Ctrl c;
    c.t = v.t0;
    c.dt = 0.0;
    c.nsteps = c.nacc = c.nrej = 0;
    c.status = v.max_steps <= 0 && v.t0 < v.t1 ? DQH_MAX_STEPS : 0;
    c.isave = is0;
    c.rot = 0;
    v.ctrl[inst] = c;
    v.ctrl[v.ninst + inst] = c;
  }
}

// pass 0: k_0 = f(y0), partials |y0|^2 / sc^2 and |k_0|^2 / sc^2;
// pass 1: f(y0 + h0 k_0), partials |f - k_0|^2 / sc^2 (sc = atol + rtol |y0|).
// Hairer-Norsett-Wanner II.4 starting step as the dense kernel computes it.
template <int HALO>
__global__ __launch_bounds__(nthreads<HALO>()) void k_init_rhs(V v, dqh_dia_ops ops, int pass) {
  constexpr int NT = nthreads<HALO>();
  __shared__ double sh[NT / 64];
  float2 *sx = nullptr;
  if constexpr (HALO > 0) {
    __shared__ float2 lds[diat::Tile<HALO>::NX];
    sx = lds;
  }
  const size_t inst = blockIdx.y;
  if (is_done(v.ctrl[inst], v.t1)) return;
  const int64_t n = v.n;
  const size_t nn = (size_t)n * n;
  const double2 *y = v.y + inst * nn;
  float2 *k0 = kslot(v, 0, 0, inst);
  double a0 = 0.0, a1 = 0.0;
  sweep<HALO>(v, ops, ysbuf(v, pass, inst), sx, [&](int i, int j, float2 f) {
    const size_t e = (size_t)i * n + j;
    const double2 yv = y[e];
    const double sc = v.atol + v.rtol * sqrt(sq(yv));
    if (pass == 0) {
      k0[e] = f;
      if (i < v.n_true && j < v.n_true) {
        a0 += sq(yv) / (sc * sc);
        a1 += ((double)f.x * f.x + (double)f.y * f.y) / (sc * sc);
      }
    } else if (i < v.n_true && j < v.n_true) {
      const float2 kq = k0[e];
      const double dr = (double)f.x - kq.x, di = (double)f.y - kq.y;
      a0 += (dr * dr + di * di) / (sc * sc);
    }
  }, v.hcur ? v.hcur + inst * (size_t)v.noff0 * n : nullptr);
  const double t0s = block_sum<NT>(a0, sh);
  const double t1s = block_sum<NT>(a1, sh);
  if (threadIdx.x == 0) {
    partial(v, inst, 0)[blockIdx.x] = t0s;
    if (pass == 0) partial(v, inst, 1)[blockIdx.x] = t1s;
  }
}

// h0 from d0, d1; ys_1 = y0 + h0 k_0; h0 parked in control block 1.
__global__ __launch_bounds__(256) void k_init_h0(V v) {
  __shared__ double sh[4];
  const size_t inst = blockIdx.y;
  const Ctrl c = v.ctrl[inst];
  if (is_done(c, v.t1)) return;
  const double d0 = sqrt(reduce_partials<256>(partial(v, inst, 0), v.npart, sh) * v.inv_nn);
  const double d1 = sqrt(reduce_partials<256>(partial(v, inst, 1), v.npart, sh) * v.inv_nn);
  const double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
  const int64_t n = v.n;
  const size_t nn = (size_t)n * n;
  int i, j;
  if (diat::row_elem(n, i, j)) {
    const int64_t e = (int64_t)i * n + j;
    const double2 yv = v.y[inst * nn + e];
    const float2 kq = kslot(v, 0, 0, inst)[e];
    ysbuf(v, 1, inst)[e] = make_float2((float)(yv.x + h0 * kq.x), (float)(yv.y + h0 * kq.y));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Ctrl o = c;
    o.dt = h0;
    v.ctrl[v.ninst + inst] = o;
  }
}

// dt0 = min(100 h0, (0.01 / max(d1, d2))^kInitExp); control block 0 and ys_1 of the
// first attempt.
__global__ __launch_bounds__(256) void k_init_dt(V v) {
  const RKTab &rk = rk_tab(v.flags);
  __shared__ double sh[4];
  const size_t inst = blockIdx.y;
  const Ctrl c = v.ctrl[v.ninst + inst];
  if (is_done(c, v.t1)) return;
  const double h0 = c.dt;
  const double d1 = sqrt(reduce_partials<256>(partial(v, inst, 1), v.npart, sh) * v.inv_nn);
  const double d2 = sqrt(reduce_partials<256>(partial(v, inst, 0), v.npart, sh) * v.inv_nn) / h0;
  const double md = fmax(d1, d2);
  const double h1 = md <= 1e-15 ? fmax(1e-6, h0 * 1e-3) : pow(0.01 / md, kInitExp);
  Ctrl o = c;
  o.dt = fmin(100.0 * h0, h1);
  const double hn = step_h(v, o);
  if (!(hn > 0.0) || !(hn > 1e-14 * fmax(1.0, fabs(o.t)))) o.status = DQH_NONFINITE;
  const int64_t n = v.n;
  const size_t nn = (size_t)n * n;
  int i, j;
  if (diat::row_elem(n, i, j) && o.status == 0) {
    const int64_t e = (int64_t)i * n + j;
    const double2 yv = v.y[inst * nn + e];
    const float2 kq = kslot(v, 0, 0, inst)[e];
    ysbuf(v, 0, inst)[e] = make_float2((float)(yv.x + hn * rk.A[1][0] * kq.x), (float)(yv.y + hn * rk.A[1][0] * kq.y));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) v.ctrl[inst] = o;
}

// f_j(tau) of modulated table j (Horner on the interval that contains tau)
__device__ __forceinline__ double2 mod_at(const V &v, int j, double tau) {
  int64_t i = (int64_t)floor((tau - v.mod_t0) / v.mod_dt);
  i = i < 0 ? 0 : (i >= v.mod_nint ? v.mod_nint - 1 : i);
  const double x = tau - (v.mod_t0 + (double)i * v.mod_dt);
  const double2 *q = v.mcoef + ((size_t)j * v.mod_nint + i) * 4;
  double re = q[0].x, im = q[0].y;
#pragma unroll
  for (int m = 1; m < 4; ++m) {
    re = re * x + q[m].x;
    im = im * x + q[m].y;
  }
  return make_double2(re, im);
}

// hcur <- diagonals of Hnh(tau) = H0 + sum_j c_j Hterms_j per instance, PWC values
// on the segment of the step start. mode 0..6: stage mode of the attempt in control
// block cur (0: the k_0 refresh after a breakpoint, only then); 7: t0 (k_0 of the
// initial step); 8: t0 + h0 (the initial step's probe, h0 parked in block 1).
// grid (ceil(noff0 n / 256), ninst)
__global__ __launch_bounds__(256) void k_hupd(V v, int cur, int mode) {
  const RKTab &rk = rk_tab(v.flags);
  const size_t inst = blockIdx.y;
  double tseg, tau;
  if (mode <= 6) {
    const Ctrl c = v.ctrl[(size_t)cur * v.ninst + inst];
    if (is_done(c, v.t1) || (mode == 0 && !(c.rot & kJump))) return;
    tseg = c.t;
    tau = c.t + rk.C[mode] * step_h(v, c);
  } else {
    const Ctrl c = v.ctrl[(mode == 7 ? 0 : v.ninst) + inst];
    if (is_done(c, v.t1)) return;
    tseg = c.t;
    tau = mode == 7 ? c.t : c.t + c.dt;
  }
  const size_t m = (size_t)v.noff0 * v.n;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= m) return;
  const int nv = (int)v.nseg - 1;
  const int seg = v.npwc > 0 ? seg_of(v, tseg) : -1;
  const float2 h0 = v.H0[e];
  double re = h0.x, im = h0.y;
  if (seg >= 0 && seg < nv)
    for (int j = 0; j < (int)v.npwc; ++j) {
      const double cr = v.pvals[2 * ((size_t)j * nv + seg)], ci = v.pvals[2 * ((size_t)j * nv + seg) + 1];
      const float2 o = v.Hterms[(size_t)j * m + e];
      re += cr * o.x - ci * o.y;
      im += cr * o.y + ci * o.x;
    }
  for (int j = 0; j < (int)v.nmod; ++j) {
    const double2 f = mod_at(v, j, tau);
    const float2 o = v.Hterms[(size_t)(v.npwc + j) * m + e];
    re += f.x * o.x - f.y * o.y;
    im += f.x * o.y + f.y * o.x;
  }
  v.hcur[inst * m + e] = make_float2((float)re, (float)im);
}

// saves not reached (an instance that stopped early keeps its last state) and stats
__global__ __launch_bounds__(256) void k_finish(V v) {
  const size_t inst = blockIdx.y;
  const Ctrl c = v.ctrl[inst];
  const int64_t n = v.n;
  const size_t nn = (size_t)n * n;
  int i, j;
  if (diat::row_elem(n, i, j)) {
    const int64_t e = (int64_t)i * n + j;
    const int64_t ntn = v.n_true * v.n_true;
    if (i < v.n_true && j < v.n_true) {
      const double2 yv = v.y[inst * nn + e];
      for (int64_t is = c.isave; is < v.nsave; ++is)
        v.states[(inst * v.nsave + is) * ntn + (size_t)i * v.n_true + j] = make_float2((float)yv.x, (float)yv.y);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < 8) {
    int64_t *st = v.stats + inst * 8;
    const int64_t vals[8] = {c.nsteps, c.nacc, c.nrej, c.status, 2 + 6 * c.nsteps, 0, 0, 0};
    st[threadIdx.x] = vals[threadIdx.x];
  }
}

struct Layout {
  size_t k, ys, part, ctrl, flag, total;
};

inline size_t align256(size_t x) { return (x + 255) / 256 * 256; }

inline Layout layout(int64_t n, int64_t ninst) {
  const size_t nn = (size_t)n * n;
  const size_t npart = (size_t)n * ((n + 255) / 256);  // workgroups of a row-blocked launch
  Layout L;
  L.k = 0;
  L.ys = L.k + align256(7 * (size_t)ninst * nn * sizeof(float2));
  L.part = L.ys + align256(2 * (size_t)ninst * nn * sizeof(float2));
  L.ctrl = L.part + align256((size_t)ninst * 2 * npart * sizeof(double));
  L.flag = L.ctrl + align256(2 * (size_t)ninst * sizeof(Ctrl));
  L.total = L.flag + 256;
  return L;
}

}  // namespace diats
}  // namespace dqh

using namespace dqh;

extern "C" int64_t dqh_sizeof_dia_tsit5(void) { return (int64_t)sizeof(dqh_dia_tsit5); }

extern "C" int64_t dqh_dia_tsit5_work_bytes(int64_t n, int64_t ninst) {
  if (n <= 0 || ninst <= 0) return -1;
  return (int64_t)diats::layout(n, ninst).total;
}

namespace {

// grid of k_hupd: the noff0 x n diagonal elements per instance
inline dim3 hupd_grid(const diats::V &v) {
  return dim3((unsigned)(((size_t)v.noff0 * v.n + 255) / 256), (unsigned)v.ninst);
}

template <int HALO>
void launch_step(const diats::V &v, const dqh_dia_ops &ops, dim3 gstage, dim3 gflat, int cur, hipStream_t s) {
  const bool td = v.hcur != nullptr;
  if (v.npwc > 0) {  // k_0 anew after a breakpoint (instances that did not land on one return at once)
    hipLaunchKernelGGL(diats::k_hupd, hupd_grid(v), dim3(256), 0, s, v, cur, 0);
    hipLaunchKernelGGL(diats::k_stage<HALO>, gstage, dim3(diats::nthreads<HALO>()), 0, s, v, ops, 0, cur);
  }
  for (int sg = 1; sg <= 6; ++sg) {
    if (td) hipLaunchKernelGGL(diats::k_hupd, hupd_grid(v), dim3(256), 0, s, v, cur, sg);
    hipLaunchKernelGGL(diats::k_stage<HALO>, gstage, dim3(diats::nthreads<HALO>()), 0, s, v, ops, sg, cur);
  }
  hipLaunchKernelGGL(diats::k_commit, gflat, dim3(256), 0, s, v, cur);
}

template <int HALO>
int drive(const dqh_dia_tsit5 &p, const diats::V &v, hipStream_t st, dim3 gstage, dim3 gflat) {
  const dqh_dia_ops &ops = p.ops;
  const bool td = v.hcur != nullptr;
  hipLaunchKernelGGL(diats::k_init, gflat, dim3(256), 0, st, v);
  if (td) hipLaunchKernelGGL(diats::k_hupd, hupd_grid(v), dim3(256), 0, st, v, 0, 7);
  hipLaunchKernelGGL(diats::k_init_rhs<HALO>, gstage, dim3(diats::nthreads<HALO>()), 0, st, v, ops, 0);
  hipLaunchKernelGGL(diats::k_init_h0, gflat, dim3(256), 0, st, v);
  if (td) hipLaunchKernelGGL(diats::k_hupd, hupd_grid(v), dim3(256), 0, st, v, 0, 8);
  hipLaunchKernelGGL(diats::k_init_rhs<HALO>, gstage, dim3(diats::nthreads<HALO>()), 0, st, v, ops, 1);
  hipLaunchKernelGGL(diats::k_init_dt, gflat, dim3(256), 0, st, v);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_tsit5_dia init: ") + hipGetErrorString(err));

  const int64_t G = p.graph_steps > 0 ? p.graph_steps : 2;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cs = nullptr;
  auto cleanup = [&]() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cs) (void)hipStreamDestroy(cs);
  };
  if (p.graph_steps > 0) {
    // capture G steps + the completion count once; replay until every instance is done
    if ((err = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess ||
        (err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
      cleanup();
      return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_tsit5_dia capture: ") + hipGetErrorString(err));
    }
    for (int64_t g = 0; g < G; ++g) launch_step<HALO>(v, ops, gstage, gflat, (int)(g & 1), cs);
    hipLaunchKernelGGL(diats::k_count, dim3(1), dim3(256), 0, cs, v);
    err = hipStreamEndCapture(cs, &graph);
    if (err == hipSuccess) err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (err != hipSuccess) {
      cleanup();
      return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_tsit5_dia graph: ") + hipGetErrorString(err));
    }
  }
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (int q = 0; q < 2; ++q)
    if ((err = hipEventCreateWithFlags(&ev[q], hipEventDisableTiming)) != hipSuccess) {
      cleanup();
      return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_tsit5_dia event: ") + hipGetErrorString(err));
    }
  volatile int64_t *hf = reinterpret_cast<volatile int64_t *>(p.host_flags);
  hf[0] = hf[1] = -1;
  // every instance stops within max_steps steps, so this many replays always suffice
  const int64_t max_iters = p.max_steps / G + 3;
  bool finished = false;
  for (int64_t it = 0; it < max_iters && err == hipSuccess; ++it) {
    if (exec) {
      err = hipGraphLaunch(exec, st);
    } else {
      for (int64_t g = 0; g < G; ++g) launch_step<HALO>(v, ops, gstage, gflat, (int)(g & 1), st);
      hipLaunchKernelGGL(diats::k_count, dim3(1), dim3(256), 0, st, v);
      err = hipGetLastError();
    }
    if (err == hipSuccess)
      err = hipMemcpyAsync(const_cast<int64_t *>(hf + (it & 1)), v.flag, sizeof(int64_t), hipMemcpyDeviceToHost, st);
    if (err == hipSuccess) err = hipEventRecord(ev[it & 1], st);
    if (err == hipSuccess && it > 0) {
      err = hipEventSynchronize(ev[(it - 1) & 1]);
      if (err == hipSuccess && hf[(it - 1) & 1] >= v.ninst) {
        finished = true;
        break;
      }
    }
  }
  if (err == hipSuccess && !finished) {
    // the last replay may have finished the batch
    err = hipStreamSynchronize(st);
    int64_t last = -1;
    if (err == hipSuccess) err = hipMemcpy(&last, v.flag, sizeof(int64_t), hipMemcpyDeviceToHost);
    finished = err == hipSuccess && last >= v.ninst;
  }
  for (int q = 0; q < 2; ++q) (void)hipEventDestroy(ev[q]);
  cleanup();
  if (err != hipSuccess) return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_tsit5_dia: ") + hipGetErrorString(err));
  hipLaunchKernelGGL(diats::k_finish, gflat, dim3(256), 0, st, v);
  if (p.nE > 0)
    hipLaunchKernelGGL(k_expect_states, dim3((unsigned)p.nsave, (unsigned)p.ninst), dim3(256), 0, st, p.n_true,
                       p.nsave, p.nE, reinterpret_cast<const float2 *>(p.states),
                       reinterpret_cast<const float2 *>(p.EopsT), reinterpret_cast<float2 *>(p.expects));
  err = hipGetLastError();
  if (err != hipSuccess) return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_tsit5_dia: ") + hipGetErrorString(err));
  if (!finished) return fail(DQH_HIP_ERROR, "dqh_mesolve_tsit5_dia: instances still running after max_steps steps");
  return DQH_OK;
}

}  // namespace

extern "C" int dqh_mesolve_tsit5_dia(const dqh_dia_tsit5 *pb, void *stream) {
  if (!pb) return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: null problem");
  const dqh_dia_tsit5 &p = *pb;
  if (p.n <= 0 || p.ninst <= 0 || p.n_true <= 0 || p.n_true > p.n || p.n > (1 << 16) || p.nsave < 1)
    return fail(DQH_BAD_SHAPE, "dqh_mesolve_tsit5_dia: need 0 < n_true <= n <= 65536, ninst >= 1, nsave >= 1");
  if (!p.y || !p.work || !p.tsave || !p.states || !p.stats || !p.host_flags)
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: null buffer");
  if (p.nE < 0 || (p.nE > 0 && (!p.EopsT || !p.expects)))
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: nE > 0 needs EopsT and expects");
  if (p.graph_steps < 0 || (p.graph_steps & 1))
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: graph_steps must be 0 or a positive even number");
  if (!(p.t1 >= p.t0) || !(p.rtol >= 0) || !(p.atol >= 0) || !(p.rtol + p.atol > 0))
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: need t1 >= t0 and positive tolerances");
  if (p.ops.nops < 1 || p.ops.nops > DQH_DIA_MAXOPS)
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: 1 <= nops <= 8 (Hnh, then the jump operators)");
  for (int k = 0; k < p.ops.nops; ++k) {
    if (p.ops.noff[k] < 1 || p.ops.noff[k] > DQH_DIA_MAXOFF || !p.ops.diags[k])
      return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: bad number of diagonals or null diagonals");
    for (int u = 0; u < p.ops.noff[k]; ++u)
      if (p.ops.off[k][u] <= -p.n || p.ops.off[k][u] >= p.n)
        return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: offset outside (-n, n)");
  }
  if (p.variant == DQH_SWEEP_SPLIT)
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: the split sweep applies to the Rouchon1 Kraus map only");
  if (p.npwc < 0 || p.nmod < 0 || p.npwc + p.nmod > 64)
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: 0 <= npwc, nmod and npwc + nmod <= 64");
  if (p.npwc + p.nmod > 0 && (!p.Hterms || !p.hcur))
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: time-dependent terms need Hterms and hcur");
  if (p.npwc > 0 && (p.nseg < 2 || !p.pwc_times || !p.pwc_vals))
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: PWC terms need nseg >= 2 breakpoints and their values");
  if (p.nmod > 0 && (!p.mod_coef || p.mod_nint < 1 || !(p.mod_dt > 0)))
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: modulated terms need their tables");
  const int halo = diat::resolve_halo(p.variant, p.n, diat::max_abs_offset(p.ops));
  if (halo < 0)
    return fail(DQH_BAD_ARG, "dqh_mesolve_tsit5_dia: the tiled sweep needs n % 64 == 0 and every |offset| <= its halo");
  const size_t nn = (size_t)p.n * p.n;
  const diats::Layout L = diats::layout(p.n, p.ninst);
  char *w = reinterpret_cast<char *>(p.work);
  diats::V v;
  v.n = p.n;
  v.n_true = p.n_true;
  v.ninst = p.ninst;
  v.nsave = p.nsave;
  v.max_steps = p.max_steps;
  v.flags = p.flags;
  v.rtol = p.rtol;
  v.atol = p.atol;
  v.safety = p.safety;
  v.fmin_ = p.factormin;
  v.fmax_ = p.factormax;
  v.t0 = p.t0;
  v.t1 = p.t1;
  v.inv_nn = 1.0 / ((double)p.n_true * (double)p.n_true);
  v.k = reinterpret_cast<float2 *>(w + L.k);
  v.ys = reinterpret_cast<float2 *>(w + L.ys);
  v.part = reinterpret_cast<double *>(w + L.part);
  v.ctrl = reinterpret_cast<diats::Ctrl *>(w + L.ctrl);
  v.flag = reinterpret_cast<int64_t *>(w + L.flag);
  v.y = reinterpret_cast<double2 *>(p.y);
  v.tsave = reinterpret_cast<const double *>(p.tsave);
  v.states = reinterpret_cast<float2 *>(p.states);
  v.stats = reinterpret_cast<int64_t *>(p.stats);
  const bool td = p.npwc + p.nmod > 0;
  v.npwc = td ? p.npwc : 0;
  v.nseg = p.npwc > 0 ? p.nseg : 0;
  v.nmod = td ? p.nmod : 0;
  v.noff0 = p.ops.noff[0];
  v.mod_nint = p.mod_nint;
  v.mod_t0 = p.mod_t0;
  v.mod_dt = p.mod_dt;
  v.ptimes = reinterpret_cast<const double *>(p.pwc_times);
  v.pvals = reinterpret_cast<const double *>(p.pwc_vals);
  v.mcoef = reinterpret_cast<const double2 *>(p.mod_coef);
  v.H0 = reinterpret_cast<const float2 *>(p.ops.diags[0]);
  v.Hterms = td ? reinterpret_cast<const float2 *>(p.Hterms) : nullptr;
  v.hcur = td ? reinterpret_cast<float2 *>(p.hcur) : nullptr;
  const dim3 gflat = diat::row_grid(p.n, p.ninst);
  hipStream_t st = (hipStream_t)stream;
  const dim3 gtile((unsigned)((p.n / diat::kTI) * (p.n / diat::kTJ)), (unsigned)p.ninst);
  v.npart = halo > 0 ? gtile.x : gflat.x;
  if (halo == 8) return drive<8>(p, v, st, gtile, gflat);
  if (halo == 16) return drive<16>(p, v, st, gtile, gflat);
  if (halo == 32) return drive<32>(p, v, st, gtile, gflat);
  return drive<0>(p, v, st, gflat, gflat);
}
