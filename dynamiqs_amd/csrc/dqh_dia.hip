// Sparse-DIA Rouchon1 (gfx950): Kraus map with Cholesky normalisation for large
// Hilbert spaces whose operators are stored by diagonals.
//
// One step (rouchon_integrator.py:250-263, 394-407, 552-582) is
//   X = T^{-dag} rho T^{-1},   rho' = M0 X M0^dag + dt sum_k L_k X L_k^dag,
// with S = M0^dag M0 + dt sum L^dag L = T T^dag banded (half-bandwidth w <= nb).
// The state rho is kept as two FP32 planes, rho = hi + lo (hi = fl(rho),
// lo = fl(rho - hi): about 48 significant bits), and every step adds an FP32
// increment to it in FP64 arithmetic:
//   K1  D1 = T^{-dag} hi - hi                            (banded left solve, MFMA)
//   K2  E  = D1 + (T^{-dag} Z - Z)^dag,  Z = (hi + D1)^dag (same solve on Z), and
//       X  = hi + E in FP32 (written over D1), so that hi + E = (hi + D1) T^{-1}
//       is the normalised state for any rho, Hermitian or not,
//   K3  rho += E + dM X + X dM^dag + dM X dM^dag + sum_k A_k X A_k^dag
//       (dM = M0 - I = -i dt Hnh, A_k = sqrt(dt) L_k): one fused stencil sweep
//       that reads X at the few neighbours the diagonals name and hi, lo, E at the
//       element, and writes hi, lo in place (40 bytes of HBM per element).
// The banded solve T^dag x = s runs over blocks of nb rows from the bottom:
//   x_i = V_i s_i - V_i U_{i,i+1} x_{i+1},  V_i = U_ii^{-1},  U = T^dag,
// i.e. d_i = x_i - s_i = V'_i s_i + Wn_i x_{i+1} with V' = V - I, Wn = -V U_{i,i+1}
// (nb x nb dense blocks precomputed in FP64 on the host). Columns are
// independent, so one workgroup owns one panel of nb columns of one instance and
// walks the block rows with its previous block kept in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "dqh.h"
#include "dqh_common.h"
#include "dqh_dia_tile.h"
#include "dqh_expect.h"

namespace dqh {
namespace dia {

typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ T *gp(uint64_t a) {
  return (T *)((__attribute__((address_space(1))) T *)a);
}

// Complex product tile accumulated with three real MFMAs per k-slice (Gauss):
// p1 = sum Ar Br, p2 = sum Ai Bi, p3 = sum (Ar + Ai)(Br + Bi); then
// Re = p1 - p2 and Im = p3 - p1 - p2.
struct Acc {
  f4 p1, p2, p3;
};

__device__ __forceinline__ void cmfma(Acc &a, float ar, float ai, float2 b) {
  a.p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b.x, a.p1, 0, 0, 0);
  a.p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b.y, a.p2, 0, 0, 0);
  a.p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar + ai, b.x + b.y, a.p3, 0, 0, 0);
}

__device__ __forceinline__ float2 acc_value(const Acc &a, int v) {
  return make_float2(a.p1[v] - a.p2[v], a.p3[v] - a.p1[v] - a.p2[v]);
}

constexpr int kNB = 64;          // block rows of the banded solve = panel width
constexpr int kLD = kNB + 2;     // padded LDS row stride (float2)
constexpr int kWaves = 8;
constexpr int kThreads = kWaves * 64;

// acc[t] += A[row0:+16, kbeg:64] (global, row-major lda=64) @ B[kbeg:64, col0+16t:+16]
// (LDS, row stride ldb); kbeg (a multiple of 16) skips the columns where A is zero.
template <int TJ>
__device__ __forceinline__ void mac_strip(Acc (&acc)[TJ], const float2 *__restrict__ A, const float2 *B, int ldb,
                                          int row0, int col0, int lane, int kbeg = 0) {
  const int r = lane & 15, g = lane >> 4;
  const float2 *Ap = A + (row0 + r) * kNB + 4 * g;
  const float2 *Bp = B + (4 * g) * ldb + col0 + r;
#pragma unroll 4
  for (int k0 = kbeg; k0 < kNB; k0 += 16) {
    const float4 a01 = *reinterpret_cast<const float4 *>(Ap + k0);
    const float4 a23 = *reinterpret_cast<const float4 *>(Ap + k0 + 2);
    const float2 *bp = Bp + k0 * ldb;
#pragma unroll
    for (int t = 0; t < TJ; ++t) {
      cmfma(acc[t], a01.x, a01.y, bp[16 * t]);
      cmfma(acc[t], a01.z, a01.w, bp[ldb + 16 * t]);
      cmfma(acc[t], a23.x, a23.y, bp[2 * ldb + 16 * t]);
      cmfma(acc[t], a23.z, a23.w, bp[3 * ldb + 16 * t]);
    }
  }
}

// Banded left solve of one column panel of PW columns (64 or 32). mode 0:
// s = rho32, output D1 = x - s. mode 1: s = Z = (rho32 + D1)^dag, output
// E = D1 + (x - s)^dag and X = rho32 + E (over D1, in place: the panel's
// workgroup is the only reader of those D1 rows, each block read before it is
// written), both written transposed (through LDS, coalesced), so that
// rho + E = (rho + D1) T^-1 = T^-dag rho T^-1 holds for any rho, Hermitian or not:
// rounding that leaves an anti-Hermitian part in the state is then contracted by
// the map instead of amplified. grid: (n / PW panels, ninst); block: kThreads. The
// 64 x PW output of a block row is 4 x PW/16 tiles of 16 x 16, TJ = PW / 32 per
// wave; narrower panels give twice the workgroups (two per CU) for small batches.
template <int PW>
__global__ __launch_bounds__(kThreads) void k_lsolve(int mode, int64_t n, const float2 *__restrict__ Vp,
                                                     const float2 *__restrict__ Wn, const float2 *__restrict__ rho32,
                                                     float2 *D1, float2 *__restrict__ out) {
  constexpr int LDP = PW + 2;  // padded LDS row stride of a panel tile
  constexpr int TJ = PW / 32;  // output tiles per wave
  constexpr int NQ = kNB * PW / kThreads;  // panel (and transposed tile) elements per thread
  __shared__ float2 s_lds[kNB * LDP];  // s_i panel (64 x PW)
  __shared__ float2 prev[kNB * LDP];   // x_{i+1} panel
  // mode 1: tr holds (rho + D1)[c0 + j][r0 + i] (j < PW, i < 64), trd holds
  // D1[c0 + j][r0 + i] and then the transposed output block E[c0 + j][r0 + i]
  __shared__ float2 tr[PW * kLD];
  __shared__ float2 trd[PW * kLD];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int panel = blockIdx.x;
  const size_t inst = blockIdx.y;
  const size_t nn = (size_t)n * n;
  const float2 *R = rho32 + inst * nn;
  const float2 *Dt = D1 + inst * nn;  // mode 1: D1 rows of this panel (read, then X written over them)
  float2 *O = out + inst * nn;
  const int c0 = panel * PW;
  const int NB = (int)(n / kNB);
  // wave w owns row tile w / 2 and column tiles TJ * (w % 2) .. + TJ - 1
  const int ti = wave >> 1, tj0 = (wave & 1) * TJ;
  // the inputs of block row i, loaded into registers one block row ahead so that
  // the next ones are in flight while the current block's products run: mode 0 the
  // s_i panel of rho, mode 1 the transposed tiles rho[c0+j][r0+i], D1[c0+j][r0+i]
  // (coalesced rows of 64)
  float2 sreg[NQ], treg[NQ];
  auto load_panel = [&](int bi) {
    const int r0 = bi * kNB;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int e = threadIdx.x + kThreads * q;
      if (mode == 0) {
        sreg[q] = R[(size_t)(r0 + e / PW) * n + c0 + e % PW];
      } else {
        const size_t at = (size_t)(c0 + e / kNB) * n + r0 + e % kNB;
        sreg[q] = R[at];
        treg[q] = Dt[at];
      }
    }
  };
  for (int e = threadIdx.x; e < kNB * LDP; e += kThreads) prev[e] = make_float2(0.f, 0.f);
  load_panel(NB - 1);
  for (int bi = NB - 1; bi >= 0; --bi) {
    const int r0 = bi * kNB;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int e = threadIdx.x + kThreads * q;
      if (mode == 0) {
        s_lds[(e / PW) * LDP + e % PW] = sreg[q];
      } else {
        const int at = (e / kNB) * kLD + e % kNB;
        tr[at] = make_float2(sreg[q].x + treg[q].x, sreg[q].y + treg[q].y);
        trd[at] = treg[q];
      }
    }
    if (bi > 0) load_panel(bi - 1);
    __syncthreads();
    if (mode == 1) {
      // s[i][j] = Z[r0+i][c0+j] = conj((rho + D1)[c0+j][r0+i])
      for (int e = threadIdx.x; e < kNB * PW; e += kThreads) {
        const int i = e / PW, j = e % PW;
        const float2 z = tr[j * kLD + i];
        s_lds[i * LDP + j] = make_float2(z.x, -z.y);
      }
      __syncthreads();
    }
    Acc acc[TJ];
#pragma unroll
    for (int t = 0; t < TJ; ++t) {
      acc[t].p1 = f4{0.f, 0.f, 0.f, 0.f};
      acc[t].p2 = f4{0.f, 0.f, 0.f, 0.f};
      acc[t].p3 = f4{0.f, 0.f, 0.f, 0.f};
    }
    // V'_i is upper triangular: rows ti*16.. of it are zero left of column ti*16
    mac_strip<TJ>(acc, Vp + (size_t)bi * kNB * kNB, s_lds, LDP, ti * 16, tj0 * 16, lane, ti * 16);
    if (bi + 1 < NB) mac_strip<TJ>(acc, Wn + (size_t)bi * kNB * kNB, prev, LDP, ti * 16, tj0 * 16, lane);
    __syncthreads();  // every wave is done reading prev
    const int r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int t = 0; t < TJ; ++t) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = ti * 16 + 4 * g + v, j = (tj0 + t) * 16 + r;
        const float2 d = acc_value(acc[t], v);
        const float2 sv = s_lds[i * LDP + j];
        prev[i * LDP + j] = make_float2(sv.x + d.x, sv.y + d.y);
        if (mode == 0) {
          O[(size_t)(r0 + i) * n + c0 + j] = d;
        } else {
          // E[c0+j][r0+i] = D1[c0+j][r0+i] + conj(d[i][j]) and
          // X[c0+j][r0+i] = (rho + D1)[c0+j][r0+i] + conj(d[i][j]), staged in place
          const float2 dd = trd[j * kLD + i];
          trd[j * kLD + i] = make_float2(dd.x + d.x, dd.y - d.y);
          const float2 tt = tr[j * kLD + i];
          tr[j * kLD + i] = make_float2(tt.x + d.x, tt.y - d.y);
        }
      }
    }
    __syncthreads();
    if (mode == 1) {
      float2 *Xo = D1 + inst * nn;
      for (int e = threadIdx.x; e < PW * kNB; e += kThreads) {
        const int j = e / kNB, i = e % kNB;
        const size_t at = (size_t)(c0 + j) * n + r0 + i;
        O[at] = trd[j * kLD + i];
        Xo[at] = tr[j * kLD + i];
      }
      __syncthreads();  // tr and trd are reloaded by the next block row
    }
  }
}

// Save times inside one step: up to kMaxSaves interpolated states
// old + theta (new - old) written by one sweep; update = 0 leaves the state as it
// was (a step holding more save times runs the sweep once per group of saves).
constexpr int kMaxSaves = 4;
struct Saves {
  int32_t count;
  int32_t update;
  int64_t idx[kMaxSaves];
  double theta[kMaxSaves];
};

// FP64 update of one element of the hi / lo planes with the increment d, and its
// saves. new = (hi + lo) + d in FP64, split back into hi = fl(new), lo = fl(new - hi).
__device__ __forceinline__ void update_elem(size_t e, int i, int j, int64_t n_true, double dr, double di,
                                            float2 *__restrict__ hi, float2 *__restrict__ lo,
                                            float2 *__restrict__ st_inst, int64_t nsave, const Saves &sv) {
  const float2 h = hi[e], l = lo[e];
  const double orr = (double)h.x + (double)l.x, oi = (double)h.y + (double)l.y;
  const double nr = orr + dr, ni = oi + di;
  if (i < n_true && j < n_true) {
    for (int k = 0; k < sv.count; ++k) {
      const double th = sv.theta[k];
      st_inst[(sv.idx[k] * n_true + i) * n_true + j] = make_float2((float)(orr + th * dr), (float)(oi + th * di));
    }
  }
  if (sv.update) {
    const float hr = (float)nr, hi_ = (float)ni;
    hi[e] = make_float2(hr, hi_);
    lo[e] = make_float2((float)(nr - (double)hr), (float)(ni - (double)hi_));
  }
}

// Fused Kraus sweep: increment E + dM X + X dM^dag + dM X dM^dag + sum A X A^dag at
// every element, X (the FP32 normalised state) read at the neighbours the
// diagonals name through the caches (any offsets, any n), then the FP64 update of
// the hi / lo planes and the saves. One thread per element, rows coalesced.
__global__ __launch_bounds__(256) void k_kraus(int64_t n, int64_t n_true, dqh_dia_ops ops,
                                               const float2 *__restrict__ X, const float2 *__restrict__ E,
                                               float2 *__restrict__ hi, float2 *__restrict__ lo,
                                               float2 *__restrict__ states, int64_t nsave, Saves sv) {
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  int i, j;
  if (!diat::row_elem(n, i, j)) return;
  const size_t e = inst * nn + (size_t)i * n + j;
  const diat::GlobalX Xg{X + inst * nn, nullptr, n, i, j};
  const float2 acc = diat::stencil(Xg, ops, true, n, i, j);
  const float2 eij = E[e];
  update_elem(e, i, j, n_true, (double)eij.x + acc.x, (double)eij.y + acc.y, hi, lo,
              states + inst * nsave * n_true * n_true, nsave, sv);
}

// LDS-tiled form of k_kraus (all offsets within the halo, n % 64 == 0): one
// workgroup per 64 x 64 tile of one instance, X staged with its halo.
template <int HALO>
__global__ __launch_bounds__(diat::kThreads) void k_kraus_tiled(int64_t n, int64_t n_true, dqh_dia_ops ops,
                                                               const float2 *__restrict__ X,
                                                               const float2 *__restrict__ E, float2 *__restrict__ hi,
                                                               float2 *__restrict__ lo, float2 *__restrict__ states,
                                                               int64_t nsave, Saves sv) {
  __shared__ float2 sx[diat::Tile<HALO>::NX];
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  int i0, j0;
  diat::tile_of(n, i0, j0);
  diat::load_halo<HALO>(sx, X + inst * nn, nullptr, n, i0, j0);
  __syncthreads();
  const int jj = threadIdx.x & (diat::kTJ - 1);
  float2 *st_inst = states + inst * nsave * n_true * n_true;
#pragma unroll 1
  for (int m = 0; m < diat::kRowsPerThread; ++m) {
    const int ii = (threadIdx.x >> 6) + (diat::kThreads / diat::kTJ) * m;
    const int i = i0 + ii, j = j0 + jj;
    const float2 acc = diat::stencil(diat::LdsX<HALO>::at(sx, ii, jj), ops, true, n, i, j);
    const size_t e = inst * nn + (size_t)i * n + j;
    const float2 eij = E[e];
    update_elem(e, i, j, n_true, (double)eij.x + acc.x, (double)eij.y + acc.y, hi, lo, st_inst, nsave, sv);
  }
}

// states[inst][save_idx] <- rho (FP64 state) for saves at the start or end time
__global__ __launch_bounds__(256) void k_save(int64_t n, int64_t n_true, const double2 *__restrict__ r64,
                                              float2 *__restrict__ states, int64_t nsave, int64_t save_idx) {
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  int i, j;
  if (!diat::row_elem(n, i, j)) return;
  const int64_t e = (int64_t)i * n + j;
  if (i < n_true && j < n_true) {
    const double2 v = r64[inst * nn + e];
    states[((inst * nsave + save_idx) * n_true + i) * n_true + j] = make_float2((float)v.x, (float)v.y);
  }
}

// hi = fl(rho), lo = fl(rho - hi) from the FP64 state (start of the solve)
__global__ __launch_bounds__(256) void k_split(int64_t total, const double2 *__restrict__ r64, float2 *__restrict__ hi,
                                               float2 *__restrict__ lo) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const double2 v = r64[e];
  const float hr = (float)v.x, hi_ = (float)v.y;
  hi[e] = make_float2(hr, hi_);
  lo[e] = make_float2((float)(v.x - (double)hr), (float)(v.y - (double)hi_));
}

// rho = hi + lo in FP64 (end of the solve)
__global__ __launch_bounds__(256) void k_combine(int64_t total, const float2 *__restrict__ hi,
                                                 const float2 *__restrict__ lo, double2 *__restrict__ r64) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const float2 h = hi[e], l = lo[e];
  r64[e] = make_double2((double)h.x + (double)l.x, (double)h.y + (double)l.y);
}

}  // namespace dia
}  // namespace dqh

using namespace dqh;

extern "C" int64_t dqh_sizeof_dia_rouchon1(void) { return (int64_t)sizeof(dqh_dia_rouchon1); }

extern "C" int dqh_mesolve_rouchon1_dia(const dqh_dia_rouchon1 *pb, void *stream) {
  if (!pb) return fail(DQH_BAD_ARG, "dqh_mesolve_rouchon1_dia: null problem");
  const dqh_dia_rouchon1 &p = *pb;
  if (p.nb != dia::kNB || p.n <= 0 || p.n % dia::kNB != 0 || p.ninst <= 0 || p.n_true > p.n)
    return fail(DQH_BAD_SHAPE, "dqh_mesolve_rouchon1_dia: n must be a positive multiple of 64 and nb == 64");
  if (p.nE < 0 || (p.nE > 0 && (!p.EopsT || !p.expects)))
    return fail(DQH_BAD_ARG, "dqh_mesolve_rouchon1_dia: nE > 0 needs EopsT and expects");
  if (p.ops[0].nops < 1 || p.ops[0].nops > DQH_DIA_MAXOPS || p.nsets < 1 || p.nsets > 2)
    return fail(DQH_BAD_ARG, "dqh_mesolve_rouchon1_dia: bad operator description");
  for (int s = 0; s < p.nsets; ++s)
    for (int k = 0; k < p.ops[s].nops; ++k)
      if (p.ops[s].noff[k] < 1 || p.ops[s].noff[k] > DQH_DIA_MAXOFF)
        return fail(DQH_BAD_ARG, "dqh_mesolve_rouchon1_dia: bad number of diagonals");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = p.n, nn = n * n;
  // state planes hi, lo (rho32) and the stage buffers X (D1 before the second
  // solve) and E (work)
  float2 *hi = reinterpret_cast<float2 *>(p.rho32), *lo = hi + p.ninst * nn;
  double2 *r64 = reinterpret_cast<double2 *>(p.rho64);
  float2 *X = reinterpret_cast<float2 *>(p.work);
  float2 *E = X + p.ninst * nn;
  float2 *states = reinterpret_cast<float2 *>(p.states);
  const dim3 gel = diat::row_grid(n, p.ninst);
  // panel width of the block solves: 32 columns when 64-column panels would leave
  // fewer than two workgroups per CU (256 CUs)
  if (p.solve_panel != 0 && p.solve_panel != 32 && p.solve_panel != 64)
    return fail(DQH_BAD_ARG, "dqh_mesolve_rouchon1_dia: solve_panel must be 0, 32 or 64");
  const int panel = p.solve_panel != 0 ? (int)p.solve_panel : ((n / dia::kNB) * p.ninst < 512 ? 32 : 64);
  const dim3 gsolve((unsigned)(n / panel), (unsigned)p.ninst);
  const dim3 gtile((unsigned)((n / diat::kTI) * (n / diat::kTJ)), (unsigned)p.ninst);
  int maxoff = 0;
  for (int s = 0; s < p.nsets; ++s) maxoff = std::max(maxoff, diat::max_abs_offset(p.ops[s]));
  // the fused sweep has a global form (GLOBAL, and SPLIT, which names the former
  // two-pass form) and LDS-tiled forms
  const int halo = p.variant == DQH_SWEEP_SPLIT ? 0 : diat::resolve_halo(p.variant, n, maxoff);
  if (halo < 0)
    return fail(DQH_BAD_ARG, "dqh_mesolve_rouchon1_dia: the tiled sweep needs every |offset| <= its halo");
  const int64_t total = p.ninst * nn;
  const unsigned gflat = (unsigned)((total + 255) / 256);
  hipLaunchKernelGGL(dia::k_split, dim3(gflat), dim3(256), 0, st, total, r64, hi, lo);
  const int64_t *hs_step = reinterpret_cast<const int64_t *>(p.save_step);
  const double *hs_theta = reinterpret_cast<const double *>(p.save_theta);
  int64_t isave = 0;
  while (isave < p.nsave && hs_step[isave] < 0) {
    hipLaunchKernelGGL(dia::k_save, gel, dim3(256), 0, st, n, p.n_true, r64, states, p.nsave, isave);
    ++isave;
  }
  // optional phase timing: events around the solves and the sweep of a few steps
  double *timing = reinterpret_cast<double *>(p.timing);
  const int64_t t_first = 1, t_count = timing ? std::min<int64_t>(8, p.nsteps - 1) : 0;
  hipEvent_t ev[3 * 8] = {};
  for (int64_t k = 0; k < 3 * t_count; ++k)
    if (hipEventCreate(&ev[k]) != hipSuccess) return fail(DQH_HIP_ERROR, "dqh_mesolve_rouchon1_dia: hipEventCreate");
  auto sweep = [&](const dqh_dia_ops &ops, const dia::Saves &sv) {
    if (halo == 8)
      hipLaunchKernelGGL(dia::k_kraus_tiled<8>, gtile, dim3(diat::kThreads), 0, st, n, p.n_true, ops, X, E, hi, lo,
                         states, p.nsave, sv);
    else if (halo == 16)
      hipLaunchKernelGGL(dia::k_kraus_tiled<16>, gtile, dim3(diat::kThreads), 0, st, n, p.n_true, ops, X, E, hi, lo,
                         states, p.nsave, sv);
    else if (halo == 32)
      hipLaunchKernelGGL(dia::k_kraus_tiled<32>, gtile, dim3(diat::kThreads), 0, st, n, p.n_true, ops, X, E, hi, lo,
                         states, p.nsave, sv);
    else
      hipLaunchKernelGGL(dia::k_kraus, gel, dim3(256), 0, st, n, p.n_true, ops, X, E, hi, lo, states, p.nsave, sv);
  };
  for (int64_t m = 0; m < p.nsteps; ++m) {
    const int set = (m == p.nsteps - 1) ? p.nsets - 1 : 0;
    const float2 *Vp = reinterpret_cast<const float2 *>(p.Vp[set]);
    const float2 *Wn = reinterpret_cast<const float2 *>(p.Wn[set]);
    const int64_t tk = m - t_first;
    const bool timed = tk >= 0 && tk < t_count;
    if (timed) hipEventRecord(ev[3 * tk], st);
    // D1 = T^-dag hi - hi into the X buffer, then E and X = hi + E (over D1)
    if (panel == 32) {
      hipLaunchKernelGGL(dia::k_lsolve<32>, gsolve, dim3(dia::kThreads), 0, st, 0, n, Vp, Wn, hi, X, X);
      hipLaunchKernelGGL(dia::k_lsolve<32>, gsolve, dim3(dia::kThreads), 0, st, 1, n, Vp, Wn, hi, X, E);
    } else {
      hipLaunchKernelGGL(dia::k_lsolve<64>, gsolve, dim3(dia::kThreads), 0, st, 0, n, Vp, Wn, hi, X, X);
      hipLaunchKernelGGL(dia::k_lsolve<64>, gsolve, dim3(dia::kThreads), 0, st, 1, n, Vp, Wn, hi, X, E);
    }
    if (timed) hipEventRecord(ev[3 * tk + 1], st);
    // the save times inside this step, at most kMaxSaves per sweep: groups beyond the
    // last one are written by sweeps that leave the state unchanged
    int64_t nin = 0;
    while (isave + nin < p.nsave && hs_step[isave + nin] == m) ++nin;
    do {
      dia::Saves sv{};
      sv.count = (int32_t)std::min<int64_t>(nin, dia::kMaxSaves);
      for (int k = 0; k < sv.count; ++k) {
        sv.idx[k] = isave + k;
        sv.theta[k] = hs_theta[isave + k];
      }
      isave += sv.count;
      nin -= sv.count;
      sv.update = nin == 0;
      sweep(p.ops[set], sv);
    } while (nin > 0);
    if (timed) hipEventRecord(ev[3 * tk + 2], st);
  }
  hipLaunchKernelGGL(dia::k_combine, dim3(gflat), dim3(256), 0, st, total, hi, lo, r64);
  while (isave < p.nsave) {
    hipLaunchKernelGGL(dia::k_save, gel, dim3(256), 0, st, n, p.n_true, r64, states, p.nsave, isave);
    ++isave;
  }
  if (p.nE > 0)
    hipLaunchKernelGGL(k_expect_states, dim3((unsigned)p.nsave, (unsigned)p.ninst), dim3(256), 0, st, p.n_true, p.nsave,
                       p.nE, states, reinterpret_cast<const float2 *>(p.EopsT),
                       reinterpret_cast<float2 *>(p.expects));
  hipError_t err = hipGetLastError();
  if (err == hipSuccess && t_count > 0) {
    err = hipEventSynchronize(ev[3 * t_count - 1]);
    double solve = 0.0, sweep_ms = 0.0;
    for (int64_t k = 0; k < t_count && err == hipSuccess; ++k) {
      float a = 0.f, b = 0.f;
      err = hipEventElapsedTime(&a, ev[3 * k], ev[3 * k + 1]);
      if (err == hipSuccess) err = hipEventElapsedTime(&b, ev[3 * k + 1], ev[3 * k + 2]);
      solve += a;
      sweep_ms += b;
    }
    timing[0] = solve / (double)t_count;
    timing[1] = sweep_ms / (double)t_count;
  } else if (timing) {
    timing[0] = timing[1] = -1.0;
  }
  for (int64_t k = 0; k < 3 * t_count; ++k) hipEventDestroy(ev[k]);
  if (err != hipSuccess) return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_rouchon1_dia: ") + hipGetErrorString(err));
  return DQH_OK;
}
