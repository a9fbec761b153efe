// Sparse-DIA Rouchon1 (gfx950): Kraus map with Cholesky normalisation for large
// Hilbert spaces whose operators are stored by diagonals.
//
// One step (rouchon_integrator.py:250-263, 394-407, 552-582) is
//   X = T^{-dag} rho T^{-1},   rho' = M0 X M0^dag + dt sum_k L_k X L_k^dag,
// with S = M0^dag M0 + dt sum L^dag L = T T^dag banded (half-bandwidth w <= nb).
// The state rho is kept in FP64 and every step adds an FP32 increment:
//   K1  D1 = T^{-dag} rho - rho                          (banded left solve, MFMA)
//   K2  E  = D1 + (T^{-dag} Z - Z)^dag,  Z = (rho + D1)^dag (same solve on Z)
//       so that X = rho + E = (rho + D1) T^{-1} is the normalised state for any
//       rho, Hermitian or not,
//   K3  rho += E + dM X + X dM^dag + dM X dM^dag + sum_k A_k X A_k^dag
//       (dM = M0 - I = -i dt Hnh, A_k = sqrt(dt) L_k; a DIA stencil sweep).
// The banded solve T^dag x = s runs over blocks of nb rows from the bottom:
//   x_i = V_i s_i - V_i U_{i,i+1} x_{i+1},  V_i = U_ii^{-1},  U = T^dag,
// i.e. d_i = x_i - s_i = V'_i s_i + Wn_i x_{i+1} with V' = V - I, Wn = -V U_{i,i+1}
// (nb x nb dense blocks precomputed in FP64 on the host). Columns are
// independent, so one workgroup owns one panel of nb columns of one instance and
// walks the block rows with its previous block kept in LDS.
#include <hip/hip_runtime.h>

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

struct Acc {
  f4 re, im;
};

__device__ __forceinline__ void cmfma(Acc &a, float ar, float ai, float2 b) {
  a.re = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b.x, a.re, 0, 0, 0);
  a.re = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, b.y, a.re, 0, 0, 0);
  a.im = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, b.y, a.im, 0, 0, 0);
  a.im = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, b.x, a.im, 0, 0, 0);
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
// E = D1 + (x - s)^dag, written transposed (through LDS, coalesced), so that
// rho + E = (rho + D1) T^-1 = T^-dag rho T^-1 holds for any rho, Hermitian or not:
// rounding that leaves an anti-Hermitian part in the state is then contracted by
// the map instead of amplified. grid: (n / PW panels, ninst); block: kThreads. The
// 64 x PW output of a block row is 4 x PW/16 tiles of 16 x 16, TJ = PW / 32 per
// wave; narrower panels give twice the workgroups (two per CU) for small batches.
template <int PW>
__global__ __launch_bounds__(kThreads) void k_lsolve(int mode, int64_t n, const float2 *__restrict__ Vp,
                                                     const float2 *__restrict__ Wn, const float2 *__restrict__ rho32,
                                                     const float2 *__restrict__ D1, float2 *__restrict__ out) {
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
  const float2 *Dt = D1 + inst * nn;
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
      acc[t].re = f4{0.f, 0.f, 0.f, 0.f};
      acc[t].im = f4{0.f, 0.f, 0.f, 0.f};
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
        const float2 d = make_float2(acc[t].re[v], acc[t].im[v]);
        const float2 sv = s_lds[i * LDP + j];
        prev[i * LDP + j] = make_float2(sv.x + d.x, sv.y + d.y);
        if (mode == 0) {
          O[(size_t)(r0 + i) * n + c0 + j] = d;
        } else {
          // E[c0+j][r0+i] = D1[c0+j][r0+i] + conj(d[i][j]), staged in place
          const float2 dd = trd[j * kLD + i];
          trd[j * kLD + i] = make_float2(dd.x + d.x, dd.y - d.y);
        }
      }
    }
    __syncthreads();
    if (mode == 1) {
      for (int e = threadIdx.x; e < PW * kNB; e += kThreads) {
        const int j = e / kNB, i = e % kNB;
        O[(size_t)(c0 + j) * n + r0 + i] = trd[j * kLD + i];
      }
      __syncthreads();  // trd is reloaded by the next block row
    }
  }
}

// Kraus increment + FP64 update + save. One thread per element, neighbours of
// X = rho32 + E read through the caches (any offsets, any n).
__global__ __launch_bounds__(256) void k_kraus(int64_t n, int64_t n_true, dqh_dia_ops ops,
                                               const float2 *__restrict__ r32_in, const float2 *__restrict__ E,
                                               double2 *__restrict__ r64, float2 *__restrict__ r32_out,
                                               float2 *__restrict__ states, int64_t nsave, int64_t save_idx,
                                               double theta) {
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  int i, j;
  if (!diat::row_elem(n, i, j)) return;
  const int64_t e = (int64_t)i * n + j;
  const float2 *Ei = E + inst * nn;
  const diat::GlobalX X{r32_in + inst * nn, Ei, n, i, j};
  const float2 acc = diat::stencil(X, ops, true, n, i, j);
  const float2 eij = Ei[e];
  const double dr = (double)eij.x + acc.x, di = (double)eij.y + acc.y;
  double2 *y = r64 + inst * nn;
  const double2 old = y[e];
  const double2 nw = make_double2(old.x + dr, old.y + di);
  y[e] = nw;
  r32_out[inst * nn + e] = make_float2((float)nw.x, (float)nw.y);
  if (save_idx >= 0 && i < n_true && j < n_true) {
    const double sr = old.x + theta * dr, si = old.y + theta * di;
    states[((inst * nsave + save_idx) * n_true + i) * n_true + j] = make_float2((float)sr, (float)si);
  }
}

// Split form of the Kraus sweep, pass 1: dY = dM X with X = rho32 + E (op 0 by
// diagonals, 3 neighbours along the column for the two-mode operators). Writes
// dY to scratch; one thread per element.
__global__ __launch_bounds__(256) void k_kraus_left(int64_t n, dqh_dia_ops ops, const float2 *__restrict__ r32,
                                                    const float2 *__restrict__ E, float2 *__restrict__ dY) {
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  int i, j;
  if (!diat::row_elem(n, i, j)) return;
  const int64_t e = (int64_t)i * n + j;
  const float2 *R = r32 + inst * nn, *Ei = E + inst * nn;
  const float2 *D = diat::cgp<float2>(ops.diags[0]);
  float ar = 0.f, ai = 0.f;
  for (int u = 0; u < ops.noff[0]; ++u) {
    const int64_t a = (int64_t)i + ops.off[0][u];
    if (a < 0 || a >= n) continue;
    const size_t o = (size_t)a * n + j;
    const float2 m = D[(size_t)u * n + a];
    const float2 r = R[o], x = Ei[o];
    const float xr = r.x + x.x, xi = r.y + x.y;
    ar += m.x * xr - m.y * xi;
    ai += m.x * xi + m.y * xr;
  }
  dY[inst * nn + e] = make_float2(ar, ai);
}

// Split form, pass 2: increment E + dY + (X + dY) dM^dag + sum_{s>=1} A_s X A_s^dag
// (= E + dM X + X dM^dag + dM X dM^dag + sum A X A^dag), FP64 update and save.
__global__ __launch_bounds__(256) void k_kraus_right(int64_t n, int64_t n_true, dqh_dia_ops ops,
                                                     const float2 *__restrict__ r32_in, const float2 *__restrict__ E,
                                                     const float2 *__restrict__ dY, double2 *__restrict__ r64,
                                                     float2 *__restrict__ r32_out, float2 *__restrict__ states,
                                                     int64_t nsave, int64_t save_idx, double theta) {
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  int i, j;
  if (!diat::row_elem(n, i, j)) return;
  const int64_t e = (int64_t)i * n + j;
  const float2 *R = r32_in + inst * nn, *Ei = E + inst * nn, *Y = dY + inst * nn;
  const float2 *D = diat::cgp<float2>(ops.diags[0]);
  float ar = 0.f, ai = 0.f;
  for (int u = 0; u < ops.noff[0]; ++u) {
    const int64_t b = (int64_t)j + ops.off[0][u];
    if (b < 0 || b >= n) continue;
    const size_t o = (size_t)i * n + b;
    const float2 m = D[(size_t)u * n + b];  // conj(dM[j][b])
    const float2 r = R[o], x = Ei[o], y = Y[o];
    const float wr = r.x + x.x + y.x, wi = r.y + x.y + y.y;
    ar += wr * m.x + wi * m.y;
    ai += wi * m.x - wr * m.y;
  }
  const diat::GlobalX X{R, Ei, n, i, j};
  for (int s = 1; s < ops.nops; ++s) {
    const float2 *A = diat::cgp<float2>(ops.diags[s]);
    const int no = ops.noff[s];
    for (int u = 0; u < no; ++u) {
      const int p = ops.off[s][u];
      const float2 ma = A[(size_t)u * n + diat::clampi((int64_t)i + p, n)];
      float zr = 0.f, zi = 0.f;
      for (int v = 0; v < no; ++v) {
        const int q = ops.off[s][v];
        const float2 mb = A[(size_t)v * n + diat::clampi((int64_t)j + q, n)];
        const float2 x = X(p, q);
        zr += x.x * mb.x + x.y * mb.y;
        zi += x.y * mb.x - x.x * mb.y;
      }
      ar += ma.x * zr - ma.y * zi;
      ai += ma.x * zi + ma.y * zr;
    }
  }
  const float2 eij = Ei[e], yij = Y[e];
  const double dr = (double)eij.x + yij.x + ar, di = (double)eij.y + yij.y + ai;
  double2 *yv = r64 + inst * nn;
  const double2 old = yv[e];
  const double2 nw = make_double2(old.x + dr, old.y + di);
  yv[e] = nw;
  r32_out[inst * nn + e] = make_float2((float)nw.x, (float)nw.y);
  if (save_idx >= 0 && i < n_true && j < n_true) {
    const double sr = old.x + theta * dr, si = old.y + theta * di;
    states[((inst * nsave + save_idx) * n_true + i) * n_true + j] = make_float2((float)sr, (float)si);
  }
}

// LDS-tiled form of k_kraus (all offsets within the halo, n % 64 == 0): one
// workgroup per 64 x 64 tile of one instance; X = rho32 + E staged with its halo.
template <int HALO>
__global__ __launch_bounds__(diat::kThreads) void k_kraus_tiled(int64_t n, int64_t n_true, dqh_dia_ops ops,
                                                               const float2 *__restrict__ r32_in,
                                                               const float2 *__restrict__ E, double2 *__restrict__ r64,
                                                               float2 *__restrict__ r32_out,
                                                               float2 *__restrict__ states, int64_t nsave,
                                                               int64_t save_idx, double theta) {
  __shared__ float2 sx[diat::Tile<HALO>::NX];
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  int i0, j0;
  diat::tile_of(n, i0, j0);
  diat::load_halo<HALO>(sx, r32_in + inst * nn, E + inst * nn, n, i0, j0);
  __syncthreads();
  const int jj = threadIdx.x & (diat::kTJ - 1);
  const float2 *Ei = E + inst * nn;
  double2 *y = r64 + inst * nn;
#pragma unroll 1
  for (int m = 0; m < diat::kRowsPerThread; ++m) {
    const int ii = (threadIdx.x >> 6) + (diat::kThreads / diat::kTJ) * m;
    const int i = i0 + ii, j = j0 + jj;
    const float2 acc = diat::stencil(diat::LdsX<HALO>::at(sx, ii, jj), ops, true, n, i, j);
    const size_t e = (size_t)i * n + j;
    const float2 eij = Ei[e];
    const double dr = (double)eij.x + acc.x, di = (double)eij.y + acc.y;
    const double2 old = y[e];
    const double2 nw = make_double2(old.x + dr, old.y + di);
    y[e] = nw;
    r32_out[inst * nn + e] = make_float2((float)nw.x, (float)nw.y);
    if (save_idx >= 0 && i < n_true && j < n_true) {
      const double sr = old.x + theta * dr, si = old.y + theta * di;
      states[((inst * nsave + save_idx) * n_true + i) * n_true + j] = make_float2((float)sr, (float)si);
    }
  }
}

// states[inst][save_idx] <- rho (FP64 state) for saves at the start time
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

// states[inst][save_idx] <- old + theta (new - old) for a second save time inside one
// step: the FP64 start-of-step state is gone, its FP32 copy r32_old is still there.
__global__ __launch_bounds__(256) void k_save_interp(int64_t n, int64_t n_true, const float2 *__restrict__ r32_old,
                                                     const double2 *__restrict__ r64, float2 *__restrict__ states,
                                                     int64_t nsave, int64_t save_idx, double theta) {
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  int i, j;
  if (!diat::row_elem(n, i, j)) return;
  const int64_t e = (int64_t)i * n + j;
  if (i < n_true && j < n_true) {
    const double2 v = r64[inst * nn + e];
    const float2 o = r32_old[inst * nn + e];
    const double sr = o.x + theta * (v.x - o.x), si = o.y + theta * (v.y - o.y);
    states[((inst * nsave + save_idx) * n_true + i) * n_true + j] = make_float2((float)sr, (float)si);
  }
}

__global__ __launch_bounds__(256) void k_round(int64_t total, const double2 *__restrict__ r64, float2 *__restrict__ r32) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < total) r32[e] = make_float2((float)r64[e].x, (float)r64[e].y);
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
  float2 *r32[2] = {reinterpret_cast<float2 *>(p.rho32), reinterpret_cast<float2 *>(p.rho32) + p.ninst * nn};
  double2 *r64 = reinterpret_cast<double2 *>(p.rho64);
  float2 *D1 = reinterpret_cast<float2 *>(p.work);
  float2 *E = D1 + p.ninst * nn;
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
  const bool split = p.variant == DQH_SWEEP_SPLIT;
  const int halo = split ? 0 : diat::resolve_halo(p.variant, n, maxoff);
  if (halo < 0)
    return fail(DQH_BAD_ARG, "dqh_mesolve_rouchon1_dia: the tiled sweep needs every |offset| <= its halo");
  int cur = 0;
  hipLaunchKernelGGL(dia::k_round, dim3((unsigned)((p.ninst * nn + 255) / 256)), dim3(256), 0, st, p.ninst * nn, r64,
                     r32[0]);
  const int64_t *hs_step = reinterpret_cast<const int64_t *>(p.save_step);
  const double *hs_theta = reinterpret_cast<const double *>(p.save_theta);
  int64_t isave = 0;
  while (isave < p.nsave && hs_step[isave] < 0) {
    hipLaunchKernelGGL(dia::k_save, gel, dim3(256), 0, st, n, p.n_true, r64, states, p.nsave, isave);
    ++isave;
  }
  for (int64_t m = 0; m < p.nsteps; ++m) {
    const int set = (m == p.nsteps - 1) ? p.nsets - 1 : 0;
    const float2 *Vp = reinterpret_cast<const float2 *>(p.Vp[set]);
    const float2 *Wn = reinterpret_cast<const float2 *>(p.Wn[set]);
    if (panel == 32) {
      hipLaunchKernelGGL(dia::k_lsolve<32>, gsolve, dim3(dia::kThreads), 0, st, 0, n, Vp, Wn, r32[cur], D1, D1);
      hipLaunchKernelGGL(dia::k_lsolve<32>, gsolve, dim3(dia::kThreads), 0, st, 1, n, Vp, Wn, r32[cur], D1, E);
    } else {
      hipLaunchKernelGGL(dia::k_lsolve<64>, gsolve, dim3(dia::kThreads), 0, st, 0, n, Vp, Wn, r32[cur], D1, D1);
      hipLaunchKernelGGL(dia::k_lsolve<64>, gsolve, dim3(dia::kThreads), 0, st, 1, n, Vp, Wn, r32[cur], D1, E);
    }
    // saves inside this step: the kernel writes one interpolated state per launch
    int64_t sidx = -1;
    double th = 0.0;
    if (isave < p.nsave && hs_step[isave] == m) {
      sidx = isave;
      th = hs_theta[isave];
      ++isave;
    }
    if (split) {
      // D1 is consumed by the second solve: its buffer holds dY
      hipLaunchKernelGGL(dia::k_kraus_left, gel, dim3(256), 0, st, n, p.ops[set], r32[cur], E, D1);
      hipLaunchKernelGGL(dia::k_kraus_right, gel, dim3(256), 0, st, n, p.n_true, p.ops[set], r32[cur], E, D1, r64,
                         r32[1 - cur], states, p.nsave, sidx, th);
    } else if (halo == 8)
      hipLaunchKernelGGL(dia::k_kraus_tiled<8>, gtile, dim3(diat::kThreads), 0, st, n, p.n_true, p.ops[set], r32[cur],
                         E, r64, r32[1 - cur], states, p.nsave, sidx, th);
    else if (halo == 16)
      hipLaunchKernelGGL(dia::k_kraus_tiled<16>, gtile, dim3(diat::kThreads), 0, st, n, p.n_true, p.ops[set], r32[cur],
                         E, r64, r32[1 - cur], states, p.nsave, sidx, th);
    else if (halo == 32)
      hipLaunchKernelGGL(dia::k_kraus_tiled<32>, gtile, dim3(diat::kThreads), 0, st, n, p.n_true, p.ops[set], r32[cur],
                         E, r64, r32[1 - cur], states, p.nsave, sidx, th);
    else
      hipLaunchKernelGGL(dia::k_kraus, gel, dim3(256), 0, st, n, p.n_true, p.ops[set], r32[cur], E, r64,
                         r32[1 - cur], states, p.nsave, sidx, th);
    // further save times inside the same step, each at its own interpolation weight
    while (isave < p.nsave && hs_step[isave] == m) {
      hipLaunchKernelGGL(dia::k_save_interp, gel, dim3(256), 0, st, n, p.n_true, r32[cur], r64, states, p.nsave,
                         isave, hs_theta[isave]);
      ++isave;
    }
    cur = 1 - cur;
  }
  while (isave < p.nsave) {
    hipLaunchKernelGGL(dia::k_save, gel, dim3(256), 0, st, n, p.n_true, r64, states, p.nsave, isave);
    ++isave;
  }
  if (p.nE > 0)
    hipLaunchKernelGGL(k_expect_states, dim3((unsigned)p.nsave, (unsigned)p.ninst), dim3(256), 0, st, p.n_true, p.nsave,
                       p.nE, states, reinterpret_cast<const float2 *>(p.EopsT),
                       reinterpret_cast<float2 *>(p.expects));
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(DQH_HIP_ERROR, std::string("dqh_mesolve_rouchon1_dia: ") + hipGetErrorString(err));
  return DQH_OK;
}
