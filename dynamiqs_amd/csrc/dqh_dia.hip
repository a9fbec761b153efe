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
//       element, and writes hi, lo in place (48 bytes of HBM per element).
// When S decouples into small index sets, K1 and K2 are instead the two memory-bound
// passes of the block form (k_blk_left / k_blk_right, below), with the same D1, E
// and X. K3 gives each thread one element of up to four instances, which share
// the operators and so the coefficient loads.
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

// ---------------------------------------------------------------- block form
// When S couples indices only inside small sets, T (lower Cholesky factor of S) is
// block-diagonal on them and the normalisation is two passes over the matrix:
//   K1  D1[r][c] = sum_{k in set(r)} (T_b^-dag - I)[loc r][loc k] hi[k][c]
//       (rows mixed within a set; one workgroup per set and 256 columns, each
//       thread one column: the d values of its column in the set's rows are loaded
//       into registers, coalesced across the workgroup, then the d outputs written),
//   K2  Y = hi + D1, E[r][c] = D1[r][c] + sum_{k in set(c)} Y[r][k] (T_b^-1 - I)[loc k][loc c],
//       X = Y + (that sum) written over D1 (columns mixed within a set; one
//       workgroup per strip of W rows, Y of the strip in LDS, each thread one
//       column and the W rows),
// so that hi + E = (hi + D1) T^-1 = T^-dag rho T^-1, as in the banded form. Both
// passes read and write each element once (memory bound); a set of d indices costs
// d^2 complex multiply-adds per row or column instead of the 64 x 128 of a block
// row of the banded recurrence.
struct Blocks {
  const int32_t *of, *loc, *start, *idx;
  const int64_t *mat;
};

// K1 for sets of at most D indices. grid: (ceil(n / 256), nblk, ninst).
template <int D>
__global__ __launch_bounds__(256) void k_blk_left(int64_t n, Blocks bk, const float2 *__restrict__ Lm,
                                                  const float2 *__restrict__ hi, float2 *__restrict__ D1) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.z;
  const int b = blockIdx.y;
  const int s0 = bk.start[b], d = bk.start[b + 1] - s0;
  const int32_t *rows = bk.idx + s0;
  const float2 *H = hi + inst * nn + c;
  float2 *O = D1 + inst * nn + c;
  const float2 *L = Lm + bk.mat[b];  // d x d, row l zero left of l
  float2 x[D];
#pragma unroll
  for (int k = 0; k < D; ++k) x[k] = k < d ? H[(size_t)rows[k] * n] : make_float2(0.f, 0.f);
  for (int l = 0; l < d; ++l) {
    const float2 *Lr = L + (size_t)l * d;
    float ar = 0.f, ai = 0.f;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      if (k >= l && k < d) {
        const float2 m = Lr[k];
        ar += m.x * x[k].x - m.y * x[k].y;
        ai += m.x * x[k].y + m.y * x[k].x;
      }
    }
    O[(size_t)rows[l] * n] = make_float2(ar, ai);
  }
}

// K2 launch shape: threads per workgroup and the widest strip (rows of Y in LDS,
// W * n * 8 bytes), which together set how many workgroups share a CU.
#ifndef DQH_BLK_THREADS
#define DQH_BLK_THREADS 512
#endif
#ifndef DQH_BLK_WMAX
#define DQH_BLK_WMAX 8
#endif
constexpr int kBlkThreads = DQH_BLK_THREADS;
// one-pass form: on or off, and the rows of Y it publishes per barrier (at the
// two-mode n = 1024 batch of 16: 8 rows 296 us per step, 4 rows 391 us, the two
// passes 320 us with 512-thread strips and 361 us with 256)
#ifndef DQH_NORM_FUSE
#define DQH_NORM_FUSE 1
#endif
#ifndef DQH_NORM_ROWS
#define DQH_NORM_ROWS 8
#endif
constexpr bool kNormFuse = DQH_NORM_FUSE != 0;
constexpr int kNormRows = DQH_NORM_ROWS;
#ifndef DQH_NORM_BATCH
#define DQH_NORM_BATCH 4
#endif
constexpr int kNormBatch = DQH_NORM_BATCH;  // column coefficients loaded per batch
static_assert(kNormRows == 4 || kNormRows == 8, "DQH_NORM_ROWS must divide the set bounds 8, 16 and 32");
constexpr int kBlkWMax = DQH_BLK_WMAX;
static_assert(kBlkThreads == 256 || kBlkThreads == 512 || kBlkThreads == 1024, "DQH_BLK_THREADS");
static_assert(kBlkWMax == 2 || kBlkWMax == 4 || kBlkWMax == 8, "DQH_BLK_WMAX");

template <int W>
__global__ __launch_bounds__(kBlkThreads) void k_blk_right(int64_t n, Blocks bk, const float2 *__restrict__ Rm,
                                                   const float2 *__restrict__ hi, float2 *D1X,
                                                   float2 *__restrict__ E) {
  extern __shared__ float2 ystrip[];  // W x n: Y[r0 + r][c] at ystrip[r * n + c]
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  const int r0 = blockIdx.x * W;
  const float2 *H = hi + inst * nn;
  float2 *DX = D1X + inst * nn;
  float2 *Eo = E + inst * nn;
  for (int e = threadIdx.x; e < W * n; e += blockDim.x) {
    const size_t g = (size_t)r0 * n + e;
    const float2 h = H[g], d1 = DX[g];
    ystrip[e] = make_float2(h.x + d1.x, h.y + d1.y);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    const int b = bk.of[c], l = bk.loc[c];
    const int s0 = bk.start[b], d = bk.start[b + 1] - s0;
    const float2 *Rc = Rm + bk.mat[b] + (int64_t)l * d;  // column l of T_b^-1 - I: zero above l
    float ar[W], ai[W];
#pragma unroll
    for (int r = 0; r < W; ++r) ar[r] = ai[r] = 0.f;
    for (int k = l; k < d; ++k) {
      const float2 m = Rc[k];
      const int src = bk.idx[s0 + k];
#pragma unroll
      for (int r = 0; r < W; ++r) {
        const float2 y = ystrip[(int64_t)r * n + src];
        ar[r] += y.x * m.x - y.y * m.y;
        ai[r] += y.x * m.y + y.y * m.x;
      }
    }
    // each (row, column) of the strip is this thread's alone: D1 is read here once
    // more (for E) and then overwritten with X
#pragma unroll
    for (int r = 0; r < W; ++r) {
      const size_t g = (size_t)(r0 + r) * n + c;
      const float2 d1 = DX[g];
      const float2 y = ystrip[(int64_t)r * n + c];
      Eo[g] = make_float2(d1.x + ar[r], d1.y + ai[r]);
      DX[g] = make_float2(y.x + ar[r], y.y + ai[r]);
    }
  }
}

// K1 and K2 in one pass, for sets of at most D indices and n = blockDim.x <= 1024
// columns: one workgroup per (set, instance), each thread one column. The d values
// of the thread's column in the set's rows are loaded once into registers; the
// set's rows are then done G at a time: each thread forms D1 and Y = hi + D1 of the
// G rows at its column, the workgroup publishes those G rows of Y in LDS, and each
// thread mixes the columns of its own column's set for the G rows at once (one
// load of each coefficient and index for G rows) and writes E and X. HBM: hi read
// once, E and X written once (24 bytes per element against 48 for the two passes).
template <int D, int G>
__global__ __launch_bounds__(1024) void k_blk_norm(int64_t n, Blocks bk, const float2 *__restrict__ Lm,
                                                   const float2 *__restrict__ Rm, const float2 *__restrict__ hi,
                                                   float2 *__restrict__ Xo, float2 *__restrict__ Eo) {
  extern __shared__ float2 ys[];  // G x n: Y of the current rows
  const int c = threadIdx.x;      // blockDim.x == n
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  const int b = blockIdx.x;
  const int s0 = bk.start[b], d = bk.start[b + 1] - s0;
  const int32_t *rows = bk.idx + s0;
  const float2 *L = Lm + bk.mat[b];  // d x d, row l zero left of l
  const float2 *H = hi + inst * nn + c;
  float2 *X = Xo + inst * nn + c;
  float2 *E = Eo + inst * nn + c;
  // the set of this thread's column: its column of T_cb^-1 - I (zero above cl) and
  // the set's indices
  const int cb = bk.of[c], cl = bk.loc[c];
  const int cs0 = bk.start[cb], cd = bk.start[cb + 1] - cs0;
  const float2 *Rc = Rm + bk.mat[cb] + (int64_t)cl * cd;
  const int32_t *cidx = bk.idx + cs0;
  float2 x[D];
#pragma unroll
  for (int k = 0; k < D; ++k) x[k] = k < d ? H[(size_t)rows[k] * n] : make_float2(0.f, 0.f);
#pragma unroll
  for (int g0 = 0; g0 < D; g0 += G) {
    if (g0 < d) {
      float2 dl[G], yo[G];
#pragma unroll
      for (int r = 0; r < G; ++r) {
        const int l = g0 + r;
        float dr = 0.f, di = 0.f;
        if (l < d) {
          const float2 *Lr = L + (size_t)l * d;
#pragma unroll
          for (int k = l; k < D; ++k) {
            if (k < d) {
              const float2 m = Lr[k];
              dr += m.x * x[k].x - m.y * x[k].y;
              di += m.x * x[k].y + m.y * x[k].x;
            }
          }
        }
        dl[r] = make_float2(dr, di);
        yo[r] = make_float2(x[l].x + dr, x[l].y + di);
        ys[r * n + c] = yo[r];
      }
      __syncthreads();
      float ar[G], ai[G];
#pragma unroll
      for (int r = 0; r < G; ++r) ar[r] = ai[r] = 0.f;
      // the column's coefficients and indices are loaded kNormBatch at a time, all
      // issued before the first use (the loop bound differs between the threads)
      for (int k0 = cl; k0 < cd; k0 += kNormBatch) {
        float2 mk[kNormBatch];
        int sk[kNormBatch];
#pragma unroll
        for (int u = 0; u < kNormBatch; ++u) {
          const bool in = k0 + u < cd;
          mk[u] = in ? Rc[k0 + u] : make_float2(0.f, 0.f);
          sk[u] = in ? cidx[k0 + u] : c;
        }
#pragma unroll
        for (int u = 0; u < kNormBatch; ++u) {
#pragma unroll
          for (int r = 0; r < G; ++r) {
            const float2 y = ys[r * n + sk[u]];
            ar[r] += y.x * mk[u].x - y.y * mk[u].y;
            ai[r] += y.x * mk[u].y + y.y * mk[u].x;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < G; ++r) {
        if (g0 + r < d) {
          const size_t g = (size_t)rows[g0 + r] * n;
          E[g] = make_float2(dl[r].x + ar[r], dl[r].y + ai[r]);
          X[g] = make_float2(yo[r].x + ar[r], yo[r].y + ai[r]);
        }
      }
      __syncthreads();  // the next rows overwrite ys
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

// The Kraus increment dM X + X dM^dag + dM X dM^dag + sum_{s>=1} A_s X A_s^dag at
// (i, j) (op 0 = dM, ops 1.. = A_s), X(p, q) = X[i + p][j + q]. The column-side
// coefficients conj(A[j][j + q]) of each operator are loaded once into registers
// (at most DQH_DIA_MAXOFF per operator) and reused by every row offset; the
// row-side ones are the same for the whole row.
template <class XA>
__device__ __forceinline__ float2 kraus_stencil(const XA &X, const dqh_dia_ops &ops, int64_t n, int i, int j) {
  float ar = 0.f, ai = 0.f;
  for (int s = 0; s < ops.nops; ++s) {
    const float2 *A = diat::cgp<float2>(ops.diags[s]);
    const int no = ops.noff[s];
    float2 mb[DQH_DIA_MAXOFF];
#pragma unroll
    for (int v = 0; v < DQH_DIA_MAXOFF; ++v)
      mb[v] = v < no ? A[(size_t)v * n + diat::clampi((int64_t)j + ops.off[s][v], n)] : make_float2(0.f, 0.f);
    for (int u = 0; u < no; ++u) {
      const int p = ops.off[s][u];
      const float2 ma = A[(size_t)u * n + diat::clampi((int64_t)i + p, n)];
      // z = (X A^dag)[i + p][j]; for op 0 also (X)[i + p][j] for the one-sided dM X
      float zr = 0.f, zi = 0.f;
#pragma unroll
      for (int v = 0; v < DQH_DIA_MAXOFF; ++v) {
        if (v < no) {
          const float2 x = X(p, ops.off[s][v]);
          zr += x.x * mb[v].x + x.y * mb[v].y;
          zi += x.y * mb[v].x - x.x * mb[v].y;
        }
      }
      if (s == 0) {
        const float2 x0 = X(p, 0);
        zr += x0.x;
        zi += x0.y;
      }
      ar += ma.x * zr - ma.y * zi;
      ai += ma.x * zi + ma.y * zr;
    }
    if (s == 0) {  // X dM^dag
#pragma unroll
      for (int v = 0; v < DQH_DIA_MAXOFF; ++v) {
        if (v < no) {
          const float2 x = X(0, ops.off[s][v]);
          ar += x.x * mb[v].x + x.y * mb[v].y;
          ai += x.y * mb[v].x - x.x * mb[v].y;
        }
      }
    }
  }
  return make_float2(ar, ai);
}

// row_elem (diat) with an XCD-aware workgroup order: consecutive workgroup ids go
// round-robin over the 8 XCDs, so id b takes item (b % 8) * (nb / 8) + b / 8 and
// each XCD sweeps a contiguous band of rows; the neighbour rows the stencil reads
// (a few dozen rows away) then stay in that XCD's L2.
__device__ __forceinline__ bool row_elem_xcd(int64_t n, int &i, int &j) {
  const unsigned cpr = (unsigned)((n + 255) >> 8);
  const unsigned nb = gridDim.x;
  unsigned b = blockIdx.x;
  if ((nb & 7u) == 0) b = (b & 7u) * (nb >> 3) + (b >> 3);
  const unsigned r = b / cpr;
  i = (int)r;
  j = (int)((b - r * cpr) * 256u + threadIdx.x);
  return j < n;
}

// Largest number of instances one thread of the global sweep updates (4 or 8).
#ifndef DQH_KRAUS_IB
#define DQH_KRAUS_IB 4
#endif
constexpr int kKrausIB = DQH_KRAUS_IB;
static_assert(kKrausIB == 4 || kKrausIB == 8, "DQH_KRAUS_IB must be 4 or 8");

// X at one element (i, j) of IB instances (instance stride nn; nib of them valid,
// the rest read the last valid instance and are discarded by the caller). A
// neighbour outside the matrix reads as zero through a clamped address and a zero
// factor, so the loads of all IB instances issue back to back without branches.
template <int IB>
struct XBatch {
  const float2 *base;
  size_t nn;
  int64_t n;
  int i, j, nib;
  __device__ __forceinline__ void load(int p, int q, float2 (&v)[IB]) const {
    const int64_t a = (int64_t)i + p, b = (int64_t)j + q;
    const bool ok = a >= 0 && a < n && b >= 0 && b < n;
    const size_t o = ok ? (size_t)a * n + b : 0;
    const float m = ok ? 1.f : 0.f;
#pragma unroll
    for (int k = 0; k < IB; ++k) {
      const float2 x = base[(size_t)(k < nib ? k : nib - 1) * nn + o];
      v[k] = make_float2(x.x * m, x.y * m);
    }
  }
};

// kraus_stencil for IB instances that share the operators: the coefficient loads
// are made once for all of them and every X load is IB independent loads, so a
// thread keeps IB times more memory requests in flight per round trip.
template <int IB>
__device__ __forceinline__ void kraus_stencil_batch(const XBatch<IB> &X, const dqh_dia_ops &ops, int64_t n, int i,
                                                    int j, float (&ar)[IB], float (&ai)[IB]) {
#pragma unroll
  for (int k = 0; k < IB; ++k) ar[k] = ai[k] = 0.f;
  for (int s = 0; s < ops.nops; ++s) {
    const float2 *A = diat::cgp<float2>(ops.diags[s]);
    const int no = ops.noff[s];
    float2 mb[DQH_DIA_MAXOFF];
#pragma unroll
    for (int v = 0; v < DQH_DIA_MAXOFF; ++v)
      mb[v] = v < no ? A[(size_t)v * n + diat::clampi((int64_t)j + ops.off[s][v], n)] : make_float2(0.f, 0.f);
    for (int u = 0; u < no; ++u) {
      const int p = ops.off[s][u];
      const float2 ma = A[(size_t)u * n + diat::clampi((int64_t)i + p, n)];
      float zr[IB], zi[IB];
      float2 x[IB];
      if (s == 0) {  // the one-sided dM X term
        X.load(p, 0, x);
#pragma unroll
        for (int k = 0; k < IB; ++k) zr[k] = x[k].x, zi[k] = x[k].y;
      } else {
#pragma unroll
        for (int k = 0; k < IB; ++k) zr[k] = zi[k] = 0.f;
      }
#pragma unroll
      for (int v = 0; v < DQH_DIA_MAXOFF; ++v) {
        if (v < no) {
          X.load(p, ops.off[s][v], x);
#pragma unroll
          for (int k = 0; k < IB; ++k) {
            zr[k] += x[k].x * mb[v].x + x[k].y * mb[v].y;
            zi[k] += x[k].y * mb[v].x - x[k].x * mb[v].y;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < IB; ++k) {
        ar[k] += ma.x * zr[k] - ma.y * zi[k];
        ai[k] += ma.x * zi[k] + ma.y * zr[k];
      }
    }
    if (s == 0) {  // X dM^dag
#pragma unroll
      for (int v = 0; v < DQH_DIA_MAXOFF; ++v) {
        if (v < no) {
          float2 x[IB];
          X.load(0, ops.off[s][v], x);
#pragma unroll
          for (int k = 0; k < IB; ++k) {
            ar[k] += x[k].x * mb[v].x + x[k].y * mb[v].y;
            ai[k] += x[k].y * mb[v].x - x[k].x * mb[v].y;
          }
        }
      }
    }
  }
}

// Fused Kraus sweep: increment E + dM X + X dM^dag + dM X dM^dag + sum A X A^dag at
// every element, X (the FP32 normalised state) read at the neighbours the
// diagonals name through the caches (any offsets, any n), then the FP64 update of
// the hi / lo planes and the saves. One thread per element of IB instances
// (blockIdx.y * IB onwards), rows coalesced.
template <int IB>
__global__ __launch_bounds__(256) void k_kraus(int64_t n, int64_t n_true, int64_t ninst, dqh_dia_ops ops,
                                               const float2 *__restrict__ X, const float2 *__restrict__ E,
                                               float2 *__restrict__ hi, float2 *__restrict__ lo,
                                               float2 *__restrict__ states, int64_t nsave, Saves sv) {
  const size_t nn = (size_t)n * n;
  const int64_t inst0 = (int64_t)blockIdx.y * IB;
  const int nib = (int)(ninst - inst0 < IB ? ninst - inst0 : IB);
  int i, j;
  if (!row_elem_xcd(n, i, j)) return;
  const XBatch<IB> Xb{X + inst0 * nn, nn, n, i, j, nib};
  float ar[IB], ai[IB];
  kraus_stencil_batch<IB>(Xb, ops, n, i, j, ar, ai);
#pragma unroll
  for (int k = 0; k < IB; ++k) {
    if (k < nib) {
      const int64_t inst = inst0 + k;
      const size_t e = inst * nn + (size_t)i * n + j;
      const float2 eij = E[e];
      update_elem(e, i, j, n_true, (double)eij.x + ar[k], (double)eij.y + ai[k], hi, lo,
                  states + inst * nsave * n_true * n_true, nsave, sv);
    }
  }
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
    const float2 acc = kraus_stencil(diat::LdsX<HALO>::at(sx, ii, jj), ops, n, i, j);
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
  // block form of the normalisation: strip width so that a strip of n elements fits
  // 64 KiB of LDS
  const bool blocks = p.nblk > 0;
  const int bw = std::min(dia::kBlkWMax, n <= 1024 ? 8 : (n <= 2048 ? 4 : 2));
  if (blocks && (n > 4096 || p.nblk > n || p.blk_maxd < 1 || p.blk_maxd > DQH_DIA_MAXBLK || p.ninst > 65535 || !p.blk_of || !p.blk_loc || !p.blk_start || !p.blk_idx || !p.blk_mat || !p.Lm[0] ||
                 !p.Rm[0] || (p.nsets > 1 && (!p.Lm[1] || !p.Rm[1]))))
    return fail(DQH_BAD_ARG, "dqh_mesolve_rouchon1_dia: block normalisation needs n <= 4096, 1 <= blk_maxd <= DQH_DIA_MAXBLK, ninst <= 65535 and every block array");
  dia::Blocks bk{reinterpret_cast<const int32_t *>(p.blk_of), reinterpret_cast<const int32_t *>(p.blk_loc),
                 reinterpret_cast<const int32_t *>(p.blk_start), reinterpret_cast<const int32_t *>(p.blk_idx),
                 reinterpret_cast<const int64_t *>(p.blk_mat)};
  const dim3 gblk((unsigned)(n / bw), (unsigned)p.ninst);
  const dim3 gleft((unsigned)((n + 255) / 256), (unsigned)std::max<int64_t>(p.nblk, 1), (unsigned)p.ninst);
  // one pass (k_blk_norm) when a workgroup holds every column and the sets are
  // small enough for their column values to stay in registers
  const bool fused = blocks && dia::kNormFuse && n <= 1024 && p.blk_maxd <= 32;
  const dim3 gnorm((unsigned)std::max<int64_t>(p.nblk, 1), (unsigned)p.ninst);
  const size_t norm_lds = (size_t)dia::kNormRows * n * sizeof(float2);
  const size_t blk_lds = (size_t)n * bw * sizeof(float2);
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
  // global sweep: IB instances per thread (up to dia::kKrausIB)
  const int kib = p.ninst >= dia::kKrausIB ? dia::kKrausIB : (p.ninst >= 4 ? 4 : (p.ninst >= 2 ? 2 : 1));
  const dim3 gkr(gel.x, (unsigned)((p.ninst + kib - 1) / kib));
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
    else if (kib == 8)
      hipLaunchKernelGGL(dia::k_kraus<8>, gkr, dim3(256), 0, st, n, p.n_true, p.ninst, ops, X, E, hi, lo, states,
                         p.nsave, sv);
    else if (kib == 4)
      hipLaunchKernelGGL(dia::k_kraus<4>, gkr, dim3(256), 0, st, n, p.n_true, p.ninst, ops, X, E, hi, lo, states,
                         p.nsave, sv);
    else if (kib == 2)
      hipLaunchKernelGGL(dia::k_kraus<2>, gkr, dim3(256), 0, st, n, p.n_true, p.ninst, ops, X, E, hi, lo, states,
                         p.nsave, sv);
    else
      hipLaunchKernelGGL(dia::k_kraus<1>, gkr, dim3(256), 0, st, n, p.n_true, p.ninst, ops, X, E, hi, lo, states,
                         p.nsave, sv);
  };
  for (int64_t m = 0; m < p.nsteps; ++m) {
    const int set = (m == p.nsteps - 1) ? p.nsets - 1 : 0;
    const float2 *Vp = reinterpret_cast<const float2 *>(p.Vp[set]);
    const float2 *Wn = reinterpret_cast<const float2 *>(p.Wn[set]);
    const int64_t tk = m - t_first;
    const bool timed = tk >= 0 && tk < t_count;
    if (timed) hipEventRecord(ev[3 * tk], st);
    // D1 = T^-dag hi - hi into the X buffer, then E and X = hi + E (over D1)
    if (blocks) {
      const float2 *Lm = reinterpret_cast<const float2 *>(p.Lm[set]);
      const float2 *Rm = reinterpret_cast<const float2 *>(p.Rm[set]);
      if (fused) {
        if (p.blk_maxd <= 8)
          hipLaunchKernelGGL((dia::k_blk_norm<8, dia::kNormRows>), gnorm, dim3((unsigned)n), norm_lds, st, n, bk, Lm,
                             Rm, hi, X, E);
        else if (p.blk_maxd <= 16)
          hipLaunchKernelGGL((dia::k_blk_norm<16, dia::kNormRows>), gnorm, dim3((unsigned)n), norm_lds, st, n, bk, Lm,
                             Rm, hi, X, E);
        else
          hipLaunchKernelGGL((dia::k_blk_norm<32, dia::kNormRows>), gnorm, dim3((unsigned)n), norm_lds, st, n, bk, Lm,
                             Rm, hi, X, E);
      } else if (p.blk_maxd <= 8)
        hipLaunchKernelGGL(dia::k_blk_left<8>, gleft, dim3(256), 0, st, n, bk, Lm, hi, X);
      else if (p.blk_maxd <= 16)
        hipLaunchKernelGGL(dia::k_blk_left<16>, gleft, dim3(256), 0, st, n, bk, Lm, hi, X);
      else if (p.blk_maxd <= 32)
        hipLaunchKernelGGL(dia::k_blk_left<32>, gleft, dim3(256), 0, st, n, bk, Lm, hi, X);
      else
        hipLaunchKernelGGL(dia::k_blk_left<64>, gleft, dim3(256), 0, st, n, bk, Lm, hi, X);
      if (fused)
        ;
      else if (bw == 8)
        hipLaunchKernelGGL(dia::k_blk_right<8>, gblk, dim3(dia::kBlkThreads), blk_lds, st, n, bk, Rm, hi, X, E);
      else if (bw == 4)
        hipLaunchKernelGGL(dia::k_blk_right<4>, gblk, dim3(dia::kBlkThreads), blk_lds, st, n, bk, Rm, hi, X, E);
      else
        hipLaunchKernelGGL(dia::k_blk_right<2>, gblk, dim3(dia::kBlkThreads), blk_lds, st, n, bk, Rm, hi, X, E);
    } else if (panel == 32) {
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
