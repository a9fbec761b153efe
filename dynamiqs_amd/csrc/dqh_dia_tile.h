// LDS-tiled sweeps of DIA operators over a dense matrix (gfx950).
//
// For operators stored by diagonals (A[i][i + p] = diags[u][i + p] for the offset
// p = off[u], the scipy.sparse.dia_matrix convention of
// dynamiqs/qarrays/sparsedia_primitives.py:11-16) the products that the Lindblad
// RHS and the Rouchon1 Kraus map need are, element by element,
//   (A X)[i][j]       = sum_u A[i][i + p_u] X[i + p_u][j]
//   (X A^dag)[i][j]   = sum_u X[i][j + p_u] conj(A[j][j + p_u])
//   (A X A^dag)[i][j] = sum_u A[i][i + p_u] sum_v X[i + p_u][j + p_v] conj(A[j][j + p_v]).
// One workgroup owns a kTI x kTJ tile of the output. It stages X over the tile
// grown by HALO on every side in LDS (zero outside the matrix), so every
// stencil read hits LDS and X crosses HBM about once. Offsets must satisfy
// |p| <= HALO. The diagonal values are read from global memory (a few KB per
// operator, cached): the row-side value is the same for all lanes of a wave,
// the column-side one is coalesced.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "dqh.h"

namespace dqh {
namespace diat {

constexpr int kTI = 64;
constexpr int kTJ = 64;
constexpr int kThreads = 512;
constexpr int kRowsPerThread = kTI * kTJ / kThreads;  // 8

// Staged tile of a sweep whose offsets satisfy |p| <= HALO: (64 + 2 HALO)^2
// complex64 in LDS (HALO 8: 50 KiB, 3 workgroups per CU; 16: 72 KiB, 2 per CU;
// 32: 128 KiB, 1 per CU). Smaller halos re-read less of X from L2 and keep more
// waves in flight; the autotuner picks among the ones a problem admits.
template <int HALO>
struct Tile {
  static constexpr int XR = kTI + 2 * HALO;
  static constexpr int XC = kTJ + 2 * HALO;
  static constexpr int NX = XR * XC;
};

// Largest |offset| over the operators of a description.
inline int max_abs_offset(const dqh_dia_ops &ops) {
  int m = 0;
  for (int s = 0; s < ops.nops; ++s)
    for (int u = 0; u < ops.noff[s]; ++u) m = std::max(m, std::abs(ops.off[s][u]));
  return m;
}

// Sweep form of a DQH_SWEEP_* request: the LDS halo (8, 16, 32), 0 for the
// global form, -1 when a tiled form was asked for and the problem does not admit it.
inline int resolve_halo(int64_t variant, int64_t n, int maxoff) {
  const bool tileable = n % kTI == 0;
  auto fits = [&](int h) { return tileable && maxoff <= h; };
  switch (variant) {
    case DQH_SWEEP_GLOBAL:
      return 0;
    case DQH_SWEEP_TILED_H8:
      return fits(8) ? 8 : -1;
    case DQH_SWEEP_TILED_H16:
      return fits(16) ? 16 : -1;
    case DQH_SWEEP_TILED_H32:
      return fits(32) ? 32 : -1;
    case DQH_SWEEP_TILED:
      return fits(8) ? 8 : fits(16) ? 16 : fits(32) ? 32 : -1;
    default:  // DQH_SWEEP_AUTO
      return fits(8) ? 8 : fits(16) ? 16 : fits(32) ? 32 : 0;
  }
}

template <typename T>
__device__ __forceinline__ const T *cgp(uint64_t a) {
  return (const T *)((__attribute__((address_space(1))) const T *)a);
}

// Tile coordinates of this workgroup. Consecutive workgroup ids go round-robin
// over the 8 XCDs; the remap gives every XCD a contiguous run of tiles (whole
// tile rows), so the halo rows that neighbouring tiles share stay in one L2.
__device__ __forceinline__ void tile_of(int64_t n, int &i0, int &j0) {
  const int tpr = (int)(n / kTJ);
  const int ntiles = tpr * (int)(n / kTI);
  int b = blockIdx.x;
  if ((ntiles & 7) == 0) b = (b & 7) * (ntiles >> 3) + (b >> 3);
  i0 = (b / tpr) * kTI;
  j0 = (b % tpr) * kTJ;
}

// sx[r][c] = A[i0 - HALO + r][j0 - HALO + c] (+ B[...]), zero outside [0, n)^2.
template <int HALO>
__device__ __forceinline__ void load_halo(float2 *sx, const float2 *__restrict__ A, const float2 *__restrict__ B,
                                          int64_t n, int i0, int j0) {
  using T = Tile<HALO>;
  for (int e = threadIdx.x; e < T::NX; e += kThreads) {
    const int r = e / T::XC, c = e % T::XC;
    const int64_t gi = (int64_t)i0 - HALO + r, gj = (int64_t)j0 - HALO + c;
    float2 v = make_float2(0.f, 0.f);
    if (gi >= 0 && gi < n && gj >= 0 && gj < n) {
      const size_t o = (size_t)gi * n + gj;
      v = A[o];
      if (B) {
        const float2 w = B[o];
        v.x += w.x;
        v.y += w.y;
      }
    }
    sx[e] = v;
  }
}

// Elementwise launches over an n x n matrix use a row-blocked grid: grid.x =
// n * ceil(n / 256) workgroups of 256 threads, workgroup b covers columns
// [256 (b mod cpr), +256) of row b / cpr. The row is one scalar division per
// workgroup instead of a 64-bit division per element.
inline dim3 row_grid(int64_t n, int64_t ninst) { return dim3((unsigned)(n * ((n + 255) / 256)), (unsigned)ninst); }

__device__ __forceinline__ bool row_elem(int64_t n, int &i, int &j) {
  const unsigned cpr = (unsigned)((n + 255) >> 8);
  const unsigned b = blockIdx.x;
  const unsigned r = b / cpr;
  i = (int)r;
  j = (int)((b - r * cpr) * 256u + threadIdx.x);
  return j < n;
}

__device__ __forceinline__ int clampi(int64_t a, int64_t n) { return (int)(a < 0 ? 0 : (a >= n ? n - 1 : a)); }

// X read from the LDS halo tile around element (ii, jj) of the tile
template <int HALO>
struct LdsX {
  const float2 *c;
  __device__ __forceinline__ float2 operator()(int p, int q) const { return c[p * Tile<HALO>::XC + q]; }
  static __device__ __forceinline__ LdsX at(const float2 *sx, int ii, int jj) {
    return LdsX{sx + (ii + HALO) * Tile<HALO>::XC + (jj + HALO)};
  }
};

// X = A (+ B) read from global memory with bounds checks (any offsets)
struct GlobalX {
  const float2 *A, *B;
  int64_t n;
  int i, j;
  __device__ __forceinline__ float2 operator()(int p, int q) const {
    const int64_t a = (int64_t)i + p, b = (int64_t)j + q;
    if (a < 0 || a >= n || b < 0 || b >= n) return make_float2(0.f, 0.f);
    const size_t o = (size_t)a * n + b;
    float2 v = A[o];
    if (B) {
      const float2 w = B[o];
      v.x += w.x;
      v.y += w.y;
    }
    return v;
  }
};

// sum of (op0 X + X op0^dag), of (op0 X op0^dag) when sand0, and of
// (op_s X op_s^dag) for s >= 1, at element (i, j); X(p, q) = X[i + p][j + q].
// Out-of-range neighbours read as zero, so the diagonal index is only clamped to
// stay inside the array.
template <class XA>
__device__ __forceinline__ float2 stencil(const XA &X, const dqh_dia_ops &ops, bool sand0, int64_t n, int i, int j,
                                          const float2 *D0 = nullptr) {
  float ar = 0.f, ai = 0.f;
  {
    // D0: the diagonals of operator 0 at the current time (time-dependent generators)
    const float2 *D = D0 ? D0 : cgp<float2>(ops.diags[0]);
    const int no = ops.noff[0];
    for (int u = 0; u < no; ++u) {
      const int p = ops.off[0][u];
      const float2 m = D[(size_t)u * n + clampi((int64_t)i + p, n)];
      const float2 x = X(p, 0);
      ar += m.x * x.x - m.y * x.y;
      ai += m.x * x.y + m.y * x.x;
      const float2 mc = D[(size_t)u * n + clampi((int64_t)j + p, n)];
      const float2 y = X(0, p);
      ar += y.x * mc.x + y.y * mc.y;
      ai += y.y * mc.x - y.x * mc.y;
    }
  }
  for (int s = sand0 ? 0 : 1; s < ops.nops; ++s) {
    const float2 *A = cgp<float2>(ops.diags[s]);
    const int no = ops.noff[s];
    for (int u = 0; u < no; ++u) {
      const int p = ops.off[s][u];
      const float2 ma = A[(size_t)u * n + clampi((int64_t)i + p, n)];
      float zr = 0.f, zi = 0.f;
      for (int v = 0; v < no; ++v) {
        const int q = ops.off[s][v];
        const float2 mb = A[(size_t)v * n + clampi((int64_t)j + q, n)];
        const float2 x = X(p, q);
        zr += x.x * mb.x + x.y * mb.y;
        zi += x.y * mb.x - x.x * mb.y;
      }
      ar += ma.x * zr - ma.y * zi;
      ai += ma.x * zi + ma.y * zr;
    }
  }
  return make_float2(ar, ai);
}

}  // namespace diat
}  // namespace dqh
