// Register/LDS-resident Tsit5 for padded dimensions np <= 64 (gfx950).
//
// One 512-thread workgroup integrates one instance. The state, the stage inputs
// and every stage k_j are Hermitian (the vector field is tmp + tmp^dag, which is
// Hermitian in floating point too), so a thread owns upper-triangle elements only:
// element u = t + 512 q of the row-major upper triangle (diagonal included). The
// seven FP32 stages of the owned elements live in registers; the stage input
// rho_s (both triangles), the products X_k = D_k rho_s and the pre-symmetrised tmp
// live in LDS (padded row stride np + 2 keeps the MFMA fragment reads
// conflict-free). The FP64 state y is kept packed (upper triangle, in owner order)
// in two global buffers that alternate on accepted steps; the operators are
// L2-resident and shared by the batch.
#pragma once

namespace dqh {

template <int NP, int NW>
struct RCfg {
  static constexpr int NT = NP / 16;
  static constexpr int NN = NP * NP;
  static constexpr int LDP = NP + 2;
  static constexpr int MAT = NP * LDP;  // float2 per LDS matrix
  static constexpr int NTH = NW * 64;  // threads per workgroup
  static constexpr int NE = (NN + NTH - 1) / NTH;
  // output tiles sharing one A fragment (one with 16 waves, so that phase 2 still
  // has a tile strip per wave)
  static constexpr int TJ = (NT % 2 == 0 && NW <= 8) ? 2 : 1;
  static constexpr int NJC = NT / TJ;
  static constexpr int ITEMS2 = NT * NJC;
  static constexpr int MAXI2 = (ITEMS2 + NW - 1) / NW;
  static constexpr int NU = NP * (NP + 1) / 2;  // upper triangle with the diagonal
  static constexpr int NEU = (NU + NTH - 1) / NTH;
};

// Owned upper-triangle elements: ij[q] = i | j << 8 (i <= j) for u = t + 512 q < NU,
// -1 past the end.
template <int NP, int NW>
__device__ __forceinline__ void upper_map(int (&ij)[(RCfg<NP, NW>::NEU)]) {
  using R = RCfg<NP, NW>;
#pragma unroll
  for (int q = 0; q < R::NEU; ++q) {
    const int u = threadIdx.x + R::NTH * q;
    if (u >= R::NU) {
      ij[q] = -1;
    } else {
      int i = 0, rem = u;
      while (rem >= NP - i) {
        rem -= NP - i;
        ++i;
      }
      ij[q] = i | ((i + rem) << 8);
    }
  }
}
__device__ __forceinline__ int ij_row(int x) { return x & 255; }
__device__ __forceinline__ int ij_col(int x) { return x >> 8; }
// weight of an owned element in a sum over the full matrix
__device__ __forceinline__ double ij_weight(int x) { return ij_row(x) == ij_col(x) ? 1.0 : 2.0; }

// (1 + kResMaxL) matrices of 64 x 66 complex64: ys, then X_1..X_3 (X_1 reused for tmp)
constexpr int kResLdsFloat2 = (1 + kResMaxL) * 64 * 66;

// acc[t] += A[row0:row0+16, 0:K] @ B[0:K, col0+16t : col0+16t+16], t < TJ; the A
// fragment of each 16-wide k block is loaded once and reused for the TJ tiles.
template <int TJ>
__device__ __forceinline__ void strip_mac(Acc (&acc)[TJ], const float2 *__restrict__ A, int lda,
                                          const float2 *__restrict__ B, int ldb, int row0, int col0, int K,
                                          int lane) {
  const int r = lane & 15, g = lane >> 4;
  const float2 *Ap = A + (size_t)(row0 + r) * lda + 4 * g;
  const float2 *Bp = B + (size_t)(4 * g) * ldb + col0 + r;
  for (int k0 = 0; k0 < K; k0 += 16) {
    const float4 a01 = *reinterpret_cast<const float4 *>(Ap + k0);
    const float4 a23 = *reinterpret_cast<const float4 *>(Ap + k0 + 2);
    const float2 *bp = Bp + (size_t)k0 * ldb;
#pragma unroll
    for (int t = 0; t < TJ; ++t) {
      const float2 b0 = bp[16 * t];
      const float2 b1 = bp[ldb + 16 * t];
      const float2 b2 = bp[2 * ldb + 16 * t];
      const float2 b3 = bp[3 * ldb + 16 * t];
      cmfma(acc[t], a01.x, a01.y, b0);
      cmfma(acc[t], a01.z, a01.w, b1);
      cmfma(acc[t], a23.x, a23.y, b2);
      cmfma(acc[t], a23.z, a23.w, b3);
    }
  }
}

// acc[t] += A[row0:row0+16, 0:K] @ M[col0+16t : col0+16t+16, 0:K]^dag, t < TJ. Both
// operands are read along rows: lane (r, g) loads A[row0+r][k0+4g .. k0+4g+3] and
// M[col0+16t+r][k0+4g .. k0+4g+3] (32 contiguous bytes each), so every fragment is
// two 16-byte loads. For a Hermitian M this is A @ M.
template <int TJ, int K>
__device__ __forceinline__ void strip_mac_h(Acc (&acc)[TJ], const float2 *__restrict__ A, int lda, const float2 *M,
                                            int ldm, int row0, int col0, int lane) {
  const int r = lane & 15, g = lane >> 4;
  const float2 *Ap = A + (size_t)(row0 + r) * lda + 4 * g;
  const float2 *Mp = M + (size_t)(col0 + r) * ldm + 4 * g;
  // the whole A strip (global memory) is requested before the first product, so
  // one memory latency is exposed per strip rather than one per k block
  float4 af[K / 16][2];
#pragma unroll
  for (int b = 0; b < K / 16; ++b) {
    af[b][0] = *reinterpret_cast<const float4 *>(Ap + 16 * b);
    af[b][1] = *reinterpret_cast<const float4 *>(Ap + 16 * b + 2);
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int b = 0; b < K / 16; ++b) {
    const int k0 = 16 * b;
    const float4 a01 = af[b][0];
    const float4 a23 = af[b][1];
#pragma unroll
    for (int t = 0; t < TJ; ++t) {
      const float2 *mp = Mp + (size_t)(16 * t) * ldm + k0;
      const float4 m01 = *reinterpret_cast<const float4 *>(mp);
      const float4 m23 = *reinterpret_cast<const float4 *>(mp + 2);
      cmfma_cb(acc[t], a01.x, a01.y, m01.x, m01.y);
      cmfma_cb(acc[t], a01.z, a01.w, m01.z, m01.w);
      cmfma_cb(acc[t], a23.x, a23.y, m23.x, m23.y);
      cmfma_cb(acc[t], a23.z, a23.w, m23.z, m23.w);
    }
  }
}

template <int TJ>
__device__ __forceinline__ void strip_store(const Acc (&acc)[TJ], float2 *C, int ldc, int row0, int col0, int lane) {
#pragma unroll
  for (int t = 0; t < TJ; ++t) tile_store(acc[t], C, ldc, row0, col0 + 16 * t, lane);
}

// tmp = [Hnh ys] + sum_k (D_k ys) C_k into lds[MAT .. 2 MAT), published by a final
// barrier. Precondition: ys written to lds[0 .. MAT) by every thread (this routine
// starts with the barrier that publishes it).
template <int NP, int NW>
__device__ __forceinline__ void res_rhs_tmp(const float2 *__restrict__ Hnh, const float2 *__restrict__ D,
                                            const float2 *__restrict__ C, int nL, bool hasH, float2 *lds) {
  using R = RCfg<NP, NW>;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float2 *ys = lds;
  float2 *X = lds + R::MAT;
  __syncthreads();
  // phase 1: X_k = D_k ys (A from L2, B from LDS)
  const int items1 = nL * R::NT * R::NJC;
  for (int it = wave; it < items1; it += NW) {
    const int kk = it / (R::NT * R::NJC), rem = it % (R::NT * R::NJC);
    const int ti = rem / R::NJC, jc = rem % R::NJC;
    Acc acc[R::TJ];
#pragma unroll
    for (int t = 0; t < R::TJ; ++t) acc_zero(acc[t]);
    strip_mac<R::TJ>(acc, D + (size_t)kk * R::NN, NP, ys, R::LDP, ti * 16, jc * R::TJ * 16, NP, lane);
    strip_store<R::TJ>(acc, X + kk * R::MAT, R::LDP, ti * 16, jc * R::TJ * 16, lane);
  }
  __syncthreads();
  // phase 2: tmp = Hnh ys + sum_k X_k C_k, kept in registers across the barrier
  Acc acc2[R::MAXI2][R::TJ];
#pragma unroll
  for (int m = 0; m < R::MAXI2; ++m) {
    const int it = wave + NW * m;
#pragma unroll
    for (int t = 0; t < R::TJ; ++t) acc_zero(acc2[m][t]);
    if (it < R::ITEMS2) {
      const int ti = it / R::NJC, jc = it % R::NJC;
      if (hasH) strip_mac<R::TJ>(acc2[m], Hnh, NP, ys, R::LDP, ti * 16, jc * R::TJ * 16, NP, lane);
      for (int kk = 0; kk < nL; ++kk)
        strip_mac<R::TJ>(acc2[m], X + kk * R::MAT, R::LDP, C + (size_t)kk * R::NN, NP, ti * 16, jc * R::TJ * 16, NP,
                         lane);
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < R::MAXI2; ++m) {
    const int it = wave + NW * m;
    if (it < R::ITEMS2) {
      const int ti = it / R::NJC, jc = it % R::NJC;
      strip_store<R::TJ>(acc2[m], X, R::LDP, ti * 16, jc * R::TJ * 16, lane);
    }
  }
  __syncthreads();
}

// Hermitian form of res_rhs_tmp for a Hermitian ys and C_k = D_k^dag / 2 (the
// Tsit5 descriptors): X_k C_k = D_k ys D_k^dag / 2 = D_k (X_k / 2)^dag, so
//   phase 1: X'_k = (D_k ys) / 2 = D_k ys^dag / 2,
//   phase 2: tmp = Hnh ys^dag + sum_k D_k X'_k^dag,
// with every global operand (Hnh, D_k) an A fragment and every LDS operand read
// along rows. tmp goes to the first free LDS matrix after X'_1..X'_nL (slot
// nL + 1), so only with nL = 3 jump operators does it wait for the phase-2 reads
// of X'_1. Returns the tmp matrix; published by a final barrier.
template <int NP, int NW>
__device__ __forceinline__ const float2 *res_rhs_tmp_h(const float2 *__restrict__ Hnh, const float2 *__restrict__ D,
                                                       int nL, bool hasH, float2 *lds) {
  using R = RCfg<NP, NW>;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const float2 *ys = lds;
  float2 *X = lds + R::MAT;
  __syncthreads();
  const int items1 = nL * R::NT * R::NJC;
  for (int it = wave; it < items1; it += NW) {
    const int kk = it / (R::NT * R::NJC), rem = it % (R::NT * R::NJC);
    const int ti = rem / R::NJC, jc = rem % R::NJC;
    Acc acc[R::TJ];
#pragma unroll
    for (int t = 0; t < R::TJ; ++t) acc_zero(acc[t]);
    strip_mac_h<R::TJ, NP>(acc, D + (size_t)kk * R::NN, NP, ys, R::LDP, ti * 16, jc * R::TJ * 16, lane);
#pragma unroll
    for (int t = 0; t < R::TJ; ++t) {
      acc[t].re *= 0.5f;
      acc[t].im *= 0.5f;
    }
    strip_store<R::TJ>(acc, X + kk * R::MAT, R::LDP, ti * 16, jc * R::TJ * 16, lane);
  }
  __syncthreads();
  Acc acc2[R::MAXI2][R::TJ];
#pragma unroll
  for (int m = 0; m < R::MAXI2; ++m) {
    const int it = wave + NW * m;
#pragma unroll
    for (int t = 0; t < R::TJ; ++t) acc_zero(acc2[m][t]);
    if (it < R::ITEMS2) {
      const int ti = it / R::NJC, jc = it % R::NJC;
      if (hasH) strip_mac_h<R::TJ, NP>(acc2[m], Hnh, NP, ys, R::LDP, ti * 16, jc * R::TJ * 16, lane);
      for (int kk = 0; kk < nL; ++kk)
        strip_mac_h<R::TJ, NP>(acc2[m], D + (size_t)kk * R::NN, NP, X + kk * R::MAT, R::LDP, ti * 16,
                               jc * R::TJ * 16, lane);
    }
  }
  float2 *T = X + (nL < kResMaxL ? nL : 0) * R::MAT;
  if (nL >= kResMaxL) __syncthreads();
#pragma unroll
  for (int m = 0; m < R::MAXI2; ++m) {
    const int it = wave + NW * m;
    if (it < R::ITEMS2) {
      const int ti = it / R::NJC, jc = it % R::NJC;
      strip_store<R::TJ>(acc2[m], T, R::LDP, ti * 16, jc * R::TJ * 16, lane);
    }
  }
  __syncthreads();
  return T;
}

// out[q] <- (tmp + tmp^dag) at the row-major elements e = t + 512 q
template <int NP, int NW>
__device__ __forceinline__ void res_rhs(const float2 *__restrict__ Hnh, const float2 *__restrict__ D, const float2 *__restrict__ C,
                        int nL, bool hasH, float2 *lds, float2 (&out)[(RCfg<NP, NW>::NE)]) {
  using R = RCfg<NP, NW>;
  res_rhs_tmp<NP, NW>(Hnh, D, C, nL, hasH, lds);
  const float2 *X = lds + R::MAT;
#pragma unroll
  for (int q = 0; q < R::NE; ++q) {
    const int e = threadIdx.x + R::NTH * q;
    if (e < R::NN) {
      const int i = e / NP, j = e % NP;
      const float2 a = X[i * R::LDP + j], b = X[j * R::LDP + i];
      out[q] = make_float2(a.x + b.x, a.y - b.y);
    } else {
      out[q] = make_float2(0.f, 0.f);
    }
  }
}

// out[q] <- (tmp + tmp^dag) at the owned upper-triangle elements (Hermitian ys,
// C_k = D_k^dag / 2)
template <int NP, int NW>
__device__ __forceinline__ void res_rhs_upper(const float2 *__restrict__ Hnh, const float2 *__restrict__ D,
                                              int nL, bool hasH, float2 *lds,
                                              const int (&ij)[(RCfg<NP, NW>::NEU)], float2 (&out)[(RCfg<NP, NW>::NEU)]) {
  using R = RCfg<NP, NW>;
  const float2 *X = res_rhs_tmp_h<NP, NW>(Hnh, D, nL, hasH, lds);
#pragma unroll
  for (int q = 0; q < R::NEU; ++q) {
    int x = ij[q];
    asm volatile("" : "+v"(x));
    if (x >= 0) {
      const int i = ij_row(x), j = ij_col(x);
      const float2 a = X[i * R::LDP + j], b = X[j * R::LDP + i];
      out[q] = make_float2(a.x + b.x, a.y - b.y);
    } else {
      out[q] = make_float2(0.f, 0.f);
    }
  }
}

template <int NP, int NW>
__device__ __forceinline__ void res_write_ys(float2 *lds, const float2 (&v)[(RCfg<NP, NW>::NE)]) {
  using R = RCfg<NP, NW>;
#pragma unroll
  for (int q = 0; q < R::NE; ++q) {
    const int e = threadIdx.x + R::NTH * q;
    if (e < R::NN) lds[(e / NP) * R::LDP + (e % NP)] = v[q];
  }
}

// ys <- v on both triangles (v given on the owned upper-triangle elements)
template <int NP, int NW>
__device__ __forceinline__ void res_write_ys_upper(float2 *lds, const int (&ij)[(RCfg<NP, NW>::NEU)],
                                                   const float2 (&v)[(RCfg<NP, NW>::NEU)]) {
  using R = RCfg<NP, NW>;
#pragma unroll
  for (int q = 0; q < R::NEU; ++q) {
    int x = ij[q];
    asm volatile("" : "+v"(x));  // recompute the addresses here instead of keeping them live
    if (x >= 0) {
      const int i = ij_row(x), j = ij_col(x);
      lds[i * R::LDP + j] = v[q];
      if (i != j) lds[j * R::LDP + i] = make_float2(v[q].x, -v[q].y);
    }
  }
}

__device__ __forceinline__ const float2 *res_hnh(const dqh_me_inst &p) {
  return (p.npwc > 0 || p.nmod > 0) ? gp<const float2>(p.work) : gp<const float2>(p.Hnh);
}

// Hcur (first np*np of work) <- Hnh0 + sum_j c_j Hpwc_j for PWC segment seg
// + sum_j f_j(tau) Hmod_j; global memory, published to the workgroup by the
// barrier at the start of the next res_rhs.
template <int NP, int NW>
__device__ __forceinline__ void res_update_pwc(const dqh_me_inst &p, int seg, double tau) {
  using R = RCfg<NP, NW>;
  const float2 *H0 = gp<const float2>(p.Hnh);
  const float2 *Hp = gp<const float2>(p.Hpwc);
  const float2 *Hm = gp<const float2>(p.Hmod);
  const double *vals = gp<const double>(p.pwc_vals);
  float2 *Hc = gp<float2>(p.work);
  const int nv = (int)p.nseg - 1, npwc = (int)p.npwc, nmod = (int)p.nmod;
  double2 fm[DQH_MAX_MOD];
  mod_coeffs(p, tau, fm);
  __syncthreads();  // no wave still reads the previous Hcur
  for (int e = threadIdx.x; e < R::NN; e += R::NTH) {
    double re = H0[e].x, im = H0[e].y;
    if (seg >= 0 && seg < nv) {
      for (int j = 0; j < npwc; ++j) {
        const double cr = vals[2 * (j * nv + seg)], ci = vals[2 * (j * nv + seg) + 1];
        const float2 o = Hp[(size_t)j * R::NN + e];
        re += cr * o.x - ci * o.y;
        im += cr * o.y + ci * o.x;
      }
    }
#pragma unroll
    for (int j = 0; j < DQH_MAX_MOD; ++j) {
      if (j < nmod) {
        const float2 o = Hm[(size_t)j * R::NN + e];
        re += fm[j].x * o.x - fm[j].y * o.y;
        im += fm[j].x * o.y + fm[j].y * o.x;
      }
    }
    Hc[e] = make_float2((float)re, (float)im);
  }
  __threadfence_block();
}

// save y + h sum_j w_j k_j at save index isave (states on both triangles, and the
// expectation values Tr(E rho) = sum_ij E^T_ij rho_ij summed over the owned
// elements and their mirror images rho_ji = conj(rho_ij))
template <int NP, int NW>
__device__ __forceinline__ void res_save(const dqh_me_inst &p, const double2 *__restrict__ y,
                                         const int (&ij)[(RCfg<NP, NW>::NEU)], const float2 (&k)[7][(RCfg<NP, NW>::NEU)],
                                         double h, const double (&w)[7], bool interp, int isave, double *sh) {
  using R = RCfg<NP, NW>;
  const int n = (int)p.n, nE = (int)p.nE, ntsave = (int)p.ntsave;
  const bool save_states = (p.flags & DQH_FLAG_SAVE_STATES) != 0;
  float2 *states = gp<float2>(p.states) + (size_t)isave * n * n;
  const float2 *ET = gp<const float2>(p.EopsT);
  double vr[R::NEU], vi[R::NEU];
#pragma unroll
  for (int q = 0; q < R::NEU; ++q) {
    const int u = threadIdx.x + R::NTH * q;
    const double2 yq = ij[q] >= 0 ? y[u] : make_double2(0.0, 0.0);
    double re = yq.x, im = yq.y;
    if (interp) {
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        re += h * w[j] * k[j][q].x;
        im += h * w[j] * k[j][q].y;
      }
    }
    vr[q] = re;
    vi[q] = im;
    const int i = ij_row(ij[q]), jj = ij_col(ij[q]);
    if (save_states && ij[q] >= 0 && jj < n) {
      states[(size_t)i * n + jj] = make_float2((float)re, (float)im);
      if (i != jj) states[(size_t)jj * n + i] = make_float2((float)re, (float)-im);
    }
  }
  float2 *ex = gp<float2>(p.expects);
  for (int qe = 0; qe < nE; ++qe) {
    const float2 *E = ET + (size_t)qe * R::NN;
    double ar = 0, ai = 0;
#pragma unroll
    for (int q = 0; q < R::NEU; ++q) {
      if (ij[q] >= 0) {
        const int i = ij_row(ij[q]), jj = ij_col(ij[q]);
        const float2 o = E[i * NP + jj];
        ar += o.x * vr[q] - o.y * vi[q];
        ai += o.x * vi[q] + o.y * vr[q];
        if (i != jj) {
          const float2 m = E[jj * NP + i];
          ar += m.x * vr[q] + m.y * vi[q];
          ai += m.y * vr[q] - m.x * vi[q];
        }
      }
    }
    const double sr = block_sum<NW>(ar, sh), si = block_sum<NW>(ai, sh);
    if (threadIdx.x == 0) ex[(size_t)qe * ntsave + isave] = make_float2((float)sr, (float)si);
  }
}

// x is equal in every lane of the wave: return it as a wave-uniform value
__device__ __forceinline__ double wave_uniform(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int NP, int NW>
__device__ __forceinline__ void rhs_resident(const dqh_me_inst &p, float2 *lds) {
  using R = RCfg<NP, NW>;
  if (p.npwc > 0 || p.nmod > 0) {
    const int seg = p.npwc > 0 ? pwc_segment(gp<const double>(p.pwc_times), (int)p.nseg, p.t0) : -1;
    res_update_pwc<NP, NW>(p, seg, p.t0);
  }
  const double2 *yg = gp<const double2>(p.y);
  float2 v[R::NE];
#pragma unroll
  for (int q = 0; q < R::NE; ++q) {
    const int e = threadIdx.x + R::NTH * q;
    v[q] = e < R::NN ? make_float2((float)yg[e].x, (float)yg[e].y) : make_float2(0.f, 0.f);
  }
  res_write_ys<NP, NW>(lds, v);
  float2 out[R::NE];
  res_rhs<NP, NW>(res_hnh(p), gp<const float2>(p.Dops), gp<const float2>(p.Cops),
              (int)p.nL, (p.flags & DQH_FLAG_HAS_HNH) != 0, lds, out);
  float2 *fin = gp<float2>(p.final_state);
  const int n = (int)p.n;
#pragma unroll
  for (int q = 0; q < R::NE; ++q) {
    const int e = threadIdx.x + R::NTH * q;
    const int i = e / NP, j = e % NP;
    if (e < R::NN && i < n && j < n) fin[(size_t)i * n + j] = out[q];
  }
}

template <int NP, int NW, bool HD>
__device__ __forceinline__ void tsit5_resident(const dqh_me_inst &p, double *sh, float2 *lds, double2 *dg) {
  using R = RCfg<NP, NW>;
  constexpr int NE = R::NEU;
  const int n = (int)p.n, nL = (int)p.nL, ntsave = (int)p.ntsave, nseg = (int)p.nseg;
  const bool pwc = p.npwc > 0;
  const bool hasH = (p.flags & DQH_FLAG_HAS_HNH) != 0;
  const double *tsave = gp<const double>(p.tsave);
  const double *ptimes = gp<const double>(p.pwc_times);
  const float2 *D = gp<const float2>(p.Dops);
  const float2 *Hnh = res_hnh(p);
  int64_t *stats = gp<int64_t>(p.stats);
  const double rtol = p.rtol, atol = p.atol;
  const double inv_nn = 1.0 / ((double)n * n);
  const double t_end = p.t1;

  const RKTab &rk = rk_tab(p.flags);
  int ij[NE];
  upper_map<NP, NW>(ij);
  // y: packed FP64 upper triangle, element u = t + 512 q of thread t at yg[u]. rho0
  // (row-major, np x np) is in the first np^2 elements of the y buffer; the packed
  // state starts in the second half, and each step's candidate y1 goes to the
  // other buffer, which becomes y on acceptance.
  double2 *const ybase = gp<double2>(p.y);
  double2 *yg = ybase + R::NN;
  double2 *yn = ybase;
  auto yload = [&](int q) -> double2 {
    const int u = threadIdx.x + R::NTH * q;
    return ij[q] >= 0 ? yg[u] : make_double2(0.0, 0.0);
  };
#pragma unroll
  for (int q = 0; q < NE; ++q) {
    if (ij[q] >= 0) yg[threadIdx.x + R::NTH * q] = ybase[ij_row(ij[q]) * NP + ij_col(ij[q])];
  }

  // FP64 diagonal of the generator (hdiag): Hnh then holds the off-diagonal part,
  // and every evaluation adds (d_i + conj(d_j)) ys[i][j] in FP64 to the FP32
  // products, the stage input ys given in FP64 by ysq(q). Published to the
  // workgroup by the barriers at the start of the first vector field.
  if constexpr (HD) {
    const double2 *hdg = gp<const double2>(p.hdiag);
    for (int i = threadIdx.x; i < NP; i += R::NTH) dg[i] = hdg[i];
  }
  auto add_diag = [&](float2 (&out)[NE], auto ysq) {
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      if (ij[q] >= 0) {
        const double2 a = dg[ij_row(ij[q])], b = dg[ij_col(ij[q])];
        const double gr = a.x + b.x, gi = a.y - b.y;
        const double2 y = ysq(q);
        out[q] = make_float2((float)((double)out[q].x + (gr * y.x - gi * y.y)),
                             (float)((double)out[q].y + (gr * y.y + gi * y.x)));
      }
      // one element at a time: keeps the FP64 temporaries of the elements from
      // being live together (the stages already fill the register file)
      __builtin_amdgcn_sched_barrier(0);
    }
  };

  float2 k[7][NE];
  float2 tmp[NE];
#pragma unroll
  for (int q = 0; q < NE; ++q) {
#pragma unroll
    for (int j = 0; j < 7; ++j) k[j][q] = make_float2(0.f, 0.f);
  }
  double w0[7] = {0, 0, 0, 0, 0, 0, 0};

  double t = p.t0;
  int isave = 0;
  int64_t nsteps = 0, nacc = 0, nrej = 0, nrhs = 0;
  int status = 0;
  int seg = -2;

  while (isave < ntsave && tsave[isave] <= t) {
    res_save<NP, NW>(p, yg, ij, k, 0.0, w0, false, isave, sh);
    ++isave;
  }
  const bool smooth = p.nmod > 0;  // H changes within a step: updated at every stage time
  const bool lin = !smooth;        // the generator is constant over each step
  if (pwc || smooth) {
    seg = pwc ? pwc_segment(ptimes, nseg, t) : -1;
    res_update_pwc<NP, NW>(p, seg, t);
  }
  // k0 = f(y0)
#pragma unroll
  for (int q = 0; q < NE; ++q) {
    const double2 yq = yload(q);
    tmp[q] = make_float2((float)yq.x, (float)yq.y);
  }
  res_write_ys_upper<NP, NW>(lds, ij, tmp);
  res_rhs_upper<NP, NW>(Hnh, D, nL, hasH, lds, ij, k[0]);
  if constexpr (HD) add_diag(k[0], yload);
  ++nrhs;

  // initial step (Hairer-Norsett-Wanner II.4), scale atol + rtol |y0|; sums over
  // the full matrix count each off-diagonal owned element twice
  double dt;
  {
    double s0 = 0, s1 = 0;
    double h0;
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const double2 yq = yload(q);
      const double ay2 = yq.x * yq.x + yq.y * yq.y;
      const double sc = atol + rtol * sqrt(ay2);
      const double wq = ij_weight(ij[q]);
      s0 += wq * ay2 / (sc * sc);
      s1 += wq * ((double)k[0][q].x * k[0][q].x + (double)k[0][q].y * k[0][q].y) / (sc * sc);
    }
    const double d0 = sqrt(block_sum<NW>(s0, sh) * inv_nn);
    const double d1 = sqrt(block_sum<NW>(s1, sh) * inv_nn);
    h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * d0 / d1;
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const double2 yq = yload(q);
      tmp[q] = make_float2((float)(yq.x + h0 * k[0][q].x), (float)(yq.y + h0 * k[0][q].y));
    }
    if (smooth) res_update_pwc<NP, NW>(p, seg, t + h0);
    res_write_ys_upper<NP, NW>(lds, ij, tmp);
    res_rhs_upper<NP, NW>(Hnh, D, nL, hasH, lds, ij, k[1]);
    if constexpr (HD)
      add_diag(k[1], [&](int q) {
        const double2 yq = yload(q);
        return make_double2(yq.x + h0 * k[0][q].x, yq.y + h0 * k[0][q].y);
      });
    ++nrhs;
    double s2 = 0;
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const double2 yq = yload(q);
      const double sc = atol + rtol * sqrt(yq.x * yq.x + yq.y * yq.y);
      const double dr = (double)k[1][q].x - k[0][q].x, di = (double)k[1][q].y - k[0][q].y;
      s2 += ij_weight(ij[q]) * (dr * dr + di * di) / (sc * sc);
    }
    const double d2 = sqrt(block_sum<NW>(s2, sh) * inv_nn) / h0;
    const double md = fmax(d1, d2);
    const double h1 = md <= 1e-15 ? fmax(1e-6, h0 * 1e-3) : ctrl_pow(0.01 / md, kInitExp);
    dt = wave_uniform(fmin(100.0 * h0, h1));
  }

  bool need_k0 = false;
  while (t < t_end) {
    if (nsteps >= p.max_steps) {
      status = DQH_MAX_STEPS;
      break;
    }
    if (pwc) {
      const int s = pwc_segment(ptimes, nseg, t);
      if (s != seg) {
        seg = s;
        if (!smooth) res_update_pwc<NP, NW>(p, seg, t);
      }
    }
    if (need_k0) {
#pragma unroll
      for (int q = 0; q < NE; ++q) {
        const double2 yq = yload(q);
        tmp[q] = make_float2((float)yq.x, (float)yq.y);
      }
      if (smooth) res_update_pwc<NP, NW>(p, seg, t);
      res_write_ys_upper<NP, NW>(lds, ij, tmp);
      res_rhs_upper<NP, NW>(Hnh, D, nL, hasH, lds, ij, k[0]);
      if constexpr (HD) add_diag(k[0], yload);
      ++nrhs;
      need_k0 = false;
    }
    // clip the step to the end time or the next breakpoint strictly after t
    double h = dt;
    double target = t_end;
    bool target_jump = false;
    if (pwc) {
      const int m = pwc_segment(ptimes, nseg, t) + 1;
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
    // stages 1..6. The stage input y + h sum_j a_sj k_j is accumulated in FP64 (the
    // coefficients reach |a| ~ 13 and the terms cancel) and rounded once to FP32.
    // With a generator that is constant over the step (lin), stages 1..5 evaluate
    // the vector field on the increment only: k_s = k_0 + g_s, g_s = f(Delta_s),
    // Delta_s = h sum_j a_sj k_j = h (sum_j a_sj) k_0 + h sum_{j>=1} a_sj g_j, so
    // the FP32 rounding of a stage scales with |Delta_s| instead of |y|; k[1..5]
    // then hold g_s, every combination below folds k_0 into the weight of k[0],
    // and stage 6 (the FSAL stage, the next k_0) is a full evaluation at y1.
    // The stage loop is a runtime loop (one copy of the GEMM code); the
    // register-resident stages are selected with compile-time indices.
#pragma unroll 1
    for (int s = 1; s < 7; ++s) {
      double c[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) c[j] = (j < s) ? h * rk.A[s][j] : 0.0;
      const bool with_y = !lin || s == 6;
      if (lin) {
        double a0 = 0.0;
#pragma unroll
        for (int j = 0; j < 6; ++j) a0 += (j < s) ? rk.A[s][j] : 0.0;
        c[0] = h * a0;
      }
      // the stage input: its FP32 rounding goes to LDS for the products, and with
      // hdiag the rounding residual stays in registers, so that the FP64 diagonal
      // term after the products sees the FP64 input (hi from LDS + lo)
      float2 ylo[HD ? NE : 1];
#pragma unroll
      for (int q = 0; q < NE; ++q) {
        const double2 yq = with_y ? yload(q) : make_double2(0.0, 0.0);
        double re = yq.x, im = yq.y;
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          re += c[j] * k[j][q].x;
          im += c[j] * k[j][q].y;
        }
        tmp[q] = make_float2((float)re, (float)im);
        if constexpr (HD) ylo[q] = make_float2((float)(re - (double)tmp[q].x), (float)(im - (double)tmp[q].y));
      }
      if (smooth) res_update_pwc<NP, NW>(p, seg, t + rk.C[s] * h);
      res_write_ys_upper<NP, NW>(lds, ij, tmp);
      res_rhs_upper<NP, NW>(Hnh, D, nL, hasH, lds, ij, tmp);
      if constexpr (HD)
        add_diag(tmp, [&](int q) {
          const float2 hi = lds[ij_row(ij[q]) * R::LDP + ij_col(ij[q])];
          return make_double2((double)hi.x + ylo[q].x, (double)hi.y + ylo[q].y);
        });
      ++nrhs;
#pragma unroll
      for (int j = 1; j < 7; ++j) {
        if (j == s) {
#pragma unroll
          for (int q = 0; q < NE; ++q) k[j][q] = tmp[q];
        }
      }
    }
    // error norm in FP64 with the candidate y1 = y + h sum b_j k_j, which is stored
    // to the other state buffer (error and solution weights folded for lin)
    double ew[7], bw[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      ew[j] = rk.E[j];
      bw[j] = rk.A[6][j];
    }
    if (lin) fold_k0(ew), fold_k0(bw);
    double ssum = 0;
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const double2 yq = yload(q);
      double er = 0, ei = 0, yr = yq.x, yi = yq.y;
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        er += ew[j] * k[j][q].x;
        ei += ew[j] * k[j][q].y;
        yr += h * bw[j] * k[j][q].x;
        yi += h * bw[j] * k[j][q].y;
      }
      er *= h;
      ei *= h;
      const double sc = atol + rtol * fmax(sqrt(yq.x * yq.x + yq.y * yq.y), sqrt(yr * yr + yi * yi));
      ssum += ij_weight(ij[q]) * (er * er + ei * ei) / (sc * sc);
      if (ij[q] >= 0) yn[threadIdx.x + R::NTH * q] = make_double2(yr, yi);
    }
    const double norm = wave_uniform(sqrt(block_sum<NW>(ssum, sh) * inv_nn));
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
      while (isave < ntsave && tsave[isave] <= t_new) {
        const double th = (tsave[isave] - t) / h;
        double w[7];
#pragma unroll
        for (int j = 0; j < 7; ++j) w[j] = th * (rk.R[j][0] + th * (rk.R[j][1] + th * (rk.R[j][2] + th * rk.R[j][3])));
        if (lin) fold_k0(w);
        res_save<NP, NW>(p, yg, ij, k, h, w, true, isave, sh);
        ++isave;
      }
#pragma unroll
      for (int q = 0; q < NE; ++q) k[0][q] = k[6][q];
      double2 *const yo = yg;
      yg = yn;
      yn = yo;
      t = t_new;
      need_k0 = made_jump;
    } else {
      ++nrej;
    }
    dt = wave_uniform(factor * h);
  }
  while (isave < ntsave) {
    res_save<NP, NW>(p, yg, ij, k, 0.0, w0, false, isave, sh);
    ++isave;
  }
  float2 *fin = gp<float2>(p.final_state);
  float2 *st = gp<float2>(p.states);
  const bool save_states = (p.flags & DQH_FLAG_SAVE_STATES) != 0;
#pragma unroll
  for (int q = 0; q < NE; ++q) {
    const int i = ij_row(ij[q]), j = ij_col(ij[q]);
    if (ij[q] >= 0 && j < n) {
      const double2 yq = yg[threadIdx.x + R::NTH * q];
      const float2 v = make_float2((float)yq.x, (float)yq.y);
      const float2 vc = make_float2(v.x, -v.y);
      fin[(size_t)i * n + j] = v;
      if (i != j) fin[(size_t)j * n + i] = vc;
      if (!save_states) {
        st[(size_t)i * n + j] = v;
        if (i != j) st[(size_t)j * n + i] = vc;
      }
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

}  // namespace dqh
