// Lindblad vector field for sparse-DIA operators (gfx950).
//
// Replaces vector_field_unvec_hermitian (dynamiqs/integrators/core/
// diffrax_integrator.py:319-323) when H and the jump operators are stored by
// diagonals (dynamiqs/qarrays/sparsedia_primitives.py:172-261):
//   drho = Hnh rho + rho Hnh^dag + sum_k L_k rho L_k^dag,  Hnh = -iH - 1/2 sum L^dag L,
// which is the reference's tmp + tmp^dag with tmp = Hnh rho + 1/2 sum L rho L^dag,
// evaluated element by element. Every output element reads a handful of
// neighbours of rho, so the sweep is bound by HBM: rho is read once and drho
// written once. The tiled forms stage a 64 x 64 tile of rho with an 8, 16 or
// 32-wide halo in LDS; the global form reads neighbours through the caches (any
// offsets).
#include <hip/hip_runtime.h>

#include <string>

#include "dqh.h"
#include "dqh_common.h"
#include "dqh_dia_tile.h"

namespace dqh {
namespace diar {

template <int HALO>
__global__ __launch_bounds__(diat::kThreads) void k_rhs_tiled(int64_t n, dqh_dia_ops ops, const float2 *__restrict__ rho,
                                                             float2 *__restrict__ drho) {
  __shared__ float2 sx[diat::Tile<HALO>::NX];
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  int i0, j0;
  diat::tile_of(n, i0, j0);
  diat::load_halo<HALO>(sx, rho + inst * nn, nullptr, n, i0, j0);
  __syncthreads();
  const int jj = threadIdx.x & (diat::kTJ - 1);
  float2 *out = drho + inst * nn;
#pragma unroll 1
  for (int m = 0; m < diat::kRowsPerThread; ++m) {
    const int ii = (threadIdx.x >> 6) + (diat::kThreads / diat::kTJ) * m;
    const int i = i0 + ii, j = j0 + jj;
    out[(size_t)i * n + j] = diat::stencil(diat::LdsX<HALO>::at(sx, ii, jj), ops, false, n, i, j);
  }
}

__global__ __launch_bounds__(256) void k_rhs_global(int64_t n, dqh_dia_ops ops, const float2 *__restrict__ rho,
                                                    float2 *__restrict__ drho) {
  const size_t nn = (size_t)n * n;
  const size_t inst = blockIdx.y;
  int i, j;
  if (!diat::row_elem(n, i, j)) return;
  const int64_t e = (int64_t)i * n + j;
  const diat::GlobalX X{rho + inst * nn, nullptr, n, i, j};
  drho[inst * nn + e] = diat::stencil(X, ops, false, n, i, j);
}

}  // namespace diar
}  // namespace dqh

using namespace dqh;

extern "C" int64_t dqh_sizeof_dia_rhs(void) { return (int64_t)sizeof(dqh_dia_rhs); }

extern "C" int dqh_lindblad_rhs_dia(const dqh_dia_rhs *pb, void *stream) {
  if (!pb) return fail(DQH_BAD_ARG, "dqh_lindblad_rhs_dia: null problem");
  const dqh_dia_rhs &p = *pb;
  if (p.n <= 0 || p.ninst <= 0 || p.n > (1 << 20) || !p.rho || !p.drho)
    return fail(DQH_BAD_SHAPE, "dqh_lindblad_rhs_dia: bad n, ninst or null buffer");
  if (p.ops.nops < 1 || p.ops.nops > DQH_DIA_MAXOPS)
    return fail(DQH_BAD_ARG, "dqh_lindblad_rhs_dia: 1 <= nops <= 8 (Hnh, then the jump operators)");
  for (int k = 0; k < p.ops.nops; ++k) {
    if (p.ops.noff[k] < 1 || p.ops.noff[k] > DQH_DIA_MAXOFF || !p.ops.diags[k])
      return fail(DQH_BAD_ARG, "dqh_lindblad_rhs_dia: bad number of diagonals or null diagonals");
    for (int u = 0; u < p.ops.noff[k]; ++u)
      if (p.ops.off[k][u] <= -p.n || p.ops.off[k][u] >= p.n)
        return fail(DQH_BAD_ARG, "dqh_lindblad_rhs_dia: offset outside (-n, n)");
  }
  if (p.variant == DQH_SWEEP_SPLIT)
    return fail(DQH_BAD_ARG, "dqh_lindblad_rhs_dia: the split sweep applies to the Rouchon1 Kraus map only");
  const int halo = diat::resolve_halo(p.variant, p.n, diat::max_abs_offset(p.ops));
  if (halo < 0)
    return fail(DQH_BAD_ARG, "dqh_lindblad_rhs_dia: the tiled sweep needs n % 64 == 0 and every |offset| <= its halo");
  hipStream_t st = (hipStream_t)stream;
  const float2 *rho = reinterpret_cast<const float2 *>(p.rho);
  float2 *drho = reinterpret_cast<float2 *>(p.drho);
  const dim3 gtile((unsigned)((p.n / diat::kTI) * (p.n / diat::kTJ)), (unsigned)p.ninst);
  if (halo == 8)
    hipLaunchKernelGGL(diar::k_rhs_tiled<8>, gtile, dim3(diat::kThreads), 0, st, p.n, p.ops, rho, drho);
  else if (halo == 16)
    hipLaunchKernelGGL(diar::k_rhs_tiled<16>, gtile, dim3(diat::kThreads), 0, st, p.n, p.ops, rho, drho);
  else if (halo == 32)
    hipLaunchKernelGGL(diar::k_rhs_tiled<32>, gtile, dim3(diat::kThreads), 0, st, p.n, p.ops, rho, drho);
  else
    hipLaunchKernelGGL(diar::k_rhs_global, diat::row_grid(p.n, p.ninst), dim3(256), 0, st, p.n, p.ops, rho, drho);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(DQH_HIP_ERROR, std::string("dqh_lindblad_rhs_dia: ") + hipGetErrorString(err));
  return DQH_OK;
}
