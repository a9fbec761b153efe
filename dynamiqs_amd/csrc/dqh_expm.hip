// Matrix exponentials of piecewise-constant generators and their time-ordered
// products (the Expm method of the propagators), for gfx950.
//
// Every interval k of every chain b has a generator already scaled on the host:
// X[b][k] = dt_k A_k / 2^s. The library forms E = T_m(X)^(2^s) with the degree-m
// Taylor polynomial in Horner form, T <- I + (X / j) T for j = m .. 1, then s
// squarings, and chains U[b][k] = E[b][k] U[b][k-1]. Everything is a batched
// complex64 GEMM on FP32 MFMA (four real v_mfma_f32_16x16x4f32 per complex
// k-slice, FP32 accumulation): one 16x16 output tile per workgroup, split over its
// four waves along k, operands read straight from L2 (a batch of N <= 1024
// matrices stays resident there), with the Horner epilogue (scale, + I) fused
// into the store.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "dqh.h"
#include "dqh_common.h"
#include "dqh_mfma.h"

namespace dqh {
namespace ex {

constexpr int kWaves = 4;
constexpr int64_t kMaxGridY = 65535;

// C[b] = scale * (A[b] @ B[b]) (+ I when add_id), N x N row-major, N % 16 == 0.
// One workgroup per 16x16 output tile; its four waves take consecutive quarters
// of the k range (split-K, so that a small batch still spreads over the CUs and
// each wave's chain of dependent loads is four times shorter) and wave 0 adds the
// partial tiles from LDS and stores with the epilogue.
__global__ __launch_bounds__(64 * kWaves) void k_cgemm(const float2 *__restrict__ A, int64_t sA,
                                                      const float2 *__restrict__ B, int64_t sB, int64_t bdiv,
                                                      float2 *__restrict__ C, int64_t sC, int N, float scale,
                                                      int add_id, int64_t b0) {
  __shared__ float part[kWaves - 1][8][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nt = N / 16;
  // XCD-aware order: workgroup L runs on XCD L % 8, so work item
  // (L % 8) * (total / 8) + L / 8 gives each XCD a contiguous run of tiles, and the
  // tiles of one product share that XCD's L2 (identity order when total % 8 != 0)
  const int64_t total = (int64_t)gridDim.x * gridDim.y;
  const int64_t L = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
  const int64_t w = (total % 8 == 0) ? (L % 8) * (total / 8) + L / 8 : L;
  const int tile = (int)(w % gridDim.x);
  const int64_t b = b0 + w / gridDim.x;
  const int tr = tile / nt, tc = tile - tr * nt;
  const int kc = ((nt + kWaves - 1) / kWaves) * 16;  // k columns per wave, a multiple of 16
  const int kbeg = wave * kc < N ? wave * kc : N;
  const int kend = kbeg + kc < N ? kbeg + kc : N;
  Acc acc;
  acc_zero(acc);
  if (kend > kbeg)
    tile_mac(acc, A + b * sA + kbeg, N, B + (b / bdiv) * sB + (int64_t)kbeg * N, N, 16 * tr, 16 * tc, kend - kbeg,
             lane);
  if (wave > 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      part[wave - 1][v][lane] = acc.re[v];
      part[wave - 1][4 + v][lane] = acc.im[v];
    }
  }
  __syncthreads();
  if (wave != 0) return;
#pragma unroll
  for (int w = 0; w < kWaves - 1; ++w)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      acc.re[v] += part[w][v][lane];
      acc.im[v] += part[w][4 + v][lane];
    }
  const int r = lane & 15, g = lane >> 4;
  float2 *Cb = C + b * sC;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = 16 * tr + 4 * g + v, col = 16 * tc + r;
    const float id = (add_id && row == col) ? 1.f : 0.f;
    Cb[(size_t)row * N + col] = make_float2(scale * acc.re[v] + id, scale * acc.im[v]);
  }
}

// T[b] = I for every matrix of the batch (vector stores; a fill kernel keeps the
// sequence free of memset nodes)
__global__ __launch_bounds__(256) void k_eye(float2 *__restrict__ T, int N, int64_t nmat) {
  const int64_t nn = (int64_t)N * N;
  const int64_t total = nmat * nn;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t w = e % nn;
    const int row = (int)(w / N), col = (int)(w - (int64_t)row * N);
    T[e] = make_float2(row == col ? 1.f : 0.f, 0.f);
  }
}

// U[b][0] = E[b][0] for every chain (strided copy)
__global__ __launch_bounds__(256) void k_first(const float2 *__restrict__ E, float2 *__restrict__ U, int64_t nn,
                                              int64_t stride, int64_t nchain) {
  const int64_t total = nchain * nn;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / nn, w = e - b * nn;
    U[b * stride + w] = E[b * stride + w];
  }
}

// C[b] = scale A[b] B[b / bdiv] (+ I) for b < nbatch
int gemm(const float2 *A, int64_t sA, const float2 *B, int64_t sB, float2 *C, int64_t sC, int N, float scale,
         int add_id, int64_t nbatch, hipStream_t st, int64_t bdiv = 1) {
  const int nt = N / 16;
  const unsigned gx = (unsigned)(nt * nt);
  for (int64_t b0 = 0; b0 < nbatch; b0 += kMaxGridY) {
    const unsigned gy = (unsigned)(nbatch - b0 < kMaxGridY ? nbatch - b0 : kMaxGridY);
    hipLaunchKernelGGL(k_cgemm, dim3(gx, gy), dim3(64 * kWaves), 0, st, A, sA, B, sB, bdiv, C, sC, N, scale, add_id,
                       b0);
  }
  const hipError_t err = hipGetLastError();
  return err == hipSuccess ? DQH_OK : fail(DQH_HIP_ERROR, std::string("dqh batched GEMM: ") + hipGetErrorString(err));
}

}  // namespace ex
}  // namespace dqh

extern "C" int64_t dqh_expm_work_bytes(int64_t nchain, int64_t nint, int64_t N) {
  if (nchain <= 0 || nint <= 0 || N <= 0) return 0;
  return 2 * nchain * nint * N * N * 8;
}

namespace dqh {
namespace ex {

// Taylor + scaling-and-squaring exponentials and the chained products, with the
// two work matrices of every interval at w0 and w1 (nchain * nint * N * N
// complex64 elements each). Every element of w0 / w1 is written before it is
// read: k_eye fills w0, each GEMM fills its whole output.
int expm_run(const float2 *x, float2 *w0, float2 *w1, float2 *u, int64_t nchain, int64_t nint, int64_t N, int64_t m,
             int64_t s, hipStream_t st) {
  const int64_t nn = N * N, nmat = nchain * nint;
  const unsigned fill_blocks = (unsigned)(nmat * nn / 256 + 1 < 8192 ? nmat * nn / 256 + 1 : 8192);
  hipLaunchKernelGGL(k_eye, dim3(fill_blocks), dim3(256), 0, st, w0, (int)N, nmat);
  // Horner: T <- I + (X / j) T, j = m .. 1
  for (int64_t j = m; j >= 1; --j) {
    const int rc = gemm(x, nn, w0, nn, w1, nn, (int)N, (float)(1.0 / (double)j), 1, nmat, st);
    if (rc != DQH_OK) return rc;
    float2 *t = w0;
    w0 = w1;
    w1 = t;
  }
  // squarings: T <- T T
  for (int64_t i = 0; i < s; ++i) {
    const int rc = gemm(w0, nn, w0, nn, w1, nn, (int)N, 1.f, 0, nmat, st);
    if (rc != DQH_OK) return rc;
    float2 *t = w0;
    w0 = w1;
    w1 = t;
  }
  // chains: U[b][0] = E[b][0]; U[b][k] = E[b][k] U[b][k-1]
  hipLaunchKernelGGL(k_first, dim3((unsigned)(nchain * nn / 256 + 1 < 8192 ? nchain * nn / 256 + 1 : 8192)), dim3(256),
                     0, st, w0, u, nn, nint * nn, nchain);
  for (int64_t k = 1; k < nint; ++k) {
    const int rc = gemm(w0 + k * nn, nint * nn, u + (k - 1) * nn, nint * nn, u + k * nn, nint * nn, (int)N, 1.f, 0,
                        nchain, st);
    if (rc != DQH_OK) return rc;
  }
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) return fail(DQH_HIP_ERROR, std::string("dqh Expm: ") + hipGetErrorString(err));
  return DQH_OK;
}

bool expm_args_ok(const void *X, const void *U, int64_t nchain, int64_t nint, int64_t N, int64_t m, int64_t s) {
  return X && U && nchain > 0 && nint > 0 && N > 0 && N % 16 == 0 && N <= 4096 && m >= 1 && m <= 30 && s >= 0 &&
         s <= 60 && nchain * nint <= ((int64_t)1 << 31);
}

// Work memory of dqh_expm_propagators_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends. Host code never passes the array's name
// to a kernel (on the host it is a stand-in, not a device address): each device's
// address comes from hipGetSymbolAddress on the first call made with that device
// current, and only that address is used.
constexpr int64_t kNoAllocBytes = DQH_EXPM_NOALLOC_MAX_BYTES;
__device__ __align__(256) float2 scratch[kNoAllocBytes / (int64_t)sizeof(float2)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of ``scratch`` on the current device (looked up once per device).
int scratch_ptr(float2 **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(DQH_HIP_ERROR, std::string("dqh_expm_propagators_noalloc: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(DQH_UNSUPPORTED, "dqh_expm_propagators_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(DQH_HIP_ERROR, std::string("dqh_expm_propagators_noalloc: hipGetSymbolAddress: ") +
                                     hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = (float2 *)g_scratch_ptr[dev];
  return DQH_OK;
}

}  // namespace ex
}  // namespace dqh

extern "C" int dqh_expm_propagators(const void *X, void *work, void *U, int64_t nchain, int64_t nint, int64_t N,
                                    int64_t m, int64_t s, void *stream) {
  using namespace dqh;
  if (!work || !ex::expm_args_ok(X, U, nchain, nint, N, m, s))
    return fail(DQH_BAD_ARG, "dqh_expm_propagators: bad arguments");
  const int64_t nmat = nchain * nint;
  float2 *w0 = (float2 *)work, *w1 = w0 + nmat * N * N;
  return ex::expm_run((const float2 *)X, w0, w1, (float2 *)U, nchain, nint, N, m, s, (hipStream_t)stream);
}

extern "C" int dqh_expm_propagators_noalloc(const void *X, void *U, int64_t nchain, int64_t nint, int64_t N,
                                            int64_t m, int64_t s, void *stream) {
  using namespace dqh;
  if (!ex::expm_args_ok(X, U, nchain, nint, N, m, s))
    return fail(DQH_BAD_ARG, "dqh_expm_propagators_noalloc: bad arguments");
  const int64_t need = dqh_expm_work_bytes(nchain, nint, N);
  if (need > ex::kNoAllocBytes)
    return fail(DQH_BAD_SHAPE, "dqh_expm_propagators_noalloc: needs " + std::to_string(need) +
                                   " bytes of work memory, above its fixed " + std::to_string(ex::kNoAllocBytes) +
                                   " (2 * nchain * nint * N * N * 8 must not exceed it)");
  float2 *w0 = nullptr;
  const int rc = ex::scratch_ptr(&w0);
  if (rc != DQH_OK) return rc;
  float2 *w1 = w0 + nchain * nint * N * N;
  return ex::expm_run((const float2 *)X, w0, w1, (float2 *)U, nchain, nint, N, m, s, (hipStream_t)stream);
}

extern "C" int dqh_cgemm_batched(const void *A, int64_t sA, const void *B, int64_t sB, int64_t bdiv, void *C,
                                 int64_t sC, int64_t N, int64_t nbatch, void *stream) {
  if (!A || !B || !C || N <= 0 || N % 16 != 0 || N > 4096 || nbatch <= 0 || bdiv <= 0 || sA < 0 || sB < 0 ||
      sC < N * N)
    return dqh::fail(DQH_BAD_ARG, "dqh_cgemm_batched: bad arguments");
  return dqh::ex::gemm((const float2 *)A, sA, (const float2 *)B, sB, (float2 *)C, sC, (int)N, 1.f, 0, nbatch,
                       (hipStream_t)stream, bdiv);
}
