// Weighted ensemble averages behind the perturbation builders (Modules/Perturbations.py,
// get_d1M_dR_av :195-238, get_d1alpha_dR_av :240-273, get_d2M_dR_av :275-328,
// get_d2alpha_dR_av :330-375): every one is a contraction over the configurations,
//
//   C[m][n] = scale * sum_k A[k][m] w[k] B[k][n]
//
// (A = the Upsilon-weighted displacements or a column of ones, B = the per-configuration
// effective charges or Raman tensors, w = rho, scale = 1 / N_eff). FP64 MFMA
// (v_mfma_f64_16x16x4f64), 64 x 64 tile per workgroup, k slices of 16 through LDS as [k][m],
// [k][n]; the next slice is loaded (branch-free, clamped) during the MFMAs of the current one.
#include <string>

#include "tdscha_internal.h"
#include "../../include/tdscha_hip.h"

namespace tdscha {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int WM = 64, WN = 64, WK = 16;
constexpr int WLD = 64 + 16;

struct WtnArgs {
  const double *A, *w, *B;
  double *C;
  int K, M, N, lda, ldb, ldc;
  double scale;
};

// one slice: thread t holds rows k0 + t/16 of A and B, columns m0 / n0 + 4 (t%16) .. +3
struct WtnRegs {
  double a[4], b[4], w;
};

__device__ inline void wtn_load(const WtnArgs &g, int m0, int n0, int k0, WtnRegs &s) {
  const int t = threadIdx.x;
  const int k = min(k0 + (t >> 4), g.K - 1), c = (t & 15) * 4;
  s.w = g.w ? g.w[k] : 1.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    s.a[j] = g.A[(size_t)k * g.lda + min(m0 + c + j, g.M - 1)];
    s.b[j] = g.B[(size_t)k * g.ldb + min(n0 + c + j, g.N - 1)];
  }
}

__device__ inline void wtn_store(const WtnArgs &g, int m0, int n0, int k0, double *As,
                                 double *Bs, const WtnRegs &s) {
  const int t = threadIdx.x;
  const int kl = t >> 4, c = (t & 15) * 4;
  const bool kin = k0 + kl < g.K;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    As[kl * WLD + c + j] = (kin && m0 + c + j < g.M) ? s.a[j] * s.w : 0.0;
    Bs[kl * WLD + c + j] = (kin && n0 + c + j < g.N) ? s.b[j] : 0.0;
  }
}

__global__ __launch_bounds__(256, 2) void k_weighted_tn(WtnArgs g) {
  __shared__ double As[2][WK * WLD], Bs[2][WK * WLD];
  const int m0 = blockIdx.y * WM, n0 = blockIdx.x * WN;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w >> 1, wc = w & 1;
  const int lr = l & 15, lg = l >> 4;
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};

  const int KT = (g.K + WK - 1) / WK;
  WtnRegs s;
  wtn_load(g, m0, n0, 0, s);
  wtn_store(g, m0, n0, 0, As[0], Bs[0], s);
  if (KT > 1) wtn_load(g, m0, n0, WK, s);
  __syncthreads();
  for (int kt = 0; kt < KT; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < KT) wtn_store(g, m0, n0, (kt + 1) * WK, As[cur ^ 1], Bs[cur ^ 1], s);
    if (kt + 2 < KT) wtn_load(g, m0, n0, (kt + 2) * WK, s);
    const double *A = As[cur], *B = Bs[cur];
#pragma unroll
    for (int kq = 0; kq < WK / 4; ++kq) {
      const int k = kq * 4 + lg;
      const double a0 = A[k * WLD + wr * 32 + lr], a1 = A[k * WLD + wr * 32 + 16 + lr];
      const double b0 = B[k * WLD + wc * 32 + lr], b1 = B[k * WLD + wc * 32 + 16 + lr];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  // f64 16x16x4 result: lane (lr, lg) holds column lr of rows lg + 4 r
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      const int n = n0 + wc * 32 + tj * 16 + lr;
      if (n >= g.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wr * 32 + ti * 16 + lg + 4 * r;
        if (m < g.M) g.C[(size_t)m * g.ldc + n] = g.scale * acc[ti][tj][r];
      }
    }
}

}  // namespace

void launch_weighted_tn(int K, int M, int N, const double *A, int lda, const double *w,
                        const double *B, int ldb, double scale, double *C, int ldc,
                        hipStream_t st) {
  WtnArgs g{A, w, B, C, K, M, N, lda, ldb, ldc, scale};
  hipLaunchKernelGGL(k_weighted_tn, dim3((N + WN - 1) / WN, (M + WM - 1) / WM), dim3(256), 0, st,
                     g);
}

}  // namespace tdscha

using namespace tdscha;

extern "C" int tdscha_weighted_tn_product(int K, int M, int N, const double *A, int lda,
                                          const double *w, const double *B, int ldb,
                                          double scale, double *C, int ldc, void *stream) {
  if (K <= 0 || M <= 0 || N <= 0) {
    set_error("K, M and N must be positive");
    return 2;
  }
  if (!A || !B || !C) {
    set_error("NULL argument");
    return 1;
  }
  if (lda < M || ldb < N || ldc < N) {
    set_error("leading dimensions must be at least M (A) and N (B, C)");
    return 2;
  }
  const long long gy = (M + WM - 1) / WM, gx = (N + WN - 1) / WN;
  if (gy > 65535 || gx > (1LL << 30)) {
    set_error("arrays too large");
    return 2;
  }
  launch_weighted_tn(K, M, N, A, lda, w, B, ldb, scale, C, ldc, (hipStream_t)stream);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_error(std::string("launch failed: ") + hipGetErrorString(err));
    return 10;
  }
  return 0;
}
