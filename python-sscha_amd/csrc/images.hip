// Exact image path for symmetry blocks whose modes do not share one frequency.
//
// The fast path rotates the n x n perturbation by every symmetry instead of building the
// N x N_sym images, which needs diag(f_ups) and diag(f_psi) to commute with every S_s: true
// when the frequencies are equal inside each block (degenerate modes). The reference
// evaluates f_ups / f_psi per mode inside its image loops (CModules/LanczosFunctions.c:
// 2214-2221, 2401-2407, 2708-2725) and so accepts any block. For ensembles with a block
// whose frequencies differ by more than 1e-8 relative, this path builds, one image slot
// (chain c, symmetry s) at a time, the images of every configuration as the reference does
// (C:2158-2170),
//   u = S_s x,  f = S_s y,  u~ = f_ups . u,  f~ = f_psi . f,
// and runs the FP64 MFMA contractions on them with the unrotated perturbation:
//   alpha = u^T Y1 u, beta = f~^T Y1 u, gamma = u~ . R1, delta = f . R1     (GEMM1)
//   B = U~^T (c1 F + c2/2 U~),  d2v += B + B^T                             (GEMM2)
//   f_av += F^T e1 + U~^T e2                                                (f_av product)
// FP64 only (2 N n^2 + 2 N n^2 MFMA work per image slot, as the fast path, plus the image
// construction, N n d_b per slot).
#include "tdscha_internal.h"

namespace tdscha {

namespace {

// Images of every configuration under symmetry s: U and F~ row-major [Np x npd], U~^T and
// F^T [npd x Np]. Thread x: configuration, block y: mode (padding modes and rows are zero).
__global__ __launch_bounds__(256) void k_images(const double *__restrict__ X,
                                                const double *__restrict__ YT,
                                                const double *__restrict__ S,
                                                const ModeMeta *__restrict__ meta,
                                                const double *__restrict__ fups,
                                                const double *__restrict__ fpsi, int N, int n,
                                                int Np, int npd, int sym_stride, int s,
                                                double *__restrict__ U, double *__restrict__ Ft,
                                                double *__restrict__ UtT,
                                                double *__restrict__ FT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (i >= Np) return;
  double u = 0.0, f = 0.0;
  if (m < n && i < N) {
    const ModeMeta mm = meta[m];
    const double *row = S + (size_t)s * sym_stride + mm.sym_off + (size_t)(m - mm.blk_off) * mm.blk_dim;
    for (int k = 0; k < mm.blk_dim; ++k) {
      const double sk = row[k];
      u = fma(sk, X[(size_t)i * npd + mm.blk_off + k], u);
      f = fma(sk, YT[(size_t)(mm.blk_off + k) * Np + i], f);
    }
  }
  U[(size_t)i * npd + m] = u;
  Ft[(size_t)i * npd + m] = fpsi[m] * f;
  UtT[(size_t)m * Np + i] = fups[m] * u;
  FT[(size_t)m * Np + i] = f;
}

// G of image slot 0: g1 = g2 = f_ups . R1 (gamma = u . g1 = u~ . R1, delta = f~ . g2 = f . R1)
__global__ void k_image_g(const double *__restrict__ R1p, const double *__restrict__ fups,
                          int npd, double *__restrict__ G) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= npd) return;
  const double g = fups[m] * R1p[m];
  G[m] = g;
  G[npd + m] = g;
}

// d2v_acc (one chain) += B + B^T;  f_acc (one chain) += V
__global__ void k_image_acc(const double *__restrict__ B, const double *__restrict__ V,
                            double *__restrict__ d2v, double *__restrict__ f, int npd, int do_d,
                            int do_f) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (b >= npd) return;
  if (do_d) d2v[(size_t)a * npd + b] += B[(size_t)a * npd + b] + B[(size_t)b * npd + a];
  if (do_f && a == 0) f[b] += V[b];
}

}  // namespace

void launch_images(const Ensemble &e, int s, hipStream_t st) {
  dim3 grid((e.Np + 255) / 256, e.npd);
  hipLaunchKernelGGL(k_images, grid, dim3(256), 0, st, e.X, e.YT, e.S, e.meta, e.fups, e.fpsi,
                     e.N, e.n, e.Np, e.npd, e.sym_stride, s, e.IU, e.IFt, e.IUT, e.IFT);
}

void launch_image_g(const Ensemble &e, int chain, hipStream_t st) {
  hipLaunchKernelGGL(k_image_g, dim3((e.npd + 255) / 256), dim3(256), 0, st,
                     e.R1p + (size_t)chain * e.npd, e.fups, e.npd, e.G);
}

void launch_image_acc(const Ensemble &e, int chain, bool do_d, bool do_f, hipStream_t st) {
  dim3 grid((e.npd + 255) / 256, e.npd);
  hipLaunchKernelGGL(k_image_acc, grid, dim3(256), 0, st, e.B, e.V,
                     e.d2v_acc + (size_t)chain * e.npd * e.npd, e.f_acc + (size_t)chain * e.npd,
                     e.npd, (int)do_d, (int)do_f);
}

}  // namespace tdscha
