// MFMA contractions of the ensemble L-application (gfx950).
//
// GEMM1 (quadratic forms + per-image weights), one workgroup per (64 configs, symmetry):
//   Q = X Ys_s                       (N x n) . (n x n), never stored
//   alpha_i = sum_j X_ij Q_ij,  beta_i = sum_j Ft_ij Q_ij           (row dots in the epilogue)
//   gamma_i = X_i . g1_s,       delta_i = Ft_i . g2_s
//   c1 = rho (k_Y alpha/8 - k_R gamma/3), c2 = rho (k_Y beta/4 - k_R delta/3)
//   e1 = -rho alpha/6, e2 = -rho beta/3                              (LanczosFunctions.c:2195-2231,
//                                                                     2374-2407, 2687-2725)
// GEMM2 (back-accumulation), one workgroup per (64 x 64 output tile, symmetry):
//   B_s = X~^T G_s,  G_s = diag(c1) Y + diag(c2/2) X~                 (n x N) . (N x n)
//   so that S_s (B_s + B_s^T) S_s^T = sum_i c1 (ut f^T + f ut^T) + c2 ut ut^T.
//
// Both are "NT" products C[m][c] = sum_k A[m][k] B[c][k] with K-contiguous operand rows,
// staged through LDS in double-buffered BK-deep slices. FP64 uses v_mfma_f64_16x16x4f64;
// each wave owns a 32 x 32 sub-tile (2 x 2 MFMA blocks) of the 64 x 64 workgroup tile.
#include "tdscha_internal.h"

namespace tdscha {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int LDS_LD = BM + 16;  // row stride (doubles) of the [k][m] LDS slices

struct Gemm1Args {
  const double *X, *Ft, *rho, *Ys, *G;
  double *W;
  int Np, npd, n;
  int do_q;              // compute the Y1 quadratic forms (alpha, beta)
  double kf, kr, ky;     // which reference terms are requested (0 or 1)
};

struct Gemm2Args {
  const double *XtT, *YT, *W;
  double *B;
  int Np, npd;
};

// Stage a 64 x BK slice of a row-major matrix (rows r0.., cols k0..) into LDS as [k][row].
__device__ inline void load_slice(const double *__restrict__ M, int ld, int r0, int k0, double (&reg)[4]) {
  const int t = threadIdx.x;
  const int row = t >> 2, kk = (t & 3) * 4;
  const double2 *p = reinterpret_cast<const double2 *>(M + (size_t)(r0 + row) * ld + k0 + kk);
  double2 v0 = p[0], v1 = p[1];
  reg[0] = v0.x; reg[1] = v0.y; reg[2] = v1.x; reg[3] = v1.y;
}

__device__ inline void store_slice(double *lds, const double (&reg)[4]) {
  const int t = threadIdx.x;
  const int row = t >> 2, kk = (t & 3) * 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) lds[(kk + j) * LDS_LD + row] = reg[j];
}

// 16 MFMAs of one BK slice for this wave's 2 x 2 block of 16 x 16 tiles.
__device__ inline void mma_slice(const double *As, const double *Bs, int wr, int wc, v4d (&acc)[2][2]) {
  const int l = threadIdx.x & 63;
  const int lr = l & 15, lk = l >> 4;
#pragma unroll
  for (int kq = 0; kq < BK / 4; ++kq) {
    const int k = kq * 4 + lk;
    double a0 = As[k * LDS_LD + wr * 32 + lr];
    double a1 = As[k * LDS_LD + wr * 32 + 16 + lr];
    double b0 = Bs[k * LDS_LD + wc * 32 + lr];
    double b1 = Bs[k * LDS_LD + wc * 32 + 16 + lr];
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

__global__ __launch_bounds__(256) void k_gemm1_f64(Gemm1Args g) {
  __shared__ double As[2][BK * LDS_LD];
  __shared__ double Bs[2][BK * LDS_LD];
  __shared__ double red[2][BM][4];

  const int i0 = blockIdx.x * BM;
  const int sl = blockIdx.y;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w >> 1, wc = w & 1;
  const int lr = l & 15, lg = l >> 4;
  const double *Ys = g.Ys + (size_t)sl * g.npd * g.npd;
  const double *g1 = g.G + (size_t)sl * 2 * g.npd;
  const double *g2 = g1 + g.npd;

  // per-lane partial row sums: rows wr*32 + ti*16 + 4*lg + r
  double pa[2][4], pb[2][4], pc[2][4], pd[2][4];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int r = 0; r < 4; ++r) pa[ti][r] = pb[ti][r] = pc[ti][r] = pd[ti][r] = 0.0;

  const int KT = g.npd / BK;
  for (int c0 = 0; c0 < g.npd; c0 += BN) {
    v4d acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};

    if (g.do_q) {
      double ra[4], rb[4];
      load_slice(g.X, g.npd, i0, 0, ra);
      load_slice(Ys, g.npd, c0, 0, rb);
      store_slice(As[0], ra);
      store_slice(Bs[0], rb);
      __syncthreads();
      for (int kt = 0; kt < KT; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < KT) {
          load_slice(g.X, g.npd, i0, (kt + 1) * BK, ra);
          load_slice(Ys, g.npd, c0, (kt + 1) * BK, rb);
        }
        mma_slice(As[cur], Bs[cur], wr, wc, acc);
        if (kt + 1 < KT) {
          store_slice(As[cur ^ 1], ra);
          store_slice(Bs[cur ^ 1], rb);
        }
        __syncthreads();
      }
    }

    // epilogue: row dots of this column tile
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      const int c = c0 + wc * 32 + tj * 16 + lr;
      const double gg1 = g1[c], gg2 = g2[c];
#pragma unroll
      for (int ti = 0; ti < 2; ++ti) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = wr * 32 + ti * 16 + 4 * lg + r;
          const size_t o = (size_t)(i0 + m) * g.npd + c;
          const double x = g.X[o], ft = g.Ft[o];
          const double q = acc[ti][tj][r];
          pa[ti][r] = fma(x, q, pa[ti][r]);
          pb[ti][r] = fma(ft, q, pb[ti][r]);
          pc[ti][r] = fma(x, gg1, pc[ti][r]);
          pd[ti][r] = fma(ft, gg2, pd[ti][r]);
        }
      }
    }
  }

  // reduce over the 16 lanes holding the same rows (different columns)
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        pa[ti][r] += __shfl_xor(pa[ti][r], off, 64);
        pb[ti][r] += __shfl_xor(pb[ti][r], off, 64);
        pc[ti][r] += __shfl_xor(pc[ti][r], off, 64);
        pd[ti][r] += __shfl_xor(pd[ti][r], off, 64);
      }
      if (lr == 0) {
        const int m = wr * 32 + ti * 16 + 4 * lg + r;
        red[wc][m][0] = pa[ti][r];
        red[wc][m][1] = pb[ti][r];
        red[wc][m][2] = pc[ti][r];
        red[wc][m][3] = pd[ti][r];
      }
    }
  __syncthreads();
  if (threadIdx.x < BM) {
    const int m = threadIdx.x, i = i0 + m;
    const double alpha = red[0][m][0] + red[1][m][0];
    const double beta = red[0][m][1] + red[1][m][1];
    const double gamma = red[0][m][2] + red[1][m][2];
    const double delta = red[0][m][3] + red[1][m][3];
    const double rho = g.rho[i];
    double *Ws = g.W + (size_t)sl * 4 * g.Np;
    Ws[i] = rho * (g.ky * alpha / 8 - g.kr * gamma / 3);
    Ws[(size_t)g.Np + i] = rho * (g.ky * beta / 4 - g.kr * delta / 3);
    Ws[2 * (size_t)g.Np + i] = -g.kf * rho * alpha / 6;
    Ws[3 * (size_t)g.Np + i] = -g.kf * rho * beta / 3;
  }
}

// Stage the computed operand G^T[c][k] = c1[k] YT[c][k] + c2[k]/2 XtT[c][k].
__device__ inline void load_g_slice(const Gemm2Args &g, const double *c1, const double *c2, int r0,
                                    int k0, double (&reg)[4]) {
  const int t = threadIdx.x;
  const int row = t >> 2, kk = (t & 3) * 4;
  const size_t o = (size_t)(r0 + row) * g.Np + k0 + kk;
  const double2 *py = reinterpret_cast<const double2 *>(g.YT + o);
  const double2 *px = reinterpret_cast<const double2 *>(g.XtT + o);
  const double2 *pc1 = reinterpret_cast<const double2 *>(c1 + k0 + kk);
  const double2 *pc2 = reinterpret_cast<const double2 *>(c2 + k0 + kk);
  double2 y0 = py[0], y1 = py[1], x0 = px[0], x1 = px[1];
  double2 a0 = pc1[0], a1 = pc1[1], b0 = pc2[0], b1 = pc2[1];
  reg[0] = fma(a0.x, y0.x, 0.5 * b0.x * x0.x);
  reg[1] = fma(a0.y, y0.y, 0.5 * b0.y * x0.y);
  reg[2] = fma(a1.x, y1.x, 0.5 * b1.x * x1.x);
  reg[3] = fma(a1.y, y1.y, 0.5 * b1.y * x1.y);
}

__global__ __launch_bounds__(256) void k_gemm2_f64(Gemm2Args g) {
  __shared__ double As[2][BK * LDS_LD];
  __shared__ double Bs[2][BK * LDS_LD];
  const int m0 = blockIdx.x * BM, c0 = blockIdx.y * BN, sl = blockIdx.z;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w >> 1, wc = w & 1;
  const int lr = l & 15, lg = l >> 4;
  const double *c1 = g.W + (size_t)sl * 4 * g.Np;
  const double *c2 = c1 + g.Np;

  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};

  const int KT = g.Np / BK;
  double ra[4], rb[4];
  load_slice(g.XtT, g.Np, m0, 0, ra);
  load_g_slice(g, c1, c2, c0, 0, rb);
  store_slice(As[0], ra);
  store_slice(Bs[0], rb);
  __syncthreads();
  for (int kt = 0; kt < KT; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < KT) {
      load_slice(g.XtT, g.Np, m0, (kt + 1) * BK, ra);
      load_g_slice(g, c1, c2, c0, (kt + 1) * BK, rb);
    }
    mma_slice(As[cur], Bs[cur], wr, wc, acc);
    if (kt + 1 < KT) {
      store_slice(As[cur ^ 1], ra);
      store_slice(Bs[cur ^ 1], rb);
    }
    __syncthreads();
  }

  double *Bo = g.B + (size_t)sl * g.npd * g.npd;
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wr * 32 + ti * 16 + 4 * lg + r;
        const int c = c0 + wc * 32 + tj * 16 + lr;
        Bo[(size_t)m * g.npd + c] = acc[ti][tj][r];
      }
}

}  // namespace

void launch_gemm1(const Ensemble &e, int sb, int prec, int what, hipStream_t st) {
  Gemm1Args g;
  g.X = e.X; g.Ft = e.Ft; g.rho = e.rho; g.Ys = e.Ys; g.G = e.G; g.W = e.W;
  g.Np = e.Np; g.npd = e.npd; g.n = e.n;
  g.do_q = (what & (kTermF | kTermD2vY)) ? 1 : 0;
  g.kf = (what & kTermF) ? 1.0 : 0.0;
  g.kr = (what & kTermD2vR) ? 1.0 : 0.0;
  g.ky = (what & kTermD2vY) ? 1.0 : 0.0;
  (void)prec;
  dim3 grid(e.Np / BM, sb);
  hipLaunchKernelGGL(k_gemm1_f64, grid, dim3(256), 0, st, g);
}

void launch_gemm2(const Ensemble &e, int sb, int prec, hipStream_t st) {
  Gemm2Args g;
  g.XtT = e.XtT; g.YT = e.YT; g.W = e.W; g.B = e.B;
  g.Np = e.Np; g.npd = e.npd;
  (void)prec;
  dim3 grid(e.npd / BM, e.npd / BN, sb);
  hipLaunchKernelGGL(k_gemm2_f64, grid, dim3(256), 0, st, g);
}

}  // namespace tdscha
