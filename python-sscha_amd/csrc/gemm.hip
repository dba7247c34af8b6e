// MFMA contractions of the ensemble L-application (gfx950).
//
// GEMM1 (quadratic forms + per-image weights), one workgroup per (configs tile, symmetry):
//   Q = X Ys_s                       (N x n) . (n x n), never stored
//   alpha_i = sum_j X_ij Q_ij,  beta_i = sum_j Ft_ij Q_ij           (row dots in the epilogue)
//   gamma_i = X_i . g1_s,       delta_i = Ft_i . g2_s
//   c1 = rho (k_Y alpha/8 - k_R gamma/3), c2 = rho (k_Y beta/4 - k_R delta/3)
//   e1 = -rho alpha/6, e2 = -rho beta/3                              (LanczosFunctions.c:2195-2231,
//                                                                     2374-2407, 2687-2725)
// GEMM2 (back-accumulation), one workgroup per (output tile, symmetry):
//   B_s = X~^T G_s,  G_s = diag(c1) Y + diag(c2/2) X~                 (n x N) . (N x n)
//   so that S_s (B_s + B_s^T) S_s^T = sum_i c1 (ut f^T + f ut^T) + c2 ut ut^T.
//
// Both are "NT" products C[m][c] = sum_k A[m][k] B[c][k] with K-contiguous operand rows.
//
// FP64 / FP32: v_mfma_{f64,f32}_16x16x4, 64 x 64 workgroup tile, each wave a 32 x 32
// sub-tile (2 x 2 MFMA blocks), BK = 16 slices double-buffered through LDS as [k][m].
// The BF16x3 forms are in gemm_bf16.hip. Every epilogue (row dots, weights) is FP64.
#include "tdscha_internal.h"

namespace tdscha {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <class T> struct Acc;
template <> struct Acc<double> { typedef v4d type; };
template <> struct Acc<float> { typedef v4f type; };

__device__ inline v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline v4f mfma(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of result element r held by lane group lg = lane / 16 (the column is lane % 16).
// Measured on gfx950: the f32 (and bf16) 16x16 results put rows 4 lg .. 4 lg + 3 in one
// lane group; the f64 16x16x4 result interleaves them, row lg + 4 r.
template <class T> __device__ inline int out_row(int lg, int r);
template <> __device__ inline int out_row<float>(int lg, int r) { return 4 * lg + r; }
template <> __device__ inline int out_row<double>(int lg, int r) { return lg + 4 * r; }

// ---------------------------------------------------------------- FP64 / FP32 (16x16x4)
constexpr int BM = 64, BN = 64, BK = 16;
constexpr int LDS_LD = BM + 16;  // [k][m] row stride: consecutive k rows land 16 or 32 banks apart

template <class T>
struct Gemm1Args {
  const T *A;        // X in the arithmetic type          [Np x npd]
  const T *Ys;       // rotated Y1 in the arithmetic type [sb x npd x npd]
  const double *X, *Ft, *rho, *G;
  double *W;
  size_t rho_stride;   // rho of slot sl at rho + sl rho_stride (0: one rho for all slots)
  int Np, npd;
  int do_q;
  int raw;           // write alpha, beta, gamma, delta themselves instead of the weights
  double kf, kr, ky;
};

template <class T>
struct Gemm2Args {
  const T *XtT, *YT;  // in the arithmetic type [npd x Np]
  const double *W;
  double *B;
  int Np, npd;
};

// Stage a 64 x BK slice of a row-major matrix (rows r0.., cols k0..) into registers.
template <class T>
__device__ inline void load_slice(const T *__restrict__ M, int ld, int r0, int k0, T (&reg)[4]) {
  const int t = threadIdx.x;
  const int row = t >> 2, kk = (t & 3) * 4;
  const T *p = M + (size_t)(r0 + row) * ld + k0 + kk;
#pragma unroll
  for (int j = 0; j < 4; ++j) reg[j] = p[j];
}

template <class T>
__device__ inline void store_slice(T *lds, const T (&reg)[4]) {
  const int t = threadIdx.x;
  const int row = t >> 2, kk = (t & 3) * 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) lds[(kk + j) * LDS_LD + row] = reg[j];
}

// 16 MFMAs of one BK slice for this wave's 2 x 2 block of 16 x 16 tiles.
template <class T, class V>
__device__ inline void mma_slice(const T *As, const T *Bs, int wr, int wc, V (&acc)[2][2]) {
  const int l = threadIdx.x & 63;
  const int lr = l & 15, lk = l >> 4;
#pragma unroll
  for (int kq = 0; kq < BK / 4; ++kq) {
    const int k = kq * 4 + lk;
    T a0 = As[k * LDS_LD + wr * 32 + lr];
    T a1 = As[k * LDS_LD + wr * 32 + 16 + lr];
    T b0 = Bs[k * LDS_LD + wc * 32 + lr];
    T b1 = Bs[k * LDS_LD + wc * 32 + 16 + lr];
    acc[0][0] = mfma(a0, b0, acc[0][0]);
    acc[0][1] = mfma(a0, b1, acc[0][1]);
    acc[1][0] = mfma(a1, b0, acc[1][0]);
    acc[1][1] = mfma(a1, b1, acc[1][1]);
  }
}

// Per-image weights from the four row sums (shared by every precision).
__device__ inline void write_weights(double *W, int Np, int i, double rho, double alpha,
                                     double beta, double gamma, double delta, double kf,
                                     double kr, double ky) {
  W[i] = rho * (ky * alpha / 8 - kr * gamma / 3);
  W[(size_t)Np + i] = rho * (ky * beta / 4 - kr * delta / 3);
  W[2 * (size_t)Np + i] = -kf * rho * alpha / 6;
  W[3 * (size_t)Np + i] = -kf * rho * beta / 3;
}

template <class T>
__global__ __launch_bounds__(256, 2) void k_gemm1(Gemm1Args<T> g) {
  typedef typename Acc<T>::type V;
  __shared__ T As[2][BK * LDS_LD];
  __shared__ T Bs[2][BK * LDS_LD];
  __shared__ double red[2][BM][4];

  const int i0 = blockIdx.x * BM;
  const int sl = blockIdx.y;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w >> 1, wc = w & 1;
  const int lr = l & 15, lg = l >> 4;
  const T *Ys = g.Ys + (size_t)sl * g.npd * g.npd;
  const double *g1 = g.G + (size_t)sl * 2 * g.npd;
  const double *g2 = g1 + g.npd;

  double pa[2][4], pb[2][4], pc[2][4], pd[2][4];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int r = 0; r < 4; ++r) pa[ti][r] = pb[ti][r] = pc[ti][r] = pd[ti][r] = 0.0;

  const int KT = g.npd / BK;
  for (int c0 = 0; c0 < g.npd; c0 += BN) {
    V acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[a][b][r] = 0;

    if (g.do_q) {
      T ra[4], rb[4];
      load_slice(g.A, g.npd, i0, 0, ra);
      load_slice(Ys, g.npd, c0, 0, rb);
      store_slice(As[0], ra);
      store_slice(Bs[0], rb);
      __syncthreads();
      for (int kt = 0; kt < KT; ++kt) {
        const int cur = kt & 1;
        if (kt + 1 < KT) {
          load_slice(g.A, g.npd, i0, (kt + 1) * BK, ra);
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

#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      const int c = c0 + wc * 32 + tj * 16 + lr;
      const double gg1 = g1[c], gg2 = g2[c];
#pragma unroll
      for (int ti = 0; ti < 2; ++ti) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = wr * 32 + ti * 16 + out_row<T>(lg, r);
          const size_t o = (size_t)(i0 + m) * g.npd + c;
          const double x = g.X[o], ft = g.Ft[o];
          const double q = (double)acc[ti][tj][r];
          pa[ti][r] = fma(x, q, pa[ti][r]);
          pb[ti][r] = fma(ft, q, pb[ti][r]);
          pc[ti][r] = fma(x, gg1, pc[ti][r]);
          pd[ti][r] = fma(ft, gg2, pd[ti][r]);
        }
      }
    }
  }

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
        const int m = wr * 32 + ti * 16 + out_row<T>(lg, r);
        red[wc][m][0] = pa[ti][r];
        red[wc][m][1] = pb[ti][r];
        red[wc][m][2] = pc[ti][r];
        red[wc][m][3] = pd[ti][r];
      }
    }
  __syncthreads();
  if (threadIdx.x < BM) {
    const int m = threadIdx.x, i = i0 + m;
    double *Ws = g.W + (size_t)sl * 4 * g.Np;
    if (g.raw) {
#pragma unroll
      for (int q = 0; q < 4; ++q) Ws[(size_t)q * g.Np + i] = red[0][m][q] + red[1][m][q];
    } else {
      write_weights(Ws, g.Np, i, g.rho[(size_t)sl * g.rho_stride + i],
                    red[0][m][0] + red[1][m][0], red[0][m][1] + red[1][m][1],
                    red[0][m][2] + red[1][m][2],
                    red[0][m][3] + red[1][m][3], g.kf, g.kr, g.ky);
    }
  }
}

// The computed operand G^T[c][k] = c1[k] YT[c][k] + c2[k]/2 XtT[c][k].
template <class T>
__device__ inline void load_g_slice(const Gemm2Args<T> &g, const double *c1, const double *c2,
                                    int r0, int k0, T (&reg)[4]) {
  const int t = threadIdx.x;
  const int row = t >> 2, kk = (t & 3) * 4;
  const size_t o = (size_t)(r0 + row) * g.Np + k0 + kk;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const T a = (T)c1[k0 + kk + j], b = (T)(0.5 * c2[k0 + kk + j]);
    reg[j] = a * g.YT[o + j] + b * g.XtT[o + j];
  }
}

template <class T>
__global__ __launch_bounds__(256, 2) void k_gemm2(Gemm2Args<T> g) {
  typedef typename Acc<T>::type V;
  __shared__ T As[2][BK * LDS_LD];
  __shared__ T Bs[2][BK * LDS_LD];
  const int m0 = blockIdx.x * BM, c0 = blockIdx.y * BN, sl = blockIdx.z;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w >> 1, wc = w & 1;
  const int lr = l & 15, lg = l >> 4;
  const double *c1 = g.W + (size_t)sl * 4 * g.Np;
  const double *c2 = c1 + g.Np;

  V acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[a][b][r] = 0;

  const int KT = g.Np / BK;
  T ra[4], rb[4];
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
        const int m = m0 + wr * 32 + ti * 16 + out_row<T>(lg, r);
        const int c = c0 + wc * 32 + tj * 16 + lr;
        Bo[(size_t)m * g.npd + c] = (double)acc[ti][tj][r];
      }
}

// ---------------------------------------------------------------- f_av back-projection
// V[s][j] = sum_k YT[j][k] e1_s[k] + XtT[j][k] e2_s[k] in FP64 MFMA for every precision:
// an NT product (rows j x slots s, k contiguous in both operands), 64 rows x 96 slots per
// workgroup (waves 2 x 2, each 32 rows x 48 slots; a pass's slots in one workgroup column,
// so YT and XtT are read once per pass), split over k in chunks of kFavChunk configurations;
// the partial sums Vp[chunk][s][j] are added in a fixed order afterwards. The weight loads
// are clamped and branch-free (zeroed when staged), so the next slice stays in flight under
// the MFMAs.
constexpr int FV_S = 96;
constexpr int FV_LDB = FV_S + 16;
constexpr int FV_PT = FV_S * BK / 256;   // weights per thread and slice

__global__ __launch_bounds__(256, 2) void k_fav_mfma(const double *__restrict__ YT,
                                                  const double *__restrict__ XtT,
                                                  const double *__restrict__ W,
                                                  double *__restrict__ Vp, int Np, int npd,
                                                  int sb) {
  __shared__ double As[2][BK * LDS_LD];
  __shared__ double Bs[2][BK * FV_LDB];
  const int j0 = blockIdx.x * BM, kc = blockIdx.y, q0 = blockIdx.z * FV_S;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w >> 1, wc = w & 1;
  const int lr = l & 15, lg = l >> 4;
  const int k_begin = kc * kFavChunk;
  const int k_end = min(Np, k_begin + kFavChunk);
  const int KT = (k_end - k_begin) / BK;
  // B slice: 96 slots x 16 k; thread t stages k = t % 16 of slots t / 16 + 16 p (p < 6)
  const int bk = threadIdx.x & 15, bq = threadIdx.x >> 4;

  v4d acc[2][3];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 3; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[ti][tj][r] = 0.0;

  for (int part = 0; part < 2; ++part) {
    const double *M = part ? XtT : YT;
    const double *E = W + (size_t)(part ? 3 : 2) * Np;   // e2 or e1 of slot 0
    auto load_b = [&](int k0, double (&rb)[FV_PT]) {
#pragma unroll
      for (int p = 0; p < FV_PT; ++p)
        rb[p] = E[(size_t)min(q0 + bq + 16 * p, sb - 1) * 4 * Np + k0 + bk];
    };
    auto store_b = [&](double *lds, const double (&rb)[FV_PT]) {
#pragma unroll
      for (int p = 0; p < FV_PT; ++p)
        lds[bk * FV_LDB + bq + 16 * p] = q0 + bq + 16 * p < sb ? rb[p] : 0.0;
    };
    double ra[4], rb[FV_PT];
    load_slice(M, Np, j0, k_begin, ra);
    load_b(k_begin, rb);
    __syncthreads();   // the previous part's last reads of As / Bs are done
    store_slice(As[0], ra);
    store_b(Bs[0], rb);
    __syncthreads();
    for (int kt = 0; kt < KT; ++kt) {
      const int cur = kt & 1;
      if (kt + 1 < KT) {
        load_slice(M, Np, j0, k_begin + (kt + 1) * BK, ra);
        load_b(k_begin + (kt + 1) * BK, rb);
      }
#pragma unroll
      for (int kq = 0; kq < BK / 4; ++kq) {
        const int k = kq * 4 + lg;
        const double a0 = As[cur][k * LDS_LD + wr * 32 + lr];
        const double a1 = As[cur][k * LDS_LD + wr * 32 + 16 + lr];
#pragma unroll
        for (int tj = 0; tj < 3; ++tj) {
          const double b = Bs[cur][k * FV_LDB + wc * 48 + tj * 16 + lr];
          acc[0][tj] = mfma(a0, b, acc[0][tj]);
          acc[1][tj] = mfma(a1, b, acc[1][tj]);
        }
      }
      if (kt + 1 < KT) {
        store_slice(As[cur ^ 1], ra);
        store_b(Bs[cur ^ 1], rb);
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int tj = 0; tj < 3; ++tj) {
    const int q = q0 + wc * 48 + tj * 16 + lr;
    if (q < sb) {
      double *out = Vp + ((size_t)kc * sb + q) * npd + j0;
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          out[wr * 32 + ti * 16 + out_row<double>(lg, r)] = acc[ti][tj][r];
    }
  }
}

// V[s][j] = sum over chunks of Vp[chunk][s][j], chunks in increasing order.
__global__ void k_fav_reduce(const double *__restrict__ Vp, double *__restrict__ V, int n_chunks,
                             size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  double acc = 0.0;
  for (int c = 0; c < n_chunks; ++c) acc += Vp[(size_t)c * count + i];
  V[i] = acc;
}

// ---------------------------------------------------------------- force adjoint
// dJ/dY for J = sum over chains of uR . f_av + <U, d2v> (seeds from k_post_adj), per image
//   dJ/dy_i = sum_s 2 c1 (U~ x~_i) + rho (-kf nu/3 + ky mu/4) f_psi (Ys x_i)
//             + e1 S^T uR - kr rho mu/3 S^T R1
// with U~ = S^T U S, mu = x~^T U~ x~, nu = x~ . S^T uR and the forward weights c1, e1 of
// GEMM1. One K-concatenated FP64 MFMA product per
// (configs x modes) tile: for every slot of the pass the two row-scaled products
// (2 c1 X)(U~ f_ups) and (k X)(Ys f_psi), then the 2 sb rank-1 terms as extra k columns.
// Force adjoint, in the arithmetic type T of the contraction (FP64, or FP32 with FP32
// accumulation inside a launch; the launches' sums are added in FP64): X, Us and Ys in T,
// every weight and seed in FP64, rounded to T as it is staged.
template <class T>
struct AdjArgs {
  const T *X;                   // [Np x npd]
  const T *Us, *Ys;             // [sb x npd x npd]  f_ups S^T U S f_ups, S^T Ysym S
  const double *W, *Wu;         // [sb x 4 x Np]     forward weights, (mu, -, nu, -)
  const double *G, *Gu;         // [sb x 2 x npd]    (f_ups S^T R1, S^T R1 / f_psi), (., S^T uR)
  const double *rho, *fups, *fpsi;
  double *grad;                 // [Np x npd]        accumulated
  int Np, npd, n, sb;
  double kf, kr, ky;
};

template <class T>
__device__ inline void adj_load(const AdjArgs<T> &g, int step, int i0, int c0, T (&ra)[4],
                                T (&rb)[4]) {
  const int t = threadIdx.x, row = t >> 2, kk = (t & 3) * 4;
  const int i = i0 + row, c = c0 + row;
  const int KT = g.npd / BK;
  const int main_steps = g.sb * 2 * KT;
  if (step < main_steps) {
    const int slot = step / (2 * KT), half = (step / KT) & 1, k0 = (step % KT) * BK + kk;
    const double *Ws = g.W + (size_t)slot * 4 * g.Np, *Wus = g.Wu + (size_t)slot * 4 * g.Np;
    const double ca = half == 0 ? 2.0 * Ws[i]
                                : g.rho[i] * (-g.kf * Wus[2 * (size_t)g.Np + i] / 3 +
                                              g.ky * Wus[i] / 4);
    const T *xr = g.X + (size_t)i * g.npd + k0;
    const T *mr = (half == 0 ? g.Us : g.Ys) + (size_t)slot * g.npd * g.npd +
                  (size_t)c * g.npd + k0;
    const double cb = half == 0 ? (c < g.n ? 1.0 / g.fups[c] : 0.0) : g.fpsi[c];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      ra[j] = (T)(ca * (double)xr[j]);
      rb[j] = (T)(cb * (double)mr[j]);
    }
  } else {
    const int k0 = (step - main_steps) * BK + kk;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int kap = k0 + j, slot = kap >> 1;
      ra[j] = rb[j] = (T)0;
      if (slot < g.sb) {
        const double *Ws = g.W + (size_t)slot * 4 * g.Np, *Wus = g.Wu + (size_t)slot * 4 * g.Np;
        if ((kap & 1) == 0) {
          ra[j] = (T)Ws[2 * (size_t)g.Np + i];                        // e1
          rb[j] = (T)g.Gu[(size_t)slot * 2 * g.npd + g.npd + c];      // S^T uR
        } else {
          ra[j] = (T)(-g.kr * g.rho[i] * Wus[i] / 3);                 // -kr rho mu / 3
          rb[j] = (T)(g.G[(size_t)slot * 2 * g.npd + g.npd + c] * g.fpsi[c]);   // S^T R1
        }
      }
    }
  }
}

template <class T>
__global__ __launch_bounds__(256, 2) void k_adj_grad(AdjArgs<T> g) {
  typedef typename Acc<T>::type V;
  __shared__ T As[2][BK * LDS_LD];
  __shared__ T Bs[2][BK * LDS_LD];
  const int i0 = blockIdx.x * BM, c0 = blockIdx.y * BN;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w >> 1, wc = w & 1;
  const int lr = l & 15, lg = l >> 4;
  V acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (V){0, 0, 0, 0};
  const int steps = g.sb * 2 * (g.npd / BK) + (2 * g.sb + BK - 1) / BK;
  T ra[4], rb[4];
  adj_load<T>(g, 0, i0, c0, ra, rb);
  store_slice(As[0], ra);
  store_slice(Bs[0], rb);
  __syncthreads();
  for (int st = 0; st < steps; ++st) {
    const int cur = st & 1;
    if (st + 1 < steps) adj_load<T>(g, st + 1, i0, c0, ra, rb);
    mma_slice(As[cur], Bs[cur], wr, wc, acc);
    if (st + 1 < steps) {
      store_slice(As[cur ^ 1], ra);
      store_slice(Bs[cur ^ 1], rb);
    }
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = i0 + wr * 32 + ti * 16 + out_row<T>(lg, r);
        const int c = c0 + wc * 32 + tj * 16 + lr;
        g.grad[(size_t)m * g.npd + c] += (double)acc[ti][tj][r];
      }
}

// grad[i][c] (N x n) = (or +=) gradp[i][c] (Np x npd)
__global__ void k_grad_out(const double *__restrict__ gp, double *__restrict__ grad, int N, int n,
                           int npd, int accumulate) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (c >= n || i >= N) return;
  const double v = gp[(size_t)i * npd + c];
  grad[(size_t)i * n + c] = accumulate ? grad[(size_t)i * n + c] + v : v;
}

}  // namespace

// rho_s[j][i] = rho[i] when i N_sym + sym(j) lies in [lo, hi], else 0 (the images of a
// range-sliced call, Modules/tdscha_core.jl:71-73: bigindex k -> config (k - 1) / N_sym,
// symmetry (k - 1) % N_sym)
__global__ void k_window_rho(const double *__restrict__ rho, double *__restrict__ rho_s, int Np,
                             int N_sym, Pass ps, long long lo, long long hi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= Np) return;
  const long long k = (long long)i * N_sym + ps.sym(j);
  rho_s[(size_t)j * Np + i] = (k >= lo && k <= hi) ? rho[i] : 0.0;
}

void launch_window_rho(const Ensemble &e, const Pass &ps, hipStream_t st) {
  hipLaunchKernelGGL(k_window_rho, dim3((e.Np + 255) / 256, ps.slots()), dim3(256), 0, st, e.rho,
                     e.rho_s, e.Np, e.N_sym, ps, e.win_lo, e.win_hi);
}

void launch_gemm1(const Ensemble &e, int sb, int prec, int what, hipStream_t st) {
  const int do_q = (what & (kTermF | kTermD2vY)) ? 1 : 0;
  const double kf = (what & kTermF) ? 1.0 : 0.0;
  const double kr = (what & kTermD2vR) ? 1.0 : 0.0;
  const double ky = (what & kTermD2vY) ? 1.0 : 0.0;
  if (prec == kBF16x3) {
    launch_gemm1_bf16x3(e, sb, what, st);
  } else if (prec == kFP32) {
    Gemm1Args<float> g;
    g.A = e.X_f; g.Ys = e.Ys_f32;
    g.X = e.X; g.Ft = e.Ft; g.rho = e.rho_w(); g.rho_stride = e.rho_stride(); g.G = e.G;
    g.W = e.W;
    g.Np = e.Np; g.npd = e.npd; g.do_q = do_q; g.raw = 0; g.kf = kf; g.kr = kr; g.ky = ky;
    hipLaunchKernelGGL(k_gemm1<float>, dim3(e.Np / BM, sb), dim3(256), 0, st, g);
  } else {
    Gemm1Args<double> g;
    g.A = e.X; g.Ys = e.Ys;
    g.X = e.X; g.Ft = e.Ft; g.rho = e.rho_w(); g.rho_stride = e.rho_stride(); g.G = e.G;
    g.W = e.W;
    g.Np = e.Np; g.npd = e.npd; g.do_q = do_q; g.raw = 0; g.kf = kf; g.kr = kr; g.ky = ky;
    hipLaunchKernelGGL(k_gemm1<double>, dim3(e.Np / BM, sb), dim3(256), 0, st, g);
  }
}

void launch_gemm1_raw(const Ensemble &e, int sb, const double *M, const double *G, double *out,
                      hipStream_t st) {
  Gemm1Args<double> g;
  g.A = e.X; g.Ys = M;
  g.X = e.X; g.Ft = e.Ft; g.rho = e.rho; g.rho_stride = 0; g.G = G; g.W = out;
  g.Np = e.Np; g.npd = e.npd; g.do_q = 1; g.raw = 1; g.kf = g.kr = g.ky = 1.0;
  hipLaunchKernelGGL(k_gemm1<double>, dim3(e.Np / BM, sb), dim3(256), 0, st, g);
}

void launch_gemm1_raw_f32(const Ensemble &e, int sb, const float *M, const double *G,
                          double *out, hipStream_t st) {
  Gemm1Args<float> g;
  g.A = e.X_f; g.Ys = M;
  g.X = e.X; g.Ft = e.Ft; g.rho = e.rho; g.rho_stride = 0; g.G = G; g.W = out;
  g.Np = e.Np; g.npd = e.npd; g.do_q = 1; g.raw = 1; g.kf = g.kr = g.ky = 1.0;
  hipLaunchKernelGGL(k_gemm1<float>, dim3(e.Np / BM, sb), dim3(256), 0, st, g);
}

void launch_gemm2(const Ensemble &e, int sb, int prec, hipStream_t st) {
  if (prec == kBF16x3) {
    launch_gemm2_bf16x3(e, sb, st);
  } else if (prec == kFP32) {
    Gemm2Args<float> g;
    g.XtT = e.XtT_f; g.YT = e.YT_f; g.W = e.W; g.B = e.B; g.Np = e.Np; g.npd = e.npd;
    hipLaunchKernelGGL(k_gemm2<float>, dim3(e.npd / BM, e.npd / BN, sb), dim3(256), 0, st, g);
  } else {
    Gemm2Args<double> g;
    g.XtT = e.XtT; g.YT = e.YT; g.W = e.W; g.B = e.B; g.Np = e.Np; g.npd = e.npd;
    hipLaunchKernelGGL(k_gemm2<double>, dim3(e.npd / BM, e.npd / BN, sb), dim3(256), 0, st, g);
  }
}

void launch_fav(const Ensemble &e, int sb, hipStream_t st) {
  const int n_chunks = (e.Np + kFavChunk - 1) / kFavChunk;
  dim3 grid(e.npd / BM, n_chunks, (sb + FV_S - 1) / FV_S);
  hipLaunchKernelGGL(k_fav_mfma, grid, dim3(256), 0, st, e.YT, e.XtT, e.W, e.Vp, e.Np, e.npd, sb);
  const size_t count = (size_t)sb * e.npd;
  hipLaunchKernelGGL(k_fav_reduce, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, e.Vp,
                     e.V, n_chunks, count);
}

// ---- the exact image path (images.hip): FP64 contractions of one image slot
void launch_gemm1_images(const Ensemble &e, const double *Ysym_c, int what, hipStream_t st) {
  Gemm1Args<double> g;
  g.A = e.IU; g.Ys = Ysym_c;
  g.X = e.IU; g.Ft = e.IFt; g.rho = e.rho_w(); g.rho_stride = e.rho_stride(); g.G = e.G;
  g.W = e.W;
  g.Np = e.Np; g.npd = e.npd;
  g.do_q = (what & (kTermF | kTermD2vY)) ? 1 : 0;
  g.raw = 0;
  g.kf = (what & kTermF) ? 1.0 : 0.0;
  g.kr = (what & kTermD2vR) ? 1.0 : 0.0;
  g.ky = (what & kTermD2vY) ? 1.0 : 0.0;
  hipLaunchKernelGGL(k_gemm1<double>, dim3(e.Np / BM, 1), dim3(256), 0, st, g);
}

void launch_gemm2_images(const Ensemble &e, hipStream_t st) {
  Gemm2Args<double> g;
  g.XtT = e.IUT; g.YT = e.IFT; g.W = e.W; g.B = e.B; g.Np = e.Np; g.npd = e.npd;
  hipLaunchKernelGGL(k_gemm2<double>, dim3(e.npd / BM, e.npd / BN, 1), dim3(256), 0, st, g);
}

void launch_fav_images(const Ensemble &e, hipStream_t st) {
  const int n_chunks = (e.Np + kFavChunk - 1) / kFavChunk;
  dim3 grid(e.npd / BM, n_chunks, 1);
  hipLaunchKernelGGL(k_fav_mfma, grid, dim3(256), 0, st, e.IFT, e.IUT, e.W, e.Vp, e.Np, e.npd, 1);
  hipLaunchKernelGGL(k_fav_reduce, dim3((unsigned)((e.npd + 255) / 256)), dim3(256), 0, st, e.Vp,
                     e.V, n_chunks, (size_t)e.npd);
}

void launch_adj_grad(const Ensemble &e, int sb, double kf, double kr, double ky, int prec,
                     hipStream_t st) {
  const dim3 grid(e.Np / BM, e.npd / BN);
  if (prec == kFP32) {
    AdjArgs<float> g;
    g.X = e.X_f; g.Us = e.Us_f32; g.Ys = e.Ys_f32;
    g.W = e.W; g.Wu = e.Wu; g.G = e.G; g.Gu = e.Gu;
    g.rho = e.rho; g.fups = e.fups; g.fpsi = e.fpsi; g.grad = e.gradp;
    g.Np = e.Np; g.npd = e.npd; g.n = e.n; g.sb = sb; g.kf = kf; g.kr = kr; g.ky = ky;
    hipLaunchKernelGGL(k_adj_grad<float>, grid, dim3(256), 0, st, g);
    return;
  }
  AdjArgs<double> g;
  g.X = e.X; g.Us = e.Us; g.Ys = e.Ys; g.W = e.W; g.Wu = e.Wu; g.G = e.G; g.Gu = e.Gu;
  g.rho = e.rho; g.fups = e.fups; g.fpsi = e.fpsi; g.grad = e.gradp;
  g.Np = e.Np; g.npd = e.npd; g.n = e.n; g.sb = sb; g.kf = kf; g.kr = kr; g.ky = ky;
  hipLaunchKernelGGL(k_adj_grad<double>, grid, dim3(256), 0, st, g);
}

void launch_grad_out(const Ensemble &e, double *grad, int accumulate, hipStream_t st) {
  hipLaunchKernelGGL(k_grad_out, dim3((e.n + 255) / 256, e.N), dim3(256), 0, st, e.gradp, grad,
                     e.N, e.n, e.npd, accumulate);
}

}  // namespace tdscha
