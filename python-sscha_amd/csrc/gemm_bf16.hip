// BF16x3 contractions of the ensemble L-application (gfx950).
//
// a*b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi with hi = RNE_bf16(v), lo = RNE_bf16(v - hi):
// three v_mfma_f32_16x16x32_bf16 per operand pair, FP32 accumulation, FP64 epilogues.
//
// Both contractions are "NT" products C[m][c] = sum_k A[m][k] B[c][k] with k-contiguous
// operand rows, tiled TM x TN per workgroup of 2 TN threads (waves in a 2 x TN/64 grid, each
// TM/2 x 64 = TM/32 x 4 MFMA blocks). Two shapes, picked per ensemble by the autotuner:
// 256 x 128 (GEMM1) / 192 x 128 (GEMM2, whose operand formation needs the registers a
// 256-row tile would take) with 4 waves, or 256 x 256 for both with 8 waves. k goes in
// slices of 32 staged through LDS as [row][k] with an 80-byte row stride
// (conflict-free 16-byte fragment reads). The next slice is fetched into registers while
// the current one is multiplied. Waves whose sub-tile lies past the valid rows / columns of
// a ragged tile issue no MFMAs.
//
// Workgroups are numbered so that each XCD (they are dealt out round-robin by workgroup id)
// runs a contiguous range of tiles, slot-major: the concurrent workgroups of an XCD work on
// the same symmetry image and share its operand slices in that XCD's L2.
//
// GEMM1 (per image s, configs i):  Q = X Ys_s, with the row sums of the tile
//     alpha_i = sum_c X_ic Q_ic, beta_i = sum_c Ft_ic Q_ic, gamma_i = sum_c X_ic g1_c,
//     delta_i = sum_c Ft_ic g2_c
//   written per column tile to P[s][ct][4][Np]; k_weights_bf16 adds the column tiles in
//   order and forms c1, c2, e1, e2 (LanczosFunctions.c:2195-2231, 2374-2407, 2687-2725).
// GEMM2 (per image s):  B_s = X~^T G_s with G_s^T[c][k] = c1_k Y^T[c][k] + c2_k/2 X~^T[c][k]
//   formed in FP32 from the FP32 operand copies and split while it is staged.
#include "tdscha_internal.h"
#include "bf16x3.h"

namespace tdscha {

using namespace bfx;

namespace {

constexpr int LDK = TK + 8;      // LDS row stride in bf16
constexpr int WN = 4;            // 16 x 16 column blocks per wave (64 columns)

// A workgroup tile is TM x TN with 2 TN threads: waves in a 2 x (TN / 64) grid, each
// (TM / 2) x 64. Variant 0: GEMM1 256 x 128, GEMM2 192 x 128 (256 threads); variant 1:
// both 256 x 256 (512 threads), half the operand traffic per flop of variant 0.
template <int TM, int TN> struct Tile {
  static constexpr int NT = 2 * TN;           // threads
  static constexpr int WC = TN / 64;          // waves across the columns
  static constexpr int WM = TM / 32;          // 16-row blocks per wave
};

template <int TM, int TN> struct Stage {
  uint16_t ah[TM * LDK], al[TM * LDK], bh[TN * LDK], bl[TN * LDK];
};

// R rows x 32 k of one bf16 plane with NT threads: 16-byte chunks, rows t/4 + (NT/4) h
// (zero past `rows`).
template <int R, int NT>
__device__ inline void load_rows(const uint16_t *__restrict__ P, size_t ld, int rows, int r0,
                                 int k0, uint4 (&reg)[4 * R / NT]) {
  const int t = threadIdx.x, row = t >> 2, kc = (t & 3) * 8;
#pragma unroll
  for (int h = 0; h < 4 * R / NT; ++h) {
    const int r = r0 + row + (NT / 4) * h;
    reg[h] = r < rows ? *reinterpret_cast<const uint4 *>(P + (size_t)r * ld + k0 + kc)
                      : make_uint4(0, 0, 0, 0);
  }
}

template <int R, int NT>
__device__ inline void store_rows(uint16_t *lds, const uint4 (&reg)[4 * R / NT]) {
  const int t = threadIdx.x, row = t >> 2, kc = (t & 3) * 8;
#pragma unroll
  for (int h = 0; h < 4 * R / NT; ++h)
    *reinterpret_cast<uint4 *>(lds + (row + (NT / 4) * h) * LDK + kc) = reg[h];
}

__device__ inline v8bf frag(const uint16_t *lds, int row, int kg) {
  return __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(lds + row * LDK + kg * 8));
}

// 3 MFMAs per 16 x 16 block of this wave's sub-tile, for one k slice.
template <int TM, int TN>
__device__ inline void mma_full(const Stage<TM, TN> &s, int wr, int wc,
                                v4f (&acc)[Tile<TM, TN>::WM][WN]) {
  constexpr int WM = Tile<TM, TN>::WM;
  const int l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
  v8bf bh[WN], bl[WN];
#pragma unroll
  for (int tj = 0; tj < WN; ++tj) {
    bh[tj] = frag(s.bh, wc * 64 + tj * 16 + lr, lg);
    bl[tj] = frag(s.bl, wc * 64 + tj * 16 + lr, lg);
  }
#pragma unroll
  for (int ti = 0; ti < WM; ++ti) {
    const v8bf ah = frag(s.ah, wr * (TM / 2) + ti * 16 + lr, lg);
    const v8bf al = frag(s.al, wr * (TM / 2) + ti * 16 + lr, lg);
#pragma unroll
    for (int tj = 0; tj < WN; ++tj) {
      acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[tj], acc[ti][tj], 0, 0, 0);
      acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[tj], acc[ti][tj], 0, 0, 0);
      acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[tj], acc[ti][tj], 0, 0, 0);
    }
  }
}

// mi / nj (blocks of this wave's sub-tile with valid rows / columns) are wave-uniform, read
// from the first lane, so the branch is scalar: a wave entirely past the edge of a ragged
// tile issues no MFMA; a partly valid one multiplies the zero padding (operand rows past
// the edge are loaded as zeros) rather than carrying a second, guarded copy of the body.
template <int TM, int TN>
__device__ inline void mma_slice(const Stage<TM, TN> &s, int wr, int wc, int mi, int nj,
                                 v4f (&acc)[Tile<TM, TN>::WM][WN]) {
  if (mi > 0 && nj > 0) mma_full<TM, TN>(s, wr, wc, acc);
}

// ------------------------------------------------------------------------------ GEMM1
struct Gemm1Args {
  const uint16_t *Xh, *Xl;    // [Np x npd]
  const uint16_t *Yh, *Yl;    // [s_batch x npd x npd] planes of the rotated Y1
  const float *Xf, *Ftf;      // [Np x npd] epilogue operands in FP32
  const double *G;            // [s_batch x 2 x npd]
  double *P;                  // [sb x n_ct x 4 x Np] row sums per column tile
  int Np, npd, sb, n_it, n_ct, sg, do_q;   // sg: image slots per tile group
};

template <int TM, int TN>
__global__ __launch_bounds__(2 * TN, 256 / TN) void k_gemm1_bf16x3(Gemm1Args g) {
  constexpr int NT = Tile<TM, TN>::NT, WC = Tile<TM, TN>::WC, WM = Tile<TM, TN>::WM;
  __shared__ Stage<TM, TN> sm;
  __shared__ double red[WC][TM][4];
  const int total = g.sb * g.n_it * g.n_ct;
  const int id = xcd_tile(total);
  if (id >= total) return;
  const SlotTile tt = slot_group_tile(id, g.sb, g.sg, g.n_it, g.n_ct);
  const int sl = tt.sl, it = tt.it, ct = tt.ct;
  const int i0 = it * TM, c0 = ct * TN;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w / WC, wc = w % WC, lr = l & 15, lg = l >> 4;
  const int nj = clamp_blocks(g.npd, c0 + wc * 64, WN);
  const int mi = WM;   // Np is a multiple of the row tile
  const size_t plane_off = (size_t)sl * g.npd * g.npd;

  v4f acc[WM][WN];
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};

  if (g.do_q) {
    const int KT = g.npd / TK;
    uint4 rah[4 * TM / NT], ral[4 * TM / NT], rbh[4 * TN / NT], rbl[4 * TN / NT];
    load_rows<TM, NT>(g.Xh, g.npd, g.Np, i0, 0, rah);
    load_rows<TM, NT>(g.Xl, g.npd, g.Np, i0, 0, ral);
    load_rows<TN, NT>(g.Yh + plane_off, g.npd, g.npd, c0, 0, rbh);
    load_rows<TN, NT>(g.Yl + plane_off, g.npd, g.npd, c0, 0, rbl);
    for (int kt = 0; kt < KT; ++kt) {
      __syncthreads();
      store_rows<TM, NT>(sm.ah, rah);
      store_rows<TM, NT>(sm.al, ral);
      store_rows<TN, NT>(sm.bh, rbh);
      store_rows<TN, NT>(sm.bl, rbl);
      __syncthreads();
      if (kt + 1 < KT) {
        const int k0 = (kt + 1) * TK;
        load_rows<TM, NT>(g.Xh, g.npd, g.Np, i0, k0, rah);
        load_rows<TM, NT>(g.Xl, g.npd, g.Np, i0, k0, ral);
        load_rows<TN, NT>(g.Yh + plane_off, g.npd, g.npd, c0, k0, rbh);
        load_rows<TN, NT>(g.Yl + plane_off, g.npd, g.npd, c0, k0, rbl);
      }
      mma_slice<TM, TN>(sm, wr, wc, mi, nj, acc);
    }
  }

  // epilogue: FP64 row sums over this tile's columns, one 16-row block at a time
  const double *g1 = g.G + (size_t)sl * 2 * g.npd, *g2 = g1 + g.npd;
#pragma unroll
  for (int ti = 0; ti < WM; ++ti) {
    double p[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) p[r][0] = p[r][1] = p[r][2] = p[r][3] = 0.0;
#pragma unroll
    for (int tj = 0; tj < WN; ++tj) {
      const int c = c0 + wc * 64 + tj * 16 + lr;
      if (tj < nj) {
        const double gg1 = g1[c], gg2 = g2[c];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = wr * (TM / 2) + ti * 16 + 4 * lg + r;
          const size_t o = (size_t)(i0 + m) * g.npd + c;
          const double x = g.Xf[o], ft = g.Ftf[o];
          const double q = (double)acc[ti][tj][r];
          p[r][0] = fma(x, q, p[r][0]);
          p[r][1] = fma(ft, q, p[r][1]);
          p[r][2] = fma(x, gg1, p[r][2]);
          p[r][3] = fma(ft, gg2, p[r][3]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double v = p[r][q];
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 64);
        if (lr == 0) red[wc][wr * (TM / 2) + ti * 16 + 4 * lg + r][q] = v;
      }
  }
  __syncthreads();
  double *Pt = g.P + ((size_t)sl * g.n_ct + ct) * 4 * g.Np + i0;
  for (int idx = threadIdx.x; idx < TM * 4; idx += blockDim.x) {
    const int q = idx / TM, m = idx - q * TM;
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < WC; ++c) v += red[c][m][q];
    Pt[(size_t)q * g.Np + m] = v;
  }
}

// W[s] = (c1, c2, e1, e2) from the column-tile row sums, tiles added in increasing order.
// Wf (when not null) receives c1 and c2 / 2 in FP32, the weights as the GEMM2 operand
// formation uses them.
// rho of slot sl at rho + sl rho_stride (0: one rho for every slot).
__global__ void k_weights_bf16(const double *__restrict__ P, const double *__restrict__ GD,
                               const double *__restrict__ rho, size_t rho_stride,
                               double *__restrict__ W, float *__restrict__ Wf, int Np, int n_ct,
                               int nq, int sb, int gd_stride, double kf, double kr, double ky) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int sl = blockIdx.y;
  if (i >= Np) return;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  const double *Ps = P + (size_t)sl * n_ct * nq * Np + i;
  for (int ct = 0; ct < n_ct; ++ct)
    for (int q = 0; q < nq; ++q) s[q] += Ps[((size_t)ct * nq + q) * Np];
  if (nq == 2 && GD) {
    s[2] = GD[(size_t)sl * Np + i];
    s[3] = GD[((size_t)gd_stride + sl) * Np + i];
  }
  const double r = rho[(size_t)sl * rho_stride + i];
  double *Ws = W + (size_t)sl * 4 * Np;
  const double c1 = r * (ky * s[0] / 8 - kr * s[2] / 3);
  const double c2 = r * (ky * s[1] / 4 - kr * s[3] / 3);
  Ws[i] = c1;
  Ws[(size_t)Np + i] = c2;
  Ws[2 * (size_t)Np + i] = -kf * r * s[0] / 6;
  Ws[3 * (size_t)Np + i] = -kf * r * s[1] / 3;
  if (Wf) {
    float *Wfs = Wf + (size_t)sl * 2 * Np;
    Wfs[i] = (float)c1;
    Wfs[(size_t)Np + i] = (float)(0.5 * c2);
  }
}

// ------------------------------------------------------------------------------ GEMM2
struct Gemm2Args {
  const uint16_t *Ah, *Al;    // X~^T planes [npd x Np]
  const float *YTf, *XtTf;    // [npd x Np]
  const double *W;            // [s_batch x 4 x Np]
  double *B;                  // [s_batch x npd x npd]
  int Np, npd, sb, n_mt, n_ct;
};

// G^T slice of TN rows (c0 + t/4 + (NT/4) h, h = 0, 1; k0 + 8 (t % 4) ..) with NT = 2 TN
// threads: fetched in FP32, formed and split when staged.
template <int NT>
__device__ inline void load_g(const Gemm2Args &g, int c0, int k0, float4 (&ry)[4],
                              float4 (&rx)[4]) {
  const int t = threadIdx.x, row = t >> 2, kc = (t & 3) * 8;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int r = c0 + row + (NT / 4) * h;
    if (r < g.npd) {
      const float4 *py = reinterpret_cast<const float4 *>(g.YTf + (size_t)r * g.Np + k0 + kc);
      const float4 *px = reinterpret_cast<const float4 *>(g.XtTf + (size_t)r * g.Np + k0 + kc);
      ry[2 * h] = py[0];
      ry[2 * h + 1] = py[1];
      rx[2 * h] = px[0];
      rx[2 * h + 1] = px[1];
    } else {
      ry[2 * h] = ry[2 * h + 1] = rx[2 * h] = rx[2 * h + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}

template <int NT>
__device__ inline void store_g(uint16_t *lh, uint16_t *ll, const float (&a)[8], const float (&b)[8],
                               const float4 (&ry)[4], const float4 (&rx)[4]) {
  const int t = threadIdx.x, row = t >> 2, kc = (t & 3) * 8;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const float y[8] = {ry[2 * h].x, ry[2 * h].y, ry[2 * h].z, ry[2 * h].w,
                        ry[2 * h + 1].x, ry[2 * h + 1].y, ry[2 * h + 1].z, ry[2 * h + 1].w};
    const float x[8] = {rx[2 * h].x, rx[2 * h].y, rx[2 * h].z, rx[2 * h].w,
                        rx[2 * h + 1].x, rx[2 * h + 1].y, rx[2 * h + 1].z, rx[2 * h + 1].w};
    uint32_t hi[4], lo[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      split2(fmaf(a[2 * j], y[2 * j], b[2 * j] * x[2 * j]),
             fmaf(a[2 * j + 1], y[2 * j + 1], b[2 * j + 1] * x[2 * j + 1]), hi[j], lo[j]);
    const int o = (row + (NT / 4) * h) * LDK + kc;
    *reinterpret_cast<uint4 *>(lh + o) = make_uint4(hi[0], hi[1], hi[2], hi[3]);
    *reinterpret_cast<uint4 *>(ll + o) = make_uint4(lo[0], lo[1], lo[2], lo[3]);
  }
}

__device__ inline void load_w(const double *c1, const double *c2, int k0, float (&a)[8],
                              float (&b)[8]) {
  const int kc = (threadIdx.x & 3) * 8;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    a[j] = (float)c1[k0 + kc + j];
    b[j] = (float)(0.5 * c2[k0 + kc + j]);
  }
}

template <int TM, int TN>
__global__ __launch_bounds__(2 * TN, 256 / TN) void k_gemm2_bf16x3(Gemm2Args g) {
  constexpr int NT = Tile<TM, TN>::NT, WC = Tile<TM, TN>::WC, WM = Tile<TM, TN>::WM;
  __shared__ Stage<TM, TN> sm;
  const int total = g.sb * g.n_mt * g.n_ct;
  const int id = xcd_tile(total);
  if (id >= total) return;
  const int sl = id / (g.n_mt * g.n_ct);
  const int rest = id - sl * g.n_mt * g.n_ct;
  const int mt = rest / g.n_ct, ct = rest - mt * g.n_ct;
  const int m0 = mt * TM, c0 = ct * TN;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w / WC, wc = w % WC, lr = l & 15, lg = l >> 4;
  const int mi = clamp_blocks(g.npd, m0 + wr * (TM / 2), WM);
  const int nj = clamp_blocks(g.npd, c0 + wc * 64, WN);
  const double *c1 = g.W + (size_t)sl * 4 * g.Np, *c2 = c1 + g.Np;

  v4f acc[WM][WN];
#pragma unroll
  for (int a = 0; a < WM; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};

  const int KT = g.Np / TK;
  uint4 rah[4 * TM / NT], ral[4 * TM / NT];
  float4 ry[4], rx[4];
  float wa[8], wb[8];
  load_rows<TM, NT>(g.Ah, g.Np, g.npd, m0, 0, rah);
  load_rows<TM, NT>(g.Al, g.Np, g.npd, m0, 0, ral);
  load_g<NT>(g, c0, 0, ry, rx);
  load_w(c1, c2, 0, wa, wb);
  for (int kt = 0; kt < KT; ++kt) {
    __syncthreads();
    store_rows<TM, NT>(sm.ah, rah);
    store_rows<TM, NT>(sm.al, ral);
    store_g<NT>(sm.bh, sm.bl, wa, wb, ry, rx);
    __syncthreads();
    if (kt + 1 < KT) {
      const int k0 = (kt + 1) * TK;
      load_rows<TM, NT>(g.Ah, g.Np, g.npd, m0, k0, rah);
      load_rows<TM, NT>(g.Al, g.Np, g.npd, m0, k0, ral);
      load_g<NT>(g, c0, k0, ry, rx);
      load_w(c1, c2, k0, wa, wb);
    }
    mma_slice<TM, TN>(sm, wr, wc, mi, nj, acc);
  }
  double *Bo = g.B + (size_t)sl * g.npd * g.npd;
#pragma unroll
  for (int ti = 0; ti < WM; ++ti)
#pragma unroll
    for (int tj = 0; tj < WN; ++tj)
      if (ti < mi && tj < nj) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = m0 + wr * (TM / 2) + ti * 16 + 4 * lg + r;
          const int c = c0 + wc * 64 + tj * 16 + lr;
          Bo[(size_t)m * g.npd + c] = (double)acc[ti][tj][r];
        }
      }
}

// ---------------------------------------------------- variant 2: direct global->LDS loads
// 256 x 256 tiles, 8 waves, two LDS stages of unpadded 64-byte rows (one k slice each). The
// operand planes are copied straight into the next stage with global_load_lds_dwordx4
// while the current stage is multiplied, so a slice costs one barrier and no register
// staging. A wave's copy instruction fills 1 KB = 16 rows contiguously (lane l -> row l/4,
// 16-byte slot l%4); the k chunk stored in slot j of row r is j ^ sw(r), sw(r) = (r >> 2) & 3,
// so the 16 rows read by one fragment load hit 16 distinct bank groups.
constexpr int DT = 256;                 // tile rows and columns
constexpr int DNT = 512;                // threads
constexpr int DROW = TK;                // bf16 per LDS row (64 bytes)

struct DStage {
  uint16_t ah[DT * DROW], al[DT * DROW], bh[DT * DROW], bl[DT * DROW];   // 64 KB
};

__device__ inline int swz(int row) { return (row >> 2) & 3; }

__device__ inline v8bf dfrag(const uint16_t *plane, int row, int kg) {
  return __builtin_bit_cast(
      v8bf, *reinterpret_cast<const uint4 *>(plane + row * DROW + ((kg ^ swz(row)) * 8)));
}

// This wave's share of one 256-row plane slice: groups w and w + 8 of 16 rows. Rows past
// `rows` read row rows - 1 instead (their products are never stored).
__device__ inline void dma_plane(const uint16_t *__restrict__ P, size_t ld, int rows, int r0,
                                 int k0, uint16_t *lds_plane) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int grp = w + 8 * h;
    const int row = grp * 16 + (l >> 2);
    int r = r0 + row;
    r = r < rows ? r : rows - 1;
    const uint16_t *src = P + (size_t)r * ld + k0 + (((l & 3) ^ swz(row)) * 8);
    __builtin_amdgcn_global_load_lds(
        (const void __attribute__((address_space(1))) *)src,
        (void __attribute__((address_space(3))) *)(lds_plane + grp * 16 * DROW), 16, 0, 0);
  }
}

__device__ inline void dma_mma(const DStage &s, int wr, int wc, v4f (&acc)[8][WN]) {
  const int l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
  v8bf bh[WN], bl[WN];
#pragma unroll
  for (int tj = 0; tj < WN; ++tj) {
    bh[tj] = dfrag(s.bh, wc * 64 + tj * 16 + lr, lg);
    bl[tj] = dfrag(s.bl, wc * 64 + tj * 16 + lr, lg);
  }
#pragma unroll
  for (int ti = 0; ti < 8; ++ti) {
    const v8bf ah = dfrag(s.ah, wr * 128 + ti * 16 + lr, lg);
    const v8bf al = dfrag(s.al, wr * 128 + ti * 16 + lr, lg);
#pragma unroll
    for (int tj = 0; tj < WN; ++tj) {
      acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[tj], acc[ti][tj], 0, 0, 0);
      acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[tj], acc[ti][tj], 0, 0, 0);
      acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[tj], acc[ti][tj], 0, 0, 0);
    }
  }
}

// every copy into LDS issued by this wave has landed, then all waves meet
__device__ inline void dma_sync() {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
}

__global__ __launch_bounds__(DNT, 1) void k_gemm1_bf16x3_dma(Gemm1Args g) {
  __shared__ DStage st[2];
  const int total = g.sb * g.n_it * g.n_ct;
  const int id = xcd_tile(total);
  if (id >= total) return;
  const SlotTile tt = slot_group_tile(id, g.sb, g.sg, g.n_it, g.n_ct);
  const int sl = tt.sl, it = tt.it, ct = tt.ct;
  const int i0 = it * DT, c0 = ct * DT;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w >> 2, wc = w & 3, lr = l & 15, lg = l >> 4;
  const int nj = clamp_blocks(g.npd, c0 + wc * 64, WN);
  const size_t plane_off = (size_t)sl * g.npd * g.npd;
  const uint16_t *Yh = g.Yh + plane_off, *Yl = g.Yl + plane_off;

  v4f acc[8][WN];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};

  if (g.do_q) {
    const int KT = g.npd / TK;
    dma_plane(g.Xh, g.npd, g.Np, i0, 0, st[0].ah);
    dma_plane(g.Xl, g.npd, g.Np, i0, 0, st[0].al);
    dma_plane(Yh, g.npd, g.npd, c0, 0, st[0].bh);
    dma_plane(Yl, g.npd, g.npd, c0, 0, st[0].bl);
    dma_sync();
    for (int kt = 0; kt < KT; ++kt) {
      const int cur = kt & 1;
      if (kt + 1 < KT) {
        const int k0 = (kt + 1) * TK;
        dma_plane(g.Xh, g.npd, g.Np, i0, k0, st[cur ^ 1].ah);
        dma_plane(g.Xl, g.npd, g.Np, i0, k0, st[cur ^ 1].al);
        dma_plane(Yh, g.npd, g.npd, c0, k0, st[cur ^ 1].bh);
        dma_plane(Yl, g.npd, g.npd, c0, k0, st[cur ^ 1].bl);
      }
      if (nj > 0) dma_mma(st[cur], wr, wc, acc);
      dma_sync();
    }
  }

  // epilogue as in k_gemm1_bf16x3, the row sums staged in the (now free) LDS stages
  double (*red)[DT][4] = reinterpret_cast<double (*)[DT][4]>(&st[0]);   // [4][256][4]: 32 KB
  const double *g1 = g.G + (size_t)sl * 2 * g.npd, *g2 = g1 + g.npd;
#pragma unroll
  for (int ti = 0; ti < 8; ++ti) {
    double p[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) p[r][0] = p[r][1] = p[r][2] = p[r][3] = 0.0;
#pragma unroll
    for (int tj = 0; tj < WN; ++tj) {
      const int c = c0 + wc * 64 + tj * 16 + lr;
      if (tj < nj) {
        const double gg1 = g1[c], gg2 = g2[c];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = wr * 128 + ti * 16 + 4 * lg + r;
          const size_t o = (size_t)(i0 + m) * g.npd + c;
          const double x = g.Xf[o], ft = g.Ftf[o];
          const double q = (double)acc[ti][tj][r];
          p[r][0] = fma(x, q, p[r][0]);
          p[r][1] = fma(ft, q, p[r][1]);
          p[r][2] = fma(x, gg1, p[r][2]);
          p[r][3] = fma(ft, gg2, p[r][3]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double v = p[r][q];
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 64);
        if (lr == 0) red[wc][wr * 128 + ti * 16 + 4 * lg + r][q] = v;
      }
  }
  __syncthreads();
  double *Pt = g.P + ((size_t)sl * g.n_ct + ct) * 4 * g.Np + i0;
  for (int idx = threadIdx.x; idx < DT * 4; idx += blockDim.x) {
    const int q = idx / DT, m = idx - q * DT;
    Pt[(size_t)q * g.Np + m] = red[0][m][q] + red[1][m][q] + red[2][m][q] + red[3][m][q];
  }
}

// B operand (G^T, formed) of a 256-column tile with 512 threads: rows c0 + t/4 + 128 h,
// stored into the swizzled layout.
__device__ inline void dstore_g(uint16_t *lh, uint16_t *ll, const float (&a)[8],
                                const float (&b)[8], const float4 (&ry)[4],
                                const float4 (&rx)[4]) {
  const int t = threadIdx.x, row0 = t >> 2, j = t & 3;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const float y[8] = {ry[2 * h].x, ry[2 * h].y, ry[2 * h].z, ry[2 * h].w,
                        ry[2 * h + 1].x, ry[2 * h + 1].y, ry[2 * h + 1].z, ry[2 * h + 1].w};
    const float x[8] = {rx[2 * h].x, rx[2 * h].y, rx[2 * h].z, rx[2 * h].w,
                        rx[2 * h + 1].x, rx[2 * h + 1].y, rx[2 * h + 1].z, rx[2 * h + 1].w};
    uint32_t hi[4], lo[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      split2(fmaf(a[2 * q], y[2 * q], b[2 * q] * x[2 * q]),
             fmaf(a[2 * q + 1], y[2 * q + 1], b[2 * q + 1] * x[2 * q + 1]), hi[q], lo[q]);
    const int row = row0 + 128 * h;
    const int o = row * DROW + ((j ^ swz(row)) * 8);
    *reinterpret_cast<uint4 *>(lh + o) = make_uint4(hi[0], hi[1], hi[2], hi[3]);
    *reinterpret_cast<uint4 *>(ll + o) = make_uint4(lo[0], lo[1], lo[2], lo[3]);
  }
}

__global__ __launch_bounds__(DNT, 1) void k_gemm2_bf16x3_dma(Gemm2Args g) {
  __shared__ DStage st[2];
  const int total = g.sb * g.n_mt * g.n_ct;
  const int id = xcd_tile(total);
  if (id >= total) return;
  const int sl = id / (g.n_mt * g.n_ct);
  const int rest = id - sl * g.n_mt * g.n_ct;
  const int mt = rest / g.n_ct, ct = rest - mt * g.n_ct;
  const int m0 = mt * DT, c0 = ct * DT;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w >> 2, wc = w & 3, lr = l & 15, lg = l >> 4;
  const int mi = clamp_blocks(g.npd, m0 + wr * 128, 8);
  const int nj = clamp_blocks(g.npd, c0 + wc * 64, WN);
  const double *c1 = g.W + (size_t)sl * 4 * g.Np, *c2 = c1 + g.Np;

  v4f acc[8][WN];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < WN; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};

  const int KT = g.Np / TK;
  float4 ry[4], rx[4];
  float wa[8], wb[8];
  dma_plane(g.Ah, g.Np, g.npd, m0, 0, st[0].ah);
  dma_plane(g.Al, g.Np, g.npd, m0, 0, st[0].al);
  load_g<DNT>(g, c0, 0, ry, rx);
  load_w(c1, c2, 0, wa, wb);
  dstore_g(st[0].bh, st[0].bl, wa, wb, ry, rx);
  dma_sync();
  for (int kt = 0; kt < KT; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < KT;
    if (more) {
      const int k0 = (kt + 1) * TK;
      dma_plane(g.Ah, g.Np, g.npd, m0, k0, st[cur ^ 1].ah);
      dma_plane(g.Al, g.Np, g.npd, m0, k0, st[cur ^ 1].al);
      load_g<DNT>(g, c0, k0, ry, rx);
      load_w(c1, c2, k0, wa, wb);
    }
    if (mi > 0 && nj > 0) dma_mma(st[cur], wr, wc, acc);
    if (more) dstore_g(st[cur ^ 1].bh, st[cur ^ 1].bl, wa, wb, ry, rx);
    dma_sync();
  }
  double *Bo = g.B + (size_t)sl * g.npd * g.npd;
#pragma unroll
  for (int ti = 0; ti < 8; ++ti)
#pragma unroll
    for (int tj = 0; tj < WN; ++tj)
      if (ti < mi && tj < nj) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = m0 + wr * 128 + ti * 16 + 4 * lg + r;
          const int c = c0 + wc * 64 + tj * 16 + lr;
          Bo[(size_t)m * g.npd + c] = (double)acc[ti][tj][r];
        }
      }
}

void gemm1_launch_dma(const Ensemble &e, int sb, int what, hipStream_t st) {
  Gemm1Args g;
  const size_t plane = (size_t)e.s_batch * e.npd * e.npd;
  g.Xh = e.X_hi; g.Xl = e.X_lo;
  g.Yh = e.Ys_bf; g.Yl = e.Ys_bf + plane;
  g.Xf = e.X_f; g.Ftf = e.Ft_f; g.G = e.G; g.P = e.Pw;
  g.Np = e.Np; g.npd = e.npd; g.sb = sb;
  g.n_it = e.Np / DT; g.n_ct = (e.npd + DT - 1) / DT;
  g.sg = sb < kG1SlotGroup ? sb : kG1SlotGroup;
  g.do_q = (what & (kTermF | kTermD2vY)) ? 1 : 0;
  const int total = sb * g.n_it * g.n_ct;
  hipLaunchKernelGGL(k_gemm1_bf16x3_dma, dim3(8 * ((total + 7) / 8)), dim3(DNT), 0, st, g);
  launch_weights_bf16(e, sb, g.n_ct, 4, what, st);
}

void gemm2_launch_dma(const Ensemble &e, int sb, hipStream_t st) {
  Gemm2Args g;
  g.Ah = e.XtT_hi; g.Al = e.XtT_lo;
  g.YTf = e.YT_f; g.XtTf = e.XtT_f; g.W = e.W; g.B = e.B;
  g.Np = e.Np; g.npd = e.npd; g.sb = sb;
  g.n_mt = (e.npd + DT - 1) / DT; g.n_ct = (e.npd + DT - 1) / DT;
  const int total = sb * g.n_mt * g.n_ct;
  hipLaunchKernelGGL(k_gemm2_bf16x3_dma, dim3(8 * ((total + 7) / 8)), dim3(DNT), 0, st, g);
}

template <int TM, int TN>
void gemm1_launch(const Ensemble &e, int sb, int what, hipStream_t st) {
  Gemm1Args g;
  const size_t plane = (size_t)e.s_batch * e.npd * e.npd;
  g.Xh = e.X_hi; g.Xl = e.X_lo;
  g.Yh = e.Ys_bf; g.Yl = e.Ys_bf + plane;
  g.Xf = e.X_f; g.Ftf = e.Ft_f; g.G = e.G; g.P = e.Pw;
  g.Np = e.Np; g.npd = e.npd; g.sb = sb;
  g.n_it = e.Np / TM; g.n_ct = (e.npd + TN - 1) / TN;
  g.sg = sb < kG1SlotGroup ? sb : kG1SlotGroup;
  g.do_q = (what & (kTermF | kTermD2vY)) ? 1 : 0;
  const int total = sb * g.n_it * g.n_ct;
  hipLaunchKernelGGL((k_gemm1_bf16x3<TM, TN>), dim3(8 * ((total + 7) / 8)), dim3(2 * TN), 0, st,
                     g);
  launch_weights_bf16(e, sb, g.n_ct, 4, what, st);
}

template <int TM, int TN>
void gemm2_launch(const Ensemble &e, int sb, hipStream_t st) {
  Gemm2Args g;
  g.Ah = e.XtT_hi; g.Al = e.XtT_lo;
  g.YTf = e.YT_f; g.XtTf = e.XtT_f; g.W = e.W; g.B = e.B;
  g.Np = e.Np; g.npd = e.npd; g.sb = sb;
  g.n_mt = (e.npd + TM - 1) / TM; g.n_ct = (e.npd + TN - 1) / TN;
  const int total = sb * g.n_mt * g.n_ct;
  hipLaunchKernelGGL((k_gemm2_bf16x3<TM, TN>), dim3(8 * ((total + 7) / 8)), dim3(2 * TN), 0, st,
                     g);
}

}  // namespace

int bf16_col_tiles(const Ensemble &e) { return (e.npd + kBfTileN - 1) / kBfTileN; }

void launch_weights_bf16(const Ensemble &e, int sb, int n_ct, int nq, int what, hipStream_t st) {
  const double kf = (what & kTermF) ? 1.0 : 0.0;
  const double kr = (what & kTermD2vR) ? 1.0 : 0.0;
  const double ky = (what & kTermD2vY) ? 1.0 : 0.0;
  hipLaunchKernelGGL(k_weights_bf16, dim3((e.Np + 255) / 256, sb), dim3(256), 0, st, e.Pw,
                     kr != 0.0 ? e.GD : nullptr, e.rho_w(), e.rho_stride(), e.W, e.Wf, e.Np,
                     n_ct, nq, sb, e.s_batch, kf, kr, ky);
}

void launch_gemm1_bf16x3(const Ensemble &e, int sb, int what, hipStream_t st) {
  if (e.bf_variant >= 3)
    launch_gemm1_bf16x3_rp(e, sb, what, e.bf_variant == 3 ? 2 : 1, st);   // 4-8: 96 columns
  else if (e.bf_variant == 2)
    gemm1_launch_dma(e, sb, what, st);
  else if (e.bf_variant == 1)
    gemm1_launch<256, 256>(e, sb, what, st);
  else
    gemm1_launch<256, 128>(e, sb, what, st);
}

void launch_gemm2_bf16x3(const Ensemble &e, int sb, hipStream_t st) {
  // forms 3-5 write B_s in FP32 for the fused back-rotation; the one-sided rotations (blocks
  // wider than the group tile) read the FP64 B, which form 0 writes
  if (e.bf_variant == 8 && e.fused_rot)
    launch_gemm2_bf16x3_rp(e, sb, 8, st, 0);                          // 448 x 96, 8 waves
  else if (e.bf_variant == 7 && e.fused_rot)
    launch_gemm2_bf16x3_rp(e, sb, 8, st, 4);                          // 512 x 96, 8 waves
  else if (e.bf_variant == 6 && e.fused_rot)
    launch_gemm2_bf16x3_rp(e, sb, 1, st, 4);                          // 256 x 96
  else if (e.bf_variant >= 3 && e.fused_rot)
    launch_gemm2_bf16x3_rp(e, sb, e.bf_variant == 4 ? 1 : 2, st);   // 3, 5: 192 columns
  else if (e.bf_variant >= 3)
    gemm2_launch<192, 128>(e, sb, st);
  else if (e.bf_variant == 2)
    gemm2_launch_dma(e, sb, st);
  else if (e.bf_variant == 1)
    gemm2_launch<256, 256>(e, sb, st);
  else
    gemm2_launch<192, 128>(e, sb, st);
}

}  // namespace tdscha
