// BF16x3 contractions with register-resident A operands (gfx950): tile form 3 of the
// autotuner.
//
// A workgroup tile is TM x 96 WCN with 4 row waves x WCN column waves (WCN = 2: 512 threads,
// one workgroup per CU, tile form 3; WCN = 1: 256 threads, two independent workgroups per CU,
// tile form 4). Wave (wr, wc) owns rows [wr WR, (wr + 1) WR) and columns [96 wc, 96 wc + 96):
// WR / 16 x 6 blocks of 16 x 16, FP32 accumulators in VGPRs (two waves per SIMD). A fragments go from
// global memory straight into registers in the MFMA operand layout and never touch LDS (the
// column waves of a row load the same rows; the second load hits L2). Only the B slice
// (96 WCN rows x 32 k, hi and lo planes) is shared through LDS: two stages of unpadded
// 64-byte rows, 16-byte chunk j of row r stored at chunk j ^ sw(r),
// sw(r) = ((r >> 1) ^ (r >> 3)) & 3, so that the fragment reads (16 rows, one chunk each,
// per 16 lanes) and the staging writes (4 rows x 4 chunks, or 2 rows x 8 half chunks, per 16
// lanes) are free of bank conflicts for both 8- and 16-lane LDS phases.
//
// One barrier per k slice. While slice k is multiplied, B slice k + 1 (in registers since
// the previous slice) is written to the other stage halfway through, and B slice k + 2 and A
// slice k + 2 are in flight. The loop is unrolled by two so that every register buffer has a
// fixed name (both k ranges hold an even number of slices), loads past the last slice re-read
// the last slice instead of branching, and fragment reads are software-pipelined one column
// block ahead of the MFMAs with scheduling fences that keep two blocks of fragments live. Per
// flop, LDS traffic is about half that of tile forms 0-2 (no A staging, no A fragment reads),
// which the counters showed to be their limit.
//
// GEMM1 (WR = 64, TM = 256): Q = X Ys_s with the FP64 row sums of the tile (alpha, beta,
//   gamma, delta) as k_gemm1_bf16x3; the two column waves of a row add theirs through LDS.
// GEMM2 (WR = 48, TM = 192): B_s = X~^T G_s (stored in FP32, as accumulated) with the B operand
//   G_s^T[c][k] = c1_k Y^T[c][k] + c2_k/2 X~^T[c][k] formed once per workgroup from the FP32
//   operand copies while it is staged, with the arithmetic of k_gemm2_bf16x3.
#include "tdscha_internal.h"
#include "bf16x3.h"

namespace tdscha {

using namespace bfx;

namespace {

constexpr int RWC = 96;           // columns per wave
constexpr int RNB = RWC / 16;     // 16-column blocks per wave

// WCN column waves: tile columns, threads (4 row waves x WCN column waves), threads per
// staged plane and rows per staging pass (three passes cover the tile's rows).
template <int WCN> struct RCfg {
  static constexpr int TN = RWC * WCN;
  static constexpr int NT = 256 * WCN;
  static constexpr int HALF = NT / 2;
  static constexpr int RPP = NT / 8;
};

// Native vector registers (HIP's uint4 / float4 structs are copied with memcpy, which keeps
// them in scratch memory).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

template <int TN> struct RStage {
  uint16_t bh[TN * TK], bl[TN * TK];     // 2 x 6 KB per 96 columns
};

__device__ inline int rswz(int row) { return ((row >> 1) ^ (row >> 3)) & 3; }

// ---- B slice from bf16 planes: the first half of the threads stage the hi plane, the second
// half the lo plane, each thread rows (t % HALF) / 4 + RPP h (h < 3), 16-byte chunk t % 4.
struct BSlice {
  u32x4 c[3];
};

template <int WCN>
__device__ inline void rload_b(const uint16_t *__restrict__ Ph, const uint16_t *__restrict__ Pl,
                               size_t ld, int rows, int r0, int k0, BSlice &b) {
  using C = RCfg<WCN>;
  const int t = threadIdx.x, tt = t % C::HALF;
  const uint16_t *P = t < C::HALF ? Ph : Pl;
#pragma unroll
  for (int h = 0; h < 3; ++h) {
    int r = r0 + (tt >> 2) + C::RPP * h;
    r = r < rows ? r : rows - 1;   // rows past the edge feed columns that are never used
    b.c[h] = *reinterpret_cast<const u32x4 *>(P + (size_t)r * ld + k0 + (tt & 3) * 8);
  }
}

template <int WCN>
__device__ inline void rstore_b(RStage<RCfg<WCN>::TN> &s, const BSlice &b) {
  using C = RCfg<WCN>;
  const int t = threadIdx.x, tt = t % C::HALF;
  uint16_t *dst = t < C::HALF ? s.bh : s.bl;
#pragma unroll
  for (int h = 0; h < 3; ++h) {
    const int row = (tt >> 2) + C::RPP * h;
    *reinterpret_cast<u32x4 *>(dst + row * TK + (((tt & 3) ^ rswz(row)) * 8)) = b.c[h];
  }
}

// ---- A operand from fragment-tiled bf16 planes (k_tile_bf16x3): tile (rb, ks) holds rows
// 16 rb .. + 16 and k 32 ks .. + 32 as 64 lane chunks of 8 values in the MFMA operand layout
// (lane l: row 16 rb + l % 16, k = 32 ks + 8 (l / 16) .. + 8), so a wave's load of one block
// is 1 KB of consecutive 16-byte lane chunks, which the address unit coalesces (the same
// rows read from the row-major planes cost one address cycle per lane). Row blocks past
// n_rb read the last block (their products are never stored).
template <int MB> struct AFrag {
  u32x4 h[MB], l[MB];
};

template <int MB>
__device__ inline void rload_a(const uint16_t *__restrict__ Ph, const uint16_t *__restrict__ Pl,
                               int n_ks, int n_rb, int rb0, int ks, AFrag<MB> &a) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int ti = 0; ti < MB; ++ti) {
    int rb = rb0 + ti;
    rb = rb < n_rb ? rb : n_rb - 1;
    const size_t o = (((size_t)rb * n_ks + ks) * 64 + l) * 8;
    a.h[ti] = *reinterpret_cast<const u32x4 *>(Ph + o);
    a.l[ti] = *reinterpret_cast<const u32x4 *>(Pl + o);
  }
}

// fp64 [rows x cols] row-major -> fragment-tiled bf16 hi plane at dst and lo plane at
// dst + rows * cols (rows % 16 == 0, cols % 32 == 0), with k_convert's split.
__global__ void k_tile_bf16x3(const double *__restrict__ src, int rows, int cols,
                              uint16_t *__restrict__ dst) {
  const int c8 = cols / 8;
  const size_t n8 = (size_t)rows * c8, plane = (size_t)rows * cols;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n8; i += stride) {
    const int row = (int)(i / c8), k = (int)(i % c8) * 8;
    const int lane = (row & 15) + 16 * ((k & 31) >> 3);
    const size_t o = (((size_t)(row >> 4) * (cols >> 5) + (k >> 5)) * 64 + lane) * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float f = (float)src[(size_t)row * cols + k + j];
      const uint16_t h = bf16_rne(f);
      dst[o + j] = h;
      dst[plane + o + j] = bf16_rne(f - __uint_as_float((uint32_t)h << 16));
    }
  }
}

// FP32 [rows x cols] row-major -> accumulator-tiled copy (rows % 16 == 0, cols % 16 == 0):
// 16 x 16 block (rb, cb) holds, at lane l = lr + 16 lg of a wave, the four values of rows
// 16 rb + 4 lg .. + 4 in column 16 cb + lr as one 16-byte chunk, so that the row-sum operands
// of one accumulator block are a single 1 KB wave load (in the row-major copy they are four
// scattered 4-byte loads per lane).
__global__ void k_tile_acc_f32(const float *__restrict__ src, int rows, int cols,
                               float *__restrict__ dst) {
  const int ncb = cols / 16;
  const size_t n4 = (size_t)rows * cols / 4;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    const int lane = (int)(i & 63);
    const size_t blk = i >> 6;
    const int rb = (int)(blk / ncb), cb = (int)(blk % ncb);
    const size_t o = (size_t)(16 * rb + 4 * (lane >> 4)) * cols + 16 * cb + (lane & 15);
    v4f v;
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = src[o + (size_t)r * cols];
    *reinterpret_cast<v4f *>(dst + 4 * i) = v;
  }
}

// ---- fragments and MFMAs
struct BFrag {
  v8bf h, l;
};

// B fragment of column block TJ of this wave: row 96 wc + 16 TJ + lr, k group lg, stored in
// chunk lg ^ sw(lr) ^ 2 (TJ & 1) of that row (96 wc is a multiple of 32). `base` holds the
// offsets of the two parities for this lane.
template <int TJ, int TN>
__device__ inline BFrag rfrag(const RStage<TN> &s, int2 base) {
  const int o = TJ * 16 * TK + ((TJ & 1) ? base.y : base.x);
  return {__builtin_bit_cast(v8bf, *reinterpret_cast<const u32x4 *>(s.bh + o)),
          __builtin_bit_cast(v8bf, *reinterpret_cast<const u32x4 *>(s.bl + o))};
}

__device__ inline int2 rfrag_base() {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
  const int row = (w >> 2) * RWC + lr;
  return make_int2(row * TK + ((lg ^ rswz(lr)) * 8), row * TK + ((lg ^ rswz(lr) ^ 2) * 8));
}

// 3 MFMAs per 16 x 16 block of column block TJ (a_lo b_hi, a_hi b_lo, a_hi b_hi).
template <int MB, int TJ>
__device__ inline void rmma_block(const AFrag<MB> &a, const BFrag &b, v4f (&acc)[MB][RNB]) {
#pragma unroll
  for (int ti = 0; ti < MB; ++ti) {
    const v8bf ah = __builtin_bit_cast(v8bf, a.h[ti]);
    const v8bf al = __builtin_bit_cast(v8bf, a.l[ti]);
    acc[ti][TJ] = mfma_bf16(al, b.h, acc[ti][TJ]);
    acc[ti][TJ] = mfma_bf16(ah, b.l, acc[ti][TJ]);
    acc[ti][TJ] = mfma_bf16(ah, b.h, acc[ti][TJ]);
  }
}

// Column blocks [TJ, J1), software-pipelined: the fragments of block TJ + 1 are read before
// the MFMAs of block TJ are issued (the group barriers pin that order: left alone, the
// scheduler sinks the reads below most of the MFMAs that should cover their latency), and a
// scheduling fence after each block keeps further reads from being hoisted.
template <int MB, int TJ, int J1, int TN>
__device__ inline void rmma_from(const RStage<TN> &s, const AFrag<MB> &a, const BFrag &cur,
                                 int2 base, v4f (&acc)[MB][RNB]) {
  if constexpr (TJ + 1 < J1) {
    const BFrag nxt = rfrag<TJ + 1>(s, base);
    rmma_block<MB, TJ>(a, cur, acc);
    __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);        // the two DS reads
    __builtin_amdgcn_sched_group_barrier(0x008, 3 * MB, 0);   // then the MFMAs
    __builtin_amdgcn_sched_barrier(0);
    rmma_from<MB, TJ + 1, J1, TN>(s, a, nxt, base, acc);
  } else {
    rmma_block<MB, TJ>(a, cur, acc);
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int MB, int J0, int J1, int TN>
__device__ inline void rmma(const RStage<TN> &s, const AFrag<MB> &a, int2 base,
                            v4f (&acc)[MB][RNB]) {
  const BFrag first = rfrag<J0>(s, base);
  __builtin_amdgcn_sched_barrier(0);
  rmma_from<MB, J0, J1, TN>(s, a, first, base, acc);
}

// Which halves of a wave's column blocks multiply data (wave-uniform): none when the wave's
// rows all lie in the zero padding past the last configuration (GEMM1) or mode (GEMM2), the
// first three blocks when the wave's last three columns blocks lie past the last mode. The
// skipped products are those of zero padding: every stored value stays as it is.
struct Halves {
  bool h0, h1;
};
__device__ inline Halves live_halves(int mi, int nj) {
  return Halves{mi > 0 && nj > 0, mi > 0 && nj > RNB / 2};
}

// ------------------------------------------------------------------------------ GEMM1
struct Gemm1RArgs {
  const uint16_t *Xh, *Xl;    // [Np x npd] fragment-tiled
  const uint16_t *Yh, *Yl;    // [s_batch x npd x npd] planes of the rotated Y1
  const float *Xa, *Fta;      // [Np x npd] epilogue operands in FP32, accumulator-tiled
  double *P;                  // [sb x n_ct x 2 x Np] row sums per column tile
  int Np, npd, sb, n_it, n_ct, sg, do_q;   // sg: image slots per tile group
  int N, n;                   // configurations and modes (the rest is zero padding)
};

// Row sums of block row ti of this wave over its columns below nj: alpha = sum X Q and
// beta = sum Ft Q, each lane's six products added in FP32 (Q itself carries the BF16x3
// error), then reduced over the 16 lanes of each row group in FP64 with DPP. gamma and delta
// do not involve Q and come from an FP64 product outside this kernel. The twelve operand
// chunks of the block row are loaded first, unconditionally (blocks past the last column
// read the last block and are not added), so that their latencies overlap.
template <int MB>
__device__ inline void row_sums(const Gemm1RArgs &g, int ti, int row0, int cw, int nj,
                                const v4f (&acc)[MB][RNB], double (&p)[4][2]) {
  const int l = threadIdx.x & 63;
  const int ncb = g.npd / 16, rb = (row0 >> 4) + ti, cb0 = cw >> 4;
  v4f xa[RNB], fa[RNB];
#pragma unroll
  for (int tj = 0; tj < RNB; ++tj) {
    const int cb = min(cb0 + tj, ncb - 1);
    const size_t o = (((size_t)rb * ncb + cb) * 64 + l) * 4;
    xa[tj] = *reinterpret_cast<const v4f *>(g.Xa + o);
    fa[tj] = *reinterpret_cast<const v4f *>(g.Fta + o);
  }
  // blocks at or past nj add x * 0 (the sums keep their values; no branch keeps the loads
  // above from being sunk into one)
  float pa[4] = {0.f, 0.f, 0.f, 0.f}, pb[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int tj = 0; tj < RNB; ++tj) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float q = tj < nj ? acc[ti][tj][r] : 0.f;
      pa[r] = fmaf(xa[tj][r], q, pa[r]);
      pb[r] = fmaf(fa[tj][r], q, pb[r]);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    p[r][0] = row16_sum((double)pa[r]);
    p[r][1] = row16_sum((double)pb[r]);
  }
}

template <int WCN>
__global__ __launch_bounds__(RCfg<WCN>::NT, 2 / WCN) void k_gemm1_bf16x3_rp(Gemm1RArgs g) {
  constexpr int MB = 4, TM = 256, TN = RCfg<WCN>::TN;
  __shared__ RStage<TN> st[2];
  const int total = g.sb * g.n_it * g.n_ct;
  const int id = xcd_tile(total);
  if (id >= total) return;
  // the FP32 epilogue panel of a (row tile, column tile) is the same for every slot
  const SlotTile tt = slot_group_tile(id, g.sb, g.sg, g.n_it, g.n_ct);
  const int sl = tt.sl, it = tt.it, ct = tt.ct;
  const int i0 = it * TM, c0 = ct * TN;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
  const int wr = w & 3, wc = w >> 2;
  const int cw = c0 + wc * RWC;                 // first column of this wave
  const int nj = clamp_blocks(g.n, cw, RNB);
  const int row0 = i0 + wr * 64;
  const Halves hv = live_halves(clamp_blocks(g.N, row0, MB), nj);

  v4f acc[MB][RNB];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < RNB; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};

  if (g.do_q) {
    const size_t plane_off = (size_t)sl * g.npd * g.npd;
    const uint16_t *Yh = g.Yh + plane_off, *Yl = g.Yl + plane_off;
    const int KT = g.npd / TK, last = (KT - 1) * TK;   // KT is even (npd % 64 == 0)
    const int2 base = rfrag_base();
    const int n_ks = KT, n_rb = g.Np / 16, rb0 = row0 / 16;
    BSlice b0, b1;
    AFrag<MB> a0, a1;
    rload_b<WCN>(Yh, Yl, g.npd, g.npd, c0, 0, b0);
    rload_a<MB>(g.Xh, g.Xl, n_ks, n_rb, rb0, 0, a0);
    rload_b<WCN>(Yh, Yl, g.npd, g.npd, c0, TK, b1);
    rload_a<MB>(g.Xh, g.Xl, n_ks, n_rb, rb0, 1, a1);
    rstore_b<WCN>(st[0], b0);
    __syncthreads();
    for (int kt = 0; kt < KT; kt += 2) {
      // stage 0: slice kt; b1: slice kt + 1; a0: slice kt; a1: slice kt + 1
      rload_b<WCN>(Yh, Yl, g.npd, g.npd, c0, min((kt + 2) * TK, last), b0);
      if (hv.h0) rmma<MB, 0, RNB / 2>(st[0], a0, base, acc);
      rstore_b<WCN>(st[1], b1);
      if (hv.h1) rmma<MB, RNB / 2, RNB>(st[0], a0, base, acc);
      rload_a<MB>(g.Xh, g.Xl, n_ks, n_rb, rb0, min(kt + 2, KT - 1), a0);
      __syncthreads();
      // stage 1: slice kt + 1; b0: slice kt + 2; a1: slice kt + 1; a0: slice kt + 2
      rload_b<WCN>(Yh, Yl, g.npd, g.npd, c0, min((kt + 3) * TK, last), b1);
      if (hv.h0) rmma<MB, 0, RNB / 2>(st[1], a1, base, acc);
      rstore_b<WCN>(st[0], b0);
      if (hv.h1) rmma<MB, RNB / 2, RNB>(st[1], a1, base, acc);
      rload_a<MB>(g.Xh, g.Xl, n_ks, n_rb, rb0, min(kt + 3, KT - 1), a1);
      __syncthreads();
    }
  }

  // epilogue: row sums over this wave's 96 columns, one 16-row block at a time (a 16-lane
  // DPP reduction adds the 16 columns of a block). With two column waves, both leave their
  // sums in LDS (the stages are free: the loop ended on a barrier) and the workgroup adds the
  // two halves, wave 0's first, in coalesced stores of the tile's row sums.
  double *Pt = g.P + ((size_t)sl * g.n_ct + ct) * 2 * g.Np;
  if constexpr (WCN == 1) {
#pragma unroll
    for (int ti = 0; ti < MB; ++ti) {
      double p[4][2];
      row_sums<MB>(g, ti, row0, cw, nj, acc, p);
      if (lr == 0)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int q = 0; q < 2; ++q)
            Pt[(size_t)q * g.Np + row0 + ti * 16 + 4 * lg + r] = p[r][q];
    }
  } else {
    double(*red)[TM][2] = reinterpret_cast<double(*)[TM][2]>(&st[0]);   // [2][256][2]: 8 KB
    static_assert(sizeof(double) * 2 * TM * 2 <= sizeof(st), "row-sum halves exceed the stages");
#pragma unroll
    for (int ti = 0; ti < MB; ++ti) {
      double p[4][2];
      row_sums<MB>(g, ti, row0, cw, nj, acc, p);
      if (lr == 0)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int q = 0; q < 2; ++q) red[wc][wr * 64 + ti * 16 + 4 * lg + r][q] = p[r][q];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 2 * TM; t += RCfg<WCN>::NT) {
      const int q = t / TM, m = t - q * TM;
      Pt[(size_t)q * g.Np + i0 + m] = red[0][m][q] + red[1][m][q];
    }
  }
}

// ------------------------------------------------------------------------------ GEMM2
struct Gemm2RArgs {
  const uint16_t *Ah, *Al;    // X~^T planes [npd x Np] fragment-tiled
  const float *YTf, *XtTf;    // [npd x Np] sources of the formed operand
  const float *Wf;            // [s_batch x 2 x Np]  c1, c2 / 2
  float *B;                   // [s_batch x npd x npd]
  int Np, npd, sb, n_mt, n_ct;
  int KT;                     // k slices that hold configurations (even; the rest is padding)
  int n;                      // modes (rows and columns past it are zero padding)
};

// Sources of the formed B slice: rows (t / 8) + RPP h (h < 3) of the tile,
// k = 4 (t % 8) .. + 4.
struct GSrc {
  v4f y[3], x[3], wa, wb;
};

template <int WCN>
__device__ inline void rload_g(const Gemm2RArgs &g, const float *__restrict__ wf, int c0, int k0,
                               GSrc &s) {
  const int t = threadIdx.x, kk = k0 + (t & 7) * 4;
#pragma unroll
  for (int h = 0; h < 3; ++h) {
    int r = c0 + (t >> 3) + RCfg<WCN>::RPP * h;
    r = r < g.npd ? r : g.npd - 1;
    s.y[h] = *reinterpret_cast<const v4f *>(g.YTf + (size_t)r * g.Np + kk);
    s.x[h] = *reinterpret_cast<const v4f *>(g.XtTf + (size_t)r * g.Np + kk);
  }
  s.wa = *reinterpret_cast<const v4f *>(wf + kk);
  s.wb = *reinterpret_cast<const v4f *>(wf + g.Np + kk);
}

// G^T = c1 Y^T + c2/2 X~^T in FP32, split into hi / lo and written as 8-byte halves of the
// swizzled 16-byte chunks.
template <int WCN>
__device__ inline void rstore_g(RStage<RCfg<WCN>::TN> &st, const GSrc &s) {
  const int t = threadIdx.x, q = t & 7;
#pragma unroll
  for (int h = 0; h < 3; ++h) {
    const int row = (t >> 3) + RCfg<WCN>::RPP * h;
    uint32_t hi0, lo0, hi1, lo1;
    split2(fmaf(s.wa.x, s.y[h].x, s.wb.x * s.x[h].x), fmaf(s.wa.y, s.y[h].y, s.wb.y * s.x[h].y),
           hi0, lo0);
    split2(fmaf(s.wa.z, s.y[h].z, s.wb.z * s.x[h].z), fmaf(s.wa.w, s.y[h].w, s.wb.w * s.x[h].w),
           hi1, lo1);
    const int o = row * TK + (((q >> 1) ^ rswz(row)) * 8) + (q & 1) * 4;
    *reinterpret_cast<u32x2 *>(st.bh + o) = (u32x2){hi0, hi1};
    *reinterpret_cast<u32x2 *>(st.bl + o) = (u32x2){lo0, lo1};
  }
}

template <int WCN, int MB>
__global__ __launch_bounds__(RCfg<WCN>::NT, 2 / WCN) void k_gemm2_bf16x3_rp(Gemm2RArgs g) {
  constexpr int TM = 64 * MB, TN = RCfg<WCN>::TN;
  __shared__ RStage<TN> st[2];
  const int total = g.sb * g.n_mt * g.n_ct;
  const int id = xcd_tile(total);
  if (id >= total) return;
  // slot fastest: both operands (X~^T rows and the Y^T, X~^T sources of G) are the same for
  // every image slot, only the weights differ, so the tiles an XCD runs together share
  // their operand panels in its L2
  const int sl = id % g.sb;
  const int rest = id / g.sb;
  const int mt = rest / g.n_ct, ct = rest - mt * g.n_ct;
  const int m0 = mt * TM, c0 = ct * TN;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
  const int wr = w & 3, wc = w >> 2;
  const int row0 = m0 + wr * 16 * MB, cw = c0 + wc * RWC;
  const int mi = clamp_blocks(g.n, row0, MB);
  const int nj = clamp_blocks(g.n, cw, RNB);
  const Halves hv = live_halves(mi, nj);
  const float *wf = g.Wf + (size_t)sl * 2 * g.Np;

  v4f acc[MB][RNB];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < RNB; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};

  const int KT = g.KT, last = (KT - 1) * TK;
  const int2 base = rfrag_base();
  // the fragment-tiled layout holds all Np / TK slices; the loop stops after KT of them
  const int n_ks = g.Np / TK, n_rb = g.npd / 16, rb0 = row0 / 16;
  GSrc src;
  AFrag<MB> a0, a1;
  rload_g<WCN>(g, wf, c0, 0, src);
  rload_a<MB>(g.Ah, g.Al, n_ks, n_rb, rb0, 0, a0);
  rload_a<MB>(g.Ah, g.Al, n_ks, n_rb, rb0, 1, a1);
  rstore_g<WCN>(st[0], src);
  rload_g<WCN>(g, wf, c0, TK, src);
  __syncthreads();
  if (hv.h0) {
    for (int kt = 0; kt < KT; kt += 2) {
      // stage 0: slice kt; src: slice kt + 1; a0: slice kt; a1: slice kt + 1
      rmma<MB, 0, RNB / 2>(st[0], a0, base, acc);
      rstore_g<WCN>(st[1], src);
      rload_g<WCN>(g, wf, c0, min((kt + 2) * TK, last), src);
      if (hv.h1) rmma<MB, RNB / 2, RNB>(st[0], a0, base, acc);
      rload_a<MB>(g.Ah, g.Al, n_ks, n_rb, rb0, min(kt + 2, KT - 1), a0);
      __syncthreads();
      // stage 1: slice kt + 1; src: slice kt + 2; a1: slice kt + 1; a0: slice kt + 2
      rmma<MB, 0, RNB / 2>(st[1], a1, base, acc);
      rstore_g<WCN>(st[0], src);
      rload_g<WCN>(g, wf, c0, min((kt + 3) * TK, last), src);
      if (hv.h1) rmma<MB, RNB / 2, RNB>(st[1], a1, base, acc);
      rload_a<MB>(g.Ah, g.Al, n_ks, n_rb, rb0, min(kt + 3, KT - 1), a1);
      __syncthreads();
    }
  } else {
    // a wave whose rows or columns all lie past the last mode stages its share of the
    // formed operand and keeps the barriers, but issues no MFMAs
    for (int kt = 0; kt < KT; kt += 2) {
      rstore_g<WCN>(st[1], src);
      rload_g<WCN>(g, wf, c0, min((kt + 2) * TK, last), src);
      __syncthreads();
      rstore_g<WCN>(st[0], src);
      rload_g<WCN>(g, wf, c0, min((kt + 3) * TK, last), src);
      __syncthreads();
    }
  }

  float *Bo = g.B + (size_t)sl * g.npd * g.npd;
#pragma unroll
  for (int ti = 0; ti < MB; ++ti)
#pragma unroll
    for (int tj = 0; tj < RNB; ++tj)
      if (ti < mi && tj < nj) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = row0 + ti * 16 + 4 * lg + r;
          const int c = cw + tj * 16 + lr;
          Bo[(size_t)m * g.npd + c] = acc[ti][tj][r];
        }
      }
}

// ---- GEMM2 with eight row waves sharing one formed operand panel (tile form 7). The formed
// operand is the kernel's heaviest load stream (its two FP32 sources and the weights: more
// vector-memory instructions per slice than the A fragments), and the 96-column panel formed
// per slice serves every wave of the workgroup: eight waves stacked along the rows (TM =
// 128 MB rows, one workgroup per CU) halve the formation loads per MFMA against four. Each
// wave runs the MFMA sequence of k_gemm2_bf16x3_rp (same slices, same order): the stored
// B_s are those of tile forms 4 and 6 bit for bit.
constexpr int kR8Threads = 512;

// The 96 x 32 slice as 768 items (row, four k): thread t forms item t (rows 0-63) and, in
// waves 0-3, item t + 512 (rows 64-95); both share the weights of k quad t % 8.
struct GSrc8 {
  v4f ya, xa, yb, xb, wa, wb;
};

__device__ inline void rload_g8(const Gemm2RArgs &g, const float *__restrict__ wf, int c0,
                                int k0, bool two, GSrc8 &s) {
  const int t = threadIdx.x, kk = k0 + (t & 7) * 4;
  int ra = c0 + (t >> 3), rb = ra + 64;
  ra = ra < g.npd ? ra : g.npd - 1;
  rb = rb < g.npd ? rb : g.npd - 1;
  s.ya = *reinterpret_cast<const v4f *>(g.YTf + (size_t)ra * g.Np + kk);
  s.xa = *reinterpret_cast<const v4f *>(g.XtTf + (size_t)ra * g.Np + kk);
  if (two) {
    s.yb = *reinterpret_cast<const v4f *>(g.YTf + (size_t)rb * g.Np + kk);
    s.xb = *reinterpret_cast<const v4f *>(g.XtTf + (size_t)rb * g.Np + kk);
  }
  s.wa = *reinterpret_cast<const v4f *>(wf + kk);
  s.wb = *reinterpret_cast<const v4f *>(wf + g.Np + kk);
}

__device__ inline void rstore_item(RStage<RWC> &st, int row, int q, v4f wa, v4f wb, v4f y,
                                   v4f x) {
  uint32_t hi0, lo0, hi1, lo1;
  split2(fmaf(wa.x, y.x, wb.x * x.x), fmaf(wa.y, y.y, wb.y * x.y), hi0, lo0);
  split2(fmaf(wa.z, y.z, wb.z * x.z), fmaf(wa.w, y.w, wb.w * x.w), hi1, lo1);
  const int o = row * TK + (((q >> 1) ^ rswz(row)) * 8) + (q & 1) * 4;
  *reinterpret_cast<u32x2 *>(st.bh + o) = (u32x2){hi0, hi1};
  *reinterpret_cast<u32x2 *>(st.bl + o) = (u32x2){lo0, lo1};
}

__device__ inline void rstore_g8(RStage<RWC> &st, bool two, const GSrc8 &s) {
  const int t = threadIdx.x, q = t & 7;
  rstore_item(st, t >> 3, q, s.wa, s.wb, s.ya, s.xa);
  if (two) rstore_item(st, 64 + (t >> 3), q, s.wa, s.wb, s.yb, s.xb);
}

// One wave's share of an eight-wave GEMM2 tile: MB row blocks from row0, the 96 columns
// from c0, every slice of the formed operand staged by all eight waves together. Every wave
// of the workgroup runs the same k loop with the same barriers (waves of different MB take
// different instantiations, with the same barrier sequence).
template <int MB>
__device__ inline void r8_wave(const Gemm2RArgs &g, RStage<RWC> (&st)[2], int sl, int row0,
                               int c0, const float *__restrict__ wf, bool two) {
  const int l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
  const int mi = clamp_blocks(g.n, row0, MB);
  const int nj = clamp_blocks(g.n, c0, RNB);
  const Halves hv = live_halves(mi, nj);

  v4f acc[MB][RNB];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < RNB; ++b) acc[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};

  const int KT = g.KT, last = (KT - 1) * TK;
  // fragment offsets of this lane in a stage (one column wave: rows lr of each block)
  const int2 base = make_int2(lr * TK + ((lg ^ rswz(lr)) * 8), lr * TK + ((lg ^ rswz(lr) ^ 2) * 8));
  const int n_ks = g.Np / TK, n_rb = g.npd / 16, rb0 = row0 / 16;
  GSrc8 src;
  AFrag<MB> a0, a1;
  rload_g8(g, wf, c0, 0, two, src);
  rload_a<MB>(g.Ah, g.Al, n_ks, n_rb, rb0, 0, a0);
  rload_a<MB>(g.Ah, g.Al, n_ks, n_rb, rb0, 1, a1);
  rstore_g8(st[0], two, src);
  rload_g8(g, wf, c0, TK, two, src);
  __syncthreads();
  if (hv.h0) {
    for (int kt = 0; kt < KT; kt += 2) {
      rmma<MB, 0, RNB / 2>(st[0], a0, base, acc);
      rstore_g8(st[1], two, src);
      rload_g8(g, wf, c0, min((kt + 2) * TK, last), two, src);
      if (hv.h1) rmma<MB, RNB / 2, RNB>(st[0], a0, base, acc);
      rload_a<MB>(g.Ah, g.Al, n_ks, n_rb, rb0, min(kt + 2, KT - 1), a0);
      __syncthreads();
      rmma<MB, 0, RNB / 2>(st[1], a1, base, acc);
      rstore_g8(st[0], two, src);
      rload_g8(g, wf, c0, min((kt + 3) * TK, last), two, src);
      if (hv.h1) rmma<MB, RNB / 2, RNB>(st[1], a1, base, acc);
      rload_a<MB>(g.Ah, g.Al, n_ks, n_rb, rb0, min(kt + 3, KT - 1), a1);
      __syncthreads();
    }
  } else {
    for (int kt = 0; kt < KT; kt += 2) {
      rstore_g8(st[1], two, src);
      rload_g8(g, wf, c0, min((kt + 2) * TK, last), two, src);
      __syncthreads();
      rstore_g8(st[0], two, src);
      rload_g8(g, wf, c0, min((kt + 3) * TK, last), two, src);
      __syncthreads();
    }
  }

  float *Bo = g.B + (size_t)sl * g.npd * g.npd;
#pragma unroll
  for (int ti = 0; ti < MB; ++ti)
#pragma unroll
    for (int tj = 0; tj < RNB; ++tj)
      if (ti < mi && tj < nj) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = row0 + ti * 16 + 4 * lg + r;
          const int c = c0 + tj * 16 + lr;
          Bo[(size_t)m * g.npd + c] = acc[ti][tj][r];
        }
      }
}

// Tile of 128 MB rows: eight waves of MB row blocks each.
template <int MB>
__global__ __launch_bounds__(kR8Threads, 1) void k_gemm2_bf16x3_r8(Gemm2RArgs g) {
  constexpr int TM = 128 * MB;
  __shared__ RStage<RWC> st[2];
  const int total = g.sb * g.n_mt * g.n_ct;
  const int id = xcd_tile(total);
  if (id >= total) return;
  const int sl = id % g.sb;
  const int rest = id / g.sb;
  const int mt = rest / g.n_ct, ct = rest - mt * g.n_ct;
  const int w = threadIdx.x >> 6;
  r8_wave<MB>(g, st, sl, mt * TM + w * 16 * MB, ct * RWC, g.Wf + (size_t)sl * 2 * g.Np, w < 4);
}

// Tile of 448 rows: waves 0-3 take four row blocks each, waves 4-7 three (so that 1344
// padded modes make three tiles exactly; a wave of each kind on every SIMD).
__global__ __launch_bounds__(kR8Threads, 1) void k_gemm2_bf16x3_r7(Gemm2RArgs g) {
  constexpr int TM = 448;
  __shared__ RStage<RWC> st[2];
  const int total = g.sb * g.n_mt * g.n_ct;
  const int id = xcd_tile(total);
  if (id >= total) return;
  const int sl = id % g.sb;
  const int rest = id / g.sb;
  const int mt = rest / g.n_ct, ct = rest - mt * g.n_ct;
  const int w = threadIdx.x >> 6;
  const float *wf = g.Wf + (size_t)sl * 2 * g.Np;
  if (w < 4)
    r8_wave<4>(g, st, sl, mt * TM + 64 * w, ct * RWC, wf, true);
  else
    r8_wave<3>(g, st, sl, mt * TM + 256 + 48 * (w - 4), ct * RWC, wf, false);
}

}  // namespace

void launch_tile_bf16x3(const double *src, int rows, int cols, uint16_t *dst, hipStream_t st) {
  const size_t n8 = (size_t)rows * (cols / 8);
  size_t grid = (n8 + 255) / 256;
  if (grid > 65535u * 16u) grid = 65535u * 16u;
  hipLaunchKernelGGL(k_tile_bf16x3, dim3((unsigned)grid), dim3(256), 0, st, src, rows, cols, dst);
}

void launch_tile_acc_f32(const float *src, int rows, int cols, float *dst, hipStream_t st) {
  const size_t n4 = (size_t)rows * cols / 4;
  size_t grid = (n4 + 255) / 256;
  if (grid > 65535u * 16u) grid = 65535u * 16u;
  hipLaunchKernelGGL(k_tile_acc_f32, dim3((unsigned)grid), dim3(256), 0, st, src, rows, cols, dst);
}

void launch_gemm1_bf16x3_rp(const Ensemble &e, int sb, int what, int wcn, hipStream_t st) {
  Gemm1RArgs g;
  const size_t plane = (size_t)e.s_batch * e.npd * e.npd;
  const size_t NX = (size_t)e.Np * e.npd;
  g.Xh = e.X_frag; g.Xl = e.X_frag + NX;
  g.Yh = e.Ys_bf; g.Yl = e.Ys_bf + plane;
  g.Xa = e.X_acc; g.Fta = e.Ft_acc; g.P = e.Pw;
  g.Np = e.Np; g.npd = e.npd; g.sb = sb;
  const int tn = RWC * wcn;
  g.n_it = e.Np / 256; g.n_ct = (e.npd + tn - 1) / tn;
  g.sg = sb < kG1SlotGroup ? sb : kG1SlotGroup;
  g.do_q = (what & (kTermF | kTermD2vY)) ? 1 : 0;
  g.N = e.N; g.n = e.n;
  if (what & kTermD2vR) {
    // gamma_i = X_i . (f_ups r_s) = sum_c r_s[c] X~^T[c][i] and delta_i = Ft_i . (r_s / f_psi)
    // = sum_c r_s[c] Y^T[c][i], in FP64 on the FP64 operand copies
    launch_weighted_tn(e.npd, sb, e.Np, e.Rv, e.s_batch, nullptr, e.XtT, e.Np, 1.0, e.GD, e.Np,
                       st);
    launch_weighted_tn(e.npd, sb, e.Np, e.Rv, e.s_batch, nullptr, e.YT, e.Np, 1.0,
                       e.GD + (size_t)e.s_batch * e.Np, e.Np, st);
  }
  const int total = sb * g.n_it * g.n_ct;
  const dim3 grid(8 * ((total + 7) / 8));
  if (wcn == 2)
    hipLaunchKernelGGL(k_gemm1_bf16x3_rp<2>, grid, dim3(RCfg<2>::NT), 0, st, g);
  else
    hipLaunchKernelGGL(k_gemm1_bf16x3_rp<1>, grid, dim3(RCfg<1>::NT), 0, st, g);
  launch_weights_bf16(e, sb, g.n_ct, 2, what, st);
}

void launch_gemm2_bf16x3_rp(const Ensemble &e, int sb, int wcn, hipStream_t st, int mb) {
  Gemm2RArgs g;
  const size_t NX = (size_t)e.Np * e.npd;
  g.Ah = e.XtT_frag; g.Al = e.XtT_frag + NX;
  g.YTf = e.YT_f; g.XtTf = e.XtT_f; g.Wf = e.Wf; g.B = e.B_f;
  g.Np = e.Np; g.npd = e.npd; g.sb = sb;
  // wcn 8: eight row waves, 128 mb rows (mb = 4) or 448 rows (mb = 0: waves of four and of
  // three row blocks)
  const int tn = RWC * (wcn == 8 ? 1 : wcn);
  const int tm = wcn == 8 ? (mb == 0 ? 448 : 128 * mb) : 64 * mb;
  g.n_mt = (e.npd + tm - 1) / tm; g.n_ct = (e.npd + tn - 1) / tn;
  // k slices up to the last configuration, rounded up to the loop's step of two
  g.KT = ((e.N + 2 * TK - 1) / (2 * TK)) * 2;
  if (g.KT > e.Np / TK) g.KT = e.Np / TK;
  g.n = e.n;
  const int total = sb * g.n_mt * g.n_ct;
  const dim3 grid(8 * ((total + 7) / 8));
  if (wcn == 8 && mb == 0)
    hipLaunchKernelGGL(k_gemm2_bf16x3_r7, grid, dim3(kR8Threads), 0, st, g);
  else if (wcn == 8)
    hipLaunchKernelGGL((k_gemm2_bf16x3_r8<4>), grid, dim3(kR8Threads), 0, st, g);
  else if (wcn == 2)
    hipLaunchKernelGGL((k_gemm2_bf16x3_rp<2, 3>), grid, dim3(RCfg<2>::NT), 0, st, g);
  else if (mb == 4)
    hipLaunchKernelGGL((k_gemm2_bf16x3_rp<1, 4>), grid, dim3(RCfg<1>::NT), 0, st, g);
  else
    hipLaunchKernelGGL((k_gemm2_bf16x3_rp<1, 3>), grid, dim3(RCfg<1>::NT), 0, st, g);
}

}  // namespace tdscha
