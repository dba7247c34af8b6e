// Symmetry rotations and the small per-pass kernels of the ensemble L-application.
//
// The reference builds every image u = S_s x_i explicitly (LanczosFunctions.c:2158-2170)
// and evaluates n^2 quadratic forms on it. Here the symmetry acts on the n x n operands
// instead, because S_s is block diagonal with blocks of degenerate modes and f_ups/f_psi
// are constant on each block (so they commute with S_s):
//   u^T Y1 u     = x^T (S^T Y1 S) x                       (rotate Y1 once per symmetry)
//   sum_i c ut f^T = S (X~^T diag(c) Y) S^T               (rotate the n x n result back)
// Rotations cost O(n^2 d_b) per symmetry, against O(N n^2) for the contractions.
#include "tdscha_internal.h"

namespace tdscha {

namespace {

__global__ void k_prep(const double *__restrict__ R1, const double *__restrict__ Y1, double *Ysym,
                       double *R1p, int n, int npd) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int a = blockIdx.y;
  if (b >= npd) return;
  double v = 0.0;
  if (Y1 && a < n && b < n) v = 0.5 * (Y1[(size_t)a * n + b] + Y1[(size_t)b * n + a]);
  Ysym[(size_t)a * npd + b] = v;
  if (a == 0) R1p[b] = (R1 && b < n) ? R1[b] : 0.0;
}

// out[slot][k][b] = sum_l M[k][off_B + l] * Sel_B(l, b'),  M = in (+ in^T if symmetrize)
// Sel_B(l, b') = S_B[l][b'] (transS = 0) or S_B[b'][l] (transS = 1). The input matrix is
// the slot's chain (per_chain = 1) or the slot itself.
__global__ void k_rot_right(const double *__restrict__ in, int per_chain, int symmetrize,
                            double *__restrict__ out, const double *__restrict__ S,
                            const ModeMeta *__restrict__ meta, int n, int npd, int sym_stride,
                            Pass ps, int transS) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  int sl = blockIdx.z;
  if (b >= npd) return;
  double acc = 0.0;
  if (b < n && k < n) {
    ModeMeta mb = meta[b];
    const double *Sb = S + (size_t)ps.sym(sl) * sym_stride + mb.sym_off;
    const double *M = in + (size_t)(per_chain ? ps.chain(sl) : sl) * npd * npd;
    const double *row = M + (size_t)k * npd + mb.blk_off;
    const int bp = b - mb.blk_off, d = mb.blk_dim;
    for (int l = 0; l < d; ++l) {
      double v = row[l];
      if (symmetrize) v += M[(size_t)(mb.blk_off + l) * npd + k];
      double sv = transS ? Sb[bp * d + l] : Sb[l * d + bp];
      acc = fma(v, sv, acc);
    }
  }
  out[(size_t)sl * npd * npd + (size_t)k * npd + b] = acc;
}

// out[s][a][b] (or out[a][b] += sum over s) = sum_k Sel_A(a', k) * in[s][off_A + k][b]
// Sel_A(a', k) = S_A[a'][k] (transS = 0) or S_A[k][a'] (transS = 1).
// With accumulate, out is per chain (grid z = chains) and sums the chain's slots.
__global__ void k_rot_left(const double *__restrict__ in, double *__restrict__ out,
                           const double *__restrict__ S, const ModeMeta *__restrict__ meta, int n,
                           int npd, int sym_stride, Pass ps, int transS, int accumulate) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int a = blockIdx.y;
  if (b >= npd) return;
  if (accumulate) {
    if (a >= n || b >= n) return;
    const int c = blockIdx.z;
    out += (size_t)c * npd * npd;
    ModeMeta ma = meta[a];
    const int ap = a - ma.blk_off, d = ma.blk_dim;
    double acc = out[(size_t)a * npd + b];
    for (int t = 0; t < ps.ns; ++t) {
      const int sl = c * ps.ns + t;
      const double *Sa = S + (size_t)ps.sym(sl) * sym_stride + ma.sym_off;
      const double *M = in + (size_t)sl * npd * npd + (size_t)ma.blk_off * npd + b;
      double part = 0.0;
      for (int k = 0; k < d; ++k) {
        double sv = transS ? Sa[k * d + ap] : Sa[ap * d + k];
        part = fma(sv, M[(size_t)k * npd], part);
      }
      acc += part;
    }
    out[(size_t)a * npd + b] = acc;
  } else {
    int sl = blockIdx.z;
    double acc = 0.0;
    if (a < n && b < n) {
      ModeMeta ma = meta[a];
      const int ap = a - ma.blk_off, d = ma.blk_dim;
      const double *Sa = S + (size_t)ps.sym(sl) * sym_stride + ma.sym_off;
      const double *M = in + (size_t)sl * npd * npd + (size_t)ma.blk_off * npd + b;
      for (int k = 0; k < d; ++k) {
        double sv = transS ? Sa[k * d + ap] : Sa[ap * d + k];
        acc = fma(sv, M[(size_t)k * npd], acc);
      }
    }
    out[(size_t)sl * npd * npd + (size_t)a * npd + b] = acc;
  }
}

// ---- Group-tiled rotations on FP64 MFMA ------------------------------------------------
// Used when every block has at most kRotTile modes (else the one-sided kernels above run).
// Modes are cut into groups of whole degenerate blocks, each at most G wide. On the tile
// (row group g, column group h) the rotations only involve the two groups' diagonal
// blocks, staged in LDS as dense G x G block-diagonal matrices:
//   rot1:  Ys_s[g][h]  = S_g^T Ysym[g][h] S_h
//   rot2:  d2v[g][h]  += S_g (B_s + B_s^T)[g][h] S_h^T
// Both products are symmetric in (g, h), so only tiles g <= h are computed and each is
// written to its mirror position too. Products run as v_mfma_f64_16x16x4 over 16 x 16
// sub-tiles, one wave per sub-tile.

typedef double v4d __attribute__((ext_vector_type(4)));

template <int G> struct GrpTile {
  static constexpr int LD = G + 2;               // LDS row stride in doubles
  static constexpr int TPW = (G / 16) * (G / 16) / 4;   // 16 x 16 sub-tiles per wave
};

// Dense S_g of one symmetry (zero outside g's blocks and outside the group width), staged
// software-pipelined: each of the 256 threads owns G^2 / 256 elements
// (element idx = thread + 256 q); their offsets inside one symmetry's packed blocks do not
// depend on the symmetry and are found once (-1: the element is zero), so that the values of
// the next slot can be loaded into registers, branch-free, while the current slot is
// multiplied.
template <int G> struct GrpSIdx {
  static constexpr int PT = G * G / 256;
  int off[PT];
};

template <int G>
__device__ inline void group_s_index(int2 grp, const ModeMeta *__restrict__ meta,
                                     GrpSIdx<G> &ix) {
#pragma unroll
  for (int q = 0; q < GrpSIdx<G>::PT; ++q) {
    const int idx = threadIdx.x + 256 * q, r = idx / G, c = idx - r * G;
    int o = -1;
    if (r < grp.y && c < grp.y) {
      const ModeMeta mr = meta[grp.x + r];
      const int cc = c - (mr.blk_off - grp.x);
      if (cc >= 0 && cc < mr.blk_dim) o = mr.sym_off + (grp.x + r - mr.blk_off) * mr.blk_dim + cc;
    }
    ix.off[q] = o;
  }
}

template <int G>
__device__ inline void group_s_load(const double *__restrict__ Ss, const GrpSIdx<G> &ix,
                                    double (&v)[GrpSIdx<G>::PT]) {
#pragma unroll
  for (int q = 0; q < GrpSIdx<G>::PT; ++q) {
    const int o = ix.off[q];
    const double x = Ss[o < 0 ? 0 : o];
    v[q] = o < 0 ? 0.0 : x;
  }
}

template <int G>
__device__ inline void group_s_store(double *Sl, const double (&v)[GrpSIdx<G>::PT]) {
  constexpr int LD = GrpTile<G>::LD;
#pragma unroll
  for (int q = 0; q < GrpSIdx<G>::PT; ++q) {
    const int idx = threadIdx.x + 256 * q, r = idx / G, c = idx - r * G;
    Sl[r * LD + c] = v[q];
  }
}

// acc[q] (sub-tile q of this wave) = sum_k A(i, k) B(k, j) with
//   A(i, k) = TA ? Am[k][i] : Am[i][k],   B(k, j) = TB ? Bm[j][k] : Bm[k][j]
// over k < G; the result element r of a lane is row lg + 4 r, column lr of the sub-tile.
template <int G, bool TA, bool TB>
__device__ inline void grp_mma(const double *Am, const double *Bm, v4d (&acc)[GrpTile<G>::TPW]) {
  constexpr int LD = GrpTile<G>::LD, NT = G / 16;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
#pragma unroll
  for (int q = 0; q < GrpTile<G>::TPW; ++q) {
    const int t = w + 4 * q, ti = t / NT, tj = t - ti * NT;
    const int i = ti * 16 + lr, j = tj * 16 + lr;
#pragma unroll
    for (int kq = 0; kq < G / 4; ++kq) {
      const int k = kq * 4 + lg;
      const double a = TA ? Am[k * LD + i] : Am[i * LD + k];
      const double b = TB ? Bm[j * LD + k] : Bm[k * LD + j];
      acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
    }
  }
}

template <int G>
__device__ inline void grp_store_lds(double *Tl, const v4d (&acc)[GrpTile<G>::TPW]) {
  constexpr int LD = GrpTile<G>::LD, NT = G / 16;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
#pragma unroll
  for (int q = 0; q < GrpTile<G>::TPW; ++q) {
    const int t = w + 4 * q, ti = t / NT, tj = t - ti * NT;
#pragma unroll
    for (int r = 0; r < 4; ++r) Tl[(ti * 16 + lg + 4 * r) * LD + tj * 16 + lr] = acc[q][r];
  }
}

// One output element of the rotated tile in the requested storage.
__device__ inline void store_ys(double *Ys, void *lp, int lp_prec, size_t lp_plane, size_t o,
                                double v) {
  if (Ys) Ys[o] = v;
  if (lp_prec == kBF16x3) {
    uint16_t *hi = (uint16_t *)lp;
    const float f = (float)v;
    const uint16_t h = bf16_rne(f);
    hi[o] = h;
    hi[lp_plane + o] = bf16_rne(f - bf16_to_f(h));
  } else if (lp_prec == kFP32) {
    ((float *)lp)[o] = (float)v;
  }
}

// grid (tile pairs g <= h, slot splits, chains); the workgroup rotates the slots
// (optionally scaled: out[a][b] *= scale[a] scale[b])
// t = blockIdx.y, blockIdx.y + gridDim.y, ... of its chain.
template <int G>
__global__ __launch_bounds__(256) void k_rot1_grp(const double *__restrict__ Ysym,
                                                  double *__restrict__ Ys, void *__restrict__ lp,
                                                  int lp_prec, size_t lp_plane,
                                                  const double *__restrict__ S,
                                                  const int2 *__restrict__ groups,
                                                  const int2 *__restrict__ pairs,
                                                  const ModeMeta *__restrict__ meta,
                                                  const double *__restrict__ scale, int npd,
                                                  int sym_stride, Pass ps) {
  constexpr int LD = GrpTile<G>::LD, TPW = GrpTile<G>::TPW;
  __shared__ double M[G * LD], T[G * LD], Sg[G * LD], Sh[G * LD];
  const int2 pr = pairs[blockIdx.x];
  const int2 gg = groups[pr.x], gh = groups[pr.y];
  const int ch = blockIdx.z;
  const double *Yc = Ysym + (size_t)ch * npd * npd;
  for (int idx = threadIdx.x; idx < G * G; idx += blockDim.x) {
    const int r = idx / G, c = idx - r * G;
    M[r * LD + c] = (r < gg.y && c < gh.y) ? Yc[(size_t)(gg.x + r) * npd + gh.x + c] : 0.0;
  }
  GrpSIdx<G> ig, ih;
  group_s_index<G>(gg, meta, ig);
  group_s_index<G>(gh, meta, ih);
  double sg[GrpSIdx<G>::PT], sh[GrpSIdx<G>::PT];
  if (blockIdx.y < ps.ns) {
    const double *Ss = S + (size_t)ps.sym(ch * ps.ns + blockIdx.y) * sym_stride;
    group_s_load<G>(Ss, ig, sg);
    group_s_load<G>(Ss, ih, sh);
  }
  for (int t = blockIdx.y; t < ps.ns; t += gridDim.y) {
    const int sl = ch * ps.ns + t;
    __syncthreads();   // previous slot's reads of Sg / Sh / T are done
    group_s_store<G>(Sg, sg);
    group_s_store<G>(Sh, sh);
    if (t + (int)gridDim.y < ps.ns) {   // the next slot's blocks, in flight meanwhile
      const double *Sn = S + (size_t)ps.sym(sl + gridDim.y) * sym_stride;
      group_s_load<G>(Sn, ig, sg);
      group_s_load<G>(Sn, ih, sh);
    }
    __syncthreads();
    v4d acc[TPW];
#pragma unroll
    for (int q = 0; q < TPW; ++q) acc[q] = (v4d){0.0, 0.0, 0.0, 0.0};
    grp_mma<G, false, false>(M, Sh, acc);           // T = M S_h
    grp_store_lds<G>(T, acc);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TPW; ++q) acc[q] = (v4d){0.0, 0.0, 0.0, 0.0};
    grp_mma<G, true, false>(Sg, T, acc);            // Ys = S_g^T T
    // the tile goes out through LDS (Sh is free: its last reads were before the barrier
    // above), so that both it and its mirror are written along rows by consecutive threads
    grp_store_lds<G>(Sh, acc);
    __syncthreads();
    const size_t base = (size_t)sl * npd * npd;
    for (int idx = threadIdx.x; idx < gg.y * gh.y; idx += blockDim.x) {
      const int i = idx / gh.y, j = idx - i * gh.y;
      const double v = Sh[i * LD + j];
      store_ys(Ys, lp, lp_prec, lp_plane, base + (size_t)(gg.x + i) * npd + gh.x + j,
               scale ? v * scale[gg.x + i] * scale[gh.x + j] : v);
    }
    if (pr.x != pr.y) {
      for (int idx = threadIdx.x; idx < gg.y * gh.y; idx += blockDim.x) {
        const int j = idx / gg.y, i = idx - j * gg.y;   // mirror row gh.x + j, column gg.x + i
        const double v = Sh[i * LD + j];
        store_ys(Ys, lp, lp_prec, lp_plane, base + (size_t)(gh.x + j) * npd + gg.x + i,
                 scale ? v * scale[gg.x + i] * scale[gh.x + j] : v);
      }
    }
  }
}

// grid (tile pairs g <= h, chains). The running sum of the tile is held in the MFMA
// accumulators; every image of the pass is added onto it in symmetry order, so the result
// does not depend on how the symmetries are split into passes.
template <int G, typename TB>
__global__ __launch_bounds__(256) void k_rot2_grp(const TB *__restrict__ B,
                                                  double *__restrict__ d2v,
                                                  const double *__restrict__ S,
                                                  const int2 *__restrict__ groups,
                                                  const int2 *__restrict__ pairs,
                                                  const ModeMeta *__restrict__ meta, int npd,
                                                  int sym_stride, Pass ps) {
  constexpr int LD = GrpTile<G>::LD, TPW = GrpTile<G>::TPW, NT = G / 16;
  __shared__ double M[G * LD], T[G * LD], Sg[G * LD], Sh[G * LD];
  const int2 pr = pairs[blockIdx.x];
  const int2 gg = groups[pr.x], gh = groups[pr.y];
  const int ch = blockIdx.y;
  d2v += (size_t)ch * npd * npd;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
  v4d acc[TPW];
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int tt = w + 4 * q, ti = tt / NT, tj = tt - ti * NT;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = ti * 16 + lg + 4 * r, j = tj * 16 + lr;
      acc[q][r] = (i < gg.y && j < gh.y) ? d2v[(size_t)(gg.x + i) * npd + gh.x + j] : 0.0;
    }
  }
  // the slot's B[g][h], B[h][g] and symmetry blocks are loaded into registers (clamped,
  // branch-free) one slot ahead, while the previous slot is multiplied
  constexpr int PT = GrpSIdx<G>::PT;
  GrpSIdx<G> ig, ih;
  group_s_index<G>(gg, meta, ig);
  group_s_index<G>(gh, meta, ih);
  TB bm[PT], bt[PT];
  double sg[PT], sh[PT];
  auto load_slot = [&](int t) {
    const int sl = ch * ps.ns + t;
    const double *Ss = S + (size_t)ps.sym(sl) * sym_stride;
    const TB *Bs = B + (size_t)sl * npd * npd;
#pragma unroll
    for (int q = 0; q < PT; ++q) {
      // indices clamped to the groups (never empty): the padding re-reads lines already
      // fetched instead of the neighbouring groups' elements
      const int idx = threadIdx.x + 256 * q, r = idx / G, c = idx - r * G;
      bm[q] = Bs[(size_t)(gg.x + min(r, gg.y - 1)) * npd + gh.x + min(c, gh.y - 1)];
      bt[q] = Bs[(size_t)(gh.x + min(r, gh.y - 1)) * npd + gg.x + min(c, gg.y - 1)];
    }
    group_s_load<G>(Ss, ig, sg);
    group_s_load<G>(Ss, ih, sh);
  };
  if (ps.ns > 0) load_slot(0);
  for (int t = 0; t < ps.ns; ++t) {
    __syncthreads();
    // M = B[g][h] + (B[h][g])^T: both tiles read along their rows (coalesced), the second
    // staged in T and added transposed
#pragma unroll
    for (int q = 0; q < PT; ++q) {
      const int idx = threadIdx.x + 256 * q, r = idx / G, c = idx - r * G;
      M[r * LD + c] = (r < gg.y && c < gh.y) ? (double)bm[q] : 0.0;
      T[r * LD + c] = (r < gh.y && c < gg.y) ? (double)bt[q] : 0.0;
    }
    group_s_store<G>(Sg, sg);
    group_s_store<G>(Sh, sh);
    if (t + 1 < ps.ns) load_slot(t + 1);
    __syncthreads();
    for (int idx = threadIdx.x; idx < G * G; idx += blockDim.x) {
      const int r = idx / G, c = idx - r * G;
      M[r * LD + c] += T[c * LD + r];
    }
    __syncthreads();
    v4d tq[TPW];
#pragma unroll
    for (int q = 0; q < TPW; ++q) tq[q] = (v4d){0.0, 0.0, 0.0, 0.0};
    grp_mma<G, false, true>(M, Sh, tq);             // T = (B + B^T) S_h^T
    grp_store_lds<G>(T, tq);
    __syncthreads();
    grp_mma<G, false, false>(Sg, T, acc);           // acc += S_g T
  }
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int tt = w + 4 * q, ti = tt / NT, tj = tt - ti * NT;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = ti * 16 + lg + 4 * r, j = tj * 16 + lr;
      if (i < gg.y && j < gh.y) {
        d2v[(size_t)(gg.x + i) * npd + gh.x + j] = acc[q][r];
        if (pr.x != pr.y) d2v[(size_t)(gh.x + j) * npd + gg.x + i] = acc[q][r];
      }
    }
  }
}

// r_s = S_s^T R1;  G[s][0][a] = f_ups[a] r_s[a],  G[s][1][a] = r_s[a] / f_psi[a];
// Rv[a][s] = r_s[a] when Rv is given (leading dimension ldr)
__global__ void k_rot_vec(const double *__restrict__ R1p, double *__restrict__ G,
                          const double *__restrict__ S, const ModeMeta *__restrict__ meta,
                          const double *__restrict__ fups, const double *__restrict__ fpsi, int n,
                          int npd, int sym_stride, Pass ps, double *__restrict__ Rv, int ldr) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  int sl = blockIdx.y;
  if (a >= npd) return;
  double g1 = 0.0, g2 = 0.0, r = 0.0;
  R1p += (size_t)ps.chain(sl) * npd;
  if (a < n) {
    ModeMeta ma = meta[a];
    const int ap = a - ma.blk_off, d = ma.blk_dim;
    const double *Sa = S + (size_t)ps.sym(sl) * sym_stride + ma.sym_off;
    for (int k = 0; k < d; ++k) r = fma(Sa[k * d + ap], R1p[ma.blk_off + k], r);
    g1 = fups[a] * r;
    g2 = r / fpsi[a];
  }
  G[(size_t)sl * 2 * npd + a] = g1;
  G[(size_t)sl * 2 * npd + npd + a] = g2;
  if (Rv) Rv[(size_t)a * ldr + sl] = r;
}

__global__ void k_fav_rot_part(const double *__restrict__ V, double *__restrict__ P,
                               const double *__restrict__ S, const ModeMeta *__restrict__ meta,
                               int n, int npd, int sym_stride, Pass ps) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int sl = blockIdx.y;
  if (a >= n) return;
  const ModeMeta ma = meta[a];
  const int ap = a - ma.blk_off, d = ma.blk_dim;
  const double *Sa = S + (size_t)ps.sym(sl) * sym_stride + ma.sym_off + ap * d;
  const double *v = V + (size_t)sl * npd + ma.blk_off;
  double part = 0.0;
  for (int k = 0; k < d; ++k) part = fma(Sa[k], v[k], part);
  P[(size_t)sl * npd + a] = part;
}

__global__ void k_fav_rot_sum(const double *__restrict__ P, double *__restrict__ f_acc, int n,
                              int npd, Pass ps) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (a >= n) return;
  double acc = f_acc[(size_t)c * npd + a];
  for (int t = 0; t < ps.ns; ++t) acc += P[(size_t)(c * ps.ns + t) * npd + a];
  f_acc[(size_t)c * npd + a] = acc;
}

__global__ void k_zero(double *p, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = 0.0;
}

__global__ void k_finalize(const double *__restrict__ f_acc, const double *__restrict__ d2v_acc,
                           double *f_out, double *d2v_out, int n, int npd, double inv_norm, int do_f,
                           int do_d, int d_accumulate) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int a = blockIdx.y;
  if (b >= n) return;
  if (do_d) {
    double v = d2v_acc[(size_t)a * npd + b] * inv_norm;
    size_t o = (size_t)a * n + b;
    d2v_out[o] = d_accumulate ? d2v_out[o] + v : v;
  }
  if (do_f && a == 0) f_out[b] = f_acc[b] * inv_norm;
}

// fp64 -> bf16 (hi, lo) planes (hi at out, lo at out + count) or fp32, element for element.
// hi = RNE(v), lo = RNE(v - hi) with v rounded to fp32 first: hi + lo keeps 16 bits.
__global__ void k_convert(const double *__restrict__ src, void *out, size_t count, int prec) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) {
    const float f = (float)src[i];
    if (prec == kBF16x3) {
      uint16_t *hi = (uint16_t *)out;
      const uint16_t h = bf16_rne(f);
      hi[i] = h;
      hi[count + i] = bf16_rne(f - __uint_as_float((uint32_t)h << 16));
    } else {
      ((float *)out)[i] = f;
    }
  }
}

inline dim3 grid1(size_t count, int block) {
  size_t g = (count + block - 1) / block;
  if (g > 65535u * 16u) g = 65535u * 16u;
  return dim3((unsigned)g);
}

}  // namespace

void launch_prep_inputs(const Ensemble &e, const double *R1, const double *Y1, hipStream_t st) {
  dim3 grid((e.npd + 255) / 256, e.npd);
  hipLaunchKernelGGL(k_prep, grid, dim3(256), 0, st, R1, Y1, e.Ysym, e.R1p, e.n, e.npd);
}

void launch_rot_right(const Ensemble &e, const double *in, bool per_chain, bool symmetrize,
                      double *out, const Pass &ps, bool transS, hipStream_t st) {
  dim3 grid((e.npd + 255) / 256, e.npd, ps.slots());
  hipLaunchKernelGGL(k_rot_right, grid, dim3(256), 0, st, in, (int)per_chain, (int)symmetrize, out,
                     e.S, e.meta, e.n, e.npd, e.sym_stride, ps, (int)transS);
}

void launch_rot_left(const Ensemble &e, const double *in, double *out, const Pass &ps, bool transS,
                     bool accumulate, hipStream_t st) {
  dim3 grid((e.npd + 255) / 256, e.npd, accumulate ? ps.nc : ps.slots());
  hipLaunchKernelGGL(k_rot_left, grid, dim3(256), 0, st, in, out, e.S, e.meta, e.n, e.npd,
                     e.sym_stride, ps, (int)transS, (int)accumulate);
}

namespace {

void rot1_group(const Ensemble &e, const Pass &ps, const double *in, double *out64, void *lp,
                int prec, const double *scale, hipStream_t st) {
  const size_t plane = (size_t)e.s_batch * e.npd * e.npd;
  // enough workgroups to fill the chip: split each chain's slots over gridDim.y
  const long tiles = (long)e.n_pairs * ps.nc;
  int split = (int)((2048 + tiles - 1) / tiles);
  if (split > ps.ns) split = ps.ns;
  if (split < 1) split = 1;
  dim3 grid(e.n_pairs, split, ps.nc);
  if (e.grp_width == 32)
    hipLaunchKernelGGL(k_rot1_grp<32>, grid, dim3(256), 0, st, in, out64, lp, prec, plane, e.S,
                       e.dgroups, e.dpairs, e.meta, scale, e.npd, e.sym_stride, ps);
  else
    hipLaunchKernelGGL(k_rot1_grp<64>, grid, dim3(256), 0, st, in, out64, lp, prec, plane, e.S,
                       e.dgroups, e.dpairs, e.meta, scale, e.npd, e.sym_stride, ps);
}

__global__ void k_scale_sym(double *M, const double *__restrict__ scale, int npd, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const size_t r = (i / npd) % npd, c = i % npd;
  M[i] *= scale[r] * scale[c];
}

}  // namespace

void launch_rot1_fused(const Ensemble &e, const Pass &ps, int prec, hipStream_t st) {
  rot1_group(e, ps, e.Ysym, prec == kFP64 ? e.Ys : (double *)nullptr, e.ys_lp(prec), prec,
             nullptr, st);
}

void launch_rot_sym(const Ensemble &e, const Pass &ps, const double *in, double *out,
                    const double *scale, hipStream_t st) {
  if (e.fused_rot) {
    rot1_group(e, ps, in, out, nullptr, kFP64, scale, st);
  } else {
    launch_rot_right(e, in, true, false, e.Tw, ps, false, st);
    launch_rot_left(e, e.Tw, out, ps, true, false, st);
    if (scale) {
      const size_t count = (size_t)ps.slots() * e.npd * e.npd;
      hipLaunchKernelGGL(k_scale_sym, grid1(count, 256), dim3(256), 0, st, out, scale, e.npd,
                         count);
    }
  }
}

void launch_rot2_fused(const Ensemble &e, const Pass &ps, bool b_f32, hipStream_t st) {
  dim3 grid(e.n_pairs, ps.nc);
  if (e.grp_width == 32)
    if (b_f32)
      hipLaunchKernelGGL((k_rot2_grp<32, float>), grid, dim3(256), 0, st, e.B_f, e.d2v_acc, e.S,
                         e.dgroups, e.dpairs, e.meta, e.npd, e.sym_stride, ps);
    else
      hipLaunchKernelGGL((k_rot2_grp<32, double>), grid, dim3(256), 0, st, e.B, e.d2v_acc, e.S,
                         e.dgroups, e.dpairs, e.meta, e.npd, e.sym_stride, ps);
  else if (b_f32)
    hipLaunchKernelGGL((k_rot2_grp<64, float>), grid, dim3(256), 0, st, e.B_f, e.d2v_acc, e.S,
                       e.dgroups, e.dpairs, e.meta, e.npd, e.sym_stride, ps);
  else
    hipLaunchKernelGGL((k_rot2_grp<64, double>), grid, dim3(256), 0, st, e.B, e.d2v_acc, e.S,
                       e.dgroups, e.dpairs, e.meta, e.npd, e.sym_stride, ps);
}

void launch_rot_vec(const Ensemble &e, const Pass &ps, hipStream_t st) {
  dim3 grid((e.npd + 255) / 256, ps.slots());
  hipLaunchKernelGGL(k_rot_vec, grid, dim3(256), 0, st, e.R1p, e.G, e.S, e.meta, e.fups, e.fpsi,
                     e.n, e.npd, e.sym_stride, ps, e.Rv, e.s_batch);
}

void launch_rot_vec_from(const Ensemble &e, const Pass &ps, const double *in, double *out,
                         const double *div, hipStream_t st) {
  dim3 grid((e.npd + 255) / 256, ps.slots());
  hipLaunchKernelGGL(k_rot_vec, grid, dim3(256), 0, st, in, out, e.S, e.meta, e.fups, div, e.n,
                     e.npd, e.sym_stride, ps, (double *)nullptr, 0);
}

void launch_fav_rot_acc(const Ensemble &e, const Pass &ps, hipStream_t st) {
  // the split-k partials of V are free once k_fav_reduce has summed them
  hipLaunchKernelGGL(k_fav_rot_part, dim3((e.n + 255) / 256, ps.slots()), dim3(256), 0, st, e.V,
                     e.Vp, e.S, e.meta, e.n, e.npd, e.sym_stride, ps);
  hipLaunchKernelGGL(k_fav_rot_sum, dim3((e.n + 255) / 256, ps.nc), dim3(256), 0, st, e.Vp,
                     e.f_acc, e.n, e.npd, ps);
}

void launch_zero(double *p, size_t count, hipStream_t st) {
  hipLaunchKernelGGL(k_zero, grid1(count, 256), dim3(256), 0, st, p, count);
}

void launch_finalize(const Ensemble &e, int what, bool d2v_accumulate, double *f_out, double *d2v_out,
                     hipStream_t st) {
  double inv = 1.0 / (e.N_eff * (double)e.N_sym);
  int do_f = (what & kTermF) && f_out;
  int do_d = (what & (kTermD2vR | kTermD2vY)) && d2v_out;
  dim3 grid((e.n + 255) / 256, e.n);
  hipLaunchKernelGGL(k_finalize, grid, dim3(256), 0, st, e.f_acc, e.d2v_acc, f_out, d2v_out, e.n,
                     e.npd, inv, do_f, do_d, (int)d2v_accumulate);
}

void launch_split_ys(const Ensemble &e, int sb, int prec, hipStream_t st) {
  // the planes are laid out for s_batch symmetries (the GEMM1 hi / lo offsets assume it)
  const size_t plane = (size_t)e.s_batch * e.npd * e.npd;
  const size_t count = (size_t)sb * e.npd * e.npd;
  if (prec == kBF16x3) {
    // hi plane [0, plane), lo plane [plane, 2 plane): convert hi and lo of the first `count`
    hipLaunchKernelGGL(k_convert, grid1(count, 256), dim3(256), 0, st, e.Ys, e.ys_lp(prec), plane,
                       prec);
  } else {
    hipLaunchKernelGGL(k_convert, grid1(count, 256), dim3(256), 0, st, e.Ys, e.ys_lp(prec), count,
                       prec);
  }
}

void launch_convert(const double *src, void *dst, size_t count, int prec, hipStream_t st) {
  hipLaunchKernelGGL(k_convert, grid1(count, 256), dim3(256), 0, st, src, dst, count, prec);
}

}  // namespace tdscha
