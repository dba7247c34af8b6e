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

// ---- Fused two-sided rotations on (row block A) x (column group) tiles --------------
// A column group is a run of consecutive whole blocks at most kRotTile columns wide, so
// both one-sided products of a tile stay inside LDS. Used when every block has at most
// kRotTile modes.

// Ys[s][A][grp] = S_A^T Ysym[A][grp] S_grp   (written as fp64, or as bf16 hi/lo / fp32)
__global__ __launch_bounds__(256) void k_rot1_fused(const double *__restrict__ Ysym,
                                                    double *__restrict__ Ys, void *__restrict__ lp,
                                                    int lp_prec, size_t lp_plane,
                                                    const double *__restrict__ S,
                                                    const int4 *__restrict__ blocks,
                                                    const int2 *__restrict__ groups,
                                                    const ModeMeta *__restrict__ meta, int npd,
                                                    int sym_stride, Pass ps) {
  __shared__ double M[kRotTile * kRotTile];
  __shared__ double T[kRotTile * kRotTile];
  const int4 A = blocks[blockIdx.x];
  const int2 grp = groups[blockIdx.y];
  const int sl = blockIdx.z, s = ps.sym(sl);
  Ysym += (size_t)ps.chain(sl) * npd * npd;
  const int offA = A.x, dA = A.y, g0 = grp.x, W = grp.y;
  const double *Ss = S + (size_t)s * sym_stride;
  const double *SA = Ss + A.z;
  const int cnt = dA * W;
  for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
    const int r = idx / W, c = idx - r * W;
    M[r * kRotTile + c] = Ysym[(size_t)(offA + r) * npd + g0 + c];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
    const int r = idx / W, c = idx - r * W;
    const ModeMeta mb = meta[g0 + c];
    const int cb = mb.blk_off - g0, d = mb.blk_dim, cp = g0 + c - mb.blk_off;
    const double *SB = Ss + mb.sym_off;
    double acc = 0.0;
    for (int l = 0; l < d; ++l) acc = fma(M[r * kRotTile + cb + l], SB[l * d + cp], acc);
    T[r * kRotTile + c] = acc;
  }
  __syncthreads();
  const size_t base = (size_t)sl * npd * npd;
  for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
    const int r = idx / W, c = idx - r * W;
    double acc = 0.0;
    for (int k = 0; k < dA; ++k) acc = fma(SA[k * dA + r], T[k * kRotTile + c], acc);
    const size_t o = base + (size_t)(offA + r) * npd + g0 + c;
    if (Ys) Ys[o] = acc;
    if (lp_prec == kBF16x3) {
      uint16_t *hi = (uint16_t *)lp;
      const float f = (float)acc;
      const uint16_t h = bf16_rne(f);
      hi[o] = h;
      hi[lp_plane + o] = bf16_rne(f - __uint_as_float((uint32_t)h << 16));
    } else if (lp_prec == kFP32) {
      ((float *)lp)[o] = (float)acc;
    }
  }
}

// d2v[A][grp] += sum_s S_A (B_s + B_s^T)[A][grp] S_grp^T   (one workgroup owns its tile)
__global__ __launch_bounds__(256) void k_rot2_fused(const double *__restrict__ B,
                                                    double *__restrict__ d2v,
                                                    const double *__restrict__ S,
                                                    const int4 *__restrict__ blocks,
                                                    const int2 *__restrict__ groups,
                                                    const ModeMeta *__restrict__ meta, int npd,
                                                    int sym_stride, Pass ps) {
  __shared__ double M[kRotTile * kRotTile];
  __shared__ double T[kRotTile * kRotTile];
  constexpr int kPer = kRotTile * kRotTile / 256;
  const int4 A = blocks[blockIdx.x];
  const int2 grp = groups[blockIdx.y];
  const int ch = blockIdx.z;
  d2v += (size_t)ch * npd * npd;
  const int offA = A.x, dA = A.y, g0 = grp.x, W = grp.y;
  const int cnt = dA * W;
  // the images are added one at a time onto the running sum, so the result does not depend
  // on how the symmetries are split into passes
  double acc[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int idx = threadIdx.x + 256 * j;
    acc[j] = 0.0;
    if (idx < cnt) {
      const int r = idx / W, c = idx - r * W;
      acc[j] = d2v[(size_t)(offA + r) * npd + g0 + c];
    }
  }
  for (int t = 0; t < ps.ns; ++t) {
    const int sl = ch * ps.ns + t;
    const double *Ss = S + (size_t)ps.sym(sl) * sym_stride;
    const double *SA = Ss + A.z;
    const double *Bs = B + (size_t)sl * npd * npd;
    __syncthreads();
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
      const int r = idx / W, c = idx - r * W;
      M[r * kRotTile + c] = Bs[(size_t)(offA + r) * npd + g0 + c];
    }
    // the transposed part, read along rows of B (r fastest) and staged in T
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
      const int c = idx / dA, r = idx - c * dA;
      T[c * kRotTile + r] = Bs[(size_t)(g0 + c) * npd + offA + r];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
      const int r = idx / W, c = idx - r * W;
      M[r * kRotTile + c] += T[c * kRotTile + r];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < cnt; idx += blockDim.x) {
      const int r = idx / W, c = idx - r * W;
      const ModeMeta mb = meta[g0 + c];
      const int cb = mb.blk_off - g0, d = mb.blk_dim, cp = g0 + c - mb.blk_off;
      const double *SB = Ss + mb.sym_off + cp * d;
      double t = 0.0;
      for (int l = 0; l < d; ++l) t = fma(M[r * kRotTile + cb + l], SB[l], t);
      T[r * kRotTile + c] = t;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int idx = threadIdx.x + 256 * j;
      if (idx < cnt) {
        const int r = idx / W, c = idx - r * W;
        double t = 0.0;
        for (int k = 0; k < dA; ++k) t = fma(SA[r * dA + k], T[k * kRotTile + c], t);
        acc[j] += t;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int idx = threadIdx.x + 256 * j;
    if (idx < cnt) {
      const int r = idx / W, c = idx - r * W;
      d2v[(size_t)(offA + r) * npd + g0 + c] = acc[j];
    }
  }
}

// r_s = S_s^T R1;  G[s][0][a] = f_ups[a] r_s[a],  G[s][1][a] = r_s[a] / f_psi[a]
__global__ void k_rot_vec(const double *__restrict__ R1p, double *__restrict__ G,
                          const double *__restrict__ S, const ModeMeta *__restrict__ meta,
                          const double *__restrict__ fups, const double *__restrict__ fpsi, int n,
                          int npd, int sym_stride, Pass ps) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  int sl = blockIdx.y;
  if (a >= npd) return;
  double g1 = 0.0, g2 = 0.0;
  R1p += (size_t)ps.chain(sl) * npd;
  if (a < n) {
    ModeMeta ma = meta[a];
    const int ap = a - ma.blk_off, d = ma.blk_dim;
    const double *Sa = S + (size_t)ps.sym(sl) * sym_stride + ma.sym_off;
    double r = 0.0;
    for (int k = 0; k < d; ++k) r = fma(Sa[k * d + ap], R1p[ma.blk_off + k], r);
    g1 = fups[a] * r;
    g2 = r / fpsi[a];
  }
  G[(size_t)sl * 2 * npd + a] = g1;
  G[(size_t)sl * 2 * npd + npd + a] = g2;
}

// V[s][j] = sum_i YT[j][i] e1[s][i] + XtT[j][i] e2[s][i]
// Block: 4 waves x 4 rows each = 16 rows j; lanes stride over configs i.
// f_acc[chain][a] += sum over the chain's slots of sum_k S_s[a'][k] V[slot][off_A + k]
__global__ void k_fav_rot_acc(const double *__restrict__ V, double *__restrict__ f_acc,
                              const double *__restrict__ S, const ModeMeta *__restrict__ meta, int n,
                              int npd, int sym_stride, Pass ps) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (a >= n) return;
  f_acc += (size_t)c * npd;
  ModeMeta ma = meta[a];
  const int ap = a - ma.blk_off, d = ma.blk_dim;
  double acc = f_acc[a];
  for (int t = 0; t < ps.ns; ++t) {
    const int sl = c * ps.ns + t;
    const double *Sa = S + (size_t)ps.sym(sl) * sym_stride + ma.sym_off + ap * d;
    const double *v = V + (size_t)sl * npd + ma.blk_off;
    double part = 0.0;
    for (int k = 0; k < d; ++k) part = fma(Sa[k], v[k], part);
    acc += part;
  }
  f_acc[a] = acc;
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

void launch_rot1_fused(const Ensemble &e, const Pass &ps, int prec, hipStream_t st) {
  const size_t plane = (size_t)e.s_batch * e.npd * e.npd;
  dim3 grid(e.n_blocks, e.n_groups, ps.slots());
  hipLaunchKernelGGL(k_rot1_fused, grid, dim3(256), 0, st, e.Ysym,
                     prec == kFP64 ? e.Ys : (double *)nullptr, e.ys_lp(prec), prec, plane, e.S,
                     e.dblocks, e.dgroups, e.meta, e.npd, e.sym_stride, ps);
}

void launch_rot2_fused(const Ensemble &e, const Pass &ps, hipStream_t st) {
  dim3 grid(e.n_blocks, e.n_groups, ps.nc);
  hipLaunchKernelGGL(k_rot2_fused, grid, dim3(256), 0, st, e.B, e.d2v_acc, e.S, e.dblocks,
                     e.dgroups, e.meta, e.npd, e.sym_stride, ps);
}

void launch_rot_vec(const Ensemble &e, const Pass &ps, hipStream_t st) {
  dim3 grid((e.npd + 255) / 256, ps.slots());
  hipLaunchKernelGGL(k_rot_vec, grid, dim3(256), 0, st, e.R1p, e.G, e.S, e.meta, e.fups, e.fpsi,
                     e.n, e.npd, e.sym_stride, ps);
}

void launch_fav_rot_acc(const Ensemble &e, const Pass &ps, hipStream_t st) {
  hipLaunchKernelGGL(k_fav_rot_acc, dim3((e.n + 255) / 256, ps.nc), dim3(256), 0, st, e.V, e.f_acc,
                     e.S, e.meta, e.n, e.npd, e.sym_stride, ps);
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
