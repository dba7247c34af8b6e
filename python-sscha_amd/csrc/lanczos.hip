// Device-resident biconjugate / symmetric Lanczos on several chains, one hipGraph per step.
//
// The recurrence is the reference's run_FT (Modules/DynamicalLanczos.py:5843-6372): with
// p_norm_k = s_norm_k / c_(k-1),
//   a_k   = p_norm_k <p_k, L q_k>
//   r_k   = L q_k - a_k q_k - c_(k-1) q_(k-1)
//   s_k   = L^T p_k - a_k p_k - b_(k-1) (p_norm_(k-1) / p_norm_k) p_(k-1)
//   b_k   = |r_k|,  s_norm_(k+1) = |s_k| p_norm_k,  c_k = <s_k/|s_k|, r_k/b_k> s_norm_(k+1)
//   q_(k+1) = r_k / b_k,  p_(k+1) = s_k / |s_k|
// (Wigner symmetric form: p = q, L^T p = L q, dot products weighted by the mask that counts
// off-diagonal pair entries twice, Py:5790-5839.) All scalars stay in FP64 device memory;
// each step appends a_k, b_k, c_k, s_norm_(k+1) to device arrays at a device step counter,
// so one captured graph serves every step. The host truncates at the first |b| or |c|
// below 1e-12 (the reference's stopping rule, Py:6144-6153).
//
// Optional full re-biorthogonalisation (full_gs): after q_(k+1) = r_k / b_k and
// p_(k+1) = s_k / |s_k| the new pair is projected, in two passes of classical two-sided
// Gram-Schmidt, against every stored pair j = 0..k with the duals pt_j = p_norm_j p_j:
//   q <- q - sum_j (pt_j . q) q_j,   p <- p - sum_j (q_j . p) pt_j
// then normalised, and p_norm_(k+1) = 1 / (p . q) (s_norm_(k+1) = c_k p_norm_(k+1)), as
// run_FT(full_gs=True) does on the host. The number of stored pairs is read from the device
// step counter, so the projection sits inside the same captured graph; every coefficient is
// a fixed-order FP64 sum (partials over fixed vector chunks, then over the chunks in order).
#include <string>
#include <vector>

#include "tdscha_internal.h"
#include "../../include/tdscha_hip.h"

namespace tdscha {
namespace {

constexpr int kDotBlocks = 256;  // fixed partial count: reductions have a fixed order
constexpr int kGsBlocks = 128;   // vector chunks of the Gram-Schmidt projections

// state per chain: [0] p_norm, [1] old p_norm, [2] c_(k-1), [3] b_(k-1), [4] a_k
enum { ST_PN = 0, ST_OPN, ST_CL, ST_BL, ST_A, ST_N };

struct ChainVecs {
  double *q, *p, *qp, *pp;   // current and previous vectors
  double *Lq, *pL;           // L q and L^T p (then overwritten by r and s)
};

__device__ inline double block_sum(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
  __syncthreads();
  return t;  // valid in thread 0
}

// partial[c][blk] = sum over the block's strided range of x . y (* mask)
__global__ __launch_bounds__(256) void k_dot_partial(const double *const *x, const double *const *y,
                                                     const double *__restrict__ mask, long n,
                                                     double *partial) {
  __shared__ double sh[4];
  const int c = blockIdx.y;
  const double *xc = x[c], *yc = y[c];
  double acc = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    acc = fma(xc[i], mask ? yc[i] * mask[i] : yc[i], acc);
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[c * kDotBlocks + blockIdx.x] = acc;
}

__device__ inline double sum_partials(const double *part, double *sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < kDotBlocks; i += blockDim.x) v += part[i];
  return block_sum(v, sh);
}

// a_k = p_norm <p, Lq>; record it.
__global__ __launch_bounds__(256) void k_finish_a(const double *partial, double *state,
                                                  double *a_all, const int *step, int max_steps) {
  __shared__ double sh[4];
  const int c = blockIdx.x;
  const double d = sum_partials(partial + c * kDotBlocks, sh);
  if (threadIdx.x == 0) {
    double *st = state + c * ST_N;
    const double a = d * st[ST_PN];
    st[ST_A] = a;
    const int k = *step;
    if (k < max_steps) a_all[c * max_steps + k] = a;
  }
}

// r = Lq - a q - c_last qp ; s = pL - a p - b_last (opn / pn) pp  (in place), and the
// partial sums of r.r, s.s, s.r
__global__ __launch_bounds__(256) void k_residuals(const ChainVecs *vecs, const double *state,
                                                   const double *__restrict__ mask, long n,
                                                   int symmetric, double *partial) {
  __shared__ double sh[4];
  const int c = blockIdx.y;
  const ChainVecs v = vecs[c];
  const double *st = state + c * ST_N;
  const double a = st[ST_A], cl = st[ST_CL], bl = st[ST_BL];
  const double ratio = st[ST_PN] != 0.0 ? st[ST_OPN] / st[ST_PN] : 0.0;
  double rr = 0.0, ss = 0.0, sr = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double r = v.Lq[i] - a * v.q[i] - cl * v.qp[i];
    const double s = symmetric ? r : v.pL[i] - a * v.p[i] - bl * ratio * v.pp[i];
    v.Lq[i] = r;
    if (!symmetric) v.pL[i] = s;
    const double m = mask ? mask[i] : 1.0;
    rr = fma(r, r * m, rr);
    ss = fma(s, s * m, ss);
    sr = fma(s, r * m, sr);
  }
  rr = block_sum(rr, sh);
  ss = block_sum(ss, sh);
  sr = block_sum(sr, sh);
  if (threadIdx.x == 0) {
    partial[(c * 3 + 0) * kDotBlocks + blockIdx.x] = rr;
    partial[(c * 3 + 1) * kDotBlocks + blockIdx.x] = ss;
    partial[(c * 3 + 2) * kDotBlocks + blockIdx.x] = sr;
  }
}

// b, c, s_norm; record them and advance the per-chain state. scal[c] = (1/b, 1/|s|).
__global__ __launch_bounds__(256) void k_finish_bc(const double *partial, double *state,
                                                   double *b_all, double *c_all, double *s_all,
                                                   double *scal, int *step, int n_chains,
                                                   int max_steps) {
  __shared__ double sh[4];
  const int c = blockIdx.x;
  const double rr = sum_partials(partial + (c * 3 + 0) * kDotBlocks, sh);
  const double ss = sum_partials(partial + (c * 3 + 1) * kDotBlocks, sh);
  const double sr = sum_partials(partial + (c * 3 + 2) * kDotBlocks, sh);
  if (threadIdx.x == 0) {
    double *st = state + c * ST_N;
    const double pn = st[ST_PN];
    const double s_raw = sqrt(ss);
    const double s_norm = s_raw * pn;
    const double b = sqrt(rr);
    const double cc = (sr / (s_raw * b)) * s_norm;
    const int k = *step;
    if (k < max_steps) {
      b_all[c * max_steps + k] = b;
      c_all[c * max_steps + k] = cc;
      s_all[c * max_steps + k] = s_norm;
    }
    st[ST_OPN] = pn;
    st[ST_PN] = s_norm / cc;
    st[ST_CL] = cc;
    st[ST_BL] = b;
    scal[2 * c] = 1.0 / b;
    scal[2 * c + 1] = 1.0 / s_raw;
  }
}

__global__ void k_advance_step(int *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// qp <- q, q <- r / b, pp <- p, p <- s / |s|; optionally store the new pair in the basis
__global__ __launch_bounds__(256) void k_shift(const ChainVecs *vecs, const double *scal, long n,
                                               int symmetric, double *basisQ, double *basisP,
                                               const int *step, int max_steps) {
  const int c = blockIdx.y;
  const ChainVecs v = vecs[c];
  const double ib = scal[2 * c], is = scal[2 * c + 1];
  const int k = *step;  // already advanced: index of the new vector
  double *bq = (basisQ && k <= max_steps) ? basisQ + ((size_t)c * (max_steps + 1) + k) * n : nullptr;
  double *bp = (basisP && k <= max_steps) ? basisP + ((size_t)c * (max_steps + 1) + k) * n : nullptr;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double nq = v.Lq[i] * ib;
    const double np = symmetric ? nq : v.pL[i] * is;
    v.qp[i] = v.q[i];
    v.q[i] = nq;
    if (!symmetric) {
      v.pp[i] = v.p[i];
      v.p[i] = np;
    }
    if (bq) bq[i] = nq;
    if (bp) bp[i] = np;
  }
}

// ---- full re-biorthogonalisation ------------------------------------------------------
// Stored pairs of chain c: Q_j = basisQ[c][j], P_j = basisP[c][j] (= Q_j in the symmetric
// form), j < K = min(*step, max_steps + 1); the new pair is in v.q, v.p.

// partial[c][j][blk] = (P_j . q, Q_j . p) over chunk blk of the vector (mask-weighted)
__global__ __launch_bounds__(256) void k_gs_dots(const ChainVecs *vecs, const double *basisQ,
                                                 const double *basisP,
                                                 const double *__restrict__ mask, long n,
                                                 int symmetric, const int *step, int max_steps,
                                                 double *partial) {
  __shared__ double sh[4];
  const int c = blockIdx.y;
  const ChainVecs v = vecs[c];
  const int K = min(*step, max_steps + 1);
  const size_t cb = (size_t)c * (max_steps + 1);
  const long chunk = (n + kGsBlocks - 1) / kGsBlocks;
  const long i0 = (long)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
  for (int j = 0; j < K; ++j) {
    const double *Qj = basisQ + (cb + j) * n;
    const double *Pj = symmetric ? Qj : basisP + (cb + j) * n;
    double hq = 0.0, hp = 0.0;
    for (long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
      const double m = mask ? mask[i] : 1.0;
      hq = fma(Pj[i], v.q[i] * m, hq);
      if (!symmetric) hp = fma(Qj[i], v.p[i] * m, hp);
    }
    hq = block_sum(hq, sh);
    if (!symmetric) hp = block_sum(hp, sh);
    if (threadIdx.x == 0) {
      double *pc = partial + ((size_t)c * (max_steps + 1) + j) * 2 * kGsBlocks;
      pc[blockIdx.x] = hq;
      pc[kGsBlocks + blockIdx.x] = hp;
    }
  }
}

// coef[c][j] = (p_norm_j * sum_blk P_j . q, sum_blk Q_j . p), the chunks added in order
__global__ __launch_bounds__(256) void k_gs_coeffs(const double *partial, const double *pn_all,
                                                   const int *step, int max_steps, double *coef) {
  const int c = blockIdx.x;
  const int K = min(*step, max_steps + 1);
  for (int j = threadIdx.x; j < K; j += blockDim.x) {
    const size_t cj = (size_t)c * (max_steps + 1) + j;
    const double *pc = partial + cj * 2 * kGsBlocks;
    double hq = 0.0, hp = 0.0;
    for (int b = 0; b < kGsBlocks; ++b) {
      hq += pc[b];
      hp += pc[kGsBlocks + b];
    }
    coef[2 * cj] = pn_all[cj] * hq;
    coef[2 * cj + 1] = hp;
  }
}

// q -= sum_j coef_q[j] Q_j ;  p -= sum_j coef_p[j] p_norm_j P_j  (j in order)
__global__ __launch_bounds__(256) void k_gs_update(const ChainVecs *vecs, const double *basisQ,
                                                   const double *basisP, const double *coef,
                                                   const double *pn_all, long n, int symmetric,
                                                   const int *step, int max_steps) {
  const int c = blockIdx.y;
  const ChainVecs v = vecs[c];
  const int K = min(*step, max_steps + 1);
  const size_t cb = (size_t)c * (max_steps + 1);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double q = v.q[i];
    double p = symmetric ? 0.0 : v.p[i];
    for (int j = 0; j < K; ++j) {
      const double *Qj = basisQ + (cb + j) * n;
      q = fma(-coef[2 * (cb + j)], Qj[i], q);
      if (!symmetric) {
        const double *Pj = basisP + (cb + j) * n;
        p = fma(-coef[2 * (cb + j) + 1] * pn_all[cb + j], Pj[i], p);
      }
    }
    v.q[i] = q;
    if (!symmetric) v.p[i] = p;
  }
}

// partial sums of q.q, p.p, p.q (mask-weighted) after the projection
__global__ __launch_bounds__(256) void k_gs_norms(const ChainVecs *vecs,
                                                  const double *__restrict__ mask, long n,
                                                  int symmetric, double *partial) {
  __shared__ double sh[4];
  const int c = blockIdx.y;
  const ChainVecs v = vecs[c];
  double qq = 0.0, pp = 0.0, pq = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double m = mask ? mask[i] : 1.0;
    const double q = v.q[i];
    const double p = symmetric ? q : v.p[i];
    qq = fma(q, q * m, qq);
    pp = fma(p, p * m, pp);
    pq = fma(p, q * m, pq);
  }
  qq = block_sum(qq, sh);
  pp = block_sum(pp, sh);
  pq = block_sum(pq, sh);
  if (threadIdx.x == 0) {
    partial[(c * 3 + 0) * kDotBlocks + blockIdx.x] = qq;
    partial[(c * 3 + 1) * kDotBlocks + blockIdx.x] = pp;
    partial[(c * 3 + 2) * kDotBlocks + blockIdx.x] = pq;
  }
}

// |q|, |p| and the new dual scale p_norm_(k+1) = |q| |p| / (p . q) of the normalised pair;
// s_norm_(k+1) = c_k p_norm_(k+1) replaces the value k_finish_bc recorded. gsn[k] holds
// min(|q|, |p|) (a linearly dependent new vector: the host ends the run there, as run_FT's
// next_converged does).
__global__ __launch_bounds__(256) void k_gs_finish(const double *partial, double *state,
                                                   double *pn_all, double *s_all, double *gsn,
                                                   double *scal, const int *step, int max_steps) {
  __shared__ double sh[4];
  const int c = blockIdx.x;
  const double qq = sum_partials(partial + (c * 3 + 0) * kDotBlocks, sh);
  const double pp = sum_partials(partial + (c * 3 + 1) * kDotBlocks, sh);
  const double pq = sum_partials(partial + (c * 3 + 2) * kDotBlocks, sh);
  if (threadIdx.x == 0) {
    double *st = state + c * ST_N;
    const double nq = sqrt(qq), np_ = sqrt(pp);
    const double pn = (nq * np_) / pq;
    st[ST_PN] = pn;
    const int k1 = *step;  // index of the new pair
    if (k1 <= max_steps) pn_all[(size_t)c * (max_steps + 1) + k1] = pn;
    if (k1 - 1 < max_steps) {
      s_all[c * max_steps + k1 - 1] = st[ST_CL] * pn;
      gsn[c * max_steps + k1 - 1] = fmin(nq, np_);
    }
    scal[2 * c] = 1.0 / nq;
    scal[2 * c + 1] = 1.0 / np_;
  }
}

// q /= |q|, p /= |p|, stored as pair k + 1 of the basis
__global__ __launch_bounds__(256) void k_gs_scale(const ChainVecs *vecs, const double *scal, long n,
                                                  int symmetric, double *basisQ, double *basisP,
                                                  const int *step, int max_steps) {
  const int c = blockIdx.y;
  const ChainVecs v = vecs[c];
  const double iq = scal[2 * c], ip = scal[2 * c + 1];
  const int k = *step;
  double *bq = k <= max_steps ? basisQ + ((size_t)c * (max_steps + 1) + k) * n : nullptr;
  double *bp = (basisP && k <= max_steps) ? basisP + ((size_t)c * (max_steps + 1) + k) * n : nullptr;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double q = v.q[i] * iq;
    v.q[i] = q;
    if (bq) bq[i] = q;
    if (!symmetric) {
      const double p = v.p[i] * ip;
      v.p[i] = p;
      if (bp) bp[i] = p;
    }
  }
}

// q = p = psi / sqrt(psi . psi (* mask)), previous vectors zero, state reset
__global__ __launch_bounds__(256) void k_start(const ChainVecs *vecs, const double *const *psi,
                                               const double *partial, long n, int symmetric,
                                               double *state, double *basisQ, double *basisP,
                                               double *pn_all, int max_steps) {
  __shared__ double sh[4];
  __shared__ double inv;
  const int c = blockIdx.y;
  const double d = sum_partials(partial + c * kDotBlocks, sh);
  if (threadIdx.x == 0) inv = 1.0 / sqrt(d);
  __syncthreads();
  const ChainVecs v = vecs[c];
  double *bq = basisQ ? basisQ + (size_t)c * (max_steps + 1) * n : nullptr;
  double *bp = basisP ? basisP + (size_t)c * (max_steps + 1) * n : nullptr;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double x = psi[c][i] * inv;
    v.q[i] = x;
    v.qp[i] = 0.0;
    if (!symmetric) {
      v.p[i] = x;
      v.pp[i] = 0.0;
    }
    if (bq) bq[i] = x;
    if (bp) bp[i] = x;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double *st = state + c * ST_N;
    if (pn_all) pn_all[(size_t)c * (max_steps + 1)] = 1.0;
    st[ST_PN] = 1.0;   // s_norm_0 / c_(-1) = 1
    st[ST_OPN] = 0.0;
    st[ST_CL] = 0.0;
    st[ST_BL] = 0.0;
    st[ST_A] = 0.0;
  }
}

__global__ void k_zero_int(int *p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

}  // namespace
}  // namespace tdscha

using namespace tdscha;

struct tdscha_lanczos {
  tdscha_ensemble *ens = nullptr;
  int prec = 0, nc = 0, symmetric = 0, max_steps = 0, keep_basis = 0, device = 0;
  long n_psi = 0;
  std::vector<int> flags;
  std::vector<double> shifts;
  std::vector<void *> allocs;
  double *vec = nullptr;         // 6 vectors per chain
  ChainVecs *dvecs = nullptr;
  const double **dpsi_in = nullptr, **dpsi_p = nullptr;  // pointer tables for the dots
  const double **dpsi0 = nullptr;                          // starting vectors
  double *mask = nullptr;
  double *partial = nullptr, *state = nullptr, *scal = nullptr;
  double *a_all = nullptr, *b_all = nullptr, *c_all = nullptr, *s_all = nullptr;
  double *basisQ = nullptr, *basisP = nullptr;
  int *step = nullptr;
  // full re-biorthogonalisation: dual scales p_norm_j [C x (max_steps + 1)], projection
  // partials [C x (max_steps + 1) x 2 x kGsBlocks] and coefficients [C x (max_steps + 1) x 2],
  // norm of each new pair after the projection [C x max_steps]
  int full_gs = 0;
  double *pn_all = nullptr, *gs_part = nullptr, *gs_coef = nullptr, *gsn = nullptr;
  long host_steps = 0;   // steps enqueued since the last start
  // zone-centre operator: the translation average applied to every L q (and, transposed in
  // the packed metric for the standard representation, to every L^T p) inside the step
  tdscha_projector *proj = nullptr;
  std::vector<const double *> apply_in;
  std::vector<double *> apply_out;
  std::vector<int> apply_flags;
  std::vector<double> apply_shifts;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
};

namespace {

int lz_alloc(tdscha_lanczos *z, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 8);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return 11;
  }
  z->allocs.push_back(*p);
  return 0;
}

// One Lanczos step enqueued on st (the body that is captured in the graph).
int enqueue_step(tdscha_lanczos *z, hipStream_t st) {
  const long n = z->n_psi;
  const int C = z->nc;
  const int napply = z->symmetric ? C : 2 * C;
  int rc = tdscha_apply_full_L_batched(z->ens, z->prec, napply, z->apply_flags.data(),
                                       z->apply_shifts.data(), z->apply_in.data(),
                                       z->apply_out.data(), st);
  if (rc) return rc;
  if (z->proj) {
    for (int k = 0; k < napply; ++k) {
      const int fl = z->apply_flags[k];
      const int adjoint = (fl & TDSCHA_OP_TRANSPOSE) && !(fl & TDSCHA_OP_WIGNER);
      rc = projector_apply(z->proj, adjoint, z->apply_out[k], z->apply_out[k], st);
      if (rc) return rc;
    }
  }
  // a = p_norm <p, L q>  (symmetric: <q, L q>_mask)
  hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks, C), dim3(256), 0, st,
                     (const double *const *)z->dpsi_p, (const double *const *)(z->dpsi_in + C),
                     z->symmetric ? z->mask : (const double *)nullptr, n, z->partial);
  hipLaunchKernelGGL(k_finish_a, dim3(C), dim3(256), 0, st, z->partial, z->state, z->a_all, z->step,
                     z->max_steps);
  hipLaunchKernelGGL(k_residuals, dim3(kDotBlocks, C), dim3(256), 0, st, z->dvecs, z->state,
                     z->symmetric ? z->mask : (const double *)nullptr, n, z->symmetric, z->partial);
  hipLaunchKernelGGL(k_finish_bc, dim3(C), dim3(256), 0, st, z->partial, z->state, z->b_all,
                     z->c_all, z->s_all, z->scal, z->step, C, z->max_steps);
  hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(64), 0, st, z->step);
  hipLaunchKernelGGL(k_shift, dim3(kDotBlocks, C), dim3(256), 0, st, z->dvecs, z->scal, n,
                     z->symmetric, z->full_gs ? (double *)nullptr : z->basisQ,
                     z->full_gs ? (double *)nullptr : z->basisP, z->step, z->max_steps);
  if (z->full_gs) {
    const double *m = z->symmetric ? z->mask : (const double *)nullptr;
    for (int pass = 0; pass < 2; ++pass) {
      hipLaunchKernelGGL(k_gs_dots, dim3(kGsBlocks, C), dim3(256), 0, st, z->dvecs, z->basisQ,
                         z->basisP, m, n, z->symmetric, z->step, z->max_steps, z->gs_part);
      hipLaunchKernelGGL(k_gs_coeffs, dim3(C), dim3(256), 0, st, z->gs_part, z->pn_all, z->step,
                         z->max_steps, z->gs_coef);
      hipLaunchKernelGGL(k_gs_update, dim3(kDotBlocks, C), dim3(256), 0, st, z->dvecs, z->basisQ,
                         z->basisP, z->gs_coef, z->pn_all, n, z->symmetric, z->step,
                         z->max_steps);
    }
    hipLaunchKernelGGL(k_gs_norms, dim3(kDotBlocks, C), dim3(256), 0, st, z->dvecs, m, n,
                       z->symmetric, z->partial);
    hipLaunchKernelGGL(k_gs_finish, dim3(C), dim3(256), 0, st, z->partial, z->state, z->pn_all,
                       z->s_all, z->gsn, z->scal, z->step, z->max_steps);
    hipLaunchKernelGGL(k_gs_scale, dim3(kDotBlocks, C), dim3(256), 0, st, z->dvecs, z->scal, n,
                       z->symmetric, z->basisQ, z->basisP, z->step, z->max_steps);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("Lanczos step launch failed: ") + hipGetErrorString(e));
    return 12;
  }
  return 0;
}

}  // namespace

extern "C" {

int tdscha_lanczos_create(tdscha_lanczos **out, tdscha_ensemble *ens, int precision, int n_chains,
                          const int *flags, const double *shifts, int symmetric, int max_steps,
                          int keep_basis, int full_gs, const double *mask) {
  if (!out || !ens || !flags || !shifts || n_chains < 1 || max_steps < 1) {
    set_error("bad argument to tdscha_lanczos_create");
    return 1;
  }
  if (2 * n_chains > TDSCHA_MAX_CHAINS) {
    set_error("too many chains");
    return 1;
  }
  if (symmetric && !mask) {
    set_error("the symmetric (Wigner) Lanczos needs the dot-product mask");
    return 1;
  }
  int dims[6];
  tdscha_ensemble_dims(ens, dims);
  // every buffer of the batched application exists before a step is captured
  int rc0 = tdscha_ensemble_reserve(ens, symmetric ? n_chains : 2 * n_chains);
  if (rc0) return rc0;
  const long n = dims[0];
  tdscha_lanczos *z = new tdscha_lanczos();
  z->ens = ens;
  z->prec = precision;
  z->nc = n_chains;
  z->symmetric = symmetric;
  z->max_steps = max_steps;
  z->full_gs = full_gs ? 1 : 0;
  z->keep_basis = (keep_basis || full_gs) ? 1 : 0;  // the projection reads the stored pairs
  z->n_psi = n + n * (n + 1);
  z->flags.assign(flags, flags + n_chains);
  z->shifts.assign(shifts, shifts + n_chains);
  (void)hipGetDevice(&z->device);
  const long N = z->n_psi;
  const int C = n_chains;
  int rc = 0;
  rc |= lz_alloc(z, (void **)&z->vec, sizeof(double) * 6 * C * N);
  rc |= lz_alloc(z, (void **)&z->dvecs, sizeof(ChainVecs) * C);
  rc |= lz_alloc(z, (void **)&z->dpsi_in, sizeof(double *) * 2 * C);
  rc |= lz_alloc(z, (void **)&z->dpsi_p, sizeof(double *) * C);
  rc |= lz_alloc(z, (void **)&z->dpsi0, sizeof(double *) * C);
  rc |= lz_alloc(z, (void **)&z->partial, sizeof(double) * 3 * C * kDotBlocks);
  rc |= lz_alloc(z, (void **)&z->state, sizeof(double) * ST_N * C);
  rc |= lz_alloc(z, (void **)&z->scal, sizeof(double) * 2 * C);
  rc |= lz_alloc(z, (void **)&z->a_all, sizeof(double) * C * max_steps);
  rc |= lz_alloc(z, (void **)&z->b_all, sizeof(double) * C * max_steps);
  rc |= lz_alloc(z, (void **)&z->c_all, sizeof(double) * C * max_steps);
  rc |= lz_alloc(z, (void **)&z->s_all, sizeof(double) * C * max_steps);
  rc |= lz_alloc(z, (void **)&z->step, sizeof(int));
  if (symmetric) rc |= lz_alloc(z, (void **)&z->mask, sizeof(double) * N);
  if (z->keep_basis) {
    rc |= lz_alloc(z, (void **)&z->basisQ, sizeof(double) * C * (max_steps + 1) * N);
    if (!symmetric) rc |= lz_alloc(z, (void **)&z->basisP, sizeof(double) * C * (max_steps + 1) * N);
  }
  if (z->full_gs) {
    const size_t nv = (size_t)C * (max_steps + 1);
    rc |= lz_alloc(z, (void **)&z->pn_all, sizeof(double) * nv);
    rc |= lz_alloc(z, (void **)&z->gs_part, sizeof(double) * nv * 2 * kGsBlocks);
    rc |= lz_alloc(z, (void **)&z->gs_coef, sizeof(double) * nv * 2);
    rc |= lz_alloc(z, (void **)&z->gsn, sizeof(double) * C * max_steps);
  }
  if (rc) { tdscha_lanczos_destroy(z); return rc; }
  std::vector<ChainVecs> hv(C);
  std::vector<const double *> hin(2 * C), hp(C);
  z->apply_in.resize(symmetric ? C : 2 * C);
  z->apply_out.resize(symmetric ? C : 2 * C);
  z->apply_flags.resize(symmetric ? C : 2 * C);
  z->apply_shifts.resize(symmetric ? C : 2 * C);
  for (int c = 0; c < C; ++c) {
    double *base = z->vec + (size_t)6 * c * N;
    hv[c] = ChainVecs{base, symmetric ? base : base + N, base + 2 * N, symmetric ? base + 2 * N : base + 3 * N,
                      base + 4 * N, base + 5 * N};
    z->apply_in[c] = hv[c].q;
    z->apply_out[c] = hv[c].Lq;
    z->apply_flags[c] = flags[c] & ~TDSCHA_OP_TRANSPOSE;
    z->apply_shifts[c] = shifts[c];
    if (!symmetric) {
      z->apply_in[C + c] = hv[c].p;
      z->apply_out[C + c] = hv[c].pL;
      z->apply_flags[C + c] = flags[c] | TDSCHA_OP_TRANSPOSE;
      z->apply_shifts[C + c] = shifts[c];
    }
    hp[c] = hv[c].p;
    hin[c] = hv[c].q;
    hin[C + c] = hv[c].Lq;
  }
  hipError_t e = hipMemcpy(z->dvecs, hv.data(), sizeof(ChainVecs) * C, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(z->dpsi_in, hin.data(), sizeof(double *) * 2 * C, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(z->dpsi_p, hp.data(), sizeof(double *) * C, hipMemcpyHostToDevice);
  if (e == hipSuccess && symmetric) e = hipMemcpy(z->mask, mask, sizeof(double) * N, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("Lanczos set-up copy failed: ") + hipGetErrorString(e));
    tdscha_lanczos_destroy(z);
    return 10;
  }
  *out = z;
  return 0;
}

int tdscha_lanczos_set_projector(tdscha_lanczos *z, tdscha_projector *p) {
  if (!z) { set_error("NULL Lanczos"); return 1; }
  if (z->exec) {
    set_error("the projector must be set before the first graph is captured");
    return 1;
  }
  z->proj = p;
  return 0;
}

int tdscha_lanczos_destroy(tdscha_lanczos *z) {
  if (!z) return 0;
  if (z->exec) (void)hipGraphExecDestroy(z->exec);
  if (z->graph) (void)hipGraphDestroy(z->graph);
  if (z->graph_stream) (void)hipStreamDestroy(z->graph_stream);
  for (void *p : z->allocs) (void)hipFree(p);
  delete z;
  return 0;
}

int tdscha_lanczos_start(tdscha_lanczos *z, const double *const *psi0, void *stream) {
  if (!z || !psi0) { set_error("NULL argument"); return 1; }
  hipStream_t st = (hipStream_t)stream;
  const int C = z->nc;
  // the starting vectors are read through a device pointer table
  std::vector<const double *> hp(psi0, psi0 + C);
  const double **dpsi = z->dpsi0;
  z->host_steps = 0;
  if (hipMemcpy(dpsi, hp.data(), sizeof(double *) * C, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("pointer table copy failed");
    return 10;
  }
  hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks, C), dim3(256), 0, st,
                     (const double *const *)dpsi, (const double *const *)dpsi,
                     z->symmetric ? z->mask : (const double *)nullptr, z->n_psi, z->partial);
  hipLaunchKernelGGL(k_start, dim3(kDotBlocks, C), dim3(256), 0, st, z->dvecs,
                     (const double *const *)dpsi, z->partial, z->n_psi, z->symmetric, z->state,
                     z->basisQ, z->basisP, z->pn_all, z->max_steps);
  if (z->gsn)
    (void)hipMemsetAsync(z->gsn, 0x7f, sizeof(double) * C * z->max_steps, st);  // large, finite
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, z->step);
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    set_error(std::string("Lanczos start failed: ") + hipGetErrorString(e));
    return 10;
  }
  return 0;
}

int tdscha_lanczos_run(tdscha_lanczos *z, int n_steps, int use_graph, void *stream) {
  if (!z || n_steps < 0) { set_error("bad argument"); return 1; }
  if (z->host_steps + n_steps > z->max_steps) {
    set_error("more steps than max_steps since the last start (" +
              std::to_string(z->host_steps) + " + " + std::to_string(n_steps) + " > " +
              std::to_string(z->max_steps) + ")");
    return 1;
  }
  z->host_steps += n_steps;
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) {
    for (int i = 0; i < n_steps; ++i) {
      int rc = enqueue_step(z, st);
      if (rc) return rc;
    }
    return 0;
  }
  if (!z->exec) {
    // capture one step on a private stream (default priority), instantiate once
    if (hipStreamCreateWithFlags(&z->graph_stream, hipStreamNonBlocking) != hipSuccess) {
      set_error("stream creation failed");
      return 10;
    }
    const bool prof = profile_swap(false);  // no event records inside a captured graph
    hipError_t e = hipStreamBeginCapture(z->graph_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      profile_swap(prof);
      set_error(std::string("stream capture failed: ") + hipGetErrorString(e));
      return 10;
    }
    int rc = enqueue_step(z, z->graph_stream);
    e = hipStreamEndCapture(z->graph_stream, &z->graph);
    profile_swap(prof);
    if (rc) return rc;
    if (e != hipSuccess) {
      set_error(std::string("end capture failed: ") + hipGetErrorString(e));
      return 10;
    }
    e = hipGraphInstantiate(&z->exec, z->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      set_error(std::string("graph instantiation failed: ") + hipGetErrorString(e));
      return 10;
    }
  }
  for (int i = 0; i < n_steps; ++i) {
    hipError_t e = hipGraphLaunch(z->exec, st);
    if (e != hipSuccess) {
      set_error(std::string("graph launch failed: ") + hipGetErrorString(e));
      return 10;
    }
  }
  return 0;
}

// ---- several Lanczos runs (one per ensemble) in one graph per step --------------------
// The step of every member is captured on a branch of its own (fork / join through events
// on default-priority streams), so the kernels of a small system can run beside those of a
// large one; one graph launch advances every member by one step.
struct tdscha_lanczos_group {
  std::vector<tdscha_lanczos *> members;
  std::vector<hipStream_t> branches;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int tdscha_lanczos_group_destroy(tdscha_lanczos_group *g) {
  if (!g) return 0;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  for (hipStream_t s : g->branches) (void)hipStreamDestroy(s);
  for (hipEvent_t e : g->joins) (void)hipEventDestroy(e);
  if (g->fork) (void)hipEventDestroy(g->fork);
  if (g->cap) (void)hipStreamDestroy(g->cap);
  delete g;
  return 0;
}

int tdscha_lanczos_group_create(tdscha_lanczos_group **out, tdscha_lanczos *const *members,
                                int n_members) {
  if (!out || !members || n_members < 1) { set_error("bad argument"); return 1; }
  for (int i = 0; i < n_members; ++i) {
    if (!members[i]) { set_error("NULL member"); return 1; }
    for (int j = 0; j < i; ++j)
      if (members[j]->ens == members[i]->ens) {
        set_error("members of a group must use different ensembles (they share its workspace)");
        return 1;
      }
    if (members[i]->device != members[0]->device) {
      set_error("members of a group must be on one device");
      return 1;
    }
    for (int j = 0; j < i; ++j)
      if (members[i]->proj && members[j]->proj == members[i]->proj) {
        set_error("members of a group must use different projectors (they share its workspace)");
        return 1;
      }
  }
  tdscha_lanczos_group *g = new (std::nothrow) tdscha_lanczos_group();
  if (!g) { set_error("out of host memory"); return 11; }
  g->members.assign(members, members + n_members);
  auto fail = [&](const char *what, hipError_t e) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    tdscha_lanczos_group_destroy(g);
    return 10;
  };
  hipError_t e = hipSetDevice(members[0]->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&g->fork, hipEventDisableTiming);
  for (int i = 0; i < n_members && e == hipSuccess; ++i) {
    hipStream_t s = nullptr;
    hipEvent_t j = nullptr;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) g->branches.push_back(s);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
    if (e == hipSuccess) g->joins.push_back(j);
  }
  if (e != hipSuccess) return fail("stream / event creation failed", e);
  const bool prof = profile_swap(false);  // no timing events inside a captured graph
  e = hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    profile_swap(prof);
    return fail("stream capture failed", e);
  }
  int rc = 0;
  e = hipEventRecord(g->fork, g->cap);
  for (int i = 0; i < n_members && e == hipSuccess && !rc; ++i) {
    e = hipStreamWaitEvent(g->branches[i], g->fork, 0);
    if (e == hipSuccess) rc = enqueue_step(g->members[i], g->branches[i]);
    if (e == hipSuccess && !rc) e = hipEventRecord(g->joins[i], g->branches[i]);
    if (e == hipSuccess && !rc) e = hipStreamWaitEvent(g->cap, g->joins[i], 0);
  }
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(g->cap, &graph);
  profile_swap(prof);
  g->graph = graph;
  if (rc) { tdscha_lanczos_group_destroy(g); return rc; }
  if (e != hipSuccess) return fail("capturing the group step failed", e);
  if (e2 != hipSuccess) return fail("end capture failed", e2);
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return fail("graph instantiation failed", e);
  *out = g;
  return 0;
}

int tdscha_lanczos_group_run(tdscha_lanczos_group *g, int n_steps, void *stream) {
  if (!g || n_steps < 0) { set_error("bad argument"); return 1; }
  for (tdscha_lanczos *m : g->members)
    if (m->host_steps + n_steps > m->max_steps) {
      set_error("a member of the group would run past its max_steps");
      return 1;
    }
  for (tdscha_lanczos *m : g->members) m->host_steps += n_steps;
  for (int i = 0; i < n_steps; ++i) {
    hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
    if (e != hipSuccess) {
      set_error(std::string("graph launch failed: ") + hipGetErrorString(e));
      return 10;
    }
  }
  return 0;
}

int tdscha_lanczos_read(tdscha_lanczos *z, int chain, double *a, double *b, double *c,
                        double *s_norm, int *steps_done) {
  if (!z || chain < 0 || chain >= z->nc || !steps_done) { set_error("bad argument"); return 1; }
  int k = 0;
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(&k, z->step, sizeof(int), hipMemcpyDeviceToHost);
  if (k > z->max_steps) k = z->max_steps;
  const size_t off = (size_t)chain * z->max_steps, bytes = sizeof(double) * k;
  if (e == hipSuccess && a) e = hipMemcpy(a, z->a_all + off, bytes, hipMemcpyDeviceToHost);
  if (e == hipSuccess && b) e = hipMemcpy(b, z->b_all + off, bytes, hipMemcpyDeviceToHost);
  if (e == hipSuccess && c) e = hipMemcpy(c, z->c_all + off, bytes, hipMemcpyDeviceToHost);
  if (e == hipSuccess && s_norm) e = hipMemcpy(s_norm, z->s_all + off, bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_error(std::string("reading the coefficients failed: ") + hipGetErrorString(e));
    return 10;
  }
  *steps_done = k;
  return 0;
}

int tdscha_lanczos_read_gs(tdscha_lanczos *z, int chain, double *norms, int *steps_done) {
  if (!z || chain < 0 || chain >= z->nc || !norms || !steps_done) {
    set_error("bad argument");
    return 1;
  }
  if (!z->full_gs) { set_error("the Lanczos was created without full_gs"); return 1; }
  int k = 0;
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(&k, z->step, sizeof(int), hipMemcpyDeviceToHost);
  if (k > z->max_steps) k = z->max_steps;
  if (e == hipSuccess)
    e = hipMemcpy(norms, z->gsn + (size_t)chain * z->max_steps, sizeof(double) * k,
                  hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_error(std::string("reading the projection norms failed: ") + hipGetErrorString(e));
    return 10;
  }
  *steps_done = k;
  return 0;
}

int tdscha_lanczos_basis(tdscha_lanczos *z, int chain, const double **Q, const double **P) {
  if (!z || chain < 0 || chain >= z->nc) { set_error("bad argument"); return 1; }
  const size_t off = (size_t)chain * (z->max_steps + 1) * z->n_psi;
  if (Q) *Q = z->basisQ ? z->basisQ + off : nullptr;
  if (P) *P = z->basisP ? z->basisP + off : (z->symmetric && z->basisQ ? z->basisQ + off : nullptr);
  return 0;
}

}  // extern "C"
