// The TD-SCHA L operator around the ensemble kernel, on the device, in FP64:
//   psi -> (R1, Y1)            pre-transform   (Modules/DynamicalLanczos.py:2946-2993, 2713-2800)
//   (f_av, d2v) -> psi'        post-transform  (DynamicalLanczos.py:3200-3281)
//   harmonic part L1           elementwise     (DynamicalLanczos.py:2227-2385)
//   reverse / shift            (DynamicalLanczos.py:3553-3560)
// psi layout: [R1 (n) | Upsilon^(1) upper triangle, row-major | ReA^(1) upper triangle].
// Pair (i <= j) sits at p(i, j) = i n - i (i - 1) / 2 + (j - i) inside each triangle; the
// reference calls w_a = w[j] (column) and w_b = w[i] (row).
#include "tdscha_internal.h"

namespace tdscha {

namespace {

__device__ inline long pair_index(int i, int j, int n) {
  return (long)i * n - (long)i * (i - 1) / 2 + (j - i);
}

// Build the symmetric, padded Y1 handed to the ensemble kernel, and R1.
__global__ void k_unpack(const double *__restrict__ psi, double *__restrict__ Ysym,
                         double *__restrict__ R1p, const double *__restrict__ w,
                         const double *__restrict__ nocc, int n, int npd, int flags) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int a = blockIdx.y;
  if (b >= npd) return;
  if (a == 0) R1p[b] = (b < n) ? psi[b] : 0.0;
  double v = 0.0;
  if (a < n && b < n) {
    const int i = a < b ? a : b, j = a < b ? b : a;
    const long nw2 = (long)n * (n + 1) / 2;
    const long p = pair_index(i, j, n);
    const double yv = psi[n + p], rv = psi[n + nw2 + p];
    const double wa = w[j], wb = w[i], na = nocc[j], nb = nocc[i];
    if (flags & kOpWigner) {
      // alpha^(1) from (a', b'), DynamicalLanczos.py:2780-2800
      const double Xf = (1 + 2 * na) * (1 + 2 * nb) / 8;
      const double w2_on_X = (wa * wb) / Xf;
      const double chi_m = ((wa - wb) * (na - nb)) / (2 * wa * wb);
      const double chi_p = ((wa + wb) * (1 + na + nb)) / (2 * wa * wb);
      v = w2_on_X * sqrt(-0.5 * chi_m) * yv - w2_on_X * sqrt(0.5 * chi_p) * rv;
    } else if (flags & kOpTranspose) {
      // half off-diagonal, then Y1 coeff_Y + ReA1 coeff_RA (DynamicalLanczos.py:2976-2993)
      const double h = (i == j) ? 1.0 : 0.5;
      const double ywa = 2 * wa / (2 * na + 1), ywb = 2 * wb / (2 * nb + 1);
      const double rwa = 2 * wa * na * (na + 1) / (2 * na + 1);
      const double rwb = 2 * wb * nb * (nb + 1) / (2 * nb + 1);
      // coeff[a][b] with a = row index: 2 (Y_a Y_b^2 + Y_a^2 Y_b), symmetric in (a, b)
      const double cY = 2 * (ywb * ywa * ywa + ywb * ywb * ywa);
      const double cR = 2 * (ywb * rwa * ywa + ywb * rwb * ywa);
      v = h * yv * cY + h * rv * cR;
    } else {
      v = yv;
    }
  }
  Ysym[(size_t)a * npd + b] = v;
}

// out = [L1 psi + sign * anharmonic(f, d2v)] * (reverse ? -1 : 1) + shift * psi
__global__ void k_post(const double *__restrict__ psi, double *__restrict__ out,
                       const double *__restrict__ f_acc, const double *__restrict__ d2v_acc,
                       double inv_norm, const double *__restrict__ w,
                       const double *__restrict__ nocc, int n, int npd, int flags, double shift) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= n) return;
  const bool wig = flags & kOpWigner, tr = flags & kOpTranspose;
  const bool harm = !(flags & kOpIgnoreHarmonic), anh = !(flags & kOpIgnoreV3);
  const double rev = (flags & kOpReverse) ? -1.0 : 1.0;
  if (i == 0) {
    // R^(1) entry j
    const double pj = psi[j], wj = w[j];
    double v = 0.0;
    if (harm) v = wig ? -(pj * wj) * wj : (pj * wj) * wj;
    if (anh) v += (wig ? 1.0 : -1.0) * f_acc[j] * inv_norm;
    out[j] = rev * v + shift * pj;
  }
  if (j < i) return;
  const long nw2 = (long)n * (n + 1) / 2;
  const long p = pair_index(i, j, n);
  const long iy = n + p, ia = n + nw2 + p;
  const double y = psi[iy], r = psi[ia];
  const double wa = w[j], wb = w[i], na = nocc[j], nb = nocc[i];
  double oy = 0.0, oa = 0.0;
  if (harm) {
    if (!wig) {
      const double nn = (2 * na + 1) * (2 * nb + 1);
      const double X_ab = -wa * wa - wb * wb - (2 * wa * wb) / nn;
      const double Y_ab = -(8 * wa * wb) / nn;
      const double s = 2 * na * nb + na + nb;
      const double X1_ab = -s * (s + 1) * (2 * wa * wb) / nn;
      const double Y1_ab = -wa * wa - wb * wb + (2 * wa * wb) / nn;
      if (!tr) {
        oy = -X_ab * y - Y_ab * r;
        oa = -X1_ab * y - Y1_ab * r;
      } else {
        oy = -X_ab * y - X1_ab * r;
        oa = -Y_ab * y - Y1_ab * r;
      }
    } else {
      oy = -(wa * wa + wb * wb - 2.0 * wa * wb) * y;
      oa = -(wa * wa + wb * wb + 2.0 * wa * wb) * r;
    }
  }
  if (anh) {
    const double d = d2v_acc[(size_t)i * npd + j] * inv_norm;
    if (wig) {
      const double chi_m = ((wa - wb) * (na - nb)) / (2 * wa * wb);
      const double chi_p = ((wa + wb) * (1 + na + nb)) / (2 * wa * wb);
      oy += sqrt(-0.5 * chi_m) * d;
      oa += -sqrt(0.5 * chi_p) * d;
    } else {
      const double ywa = 2 * wa / (2 * na + 1), ywb = 2 * wb / (2 * nb + 1);
      if (!tr) {
        const double rwa = 2 * wa * na * (na + 1) / (2 * na + 1);
        const double rwb = 2 * wb * nb * (nb + 1) / (2 * nb + 1);
        oy -= d * ywb + d * ywa;
        oa -= d * rwb + d * rwa;
      } else {
        const double m = (i == j) ? 1.0 : 2.0;
        oy -= m * (0.5 * (1 / ywb) * d * (1 / ywa));
      }
    }
  }
  out[iy] = rev * oy + shift * y;
  out[ia] = rev * oa + shift * r;
}

// Adjoint seeds of the anharmonic part of k_post for a cotangent u:
//   u . post(f_acc, d2v_acc) = uR . f_acc + <U, d2v_acc>  (U symmetric, zero padded),
// with the sign, reverse and 1 / (N_eff N_sym) factors of k_post folded in.
__global__ void k_post_adj(const double *__restrict__ u, double *__restrict__ U,
                           double *__restrict__ uR, double inv_norm, const double *__restrict__ w,
                           const double *__restrict__ nocc, int n, int npd, int flags) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (b >= npd) return;
  const bool wig = flags & kOpWigner, tr = flags & kOpTranspose, anh = !(flags & kOpIgnoreV3);
  const double rev = (flags & kOpReverse) ? -1.0 : 1.0;
  if (a == 0) uR[b] = (anh && b < n) ? rev * (wig ? 1.0 : -1.0) * inv_norm * u[b] : 0.0;
  double v = 0.0;
  if (anh && a < n && b < n) {
    const int i = a < b ? a : b, j = a < b ? b : a;
    const long nw2 = (long)n * (n + 1) / 2;
    const long p = pair_index(i, j, n);
    const double uy = u[n + p], ua = u[n + nw2 + p];
    const double wa = w[j], wb = w[i], na = nocc[j], nb = nocc[i];
    double cy, ca;
    if (wig) {
      const double chi_m = ((wa - wb) * (na - nb)) / (2 * wa * wb);
      const double chi_p = ((wa + wb) * (1 + na + nb)) / (2 * wa * wb);
      cy = sqrt(-0.5 * chi_m);
      ca = -sqrt(0.5 * chi_p);
    } else {
      const double ywa = 2 * wa / (2 * na + 1), ywb = 2 * wb / (2 * nb + 1);
      if (!tr) {
        const double rwa = 2 * wa * na * (na + 1) / (2 * na + 1);
        const double rwb = 2 * wb * nb * (nb + 1) / (2 * nb + 1);
        cy = -(ywb + ywa);
        ca = -(rwb + rwa);
      } else {
        const double m = (i == j) ? 1.0 : 2.0;
        cy = -m * (0.5 * (1 / ywb) * (1 / ywa));
        ca = 0.0;
      }
    }
    v = rev * inv_norm * (uy * cy + ua * ca);
    if (i != j) v *= 0.5;   // the pair's entry is shared by (i, j) and (j, i)
  }
  U[(size_t)a * npd + b] = v;
}

}  // namespace

void launch_post_adj(const Ensemble &e, const double *u, double *U, double *uR, int flags,
                     hipStream_t st) {
  double inv = 1.0 / (e.N_eff * (double)e.N_sym);
  dim3 grid((e.npd + 255) / 256, e.npd);
  hipLaunchKernelGGL(k_post_adj, grid, dim3(256), 0, st, u, U, uR, inv, e.w, e.nocc, e.n, e.npd,
                     flags);
}

void launch_unpack(const Ensemble &e, int chain, const double *psi, int flags, hipStream_t st) {
  dim3 grid((e.npd + 255) / 256, e.npd);
  hipLaunchKernelGGL(k_unpack, grid, dim3(256), 0, st, psi, e.Ysym + (size_t)chain * e.npd * e.npd,
                     e.R1p + (size_t)chain * e.npd, e.w, e.nocc, e.n, e.npd, flags);
}

void launch_post(const Ensemble &e, int chain, const double *psi, double *out, int flags,
                 double shift, hipStream_t st) {
  double inv = 1.0 / (e.N_eff * (double)e.N_sym);
  dim3 grid((e.n + 255) / 256, e.n);
  hipLaunchKernelGGL(k_post, grid, dim3(256), 0, st, psi, out, e.f_acc + (size_t)chain * e.npd,
                     e.d2v_acc + (size_t)chain * e.npd * e.npd, inv, e.w, e.nocc, e.n, e.npd,
                     flags, shift);
}

}  // namespace tdscha
