// Projection of an ensemble on the SCHA modes (the arrays Lanczos.__init__ builds,
// Modules/DynamicalLanczos.py:380-442), one fused FP64 MFMA product:
//
//   X = (u / bohr * sqrt(m)) P                                   Py:385, 425-426, 438
//   Y = ((f - f_sscha) * bohr - f_mean * bohr) / sqrt(m) P       Py:387-388, 404-422, 426, 439
//
// u, f, f_sscha: [N x n_cart] (Angstrom, Ry/Angstrom); f_mean: [n_cart] (the symmetrized
// average force already tiled over the supercell, Ry/Angstrom); m: [n_cart] masses per
// Cartesian component; P: [n_cart x n_modes] polarization vectors of the kept modes
// (translations removed). X, Y: [N x n_modes] row-major, Bohr sqrt(m) and Ry/Bohr/sqrt(m).
//
// The per-component scales sqrt(m)/bohr and bohr/sqrt(m) are formed once per column and slice
// and applied when the slice is written to LDS (no scaled copy of u or f is stored);
// multiplying by them instead of dividing, as the reference does, moves each element by at
// most an ulp or two. Both products share the P slice: one workgroup computes a 64 x 64 tile of X and the
// same tile of Y.
// v_mfma_f64_16x16x4f64, BK = 16 slices double-buffered through LDS.
#include <string>

#include "tdscha_internal.h"
#include "../../include/tdscha_hip.h"

namespace tdscha {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int PM = 64, PN = 64, PK = 16;
constexpr int ALD = PK + 1;    // u, f slices as [row][k]: stores and fragment reads spread over banks
constexpr int PLD = PN + 16;   // P slice as [k][mode] (rows of consecutive k land 32 banks apart)

struct ProjArgs {
  const double *u, *f, *fs, *fmean, *m, *P;
  double *X, *Y;
  int N, nc, nm;
  double bohr;
};

// Workgroup id -> logical tile id, XCD x (workgroups are dealt out to the 8 XCDs round-robin)
// running logical ids [x * per, (x + 1) * per); the grid has 8 * per workgroups.
__device__ inline int xcd_id(int total) {
  const int per = (total + 7) / 8;
  return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

__device__ inline v4d mfma64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Registers of one slice, loaded without branches (indices past the ends are clamped and the
// values zeroed when stored) so that the loads of the next slices stay in flight under the
// MFMAs. A side: column k0 + t%16 of rows r0 + t/16 + 16 j (j < 4) of u, f and f_sscha, and
// the mass and mean force of that column; P side: row k0 + t/16, modes c0 + 4 (t%16) .. +3.
struct ProjRegs {
  double u[4], f[4], fs[4], m, fm, b[4];
};

template <bool FS, bool FM>
__device__ inline void proj_load(const ProjArgs &g, int r0, int c0, int k0, ProjRegs &s) {
  const int t = threadIdx.x;
  const int k = min(k0 + (t & 15), g.nc - 1), rl = r0 + (t >> 4);
  s.m = g.m[k];
  s.fm = FM ? g.fmean[k] : 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const size_t o = (size_t)min(rl + 16 * j, g.N - 1) * g.nc + k;
    s.u[j] = g.u[o];
    s.f[j] = g.f[o];
    s.fs[j] = FS ? g.fs[o] : 0.0;
  }
  const int kr = min(k0 + (t >> 4), g.nc - 1), cc = c0 + (t & 15) * 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) s.b[j] = g.P[(size_t)kr * g.nm + min(cc + j, g.nm - 1)];
}

// Scale and store a slice: x = u sqrt(m) / bohr, y = (f - f_sscha - f_mean) bohr / sqrt(m);
// entries past the ends of the arrays are zero.
__device__ inline void proj_store(const ProjArgs &g, int r0, int c0, int k0, double *Ax,
                                  double *Ay, double *Bs, const ProjRegs &s) {
  const int t = threadIdx.x;
  const int kl = t & 15, rl = t >> 4;
  const bool kin = k0 + kl < g.nc;
  const double sm = sqrt(s.m);
  const double sx = kin ? sm / g.bohr : 0.0, sy = kin ? g.bohr / sm : 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const bool in = r0 + rl + 16 * j < g.N;
    Ax[(rl + 16 * j) * ALD + kl] = in ? s.u[j] * sx : 0.0;
    Ay[(rl + 16 * j) * ALD + kl] = in ? (s.f[j] - s.fs[j] - s.fm) * sy : 0.0;
  }
  const int kr = t >> 4, cc = (t & 15) * 4;
  const bool krin = k0 + kr < g.nc;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    Bs[kr * PLD + cc + j] = (krin && c0 + cc + j < g.nm) ? s.b[j] : 0.0;
}

// The 32 MFMAs of one slice for this wave's 32 x 32 blocks of X and of Y.
__device__ inline void proj_mma(const double *A0, const double *A1, const double *B, int wr,
                                int wc, v4d (&ax)[2][2], v4d (&ay)[2][2]) {
  const int l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
#pragma unroll
  for (int kq = 0; kq < PK / 4; ++kq) {
    const int k = kq * 4 + lg;
    const int a0 = (wr * 32 + lr) * ALD + k, a1 = a0 + 16 * ALD;
    const double x0 = A0[a0], x1 = A0[a1], y0 = A1[a0], y1 = A1[a1];
    const double b0 = B[k * PLD + wc * 32 + lr], b1 = B[k * PLD + wc * 32 + 16 + lr];
    ax[0][0] = mfma64(x0, b0, ax[0][0]);
    ax[0][1] = mfma64(x0, b1, ax[0][1]);
    ax[1][0] = mfma64(x1, b0, ax[1][0]);
    ax[1][1] = mfma64(x1, b1, ax[1][1]);
    ay[0][0] = mfma64(y0, b0, ay[0][0]);
    ay[0][1] = mfma64(y0, b1, ay[0][1]);
    ay[1][0] = mfma64(y1, b0, ay[1][0]);
    ay[1][1] = mfma64(y1, b1, ay[1][1]);
  }
}

template <bool FS, bool FM>
__global__ __launch_bounds__(256, 2) void k_project(ProjArgs g) {
  __shared__ double Ax[2][PM * ALD], Ay[2][PM * ALD], Bs[2][PK * PLD];
  // row tiles in order, their column tiles fastest, each XCD a contiguous range of tiles:
  // the workgroups sharing a block of rows of u and f run on one L2
  const int n_ct = (g.nm + PN - 1) / PN;
  const int total = ((g.N + PM - 1) / PM) * n_ct;
  const int id = xcd_id(total);
  if (id >= total) return;
  const int r0 = (id / n_ct) * PM, c0 = (id % n_ct) * PN;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wr = w >> 1, wc = w & 1;
  const int lr = l & 15, lg = l >> 4;

  v4d ax[2][2], ay[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) ax[a][b] = ay[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};

  // slice kt + 1 is loaded into registers during the MFMAs of slice kt and written to the
  // other LDS stage at the start of the next step
  const int KT = (g.nc + PK - 1) / PK;
  ProjRegs s;
  proj_load<FS, FM>(g, r0, c0, 0, s);
  proj_store(g, r0, c0, 0, Ax[0], Ay[0], Bs[0], s);
  if (KT > 1) proj_load<FS, FM>(g, r0, c0, PK, s);
  __syncthreads();
  for (int kt = 0; kt < KT; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < KT) proj_store(g, r0, c0, (kt + 1) * PK, Ax[cur ^ 1], Ay[cur ^ 1], Bs[cur ^ 1], s);
    if (kt + 2 < KT) proj_load<FS, FM>(g, r0, c0, (kt + 2) * PK, s);
    proj_mma(Ax[cur], Ay[cur], Bs[cur], wr, wc, ax, ay);
    __syncthreads();
  }

  // f64 16x16x4 result: lane (lr, lg) holds column lr of rows lg + 4 r
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      const int c = c0 + wc * 32 + tj * 16 + lr;
      if (c >= g.nm) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = r0 + wr * 32 + ti * 16 + lg + 4 * r;
        if (i >= g.N) continue;
        g.X[(size_t)i * g.nm + c] = ax[ti][tj][r];
        g.Y[(size_t)i * g.nm + c] = ay[ti][tj][r];
      }
    }
}

}  // namespace
}  // namespace tdscha

using namespace tdscha;

extern "C" int tdscha_project_ensemble(int n_configs, int n_cart, int n_modes, const double *u,
                                       const double *forces, const double *sscha_forces,
                                       const double *f_mean, const double *masses,
                                       const double *pols, double bohr, double *X, double *Y,
                                       void *stream) {
  if (n_configs <= 0 || n_cart <= 0 || n_modes <= 0) {
    set_error("n_configs, n_cart and n_modes must be positive");
    return 2;
  }
  if (!u || !forces || !masses || !pols || !X || !Y) {
    set_error("NULL input or output pointer");
    return 1;
  }
  if (!(bohr > 0.0)) {
    set_error("bohr must be positive");
    return 2;
  }
  if ((long long)n_configs * n_cart >= (1LL << 40) || (long long)n_cart * n_modes >= (1LL << 40)) {
    set_error("arrays too large");
    return 2;
  }
  ProjArgs g{u, forces, sscha_forces, f_mean, masses, pols, X, Y, n_configs, n_cart, n_modes, bohr};
  const long long tiles =
      (long long)((n_configs + PM - 1) / PM) * ((n_modes + PN - 1) / PN);
  if (tiles > (1LL << 30)) {
    set_error("arrays too large");
    return 2;
  }
  const int grid = 8 * (int)((tiles + 7) / 8);
  hipStream_t st = (hipStream_t)stream;
  if (sscha_forces && f_mean)
    hipLaunchKernelGGL((k_project<true, true>), dim3(grid), dim3(256), 0, st, g);
  else if (sscha_forces)
    hipLaunchKernelGGL((k_project<true, false>), dim3(grid), dim3(256), 0, st, g);
  else if (f_mean)
    hipLaunchKernelGGL((k_project<false, true>), dim3(grid), dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL((k_project<false, false>), dim3(grid), dim3(256), 0, st, g);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_error(std::string("projection launch failed: ") + hipGetErrorString(err));
    return 10;
  }
  return 0;
}
