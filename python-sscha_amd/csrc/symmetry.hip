// Symmetry operations of the supercell in the basis of the SCHA modes, restricted to blocks of
// degenerate modes (what Lanczos.prepare_symmetrization stores, Modules/DynamicalLanczos.py:
// 686-778, through CellConstructor's GetSymmetriesOnModesDeg):
//
//   S_b[s][i][j] = sum_a sum_x E[3 irt_s(a) + x][b0 + i] (R_s E[a])[x][b0 + j]
//
// for the Cartesian rotation R_s of symmetry s, which carries atom a to atom irt_s(a), and the
// mass-scaled polarization vectors E [3 nat x n] of the modes b0 .. b0 + d - 1 of block b.
// FP64. One launch per block; a workgroup handles G symmetries of the block (G d^2 outputs
// spread over its 256 threads), staging CA atoms at a time of E at the permuted atoms and of
// R E through LDS.
#include <string>

#include "tdscha_internal.h"
#include "../../include/tdscha_hip.h"

namespace tdscha {
namespace {

constexpr int SB_THREADS = 256;
constexpr int SB_CA = 8;        // atoms per LDS stage (24 rows of E)
constexpr int SB_OUT = 4;       // outputs per thread and pass

struct SymArgs {
  const double *rot;   // [n_sym][3][3]
  const int *irt;      // [n_sym][nat]
  const double *E;     // [3 nat][n]
  double *out;         // this block: [n_sym][d][d]
  int n_sym, nat, n, b0, d, G;
  int o0;              // first output of this pass (per symmetry group)
};

__global__ __launch_bounds__(SB_THREADS, 2) void k_sym_block(SymArgs g) {
  extern __shared__ double sh[];
  const int R3 = 3 * SB_CA;
  double *Ep = sh;                        // [G][R3][d]: E at the permuted atoms
  double *RE = sh + g.G * R3 * g.d;       // [G][R3][d]: R_s E
  const int s0 = blockIdx.x * g.G;
  const int ng = min(g.G, g.n_sym - s0);
  const int dd = g.d * g.d, total = g.G * dd;
  double acc[SB_OUT];
#pragma unroll
  for (int q = 0; q < SB_OUT; ++q) acc[q] = 0.0;

  for (int a0 = 0; a0 < g.nat; a0 += SB_CA) {
    const int na = min(SB_CA, g.nat - a0);
    // stage: element (gi, r, c) of both arrays, r = 3 (a - a0) + x
    for (int e = threadIdx.x; e < ng * R3 * g.d; e += SB_THREADS) {
      const int gi = e / (R3 * g.d), rc = e - gi * R3 * g.d;
      const int r = rc / g.d, c = rc - r * g.d;
      const int al = r / 3, x = r - 3 * al;
      double ep = 0.0, re = 0.0;
      if (al < na) {
        const int s = s0 + gi, a = a0 + al;
        // clamped so that a bad permutation cannot read outside E (the host layer checks it)
        const int ap = min(max(g.irt[(size_t)s * g.nat + a], 0), g.nat - 1);
        ep = g.E[(size_t)(3 * ap + x) * g.n + g.b0 + c];
        const double *R = g.rot + (size_t)s * 9 + 3 * x;
        const double *Ea = g.E + (size_t)(3 * a) * g.n + g.b0 + c;
        re = R[0] * Ea[0] + R[1] * Ea[g.n] + R[2] * Ea[2 * (size_t)g.n];
      }
      Ep[e] = ep;
      RE[e] = re;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SB_OUT; ++q) {
      const int o = g.o0 + threadIdx.x + q * SB_THREADS;
      if (o < total) {
        const int gi = o / dd, ij = o - gi * dd;
        const int i = ij / g.d, j = ij - i * g.d;
        if (gi < ng) {
          const double *ep = Ep + (size_t)gi * R3 * g.d, *re = RE + (size_t)gi * R3 * g.d;
          double v = acc[q];
          for (int r = 0; r < 3 * na; ++r) v = fma(ep[r * g.d + i], re[r * g.d + j], v);
          acc[q] = v;
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < SB_OUT; ++q) {
    const int o = g.o0 + threadIdx.x + q * SB_THREADS;
    if (o < total) {
      const int gi = o / dd, ij = o - gi * dd;
      if (gi < ng) g.out[(size_t)(s0 + gi) * dd + ij] = acc[q];
    }
  }
}

}  // namespace
}  // namespace tdscha

using namespace tdscha;

extern "C" int tdscha_symmetries_on_modes(int n_sym, int nat, int n_modes, const double *rot,
                                          const int *irt, const double *pols, int n_blocks,
                                          const int *block_start, double *out, void *stream) {
  if (n_sym <= 0 || nat <= 0 || n_modes <= 0 || n_blocks <= 0) {
    set_error("n_sym, nat, n_modes and n_blocks must be positive");
    return 2;
  }
  if (!rot || !irt || !pols || !block_start || !out) {
    set_error("NULL argument");
    return 1;
  }
  if (block_start[0] != 0 || block_start[n_blocks] > n_modes) {
    set_error("block_start must begin at 0 and end at most at n_modes");
    return 2;
  }
  hipStream_t st = (hipStream_t)stream;
  size_t off = 0;
  for (int b = 0; b < n_blocks; ++b) {
    const int d = block_start[b + 1] - block_start[b];
    if (d <= 0) {
      set_error("block " + std::to_string(b) + " is empty or out of order");
      return 2;
    }
    const int dd = d * d;
    int G = dd >= SB_THREADS ? 1 : SB_THREADS / dd;
    G = G > 32 ? 32 : G;
    const size_t lds = (size_t)2 * G * 3 * SB_CA * d * sizeof(double);
    if (lds > 64 * 1024) {
      set_error("symmetry block of " + std::to_string(d) + " modes is too wide");
      return 2;
    }
    SymArgs g{rot, irt, pols, out + off, n_sym, nat, n_modes, block_start[b], d, G, 0};
    const int grid = (n_sym + G - 1) / G;
    for (int o0 = 0; o0 < G * dd; o0 += SB_THREADS * SB_OUT) {
      g.o0 = o0;
      hipLaunchKernelGGL(k_sym_block, dim3(grid), dim3(SB_THREADS), lds, st, g);
    }
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
      set_error(std::string("symmetry launch failed: ") + hipGetErrorString(err));
      return 10;
    }
    off += (size_t)n_sym * dd;
  }
  return 0;
}
