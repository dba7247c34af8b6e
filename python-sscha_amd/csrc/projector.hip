// Translation average of a psi vector on the device (the projector of the zone-centre
// operator): P0 psi, or its transpose in the packed metric, P0^T = W P0 W^-1 (W = 2 on the
// off-diagonal packed entries), for psi = (R (n), two symmetric n x n matrices as packed upper
// triangles, Modules/DynamicalLanczos.py:444-462).
//
// The average runs in Cartesian space, where a pure translation t is a permutation of the
// atoms: with E the (3 nat x n) polarization vectors of the kept modes,
//   R' = E^T avg_t P_t (E R),     M' = E^T avg_t (P_t (E M E^T) P_t^T) E.
// The translations form an abelian group acting freely on the atoms, so the average of a
// Cartesian matrix C depends only on the key of each atom pair (a, b): the two atoms'
// orbits and their relative translation tau_a^-1 tau_b (tau_x carries its orbit's first atom
// to x). Each key holds exactly n_t pairs; the average is the sum over a key's pairs (a
// fixed, ordered list: deterministic, no atomics) divided by n_t, written back to every pair
// of the key. The four products E M, (E M) E^T, C E, E^T (C E) are FP64 MFMA contractions
// (k_weighted_tn), ~8 (3 nat)^3 flop per matrix.
#include <algorithm>
#include <string>
#include <vector>

#include "tdscha_internal.h"
#include "../../include/tdscha_hip.h"

using namespace tdscha;

struct tdscha_projector {
  int device = 0, n = 0, nat = 0, nt = 0, nb = 0;
  long n_keys = 0;
  std::vector<void *> allocs;
  double *E = nullptr, *Et = nullptr;         // [3 nat x n], [n x 3 nat]
  int *pairs = nullptr;                        // [n_keys x nt] a * nat + b, grouped by key
  int *key_of = nullptr;                       // [nat x nat]
  int *orbit_members = nullptr;                // [nb x nt]
  int *orbit_of = nullptr;                     // [nat]
  double *M = nullptr, *T1 = nullptr, *C = nullptr, *T2 = nullptr, *G = nullptr;
  double *r = nullptr, *rbar = nullptr, *Rn = nullptr;
};

namespace tdscha {
namespace {

__device__ inline long packed_index(int i, int j, int n) {   // i <= j, row-major upper triangle
  return (long)i * n - (long)i * (i - 1) / 2 + (j - i);
}

// M[i][j] from the packed block (off-diagonal entries / 2 for the transposed metric)
__global__ void k_proj_unpack(const double *__restrict__ p, double *__restrict__ M, int n,
                              int adjoint) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= n) return;
  const int a = min(i, j), b = max(i, j);
  double v = p[packed_index(a, b, n)];
  if (adjoint && a != b) v *= 0.5;
  M[(size_t)i * n + j] = v;
}

__global__ void k_proj_pack(const double *__restrict__ M, double *__restrict__ p, int n,
                            int adjoint) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= n || j < i) return;
  double v = M[(size_t)i * n + j];
  if (adjoint && i != j) v *= 2.0;
  p[packed_index(i, j, n)] = v;
}

// G[key][3 x + y] = sum over the key's pairs (a, b), in list order, of C[3a + x][3b + y]
__global__ void k_proj_fold(const double *__restrict__ C, const int *__restrict__ pairs,
                            double *__restrict__ G, long n_keys, int nt, int nat) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_keys * 9) return;
  const long key = t / 9;
  const int comp = (int)(t % 9), x = comp / 3, y = comp % 3;
  const int ld = 3 * nat;
  double s = 0.0;
  for (int k = 0; k < nt; ++k) {
    const int ab = pairs[key * nt + k];
    const int a = ab / nat, b = ab - a * nat;
    s += C[(size_t)(3 * a + x) * ld + 3 * b + y];
  }
  G[t] = s;
}

// C[3a + x][3b + y] = G[key(a, b)][3 x + y] / nt
__global__ void k_proj_unfold(const double *__restrict__ G, const int *__restrict__ key_of,
                              double *__restrict__ C, int nat, double inv_nt) {
  const int cb = blockIdx.x * blockDim.x + threadIdx.x, ca = blockIdx.y;
  const int ld = 3 * nat;
  if (cb >= ld) return;
  const int a = ca / 3, x = ca % 3, b = cb / 3, y = cb % 3;
  C[(size_t)ca * ld + cb] = G[(size_t)key_of[a * nat + b] * 9 + 3 * x + y] * inv_nt;
}

// rbar[3a + x] = (1 / nt) sum over the orbit of a (in member order) of r[3a' + x]
__global__ void k_proj_vec_avg(const double *__restrict__ r, const int *__restrict__ members,
                               const int *__restrict__ orbit_of, double *__restrict__ rbar,
                               int nat, int nt, double inv_nt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * nat) return;
  const int a = t / 3, x = t % 3, o = orbit_of[a];
  double s = 0.0;
  for (int k = 0; k < nt; ++k) s += r[3 * members[o * nt + k] + x];
  rbar[t] = s * inv_nt;
}

__global__ void k_proj_copy(const double *__restrict__ src, double *__restrict__ dst, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

int proj_alloc(tdscha_projector *p, void **ptr, size_t bytes) {
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 8);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return 11;
  }
  p->allocs.push_back(*ptr);
  return 0;
}

// M (n x n, in p->M) -> E^T avg(E M E^T) E, in p->M
void project_matrix(tdscha_projector *p, hipStream_t st) {
  const int n = p->n, c3 = 3 * p->nat;
  launch_weighted_tn(n, n, c3, p->M, n, nullptr, p->Et, c3, 1.0, p->T1, c3, st);     // (E M)^T
  launch_weighted_tn(n, c3, c3, p->T1, c3, nullptr, p->Et, c3, 1.0, p->C, c3, st);   // E M E^T
  const long nk9 = p->n_keys * 9;
  hipLaunchKernelGGL(k_proj_fold, dim3((unsigned)((nk9 + 255) / 256)), dim3(256), 0, st, p->C,
                     p->pairs, p->G, p->n_keys, p->nt, p->nat);
  hipLaunchKernelGGL(k_proj_unfold, dim3((c3 + 255) / 256, c3), dim3(256), 0, st, p->G,
                     p->key_of, p->C, p->nat, 1.0 / p->nt);
  launch_weighted_tn(c3, c3, n, p->C, c3, nullptr, p->E, n, 1.0, p->T2, n, st);      // C E
  launch_weighted_tn(c3, n, n, p->E, n, nullptr, p->T2, n, 1.0, p->M, n, st);        // E^T C E
}

}  // namespace

int projector_apply(tdscha_projector *p, int adjoint, const double *psi, double *out,
                    hipStream_t st) {
  const int n = p->n, c3 = 3 * p->nat;
  const long nw2 = (long)n * (n + 1) / 2;
  // R part first: it is read before out[0, n) is written (out may be psi)
  launch_weighted_tn(n, c3, 1, p->Et, c3, nullptr, psi, 1, 1.0, p->r, 1, st);         // E R
  hipLaunchKernelGGL(k_proj_vec_avg, dim3((c3 + 255) / 256), dim3(256), 0, st, p->r,
                     p->orbit_members, p->orbit_of, p->rbar, p->nat, p->nt, 1.0 / p->nt);
  launch_weighted_tn(c3, n, 1, p->E, n, nullptr, p->rbar, 1, 1.0, p->Rn, 1, st);      // E^T rbar
  for (int k = 0; k < 2; ++k) {
    const double *in = psi + n + k * nw2;
    double *o = out + n + k * nw2;
    hipLaunchKernelGGL(k_proj_unpack, dim3((n + 255) / 256, n), dim3(256), 0, st, in, p->M, n,
                       adjoint);
    project_matrix(p, st);
    hipLaunchKernelGGL(k_proj_pack, dim3((n + 255) / 256, n), dim3(256), 0, st, p->M, o, n,
                       adjoint);
  }
  hipLaunchKernelGGL(k_proj_copy, dim3((n + 255) / 256), dim3(256), 0, st, p->Rn, out, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("projector launch failed: ") + hipGetErrorString(e));
    return 12;
  }
  return 0;
}

}  // namespace tdscha

extern "C" {

int tdscha_projector_destroy(tdscha_projector *p) {
  if (!p) return 0;
  (void)hipSetDevice(p->device);
  for (void *a : p->allocs) (void)hipFree(a);
  delete p;
  return 0;
}

int tdscha_projector_create(tdscha_projector **out, int device, int n_modes, int nat,
                            const double *pols, int n_trans, const int *irt) {
  if (!out || !pols || !irt || n_modes <= 0 || nat <= 0 || n_trans < 1) {
    set_error("bad argument to tdscha_projector_create");
    return 1;
  }
  *out = nullptr;
  const int nt = n_trans;
  // every row a permutation; the group acts freely (every orbit has nt atoms)
  std::vector<std::vector<int>> inv(nt, std::vector<int>(nat, -1));
  for (int t = 0; t < nt; ++t)
    for (int a = 0; a < nat; ++a) {
      const int b = irt[(size_t)t * nat + a];
      if (b < 0 || b >= nat || inv[t][b] >= 0) {
        set_error("row " + std::to_string(t) + " of the translations is not a permutation");
        return 2;
      }
      inv[t][b] = a;
    }
  if (nat % nt) {
    set_error("the translations do not act freely on the atoms (nat % n_trans != 0)");
    return 2;
  }
  std::vector<int> orbit(nat, -1), tau(nat, -1), reps;
  for (int a = 0; a < nat; ++a) {
    if (orbit[a] >= 0) continue;
    const int o = (int)reps.size();
    for (int t = 0; t < nt; ++t) {
      const int b = irt[(size_t)t * nat + a];
      if (orbit[b] >= 0) {
        set_error("the translations do not act freely on the atoms");
        return 2;
      }
      orbit[b] = o;
      tau[b] = t;
    }
    reps.push_back(a);
  }
  const int nb = (int)reps.size();
  const int r0 = reps[0];
  std::vector<int> trans_of(nat, -1);   // atom of orbit 0 -> the translation carrying r0 to it
  for (int t = 0; t < nt; ++t) trans_of[irt[(size_t)t * nat + r0]] = t;
  // key(a, b) = (orbit(a) nb + orbit(b)) nt + d, d the translation tau_a^-1 tau_b
  const long n_keys = (long)nb * nb * nt;
  std::vector<int> key_of((size_t)nat * nat);
  std::vector<std::vector<int>> bucket(n_keys);
  for (int a = 0; a < nat; ++a)
    for (int b = 0; b < nat; ++b) {
      const int img = inv[tau[a]][irt[(size_t)tau[b] * nat + r0]];   // (tau_a^-1 tau_b)(r0)
      const int d = trans_of[img];
      if (d < 0) {
        set_error("the translations are not a group acting freely on the atoms");
        return 2;
      }
      const long key = ((long)orbit[a] * nb + orbit[b]) * nt + d;
      key_of[(size_t)a * nat + b] = (int)key;
      bucket[key].push_back(a * nat + b);
    }
  std::vector<int> pairs((size_t)n_keys * nt);
  for (long k = 0; k < n_keys; ++k) {
    if ((int)bucket[k].size() != nt) {
      set_error("the translations are not an abelian group acting freely on the atoms");
      return 2;
    }
    std::copy(bucket[k].begin(), bucket[k].end(), pairs.begin() + k * nt);
  }
  std::vector<int> members((size_t)nb * nt);
  for (int o = 0; o < nb; ++o)
    for (int t = 0; t < nt; ++t) members[(size_t)o * nt + t] = irt[(size_t)t * nat + reps[o]];
  const int n = n_modes, c3 = 3 * nat;
  std::vector<double> Et((size_t)n * c3);
  for (int a = 0; a < c3; ++a)
    for (int j = 0; j < n; ++j) Et[(size_t)j * c3 + a] = pols[(size_t)a * n + j];

  tdscha_projector *p = new (std::nothrow) tdscha_projector();
  if (!p) { set_error("out of host memory"); return 11; }
  p->device = device;
  p->n = n; p->nat = nat; p->nt = nt; p->nb = nb; p->n_keys = n_keys;
  hipError_t err = hipSetDevice(device);
  if (err != hipSuccess) {
    set_error(std::string("hipSetDevice failed: ") + hipGetErrorString(err));
    delete p;
    return 10;
  }
  int rc = 0;
  rc |= proj_alloc(p, (void **)&p->E, sizeof(double) * c3 * n);
  rc |= proj_alloc(p, (void **)&p->Et, sizeof(double) * c3 * n);
  rc |= proj_alloc(p, (void **)&p->pairs, sizeof(int) * pairs.size());
  rc |= proj_alloc(p, (void **)&p->key_of, sizeof(int) * key_of.size());
  rc |= proj_alloc(p, (void **)&p->orbit_members, sizeof(int) * members.size());
  rc |= proj_alloc(p, (void **)&p->orbit_of, sizeof(int) * nat);
  rc |= proj_alloc(p, (void **)&p->M, sizeof(double) * n * n);
  rc |= proj_alloc(p, (void **)&p->T1, sizeof(double) * n * c3);
  rc |= proj_alloc(p, (void **)&p->C, sizeof(double) * c3 * c3);
  rc |= proj_alloc(p, (void **)&p->T2, sizeof(double) * c3 * n);
  rc |= proj_alloc(p, (void **)&p->G, sizeof(double) * n_keys * 9);
  rc |= proj_alloc(p, (void **)&p->r, sizeof(double) * c3);
  rc |= proj_alloc(p, (void **)&p->rbar, sizeof(double) * c3);
  rc |= proj_alloc(p, (void **)&p->Rn, sizeof(double) * n);
  if (rc) { tdscha_projector_destroy(p); return rc; }
  err = hipMemcpy(p->E, pols, sizeof(double) * c3 * n, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(p->Et, Et.data(), sizeof(double) * c3 * n, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(p->pairs, pairs.data(), sizeof(int) * pairs.size(), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(p->key_of, key_of.data(), sizeof(int) * key_of.size(), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(p->orbit_members, members.data(), sizeof(int) * members.size(), hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipMemcpy(p->orbit_of, orbit.data(), sizeof(int) * nat, hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    set_error(std::string("projector upload failed: ") + hipGetErrorString(err));
    tdscha_projector_destroy(p);
    return 10;
  }
  *out = p;
  return 0;
}

int tdscha_projector_apply(tdscha_projector *p, int adjoint, const double *psi, double *out,
                           void *stream) {
  if (!p || !psi || !out) { set_error("NULL argument"); return 1; }
  hipError_t err = hipSetDevice(p->device);
  if (err != hipSuccess) {
    set_error(std::string("hipSetDevice failed: ") + hipGetErrorString(err));
    return 10;
  }
  return projector_apply(p, adjoint ? 1 : 0, psi, out, (hipStream_t)stream);
}

}  // extern "C"
