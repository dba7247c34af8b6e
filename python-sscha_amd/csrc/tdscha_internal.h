// Internal declarations shared by the HIP translation units of libtdscha_hip.
//
// Device data layout of one ensemble (all row-major, zero padded):
//   Np  = N rounded up to a multiple of kPadRows (configs)
//   npd = n rounded up to a multiple of kPadCols (modes)
//   X   [Np x npd]  fp64  displacements in the mode basis          (GEMM1 A operand)
//   Ft  [Np x npd]  fp64  f_psi * Y (scaled forces)                 (GEMM1 epilogue)
//   XtT [npd x Np]  fp64  (f_ups * X)^T                              (GEMM2 A operand)
//   YT  [npd x Np]  fp64  Y^T                                        (GEMM2 B operand source)
//   rho [Np]        fp64  importance weights (0 in the padding)
//   S   [N_sym x sym_stride] fp64  block-diagonal symmetry matrices, per symmetry the
//        blocks packed one after the other, each d_b x d_b row-major.
//   Reduced-precision copies of the GEMM operands (bf16 hi/lo pairs, fp32) live
//   next to them when the ensemble was created with those precisions enabled.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <string>
#include <vector>

namespace tdscha {

constexpr int kPadRows = 256;  // config padding (GEMM1 row tiles, GEMM2 k slices)
constexpr int kPadCols = 64;   // mode padding (GEMM tiles along modes)
constexpr double kRyToK = 157887.32400374097;  // LanczosFunctions.h:9

constexpr int kRotTile = 64;   // widest block / column group of the fused rotations
constexpr int kFavChunk = 1024; // configurations per split-k chunk of the f_av product
constexpr int kBfTileN = 96;    // narrowest column tile of the BF16x3 GEMM1 row sums
constexpr int kG1SlotGroup = 14; // image slots per GEMM1 tile group (slots run fastest in it)

enum Precision { kFP64 = 0, kFP32 = 1, kBF16x3 = 2 };

// bf16 as raw 16-bit words: round to nearest even (finite inputs), and back.
__device__ inline uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
enum Terms { kTermF = 1, kTermD2vR = 2, kTermD2vY = 4 };
enum OpFlags {
  kOpTranspose = 1, kOpWigner = 2, kOpIgnoreV3 = 4, kOpIgnoreV4 = 8, kOpIgnoreHarmonic = 16,
  kOpReverse = 32
};

// One pass of the ensemble loop: nc chains x ns symmetries [s0, s0 + ns). Workspace slot j
// holds chain j / ns and symmetry s0 + j % ns.
struct Pass {
  int s0, ns, nc;
  __host__ __device__ int slots() const { return ns * nc; }
  __host__ __device__ int chain(int j) const { return j / ns; }
  __host__ __device__ int sym(int j) const { return s0 + j % ns; }
};

// Per-mode symmetry metadata on the device.
struct ModeMeta {
  int blk_off;   // first mode of the block this mode belongs to
  int blk_dim;   // block size d_b
  int sym_off;   // offset of the block inside one symmetry's packed matrices
  int pad;
};

struct Ensemble {
  int device = 0;
  int N = 0, n = 0, Np = 0, npd = 0;
  int N_sym = 0, n_blocks = 0, sym_stride = 0, max_block = 0;
  int s_batch = 0;            // symmetries processed per pass
  int prec_mask = 0;          // precisions whose operand copies exist
  double T = 0.0, N_eff = 0.0;
  std::vector<int> blk_off, blk_dim, sym_off;

  // fp64 operands
  double *X = nullptr, *Ft = nullptr, *XtT = nullptr, *YT = nullptr, *rho = nullptr;
  double *fups = nullptr, *fpsi = nullptr, *S = nullptr;
  double *w = nullptr, *nocc = nullptr;   // frequencies [n] and Bose occupations [n]
  ModeMeta *meta = nullptr;
  // group-tiled rotations (every block at most kRotTile wide): groups of whole blocks
  // (start, width <= grp_width) and the tile pairs (g, h), g <= h
  bool fused_rot = false;
  int grp_width = 64;
  int n_groups = 0, n_pairs = 0;
  int2 *dgroups = nullptr;
  int2 *dpairs = nullptr;
  // reduced-precision operand copies (bf16 stored as raw 16-bit words; lo = hi + count)
  uint16_t *X_hi = nullptr, *X_lo = nullptr, *XtT_hi = nullptr, *XtT_lo = nullptr;
  // the same bf16 hi / lo planes of X and XtT in the MFMA fragment-tiled layout (hi plane,
  // then lo plane) read by the register-operand contraction kernels (tile form 3)
  uint16_t *X_frag = nullptr, *XtT_frag = nullptr;
  float *X_f = nullptr, *Ft_f = nullptr, *XtT_f = nullptr, *YT_f = nullptr;
  // X_f and Ft_f in the accumulator-tiled layout (16 x 16 blocks of 64 lane chunks of 4
  // values), the GEMM1 row-sum operands of tile forms 3-6
  float *X_acc = nullptr, *Ft_acc = nullptr;

  // per-chain inputs / accumulators, for up to n_chains chains (grown on demand)
  int n_chains = 0;
  double *Ysym = nullptr;     // [chains x npd x npd]  (Y1 + Y1^T)/2, padded
  double *R1p = nullptr;      // [chains x npd]
  double *d2v_acc = nullptr;  // [chains x npd x npd]
  double *f_acc = nullptr;    // [chains x npd]
  // workspace (per pass of s_batch slots)
  double *Tw = nullptr;       // [s_batch x npd x npd]  rotation intermediate
  double *Ys = nullptr;       // [s_batch x npd x npd]  S_s^T Ysym S_s
  // reduced-precision copies of Ys, one buffer per precision (their layouts differ, and the
  // padding each GEMM reads as part of its k range must stay zero in its own layout)
  float *Ys_f32 = nullptr;    // [s_batch x npd x npd]
  uint16_t *Ys_bf = nullptr;  // hi plane [s_batch x npd x npd], then the lo plane
  void *ys_lp(int prec) const {
    return prec == kBF16x3 ? (void *)Ys_bf : prec == kFP32 ? (void *)Ys_f32 : nullptr;
  }
  double *G = nullptr;        // [s_batch x 2 x npd]  g1 = f_ups*S^T R1, g2 = S^T R1 / f_psi
  double *W = nullptr;        // [s_batch x 4 x Np]  c1, c2, e1, e2 per image
  float *Wf = nullptr;        // [s_batch x 2 x Np]  c1, c2 / 2 in FP32 (BF16x3 GEMM2 operand)
  double *B = nullptr;        // [s_batch x npd x npd]  X~^T G_s
  float *B_f = nullptr;       // the same in FP32 (BF16x3 tile forms 3-5: FP32 accumulators)
  double *V = nullptr;        // [s_batch x npd]  Y^T e1 + X~^T e2
  double *Vp = nullptr;       // [ceil(Np / kFavChunk) x s_batch x npd]  partial sums of V
  double *Pw = nullptr;       // [s_batch x col tiles x 4 x Np]  BF16x3 GEMM1 row sums
  double *Rv = nullptr;       // [npd x s_batch]  r_s = S_s^T R1 of every slot (BF16x3)
  double *GD = nullptr;       // [2 x s_batch x Np]  gamma, delta of every image (BF16x3)
  // force adjoint (allocated by its first call): per-chain seeds, per-slot rotated seeds,
  // their row sums and the gradient accumulator
  int adj_chains = 0, adj_slots = 0;
  int pass_slots = 0;         // slots used per pass (autotuned; 0 = s_batch)
  int bf_variant = 0;         // BF16x3 contraction kernels (autotuned): 0 = 256/192 x 128,
                              // 1 = 256 x 256, 2 = 256 x 256 with direct global->LDS copies,
                              // 3 = 256/192 x 192 with wave-private A operands,
                              // 4 = 256/192 x 96, the same with two workgroups per CU,
                              // 5 = GEMM1 of form 4 with GEMM2 of form 3,
                              // 6 = form 4 with 256 x 96 GEMM2 tiles,
                              // 7 = GEMM1 of form 4 with 512 x 96 GEMM2 tiles of
                              //     eight row waves sharing the formed operand,
                              // 8 = the same with 448-row GEMM2 tiles (waves of four
                              //     and of three row blocks)
  double *Uc = nullptr;       // [chains x npd x npd]  symmetric seed of d2v
  double *uRc = nullptr;      // [chains x npd]        seed of f_av
  double *Us = nullptr;       // [s_batch x npd x npd] f_ups S^T U S f_ups
  float *Us_f32 = nullptr;    // the same in FP32 (FP32 adjoint)
  double *Gu = nullptr;       // [s_batch x 2 x npd]   f_ups S^T uR, S^T uR
  double *Wu = nullptr;       // [s_batch x 4 x Np]    mu, -, nu, - per image
  double *ones = nullptr;     // [npd]
  // exact image path (a block holds modes of different frequencies): the images of one slot
  bool unequal = false;
  double *IU = nullptr, *IFt = nullptr;   // [Np x npd]  u = S x, f~ = f_psi S y
  double *IUT = nullptr, *IFT = nullptr;  // [npd x Np]  u~^T = (f_ups S x)^T, f^T = (S y)^T
  double *gradp = nullptr;    // [Np x npd]            dJ/dY, padded
  // range-sliced evaluation (tdscha_jl_*): only the images whose flattened index
  // i N_sym + s (configuration i, symmetry s) lies in [win_lo, win_hi] count; the weights
  // then read rho_s[slot][i], rho[i] inside the window and 0 outside (filled per pass)
  bool win = false;
  long long win_lo = 0, win_hi = -1;
  double *rho_s = nullptr;    // [s_batch x Np]
  const double *rho_w() const { return win ? rho_s : rho; }
  size_t rho_stride() const { return win ? (size_t)Np : 0; }
};

void set_error(const std::string &msg);
}  // namespace tdscha

struct tdscha_projector;

namespace tdscha {
// P0 psi (adjoint = 0) or P0^T psi (adjoint = 1) on device vectors, enqueued on st; out may
// be psi (projector.hip)
int projector_apply(tdscha_projector *p, int adjoint, const double *psi, double *out,
                    hipStream_t st);
// turn the kernel-group event timing on / off (no event records may enter a captured
// graph); returns the previous setting
bool profile_swap(bool on);

// ---- kernels / launchers (rotate.hip) ----
void launch_prep_inputs(const Ensemble &e, const double *R1, const double *Y1, hipStream_t st);
void launch_rot_right(const Ensemble &e, const double *in, bool per_chain, bool symmetrize,
                      double *out, const Pass &ps, bool transS, hipStream_t st);
void launch_rot_left(const Ensemble &e, const double *in, double *out, const Pass &ps, bool transS,
                     bool accumulate, hipStream_t st);
void launch_rot_vec(const Ensemble &e, const Pass &ps, hipStream_t st);
// out[slot] = (f_ups * S^T in[chain], S^T in[chain] / div)
// C[m][n] = scale sum_k A[k][m] w[k] B[k][n] (perturb.hip; w may be null)
void launch_weighted_tn(int K, int M, int N, const double *A, int lda, const double *w,
                        const double *B, int ldb, double scale, double *C, int ldc,
                        hipStream_t st);
void launch_rot_vec_from(const Ensemble &e, const Pass &ps, const double *in, double *out,
                         const double *div, hipStream_t st);
// out[slot] = S^T in[chain] S (FP64), optionally scaled out[a][b] *= scale[a] scale[b]
void launch_rot_sym(const Ensemble &e, const Pass &ps, const double *in, double *out,
                    const double *scale, hipStream_t st);
// Ys (fp64 when prec is FP64, else the reduced-precision planes) = S^T Ysym S, fused
void launch_rot1_fused(const Ensemble &e, const Pass &ps, int prec, hipStream_t st);
// d2v_acc[chain] += sum over the chain's slots of S (B_s + B_s^T) S^T, fused; B_s from B_f
// (FP32) when b_f32, else from B
void launch_rot2_fused(const Ensemble &e, const Pass &ps, bool b_f32, hipStream_t st);
void launch_fav_rot_acc(const Ensemble &e, const Pass &ps, hipStream_t st);
void launch_zero(double *p, size_t count, hipStream_t st);
void launch_finalize(const Ensemble &e, int what, bool d2v_accumulate, double *f_out, double *d2v_out,
                     hipStream_t st);
void launch_split_ys(const Ensemble &e, int sb, int prec, hipStream_t st);
// fp64 -> bf16 hi/lo planes (dst, dst + count) or fp32
void launch_convert(const double *src, void *dst, size_t count, int prec, hipStream_t st);

// ---- exact image path (images.hip; contractions in gemm.hip) ----
void launch_images(const Ensemble &e, int s, hipStream_t st);
void launch_image_g(const Ensemble &e, int chain, hipStream_t st);
void launch_image_acc(const Ensemble &e, int chain, bool do_d, bool do_f, hipStream_t st);
void launch_gemm1_images(const Ensemble &e, const double *Ysym_c, int what, hipStream_t st);
void launch_gemm2_images(const Ensemble &e, hipStream_t st);
void launch_fav_images(const Ensemble &e, hipStream_t st);

// ---- operator pre/post transforms and harmonic part (operator.hip) ----
void launch_unpack(const Ensemble &e, int chain, const double *psi, int flags, hipStream_t st);
void launch_post(const Ensemble &e, int chain, const double *psi, double *out, int flags,
                 double shift, hipStream_t st);
// seeds (U, uR) of u . (anharmonic part of post) for one chain
void launch_post_adj(const Ensemble &e, const double *u, double *U, double *uR, int flags,
                     hipStream_t st);

// ---- GEMMs (gemm.hip) ----
// rho_s of the pass's slots for the image window (e.win)
void launch_window_rho(const Ensemble &e, const Pass &ps, hipStream_t st);
// GEMM1 with the rotated matrices M and vectors G of the slots, writing the four row sums
// themselves (alpha, beta, gamma, delta) to out [sb x 4 x Np]
void launch_gemm1_raw(const Ensemble &e, int sb, const double *M, const double *G, double *out,
                      hipStream_t st);
// ---- force adjoint (gemm.hip) ----
void launch_adj_grad(const Ensemble &e, int sb, double kf, double kr, double ky, int prec,
                     hipStream_t st);
// launch_gemm1_raw on FP32 copies (X_f and M in FP32, FP32 accumulation, FP64 row sums)
void launch_gemm1_raw_f32(const Ensemble &e, int sb, const float *M, const double *G,
                          double *out, hipStream_t st);
void launch_grad_out(const Ensemble &e, double *grad, int accumulate, hipStream_t st);
// BF16x3 forms (gemm_bf16.hip)
int bf16_col_tiles(const Ensemble &e);
void launch_gemm1_bf16x3(const Ensemble &e, int sb, int what, hipStream_t st);
void launch_gemm2_bf16x3(const Ensemble &e, int sb, hipStream_t st);
// per-image weights W (and their FP32 copy Wf) from GEMM1's row sums over n_ct column tiles
// nq = 4: P holds alpha, beta, gamma, delta per column tile; nq = 2: alpha, beta, with gamma
// and delta in e.GD
void launch_weights_bf16(const Ensemble &e, int sb, int n_ct, int nq, int what, hipStream_t st);
// wave-private-operand forms (gemm_bf16_rp.hip); GEMM2 writes B_s^T
void launch_gemm1_bf16x3_rp(const Ensemble &e, int sb, int what, int wcn, hipStream_t st);
// fp64 [rows x cols] -> fragment-tiled bf16 hi / lo planes (dst, dst + rows * cols)
void launch_tile_bf16x3(const double *src, int rows, int cols, uint16_t *dst, hipStream_t st);
void launch_tile_acc_f32(const float *src, int rows, int cols, float *dst, hipStream_t st);
// GEMM2 row tiles of 64 mb rows (mb = 3 or, with wcn = 1, 4)
void launch_gemm2_bf16x3_rp(const Ensemble &e, int sb, int wcn, hipStream_t st, int mb = 3);
// V[s] = Y^T e1_s + X~^T e2_s for the first sb slots (FP64 MFMA, split over configurations)
void launch_fav(const Ensemble &e, int sb, hipStream_t st);
// GEMM1: per image (config i, symmetry s) the quadratic forms
//   alpha = x_i^T Ys_s x_i, beta = ft_i^T Ys_s x_i, gamma = x_i . g1_s, delta = ft_i . g2_s
// and the per-image weights (c1, c2, e1, e2) written to W.
void launch_gemm1(const Ensemble &e, int sb, int prec, int what, hipStream_t st);
// GEMM2: B_s = X~^T (c1_s * Y + c2_s/2 * X~) for each symmetry in the pass.
void launch_gemm2(const Ensemble &e, int sb, int prec, hipStream_t st);

}  // namespace tdscha
