/*
 * libtdscha_hip: MI355X (gfx950) implementation of the TD-SCHA ensemble
 * L-application, the symmetry-expanded weighted ensemble average that every
 * TD-SCHA Lanczos step calls.
 *
 * C ABI only: plain pointers and sizes. Every function returning int returns
 * 0 on success and a non-zero code on failure; tdscha_last_error() then holds
 * a message (thread-local). The reference signals bad input by printing and
 * calling exit() (CModules/odd_corr_module.c:535-547); the prefixed entry points
 * here return an error instead. The unprefixed drop-in symbols keep the
 * reference's void signatures and therefore keep its behaviour on failure
 * (message on stderr, exit(EXIT_FAILURE)).
 *
 * Units and layouts follow the reference exactly:
 *   X, Y  (n_configs, n_modes) row-major, mode basis (Bohr sqrt(m) / Ry/Bohr/sqrt(m))
 *   w     (n_modes) SCHA frequencies [Ry];  T in K;  w_is / rho (n_configs)
 *   symmetries[b] (N_sym, d_b, d_b) row-major acting on the modes
 *   degenerate_space[b][0..d_b-1], which must be contiguous and sorted
 *   (the reference indexes rows as mu - degenerate_space[b][0],
 *   LanczosFunctions.c:2162), blocks_ids[mu] = block of mode mu.
 * When the frequencies are equal (to 1e-8 relative) inside every block, as they are for
 * blocks of degenerate modes, the device path rotates the n x n operands instead of
 * building every image. An ensemble with a block of unequal frequencies is accepted, as
 * the reference accepts it: it runs on an exact image path (the images of every
 * configuration built on the device, one symmetry at a time), in FP64 only; the reduced
 * precisions and the force adjoint refuse it.
 *
 * Thread safety: calls on different ensembles may run concurrently; calls on
 * one ensemble share its workspace and must not overlap.
 */
#ifndef TDSCHA_HIP_H
#define TDSCHA_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDSCHA_ABI_VERSION 2

/* Arithmetic of the two ensemble contractions. */
#define TDSCHA_PREC_FP64 0   /* FP64 MFMA, the plain path */
#define TDSCHA_PREC_FP32 1   /* FP32 MFMA, FP32 accumulation */
#define TDSCHA_PREC_BF16X3 2 /* hi*hi + hi*lo + lo*hi BF16 MFMA, FP32 accumulation */

/* Terms of the perturbed averages (bit mask). */
#define TDSCHA_TERM_F 1     /* <f>^(1) from Y1            (LanczosFunctions.c:2082) */
#define TDSCHA_TERM_D2V_R 2 /* <d2V/dR2>^(1), D3 from R1   (LanczosFunctions.c:2257) */
#define TDSCHA_TERM_D2V_Y 4 /* <d2V/dR2>^(1), D4 from Y1   (LanczosFunctions.c:2615) */

/* Flags of tdscha_apply_full_L (Lanczos attributes, Modules/DynamicalLanczos.py:235-290). */
#define TDSCHA_OP_TRANSPOSE 1
#define TDSCHA_OP_WIGNER 2
#define TDSCHA_OP_IGNORE_V3 4
#define TDSCHA_OP_IGNORE_V4 8
#define TDSCHA_OP_IGNORE_HARMONIC 16
#define TDSCHA_OP_REVERSE 32

typedef struct tdscha_ensemble tdscha_ensemble;

const char *tdscha_last_error(void);
int tdscha_abi_version(void);

/* ---- Drop-in replacements of the reference's C trio (host pointers) ---------------- */

/* Replaces get_f_average_from_Y_pert_sym_fast, CModules/LanczosFunctions.h:264-266.
 * f_average (n_modes) is overwritten. */
void get_f_average_from_Y_pert_sym_fast(double *X, double *Y, double *w, double *Y1, double T,
                                        int n_modes, int n_configs, double *w_is,
                                        double **symmetries, int N_sym, int *N_degeneracy,
                                        int **degenerate_space, int *blocks_ids,
                                        double *f_average);
/* Replaces get_d2v_dR2_from_R_pert_sym_fast, CModules/LanczosFunctions.h:267-270.
 * d2v_dR2 (n_modes x n_modes) is overwritten. */
void get_d2v_dR2_from_R_pert_sym_fast(double *X, double *Y, double *w, double *R1, double T,
                                      int n_modes, int n_configs, double *w_is,
                                      double **symmetries, int N_sym, int *N_degeneracy,
                                      int **degenerate_space, int *blocks_ids, double *d2v_dR2);
/* Replaces get_d2v_dR2_from_Y_pert_sym_fast, CModules/LanczosFunctions.h:271-273.
 * The D4 part is ADDED to d2v_dR2_out, as in the reference. */
void get_d2v_dR2_from_Y_pert_sym_fast(double *X, double *Y, double *w, double *Y1, double T,
                                      int n_modes, int n_configs, double *w_is,
                                      double **symmetries, int N_sym, int *N_degeneracy,
                                      int **degenerate_space, int *blocks_ids,
                                      double *d2v_dR2_out);

/* Error-returning forms of the three above (same arguments). */
int tdscha_get_f_average_from_Y_pert_sym_fast(const double *X, const double *Y, const double *w,
                                              const double *Y1, double T, int n_modes,
                                              int n_configs, const double *w_is,
                                              const double *const *symmetries, int N_sym,
                                              const int *N_degeneracy,
                                              const int *const *degenerate_space,
                                              const int *blocks_ids, double *f_average);
int tdscha_get_d2v_dR2_from_R_pert_sym_fast(const double *X, const double *Y, const double *w,
                                            const double *R1, double T, int n_modes,
                                            int n_configs, const double *w_is,
                                            const double *const *symmetries, int N_sym,
                                            const int *N_degeneracy,
                                            const int *const *degenerate_space,
                                            const int *blocks_ids, double *d2v_dR2);
int tdscha_get_d2v_dR2_from_Y_pert_sym_fast(const double *X, const double *Y, const double *w,
                                            const double *Y1, double T, int n_modes,
                                            int n_configs, const double *w_is,
                                            const double *const *symmetries, int N_sym,
                                            const int *N_degeneracy,
                                            const int *const *degenerate_space,
                                            const int *blocks_ids, double *d2v_dR2_out);

/* Replaces sscha_HP_odd.GetPerturbAverageSym, CModules/odd_corr_module.c:491-634
 * (argument order of its "OOOOOOdiiOOOOOO" format, with the list lengths and array
 * dimensions made explicit). force_out (n) is overwritten; d2v_out (n x n) is
 * overwritten with the D3 part, then the D4 part is added when apply_D4 is 1. */
int tdscha_get_perturb_average_sym(const double *X, const double *Y, const double *w,
                                   const double *rho, const double *R1, const double *Y1,
                                   double T, int apply_D4, int N_syms, int n_blocks,
                                   const double *const *sym_list, const int *n_deg,
                                   const int *const *sym_basis_list, const int *blocks_ids,
                                   int N_configs, int N_modes, double *force_out,
                                   double *d2v_out);

/* The host-pointer entry points above and below (not the _noalloc one) keep the device
 * ensemble of their last call, with a host copy of the inputs it was built from, and reuse
 * it when the next call's inputs are equal byte for byte (the reference passes the same
 * ensemble at every Lanczos step). No counterpart in the reference, which reads the host
 * arrays on every call. enable = 0 releases the kept ensemble (device memory and host copy)
 * and turns the reuse off, 1 turns it on (the default); returns the previous setting. */
int tdscha_drop_in_cache(int enable);

/* ---- Range-sliced entries of the Julia level (Modules/tdscha_core.jl, "JL" below) -----
 * Replace get_perturb_averages_sym (JL:202-229), get_perturb_d2v_averages_sym (JL:232-259)
 * and get_perturb_f_averages_sym (JL:262-284), with the arguments in their order and the
 * array sizes made explicit:
 *   X, Y        the (n_modes, n_configs) column-major matrices the reference passes as
 *               self.X.T, self.Y.T, i.e. the (n_configs, n_modes) row-major X and Y above;
 *   symmetries  the zero-padded (n_blocks, n_syms, d_max, d_max) row-major array sym_julia
 *               (Modules/DynamicalLanczos.py:753-768), block b in its top-left
 *               n_degeneracies[b] x n_degeneracies[b] corner;
 *   degenerate_space  (n_blocks, d_max) row-major, deg_julia (0-based mode ids);
 *   blocks      the mode -> block map (sym_block_id);
 *   start_index, end_index  1-based inclusive range of bigindex = i n_syms + s + 1 over
 *               configurations i and symmetries s (JL:71-73).
 * The images in the range are summed and divided by n_syms sum(rho) over ALL configurations,
 * so the results of disjoint ranges add up to the full average (the split of
 * Modules/DynamicalLanczos.py:3083-3103). Outputs are overwritten; d2v_dr2 holds the D3 part
 * plus, with apply_v4, the D4 part. As in the reference, the force entries (and
 * get_perturb_averages_sym) reject start_index > end_index (JL:171); the d2v entry then
 * returns zeros. A range outside [1, n_configs n_syms] is rejected (code 6). FP64 on the
 * current device, synchronous. */
int tdscha_jl_get_perturb_averages_sym(const double *X, const double *Y, const double *omega,
                                       const double *rho, const double *R1, const double *Y1,
                                       double temperature, int apply_v4, const double *symmetries,
                                       const int *n_degeneracies, const int *degenerate_space,
                                       const int *blocks, long long start_index,
                                       long long end_index, int n_modes, int n_configs,
                                       int n_syms, int n_blocks, int d_max, double *f_average,
                                       double *d2v_dr2);
int tdscha_jl_get_perturb_d2v_averages_sym(const double *X, const double *Y, const double *omega,
                                           const double *rho, const double *R1, const double *Y1,
                                           double temperature, int apply_v4,
                                           const double *symmetries, const int *n_degeneracies,
                                           const int *degenerate_space, const int *blocks,
                                           long long start_index, long long end_index,
                                           int n_modes, int n_configs, int n_syms, int n_blocks,
                                           int d_max, double *d2v_dr2);
int tdscha_jl_get_perturb_f_averages_sym(const double *X, const double *Y, const double *omega,
                                         const double *rho, const double *R1, const double *Y1,
                                         double temperature, int apply_v4,
                                         const double *symmetries, const int *n_degeneracies,
                                         const int *degenerate_space, const int *blocks,
                                         long long start_index, long long end_index, int n_modes,
                                         int n_configs, int n_syms, int n_blocks, int d_max,
                                         double *f_average);

/* ---- Measurement -------------------------------------------------------------------
 * No counterpart in the reference. The sustained MFMA rate of the current device: a stream
 * of independent MFMAs with pseudo-random operands on every CU for `seconds`, timed over
 * the second half (after the clock has settled under the board's power limit).
 * precision TDSCHA_PREC_BF16X3 (v_mfma_f32_16x16x32_bf16) or TDSCHA_PREC_FP64
 * (v_mfma_f64_16x16x4_f64); *tflops receives the rate, *ms_per_launch (may be NULL) the
 * time of one launch. Synchronous. */
int tdscha_mfma_stream(int precision, double seconds, double *tflops, double *ms_per_launch);

/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * arguments, contract and FP64 result as tdscha_get_perturb_average_sym, computed with no
 * hipMalloc / hipFree and no allocating library call. Every device buffer of the call is
 * carved from a static device scratch array of tdscha_noalloc_scratch_bytes() (256 MiB),
 * held by each GPU the library runs on until the process ends; the symmetry images are
 * processed in fewer slots per pass when the workspace would not fit (results do not
 * depend on it). An ensemble that does not fit even with one slot per pass is rejected
 * with code 5 before any upload. All calls on one GPU share that GPU's scratch: calls that
 * overlap on one GPU (from several host threads) race and must be serialised by the
 * caller. Runs on the current device, synchronously. */
int tdscha_get_perturb_average_sym_noalloc(const double *X, const double *Y, const double *w,
                                           const double *rho, const double *R1,
                                           const double *Y1, double T, int apply_D4,
                                           int N_syms, int n_blocks,
                                           const double *const *sym_list, const int *n_deg,
                                           const int *const *sym_basis_list,
                                           const int *blocks_ids, int N_configs, int N_modes,
                                           double *force_out, double *d2v_out);
/* Size in bytes of the per-GPU scratch array of the entry point above. */
size_t tdscha_noalloc_scratch_bytes(void);

/* ---- Device-resident ensemble (the hot path of a Lanczos run) ------------------------
 * No counterpart in the reference, which re-reads X, Y and the symmetries on every call
 * (CModules/odd_corr_module.c:491-634). The ensemble is uploaded once; each apply then
 * reads only device memory. */

/* Host inputs are copied to device `device`. precisions: bit mask of (1 << TDSCHA_PREC_*)
 * whose operand copies to build (FP64 is always built). s_batch: symmetries processed
 * per pass (0 = automatic). */
int tdscha_ensemble_create(tdscha_ensemble **out, int device, const double *X, const double *Y,
                           const double *w, const double *rho, double T, int n_modes,
                           int n_configs, int N_sym, int n_blocks,
                           const double *const *symmetries, const int *N_degeneracy,
                           const int *const *degenerate_space, const int *blocks_ids,
                           int precisions, int s_batch);
int tdscha_ensemble_destroy(tdscha_ensemble *ens);
/* n_modes, n_configs, N_sym, s_batch, padded modes, padded configs */
int tdscha_ensemble_dims(const tdscha_ensemble *ens, int *dims6);
/* Device bytes held by the ensemble (operands + workspace). */
long long tdscha_ensemble_bytes(const tdscha_ensemble *ens);

/* The perturbed averages on device pointers, enqueued on `stream` (a hipStream_t, NULL =
 * default stream). R1 (n), Y1 (n x n) in; f (n), d2v (n x n) out. terms selects the
 * reference terms; f is overwritten when TERM_F is set; d2v is overwritten, or added to
 * when d2v_accumulate is 1. Y1 need not be symmetric: its symmetric part is what the
 * reference's formulas use. */
int tdscha_ensemble_apply(tdscha_ensemble *ens, int precision, int terms, const double *R1,
                          const double *Y1, double *f, double *d2v, int d2v_accumulate,
                          void *stream);

/* Full TD-SCHA L on a device vector psi of length n + n (n + 1), as
 * Lanczos.apply_full_L (Modules/DynamicalLanczos.py:3467-3570): harmonic part,
 * anharmonic part through the ensemble, reverse and shift. out must not alias psi. */
int tdscha_apply_full_L(tdscha_ensemble *ens, int precision, int flags, double shift,
                        const double *psi, double *out, void *stream);

/* L on n_chains device vectors at once (no reference counterpart: the reference applies L
 * to one vector per call, Py:6026-6028). Chain c uses flags[c] and shifts[c]; all chains
 * of a batch share IGNORE_V3 / IGNORE_V4. Every pass over the symmetries serves all chains,
 * so a biconjugate Lanczos step (L q and L^T p) is one call with two chains. The first call
 * with more chains than before allocates the per-chain buffers. */
#define TDSCHA_MAX_CHAINS 64
int tdscha_apply_full_L_batched(tdscha_ensemble *ens, int precision, int n_chains,
                                const int *flags, const double *shifts,
                                const double *const *psi, double *const *out, void *stream);

/* Autotuner (no reference counterpart): times the batched L-application for n_chains chains
 * (at least four passes and 168 symmetry images, best of two) at slot counts s_batch,
 * s_batch/2, ... (at most six) and keeps the fastest per image; for BF16x3 it then times the
 * nine contraction kernel forms at that slot count (after 4 s of passes that bring the board
 * to its sustained clock, each over at least 0.8 s of passes, two interleaved rounds) and
 * keeps the fastest, or the default form 8 unless another one is faster by more than 1 %.
 * Returns
 * the chosen slot count and its time per (chain x configuration x symmetry) image in
 * microseconds (either may be NULL). Call before capturing Lanczos graphs on the ensemble. */
int tdscha_ensemble_autotune(tdscha_ensemble *ens, int precision, int n_chains, int *slots,
                             double *us_per_image);

/* The tuning the autotuner sets, set or read directly: bf16_tiles 0 (256/192 x 128
 * contraction tiles), 1 (256 x 256), 2 (256 x 256, operands copied straight into
 * double-buffered LDS), 3 (256/192 x 192, A operands in registers, only B through LDS),
 * 4 (the same with 96-column tiles, two workgroups per CU), 5 (GEMM1 of 4, GEMM2 of 3),
 * 6 (GEMM1 of 4, 256 x 96 GEMM2 tiles), 7 (GEMM1 of 4, 512 x 96 GEMM2 tiles of eight
 * row waves sharing one formed operand panel, 64 rows per wave) or 8 (the same with 448-row
 * tiles: four waves of 64 rows and four of 48); pass_slots image slots per pass (0 = all
 * s_batch). Results do not depend on pass_slots. Forms 0-2 differ from each other only in the
 * order of GEMM1's FP64 row sums, as do forms 3-8 (6, 7 and 8 equal 4 bit for bit: the same
 * MFMA sequence per output element); forms 3-8 add each lane's
 * row-sum products in FP32 and form gamma / delta in an FP64 product (agreement with forms
 * 0-2: ~1e-7 relative). */
int tdscha_ensemble_set_tuning(tdscha_ensemble *ens, int bf16_tiles, int pass_slots);
int tdscha_ensemble_get_tuning(const tdscha_ensemble *ens, int *bf16_tiles, int *pass_slots);

/* Allocate the per-chain buffers for n_chains chains now (what the first batched call with
 * that many chains would do), so that later calls allocate nothing. */
int tdscha_ensemble_reserve(tdscha_ensemble *ens, int n_chains);

/* ---- Force adjoint (no counterpart in the reference: the derivative of the
 * anharmonic part of apply_full_L, Py:3467-3570 with apply_anharmonic_FT Py:2891-3281, with
 * respect to the configuration forces Y) --------------------------------------------------
 * For pairs p = 0..n_pairs-1 of device vectors (u_p, v_p) of length n + n (n + 1):
 *   grad (+)= d/dY  sum_p  u_p . A_p(v_p; Y)
 * where A_p is the anharmonic part of the operator of tdscha_apply_full_L with flags[p]
 * (TRANSPOSE / WIGNER / REVERSE / IGNORE_V3 / IGNORE_V4 as there; IGNORE_V4 equal for all
 * pairs; the harmonic part and the shift do not depend on Y). grad: device, N x n
 * row-major FP64 (the layout of Y); written (accumulate = 0) or added to.
 * precision TDSCHA_PREC_FP64: FP64 MFMA throughout. TDSCHA_PREC_FP32 (ensemble created with
 * the FP32 copies): the three contractions per pass (the forward weights, the seeds' row
 * sums and the adjoint product itself) in FP32 MFMA with FP32 accumulation inside a pass;
 * rotations, weights and the sum over passes in FP64. */
int tdscha_ensemble_force_grad(tdscha_ensemble *ens, int precision, int n_pairs,
                               const int *flags, const double *const *u,
                               const double *const *v, double *grad, int accumulate,
                               void *stream);

/* ---- Device-resident Lanczos (the recurrence of Lanczos.run_FT, Py:5843-6372) ---------
 * n_chains starting vectors advance together: every step is one batched L application
 * (L q and L^T p of every chain, or L q only for the symmetric Wigner form) followed by the
 * FP64 recurrence on the device; a step is captured once as a hipGraph and replayed.
 * full_gs = 0: no re-orthogonalisation (the reference's default, n_rep_orth = 0).
 * full_gs = 1: every new pair is re-biorthogonalised against all stored pairs (two passes
 * of classical two-sided Gram-Schmidt with the duals p_norm_j p_j, FP64, inside the same
 * graph; implies keep_basis), as Lanczos.run_FT(full_gs=True) does on the host. The
 * reference's own loop (Py:6245-6317) is not what this computes. flags / shifts per chain
 * as for tdscha_apply_full_L (the TRANSPOSE bit is managed internally). mask (host, length
 * n + n (n + 1)) is the Wigner dot-product weight, needed when symmetric is 1. keep_basis
 * stores every q (and p) on the device. */
typedef struct tdscha_lanczos tdscha_lanczos;
int tdscha_lanczos_create(tdscha_lanczos **out, tdscha_ensemble *ens, int precision,
                          int n_chains, const int *flags, const double *shifts, int symmetric,
                          int max_steps, int keep_basis, int full_gs, const double *mask);
int tdscha_lanczos_destroy(tdscha_lanczos *lz);
/* psi0[c]: device vectors; they are normalised as in run_FT (Py:5936-5945). */
int tdscha_lanczos_start(tdscha_lanczos *lz, const double *const *psi0, void *stream);
/* Enqueue n_steps steps; at most max_steps since the last start (error otherwise). */
int tdscha_lanczos_run(tdscha_lanczos *lz, int n_steps, int use_graph, void *stream);
/* Host copies of the first steps_done a / b / c / s_norm_(k+1) of a chain (any may be NULL);
 * synchronises the device. */
int tdscha_lanczos_read(tdscha_lanczos *lz, int chain, double *a, double *b, double *c,
                        double *s_norm, int *steps_done);
/* full_gs only: min(|q|, |p|) of each step's new pair after the projection, before its
 * normalisation (below 1e-12 the new vector is linearly dependent on the stored ones and
 * the run ends after the next a, as run_FT's next_converged does, Py:6283-6303). */
int tdscha_lanczos_read_gs(tdscha_lanczos *lz, int chain, double *norms, int *steps_done);
/* Device pointers to the stored basis of a chain ((max_steps + 1) x (n + n (n + 1))). */
int tdscha_lanczos_basis(tdscha_lanczos *lz, int chain, const double **Q, const double **P);

/* ---- Translation average (the zone-centre operator's projector) -------------------------
 * No reference counterpart. For psi = (R (n), two packed symmetric n x n matrices) of a
 * supercell whose pure translations (n_trans x nat atom permutations irt, host) act freely on
 * the atoms, P0 psi = the average of psi over the translations (adjoint = 0), or P0^T psi in
 * the packed metric (adjoint = 1: the projector of transposed standard-representation
 * applications). pols: host (3 nat x n_modes) polarization vectors of the kept modes,
 * orthonormal columns. Device vectors; out may be psi. A projector's workspace is shared by
 * its calls: calls on one projector must not overlap. */
typedef struct tdscha_projector tdscha_projector;
int tdscha_projector_create(tdscha_projector **out, int device, int n_modes, int nat,
                            const double *pols, int n_trans, const int *irt);
int tdscha_projector_destroy(tdscha_projector *p);
int tdscha_projector_apply(tdscha_projector *p, int adjoint, const double *psi, double *out,
                           void *stream);
/* Run a Lanczos on the zone-centre operator: its ensemble holds one symmetry per rotation
 * and every L application of a step is followed by p (transposed for the L^T p of the
 * standard representation), inside the captured step. Set before the first run. */
int tdscha_lanczos_set_projector(tdscha_lanczos *lz, tdscha_projector *p);

/* Several Lanczos runs on different ensembles (systems of uneven size) advanced together:
 * one graph per step holds the step of every member, each on a branch of its own, so the
 * kernels of a small system run beside those of a large one. Members stay usable on their
 * own (read, basis); a group must not be run while a member is run alone. No reference
 * counterpart (the reference runs one Lanczos at a time). */
typedef struct tdscha_lanczos_group tdscha_lanczos_group;
int tdscha_lanczos_group_create(tdscha_lanczos_group **out, tdscha_lanczos *const *members,
                                int n_members);
int tdscha_lanczos_group_run(tdscha_lanczos_group *g, int n_steps, void *stream);
int tdscha_lanczos_group_destroy(tdscha_lanczos_group *g);

/* ---- Ensemble projection on the modes ------------------------------------------------
 * Replaces the projection inside Lanczos.__init__ (Modules/DynamicalLanczos.py:380-442,
 * X = u_tilde . pols and Y = f_tilde . pols at Py:438-439), fused with the unit and mass
 * rescaling of Py:385-426:
 *   X = (u / bohr * sqrt(m)) . pols
 *   Y = ((forces - sscha_forces) * bohr - f_mean * bohr) / sqrt(m) . pols
 * Device pointers, FP64, row-major: u, forces, sscha_forces [n_configs x n_cart]
 * (Angstrom, Ry/Angstrom); f_mean [n_cart] (the symmetrized average force tiled over the
 * supercell, Ry/Angstrom); masses [n_cart] (one per Cartesian component, Ry units);
 * pols [n_cart x n_modes] (kept modes only); X, Y [n_configs x n_modes]. sscha_forces and
 * f_mean may be NULL (zero). bohr: Angstrom per Bohr. Enqueued on `stream` (NULL = the
 * null stream); returns once launched. */
int tdscha_project_ensemble(int n_configs, int n_cart, int n_modes, const double *u,
                            const double *forces, const double *sscha_forces,
                            const double *f_mean, const double *masses, const double *pols,
                            double bohr, double *X, double *Y, void *stream);

/* ---- Symmetry operations in the mode basis --------------------------------------------
 * Replaces CellConstructor's GetSymmetriesOnModesDeg as Lanczos.prepare_symmetrization calls
 * it (Modules/DynamicalLanczos.py:735-746): for every symmetry s and block b of degenerate
 * modes [block_start[b], block_start[b+1]),
 *   out_b[s][i][j] = sum_a sum_x pols[3 irt[s][a] + x][b0 + i] (rot[s] pols[a])[x][b0 + j].
 * Device pointers: rot [n_sym x 3 x 3] Cartesian rotations, irt [n_sym x nat] the atom each
 * atom is carried to (a permutation; entries outside [0, nat) are clamped, not reported),
 * pols [3 nat x n_modes] mass-scaled polarization vectors, out the blocks one after the
 * other, each [n_sym x d_b x d_b] row-major. block_start: HOST array of n_blocks + 1
 * increasing offsets starting at 0 (blocks at most 170 modes wide). Enqueued on `stream`. */
int tdscha_symmetries_on_modes(int n_sym, int nat, int n_modes, const double *rot,
                               const int *irt, const double *pols, int n_blocks,
                               const int *block_start, double *out, void *stream);

/* ---- Weighted ensemble average (perturbation builders) ----------------------------------
 * C[m][n] = scale * sum_k A[k][m] w[k] B[k][n]: the contraction over configurations behind
 * Modules/Perturbations.py get_d1M_dR_av (:195), get_d1alpha_dR_av (:240), get_d2M_dR_av
 * (:275) and get_d2alpha_dR_av (:330), and the Upsilon product that follows the last two.
 * Device pointers, FP64, row-major with leading dimensions lda >= M, ldb >= N, ldc >= N;
 * w may be NULL (all ones). Enqueued on `stream`. */
int tdscha_weighted_tn_product(int K, int M, int N, const double *A, int lda, const double *w,
                               const double *B, int ldb, double scale, double *C, int ldc,
                               void *stream);

/* ---- Event timing of the kernel groups (no reference counterpart: measurement only) ---
 * When enabled, every launch of a group is bracketed by two HIP events on its stream;
 * query synchronises on them and returns the summed elapsed time and the launch count. */
#define TDSCHA_KGROUP_GEMM1 0  /* quadratic forms + weights  (k_gemm1_*) */
#define TDSCHA_KGROUP_GEMM2 1  /* back-accumulation          (k_gemm2_*) */
#define TDSCHA_KGROUP_ROTATE 2 /* symmetry rotations         (k_rot_*, k_split_ys) */
#define TDSCHA_KGROUP_FAV 3    /* force-average contraction  (k_fav*) */
int tdscha_profile_enable(int on);
int tdscha_profile_reset(void);
int tdscha_profile_query(int group, double *total_ms, long long *count);

#ifdef __cplusplus
}
#endif

#endif /* TDSCHA_HIP_H */
