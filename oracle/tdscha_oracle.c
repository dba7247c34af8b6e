/*
 * CPU oracle for the symmetry-expanded perturbed ensemble averages of TD-SCHA.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the HIP path is compared
 * against; nothing in the product library links or calls it.
 *
 * It restates, per image (configuration i, symmetry s), the arithmetic of the
 * reference's three "*_pert_sym_fast" routines in one fused pass:
 *   - get_f_average_from_Y_pert_sym_fast   CModules/LanczosFunctions.c:2082-2254
 *   - get_d2v_dR2_from_R_pert_sym_fast     CModules/LanczosFunctions.c:2257-2438
 *   - get_d2v_dR2_from_Y_pert_sym_fast     CModules/LanczosFunctions.c:2615-2754
 *   - Bose helpers f_ups / f_psi           CModules/LanczosFunctions.c:14-30
 *   - RY_TO_K                              CModules/LanczosFunctions.h:9
 * with the symmetry-image construction of LanczosFunctions.c:2158-2170
 * (block b acts on the contiguous modes degenerate_space[b][0..d_b-1]).
 *
 * Per image, with u = S_s x_i, f = S_s y_i, ut = f_ups*u, ft = f_psi*f:
 *   alpha = u^T Y1 u
 *   beta  = (u^T Y1 ft + ft^T Y1 u) / 2
 *   gamma = ut . R1,   delta = f . R1
 *   f_av  += rho_i * ( -alpha/6 * f - beta/3 * ut )
 *   d2v   += rho_i * ( c1 (ut f^T + f ut^T) + c2 ut ut^T )
 *   c1 = kappa*alpha/8 - gamma/3, c2 = kappa*beta/4 - delta/3
 * and everything is divided by N_eff * N_sym at the end, N_eff = sum(rho).
 *
 * Optionally the same sums of absolute values of every term are returned
 * (f_abs, d2v_abs): the scale against which the rounding error of any
 * summation order is bounded, used to state parity tolerances for outputs
 * that cancel (e.g. a force that vanishes by symmetry).
 *
 * Accumulation is in long double (x87 80-bit, 64-bit mantissa) so that the
 * oracle is more accurate than either the reference or the device path.
 * Config ranges are split over threads in fixed contiguous slices and the
 * per-thread partial sums are added in thread order, so results do not
 * depend on scheduling.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_RY_TO_K 157887.32400374097

/* Occupation-dependent eigenvalues, LanczosFunctions.c:14-30. */
static double bose_n(double w, double T) {
    return (T > 0) ? 1.0 / (exp(w * ORACLE_RY_TO_K / T) - 1.0) : 0.0;
}
double oracle_f_ups(double w, double T) { return 2 * w / (1 + 2 * bose_n(w, T)); }
double oracle_f_psi(double w, double T) { return (1 + 2 * bose_n(w, T)) / (2 * w); }

enum { ORACLE_F = 1, ORACLE_D2V_R = 2, ORACLE_D2V_Y = 4 };

/*
 * Fused per-image pass over configs [cfg_start, cfg_stop) and symmetries
 * [sym_start, sym_stop). `what` selects which of the three reference terms
 * are accumulated (bit mask of ORACLE_*). Outputs are overwritten with the
 * partial sums normalised by norm_neff * n_sym_norm, so that slices computed
 * with the same normalisation add up to the full result (the additivity the
 * reference's MPI split relies on, LanczosFunctions.c:2123-2136).
 *
 * symmetries[b] is (N_sym, d_b, d_b) row-major, as the reference binding
 * passes it (odd_corr_module.c:585-587).
 */
int oracle_perturb_average_sym_range(
    const double *X, const double *Y, const double *w, const double *rho,
    const double *R1, const double *Y1, double T, int n_modes, int n_configs,
    int N_sym, int n_blocks, const double *const *symmetries,
    const int *N_degeneracy, const int *const *degenerate_space,
    const int *blocks_ids, int what, int cfg_start, int cfg_stop,
    int sym_start, int sym_stop, double norm_neff, double n_sym_norm,
    int n_threads, double *f_out, double *d2v_out, double *f_abs, double *d2v_abs)
{
    const int n = n_modes;
    if (n <= 0 || n_configs < 0 || N_sym <= 0 || cfg_start < 0 || cfg_stop > n_configs ||
        sym_start < 0 || sym_stop > N_sym || n_threads <= 0)
        return 1;
    (void)n_blocks;

    double *fu = (double *)malloc(sizeof(double) * n);
    double *fp = (double *)malloc(sizeof(double) * n);
    for (int m = 0; m < n; ++m) {
        fu[m] = oracle_f_ups(w[m], T);
        fp[m] = oracle_f_psi(w[m], T);
    }

    int nt = n_threads;
    long double *f_part = (long double *)calloc((size_t)nt * n, sizeof(long double));
    long double *d_part = (long double *)calloc((size_t)nt * n * n, sizeof(long double));
    const int want_abs = (f_abs != NULL) && (d2v_abs != NULL);
    long double *fa_part = want_abs ? (long double *)calloc((size_t)nt * n, sizeof(long double)) : NULL;
    long double *da_part = want_abs ? (long double *)calloc((size_t)nt * n * n, sizeof(long double)) : NULL;
    const int want_f = (what & ORACLE_F) != 0;
    const int want_r = (what & ORACLE_D2V_R) != 0;
    const int want_y = (what & ORACLE_D2V_Y) != 0;
    const int want_d = want_r || want_y;

#pragma omp parallel num_threads(nt)
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        int total = cfg_stop - cfg_start;
        int chunk = total / nth, rem = total % nth;
        int my0 = cfg_start + tid * chunk + (tid < rem ? tid : rem);
        int my1 = my0 + chunk + (tid < rem ? 1 : 0);

        long double *u = (long double *)malloc(sizeof(long double) * n);
        long double *f = (long double *)malloc(sizeof(long double) * n);
        long double *ut = (long double *)malloc(sizeof(long double) * n);
        long double *ft = (long double *)malloc(sizeof(long double) * n);
        long double *fpart = f_part + (size_t)tid * n;
        long double *dpart = d_part + (size_t)tid * n * n;
        long double *fapart = want_abs ? fa_part + (size_t)tid * n : NULL;
        long double *dapart = want_abs ? da_part + (size_t)tid * n * n : NULL;

        for (int i = my0; i < my1; ++i) {
            const double *x = X + (size_t)i * n;
            const double *y = Y + (size_t)i * n;
            const long double r = rho[i];
            for (int s = sym_start; s < sym_stop; ++s) {
                /* Image: block-diagonal S_s applied to x_i and y_i. */
                for (int m = 0; m < n; ++m) {
                    int b = blocks_ids[m];
                    int d = N_degeneracy[b];
                    int m0 = degenerate_space[b][0];
                    const double *S = symmetries[b] + (size_t)s * d * d + (size_t)(m - m0) * d;
                    long double au = 0, af = 0;
                    for (int k = 0; k < d; ++k) {
                        int nu = degenerate_space[b][k];
                        au += (long double)S[k] * x[nu];
                        af += (long double)S[k] * y[nu];
                    }
                    u[m] = au;
                    f[m] = af;
                    ut[m] = fu[m] * au;
                    ft[m] = fp[m] * af;
                }
                long double alpha = 0, beta = 0, gamma = 0, delta = 0;
                if (want_f || want_y) {
                    for (int a = 0; a < n; ++a) {
                        long double qa = 0, qta = 0; /* (Y1 u)_a and (Y1^T u)_a */
                        const double *row = Y1 + (size_t)a * n;
                        for (int b2 = 0; b2 < n; ++b2) {
                            qa += (long double)row[b2] * u[b2];
                            qta += (long double)Y1[(size_t)b2 * n + a] * u[b2];
                        }
                        alpha += u[a] * qa;
                        beta += ft[a] * (qa + qta);
                    }
                    beta *= 0.5L;
                }
                if (want_r) {
                    for (int m = 0; m < n; ++m) {
                        gamma += ut[m] * R1[m];
                        delta += f[m] * R1[m];
                    }
                }
                if (want_f) {
                    long double e1 = -r * alpha / 6, e2 = -r * beta / 3;
                    for (int m = 0; m < n; ++m) fpart[m] += e1 * f[m] + e2 * ut[m];
                    if (want_abs)
                        for (int m = 0; m < n; ++m)
                            fapart[m] += fabsl(e1 * f[m]) + fabsl(e2 * ut[m]);
                }
                if (want_d) {
                    long double c1 = 0, c2 = 0;
                    if (want_y) { c1 += alpha / 8; c2 += beta / 4; }
                    if (want_r) { c1 -= gamma / 3; c2 -= delta / 3; }
                    c1 *= r;
                    c2 *= r;
                    for (int a = 0; a < n; ++a) {
                        long double pa = c1 * ut[a], qa = c1 * f[a], sa = c2 * ut[a];
                        long double *drow = dpart + (size_t)a * n;
                        for (int b2 = 0; b2 < n; ++b2)
                            drow[b2] += pa * f[b2] + qa * ut[b2] + sa * ut[b2];
                        if (want_abs) {
                            long double *arow = dapart + (size_t)a * n;
                            for (int b2 = 0; b2 < n; ++b2)
                                arow[b2] += fabsl(pa * f[b2]) + fabsl(qa * ut[b2]) + fabsl(sa * ut[b2]);
                        }
                    }
                }
            }
        }
        free(u); free(f); free(ut); free(ft);
    }

    long double norm = (long double)norm_neff * (long double)n_sym_norm;
    if (want_f) {
        for (int m = 0; m < n; ++m) {
            long double acc = 0;
            for (int t = 0; t < nt; ++t) acc += f_part[(size_t)t * n + m];
            f_out[m] = (double)(acc / norm);
        }
    }
    if (want_d) {
        for (size_t e = 0; e < (size_t)n * n; ++e) {
            long double acc = 0;
            for (int t = 0; t < nt; ++t) acc += d_part[(size_t)t * n * n + e];
            d2v_out[e] = (double)(acc / norm);
        }
    }
    if (want_abs) {
        for (int m = 0; m < n; ++m) {
            long double acc = 0;
            for (int t = 0; t < nt; ++t) acc += fa_part[(size_t)t * n + m];
            f_abs[m] = (double)(acc / norm);
        }
        for (size_t e = 0; e < (size_t)n * n; ++e) {
            long double acc = 0;
            for (int t = 0; t < nt; ++t) acc += da_part[(size_t)t * n * n + e];
            d2v_abs[e] = (double)(acc / norm);
        }
        free(fa_part); free(da_part);
    }
    free(f_part); free(d_part); free(fu); free(fp);
    return 0;
}

/*
 * Whole-ensemble convenience form with the output contract of the reference
 * binding GetPerturbAverageSym (odd_corr_module.c:613-627): f is overwritten,
 * d2v is overwritten with the R1 (D3) part and the Y1 (D4) part is added when
 * apply_d4 is 1.
 */
int oracle_perturb_average_sym(
    const double *X, const double *Y, const double *w, const double *rho,
    const double *R1, const double *Y1, double T, int apply_d4, int n_modes,
    int n_configs, int N_sym, int n_blocks, const double *const *symmetries,
    const int *N_degeneracy, const int *const *degenerate_space,
    const int *blocks_ids, int n_threads, double *f_out, double *d2v_out)
{
    if (apply_d4 != 0 && apply_d4 != 1) return 2;
    long double neff = 0;
    for (int i = 0; i < n_configs; ++i) neff += rho[i];
    int what = ORACLE_F | ORACLE_D2V_R | (apply_d4 ? ORACLE_D2V_Y : 0);
    return oracle_perturb_average_sym_range(
        X, Y, w, rho, R1, Y1, T, n_modes, n_configs, N_sym, n_blocks, symmetries,
        N_degeneracy, degenerate_space, blocks_ids, what, 0, n_configs, 0, N_sym,
        (double)neff, (double)N_sym, n_threads, f_out, d2v_out, NULL, NULL);
}
