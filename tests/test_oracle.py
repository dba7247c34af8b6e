"""Pins of the CPU oracle: the reference's own C build, its golden vectors, its Krylov
fixtures and known-answer test, and closed forms."""
import numpy as np
import pytest

import fixtures
from oracle import oracle as orc


def _scaled(out, ref, scale):
    return np.linalg.norm(out - ref) / np.linalg.norm(scale)


def test_oracle_matches_reference_golden_vectors(mode10):
    kv = fixtures.load_npz("ref_kernel_vectors.npz")
    for k in range(int(kv["n_cases"])):
        f, d, fa, da = orc.perturb_average_sym(mode10.X, mode10.Y, mode10.w, mode10.rho,
                                               kv["R1_%d" % k], kv["Y1_%d" % k], mode10.T,
                                               int(kv["d4_%d" % k]), mode10.sym, with_abs=True)
        assert _scaled(f, kv["f_%d" % k], fa) < 1e-14
        assert _scaled(d, kv["d2v_%d" % k], da) < 1e-14


@pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed,d4", [(1, 1), (2, 0)])
def test_oracle_matches_reference_build_random(seed, d4):
    ens = fixtures.random_ensemble(seed, 23, 9, 5)
    R1, Y1 = fixtures.random_perturbation(seed, 23, symmetric=False)
    f, d, fa, da = orc.perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, d4,
                                           ens.sym, with_abs=True)
    fr, dr = orc.ref_perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, d4, ens.sym)
    assert _scaled(f, fr, fa) < 1e-14
    assert _scaled(d, dr, da) < 1e-14


@pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("rel,d4", [(1e-6, 1), (0.05, 1), (0.05, 0)])
def test_oracle_matches_reference_build_with_split_blocks(rel, d4):
    """Blocks whose modes differ in frequency: the reference applies f_ups / f_psi per mode
    to each image (C:2214-2221, 2401-2407, 2708-2725), as the oracle does."""
    ens = fixtures.split_frequencies(fixtures.random_ensemble(95, 23, 9, 5, max_block=6), rel)
    R1, Y1 = fixtures.random_perturbation(96, 23)
    f, d, fa, da = orc.perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, d4,
                                           ens.sym, with_abs=True)
    fr, dr = orc.ref_perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, d4, ens.sym)
    assert _scaled(f, fr, fa) < 1e-14
    assert _scaled(d, dr, da) < 1e-14
    # and the split is visible: the result differs from the block-averaged frequencies'
    eq = orc.Ensemble(ens.X, ens.Y, fixtures.random_ensemble(95, 23, 9, 5, max_block=6).w,
                      ens.rho, ens.T, ens.sym)
    f0, d0 = orc.perturb_average_sym(eq.X, eq.Y, eq.w, eq.rho, R1, Y1, eq.T, d4, eq.sym)
    assert np.linalg.norm(d - d0) > 1e3 * 1e-14 * np.linalg.norm(da)


def test_symmetry_expansion_equals_unwrapped_ensemble():
    """Sym-expanded average == plain average over the explicit images
    (tests/_heavy_tests/test_fast_lanczos_snse/test_fast_lanczos_snse.py:66-76)."""
    ens = fixtures.random_ensemble(5, 12, 4, 6)
    R1, Y1 = fixtures.random_perturbation(6, 12)
    f, d = orc.perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 1, ens.sym)
    Sd = [ens.sym.dense(s) for s in range(ens.sym.n_syms)]
    Xu = np.concatenate([ens.X @ S.T for S in Sd])
    Yu = np.concatenate([ens.Y @ S.T for S in Sd])
    rho_u = np.concatenate([ens.rho] * len(Sd))
    fu, du = orc.perturb_average_sym(Xu, Yu, ens.w, rho_u, R1, Y1, ens.T, 1,
                                     orc.SymBlocks.identity(12))
    assert np.linalg.norm(fu - f) < 1e-12 * np.linalg.norm(f)
    assert np.linalg.norm(du - d) < 1e-12 * np.linalg.norm(d)


def test_invariants():
    ens = fixtures.random_ensemble(8, 10, 7, 3)
    R1, Y1 = fixtures.random_perturbation(9, 10)
    f, d = orc.perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 1, ens.sym)
    # rho -> c rho leaves the normalised average unchanged
    f2, d2 = orc.perturb_average_sym(ens.X, ens.Y, ens.w, 3.7 * ens.rho, R1, Y1, ens.T, 1, ens.sym)
    assert np.allclose(f2, f, rtol=1e-13, atol=0) and np.allclose(d2, d, rtol=1e-13, atol=0)
    # additivity over config slices (the reference's MPI split, LanczosFunctions.c:2123-2136)
    norm = (float(np.sum(ens.rho)), float(ens.sym.n_syms))
    fa, da = orc.perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 1, ens.sym,
                                     cfg_range=(0, 3), norm=norm)
    fb, db = orc.perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 1, ens.sym,
                                     cfg_range=(3, 7), norm=norm)
    assert np.allclose(fa + fb, f, rtol=1e-13, atol=1e-30)
    assert np.allclose(da + db, d, rtol=1e-12, atol=1e-30)
    # d2v symmetric; zero perturbation gives zero
    assert np.abs(d - d.T).max() <= 1e-15 * np.abs(d).max()
    f0, d0 = orc.perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, 0 * R1, 0 * Y1, ens.T, 1, ens.sym)
    assert not f0.any() and not d0.any()


def test_full_L_reproduces_reference_recurrence(mode10):
    st = fixtures.load_npz("lanczos_std.npz")
    Q, a, b, c = st["Q"], st["a"], st["b"], st["c"]
    for k in range(10):
        Lq = orc.apply_full_L(mode10, Q[k])
        rhs = a[k] * Q[k] + b[k] * Q[k + 1] + (c[k - 1] * Q[k - 1] if k else 0)
        assert np.linalg.norm(Lq - rhs) / np.linalg.norm(Lq) < 1e-10


def test_full_L_transpose_reproduces_P_recurrence(mode10):
    """L^T p_k = (s_k+1 / pn_k) p_k+1 + a_k p_k + b_k-1 (pn_k-1 / pn_k) p_k-1, with
    pn_k = s_norm_k / c_k-1 (Py:6059-6113)."""
    st = fixtures.load_npz("lanczos_std.npz")
    P, a, b, c, s = st["P"], st["a"], st["b"], st["c"], st["s_norm"]
    pn = [s[0]] + [s[k] / c[k - 1] for k in range(1, 11)]
    for k in range(10):
        Lp = orc.apply_full_L(mode10, P[k], transpose=True)
        rhs = (s[k + 1] / pn[k]) * P[k + 1] + a[k] * P[k]
        if k:
            rhs = rhs + b[k - 1] * (pn[k - 1] / pn[k]) * P[k - 1]
        assert np.linalg.norm(Lp - rhs) / np.linalg.norm(Lp) < 1e-9, k


@pytest.mark.parametrize("kw", [dict(), dict(ignore_v4=True), dict(ignore_v3=True),
                                dict(reverse_L=True, shift_value=2e-6)])
def test_transposed_L_is_the_exact_transpose(mode10, kw):
    """<p, L q> = <L^T p, q> for the packed psi basis (what rmatvec relies on, Py:470-479)."""
    rng = np.random.default_rng(5)
    n = orc.psi_len(45)
    p, q = rng.standard_normal(n), rng.standard_normal(n)
    a = p.dot(orc.apply_full_L(mode10, q, **kw))
    b = orc.apply_full_L(mode10, p, transpose=True, **kw).dot(q)
    assert abs(a - b) <= 1e-13 * abs(a)


def test_wigner_L_is_self_adjoint_in_masked_product(mode10):
    """The symmetric Lanczos (run_simm) relies on <p, L q>_mask = <L p, q>_mask."""
    rng = np.random.default_rng(6)
    n = orc.psi_len(45)
    m = orc.mask_dot_wigner(45)
    p, q = rng.standard_normal(n), rng.standard_normal(n)
    a = p.dot(m * orc.apply_full_L(mode10, q, use_wigner=True))
    b = orc.apply_full_L(mode10, p, use_wigner=True).dot(m * q)
    assert abs(a - b) <= 1e-13 * abs(a)


def test_run_FT_kat_and_coefficients(mode10):
    st = fixtures.load_npz("lanczos_std.npz")
    run = orc.run_FT(lambda q: orc.apply_full_L(mode10, q),
                     lambda p: orc.apply_full_L(mode10, p, transpose=True), st["Q"][0], 10)
    c = np.asarray(run.c)
    assert abs(c[4] - 6.59744344e-07) / 6.59744344e-07 < 1e-6
    assert np.max(np.abs(np.asarray(run.a) - st["a"]) / np.abs(st["a"])) < 1e-10
    assert np.max(np.abs(c[:8] - st["c"][:8]) / np.abs(st["c"][:8])) < 1e-7


def test_run_FT_wigner_symmetric(mode10):
    st = fixtures.load_npz("lanczos_wigner.npz")
    run = orc.run_FT(lambda q: orc.apply_full_L(mode10, q, use_wigner=True), None, st["Q"][0], 10,
                     run_simm=True, mask=orc.mask_dot_wigner(45))
    assert np.max(np.abs(np.asarray(run.a) - st["a"]) / np.abs(st["a"])) < 1e-10
    assert np.max(np.abs(np.asarray(run.b[:8]) - st["b"][:8]) / np.abs(st["b"][:8])) < 1e-6


def test_continued_fraction_equals_tridiagonal_resolvent():
    """Without terminator, G(w) = e0^T (T - z)^-1 e0 for the tridiagonal T of the a/b/c
    coefficients, z = w^2 - 2 i w eta (Py:4951-5047 normal representation)."""
    rng = np.random.default_rng(3)
    k = 7
    a = rng.uniform(1, 2, k)
    b = rng.uniform(0.1, 0.5, k - 1)
    c = rng.uniform(0.1, 0.5, k - 1)
    Tm = np.diag(a) + np.diag(b, -1) + np.diag(c, 1)
    w = np.linspace(0.1, 1.7, 31)
    eta = 0.03
    g = orc.green_function_continued_fraction(a, b, c, w, use_terminator=False, smearing=eta)
    ref = np.array([np.linalg.solve(Tm - (x * x - 2j * x * eta) * np.eye(k), np.eye(k)[0])[0]
                    for x in w])
    assert np.allclose(g, ref, rtol=1e-12, atol=0)


def test_continued_fraction_terminator_is_tail_fixed_point():
    """The terminator is the infinite constant tail: g = 1/(a - w^2 - b c g)."""
    a, b, c = np.array([1.3]), np.array([0.4]), np.array([0.25])
    w = np.linspace(0.0, 2.0, 9)
    g = orc.green_function_continued_fraction(a, b, c, w, use_terminator=True)
    assert np.allclose(g, 1 / (a[0] - w ** 2 - b[0] * c[0] * g), rtol=1e-12, atol=0)
