"""Host-side logic of the Lanczos mirror that needs no GPU: status files, psi helpers."""
import numpy as np
import pytest

import fixtures
from oracle import oracle as orc


def _lanczos(ens):
    from tdscha_amd import DynamicalLanczos as DL
    lz = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    lz.set_ensemble(ens.X, ens.Y, ens.w, ens.rho, ens.T)
    s = ens.sym
    lz.set_symmetries(s.symmetries, s.N_degeneracy, s.degenerate_space, s.block_id)
    return lz


def test_status_round_trip(tmp_path, mode10):
    st = fixtures.load_npz("lanczos_std.npz")
    lz = _lanczos(mode10)
    lz.psi = st["psi"].copy()
    lz.a_coeffs, lz.b_coeffs, lz.c_coeffs = list(st["a"]), list(st["b"]), list(st["c"])
    lz.basis_Q, lz.basis_P, lz.s_norm = list(st["Q"]), list(st["P"]), list(st["s_norm"])
    f = str(tmp_path / "status")
    lz.save_status(f)
    from tdscha_amd import DynamicalLanczos as DL
    lz2 = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    lz2.load_status(f + ".npz")
    assert np.array_equal(lz2.X, mode10.X) and np.array_equal(lz2.w, mode10.w)
    assert np.array_equal(np.asarray(lz2.basis_Q), st["Q"])
    assert np.array_equal(np.asarray(lz2.c_coeffs), st["c"])
    for b in range(mode10.sym.n_blocks):
        assert np.array_equal(lz2.symmetries[b], mode10.sym.symmetries[b])
        assert np.array_equal(lz2.degenerate_space[b], mode10.sym.degenerate_space[b])
    assert np.array_equal(lz2.sym_block_id, mode10.sym.block_id)


def test_padded_blocks_without_sizes(mode10):
    """Reference status files store sym_julia / deg_julia only (Py:753-768)."""
    from tdscha_amd.DynamicalLanczos import blocks_from_padded
    lz = _lanczos(mode10)
    sj, dj = lz._padded_symmetries()
    syms, ndeg, degs, bid = blocks_from_padded(sj, dj, 45)
    assert np.array_equal(ndeg, mode10.sym.N_degeneracy)
    assert np.array_equal(bid, mode10.sym.block_id)


def test_psi_helpers_match_oracle(mode10):
    st = fixtures.load_npz("lanczos_std.npz")
    lz = _lanczos(mode10)
    lz.psi = st["Q"][4].copy()
    n = 45
    nw2 = n * (n + 1) // 2
    for half in (False, True):
        assert np.array_equal(lz.get_Y1(half), orc.unpack_sym(lz.psi[n:n + nw2], n, half))
        assert np.array_equal(lz.get_ReA1(half), orc.unpack_sym(lz.psi[n + nw2:], n, half))
    assert np.array_equal(lz.mask_dot_wigner(), orc.mask_dot_wigner(n))


def test_lanczos_requires_init_and_rejects_cpu_modes(mode10):
    from tdscha_amd import DynamicalLanczos as DL
    with pytest.raises(ValueError):
        DL.Lanczos(mode=DL.MODE_FAST_SERIAL)
    lz = _lanczos(mode10)
    with pytest.raises(ValueError):
        lz.apply_full_L(np.zeros(len(lz.psi)))


def test_continued_fraction_host_matches_oracle():
    from tdscha_amd.DynamicalLanczos import continued_fraction
    st = fixtures.load_npz("lanczos_std.npz")
    w = np.linspace(0, 3e-3, 101)
    for term in (True, False):
        for wig in (False, True):
            g = continued_fraction(st["a"], st["b"], st["c"], w, term, 1, 1e-5, wig)
            g_o = orc.green_function_continued_fraction(st["a"], st["b"], st["c"], w, term, 1,
                                                        1e-5, wig)
            assert np.allclose(g, g_o, rtol=1e-12, atol=0)


def test_wigner_alpha_beta_and_their_inverse(mode10):
    """get_alpha1_beta1_wigner against the oracle's restatement (pinned through the full L on
    the reference's Wigner Krylov vectors), and get_a1_b1_wigner as its inverse: it must give
    back b'(1) exactly and sqrt(-chi^-/2) a'(1), at T > 0 and at T = 0 (chi^- = 0)."""
    for T in (300.0, 0.0):
        lz = _lanczos(mode10)
        lz.T = T
        lz.use_wigner = True
        n = lz.n_modes
        nw2 = n * (n + 1) // 2
        lz.psi = np.random.default_rng(11).standard_normal(n + 2 * nw2)
        alpha = lz.get_alpha1_beta1_wigner(True)
        ref = orc.wigner_alpha1(lz.psi, mode10.w, T)
        assert np.allclose(alpha, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
        assert np.array_equal(alpha, alpha.T)
        beta = lz.get_alpha1_beta1_wigner(False)
        assert np.array_equal(beta, beta.T)
        b1 = lz.get_a1_b1_wigner(False)
        b_true = lz.psi[n + nw2:]
        assert np.allclose(b1, b_true, rtol=1e-12, atol=1e-12 * np.abs(b_true).max())
        iu = np.triu_indices(n)
        chi_m, _ = orc.chi_minus_plus(mode10.w, T)
        a1 = lz.get_a1_b1_wigner(True)
        a_scaled = np.sqrt(-0.5 * chi_m[iu[1], iu[0]]) * lz.psi[n:n + nw2]
        # at T = 0 chi^- vanishes: a1 is the rounding of the cancelling b' terms
        scale = max(np.abs(a_scaled).max(), np.abs(b_true).max())
        assert np.max(np.abs(a1 - a_scaled)) <= 1e-12 * scale
        if T > 0:
            assert np.abs(a_scaled).max() > 1e-3 * scale


def test_prepare_two_ph(mode10):
    """The two-phonon perturbation on a mode pair equals the oracle's two-phonon sectors of
    D = (E_ab + E_ba) / 2 (pinned by closed forms), with the Wigner-mask modulus."""
    from tdscha_amd import DynamicalLanczos as DL
    lz = _lanczos(mode10)
    lz.verbose = False
    with pytest.raises(NotImplementedError):
        lz.prepare_two_ph(1, 2)
    lz.use_wigner = True
    n = lz.n_modes
    for a, b in ((3, 7), (5, 5)):
        lz.prepare_two_ph(a, b)
        D = np.zeros((n, n))
        D[a, b] += 0.5
        D[b, a] += 0.5
        want = orc.pack_two_phonon(np.zeros(n + n * (n + 1)),
                                   *orc.two_phonon_sectors(D, mode10.w, mode10.T))
        assert np.allclose(lz.psi, want, rtol=0, atol=1e-15 * np.abs(want).max())
        assert not np.any(lz.psi[:n])
        mask = orc.mask_dot_wigner(n)
        assert np.isclose(lz.perturbation_modulus, want.dot(want * mask), rtol=1e-14)
    with pytest.raises(ValueError):
        lz.prepare_two_ph(n, 0)


def _dense_L(ens):
    n_psi = orc.psi_len(len(ens.w))
    return np.array([orc.apply_full_L(ens, v) for v in np.eye(n_psi)]).T


@pytest.mark.parametrize("reverse", [False, True])
def test_lanczos_matrix_from_coefficients_has_the_spectrum_of_L(reverse):
    """On a complete Krylov space (2 modes, 8-dimensional psi) the matrix that
    build_lanczos_matrix_from_coeffs makes from a full-GS run's a, b, c has the eigenvalues
    of L; with set_max_frequency the run is made on 4 f^2 - L and the matrix maps back."""
    from tdscha_amd import DynamicalLanczos as DL
    ens = fixtures.random_ensemble(13, 2, 12, 2, max_block=2)
    Ld = _dense_L(ens)
    want = np.sort(np.linalg.eigvals(Ld).real)
    lz = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    lz.verbose = False
    op = Ld
    if reverse:
        lz.set_max_frequency(float(np.max(ens.w)))
        op = -Ld + lz.shift_value * np.eye(len(Ld))
    psi0 = np.random.default_rng(2).standard_normal(len(Ld))
    run = orc.run_FT(lambda q: op @ q, lambda p: op.T @ p, psi0, len(Ld), full_gs=True)
    lz.a_coeffs, lz.b_coeffs, lz.c_coeffs = list(run.a), list(run.b), list(run.c)
    T = lz.build_lanczos_matrix_from_coeffs()
    assert T.shape == (len(run.a), len(run.a))
    got = np.linalg.eigvals(T).real
    # every eigenvalue of T is one of L's, and every distinct eigenvalue of L appears
    scale = np.max(np.abs(want))
    distinct = [want[0]] + [x for p, x in zip(want, want[1:]) if x - p > 1e-8 * scale]
    assert len(got) == len(distinct)
    assert max(np.min(np.abs(want - x)) for x in got) <= 1e-8 * scale
    assert max(np.min(np.abs(got - x)) for x in distinct) <= 1e-8 * scale
    with pytest.raises(NotImplementedError):
        lz.build_lanczos_matrix_from_coeffs(use_arnoldi=True)
