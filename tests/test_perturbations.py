"""Perturbation builders (Modules/DynamicalLanczos.py:1003-1221, 1647-1691) and the oracle's
ensemble averages of effective charges and Raman tensors (Modules/Perturbations.py:195-375).

Pins: the seven unpolarized Raman vectors the reference's test checks
(tests/test_raman_unpol, in tests/golden/raman_unpol.npz); the IR perturbation against its
closed form on the SnTe modes; the second-order averages against Stein's identity, which
makes them exact derivatives on a Gaussian ensemble whose sample covariance is Upsilon^-1.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with np.load(os.path.join(HERE, "golden", name), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def test_unpolarized_raman_vectors_match_the_reference_test_data():
    from tdscha_amd import perturbations as pt
    g = load("raman_unpol.npz")
    T = g["raman_tensor"]
    for i in range(7):
        ref = g["raman_v_%d" % i]
        v = pt.unpolarized_raman_vector(T, i)
        assert np.abs(v - ref).max() <= 1e-12 * np.abs(ref).max(), i
    with pytest.raises(ValueError):
        pt.unpolarized_raman_vector(T, 7)
    assert pt.UNPOLARIZED_RAMAN_PREFACTORS == (5.0, 3.5, 3.5, 3.5, 21, 21, 21)


def test_raman_asr_removes_only_the_rigid_part():
    from tdscha_amd import perturbations as pt
    T = load("raman_unpol.npz")["raman_tensor"]
    shift = np.zeros_like(T)
    shift[:, :, 0::3] = 0.7          # the same x component on every atom
    v0 = pt.raman_vector(T, [1, 0, 0], [0, 1, 0])
    v1 = pt.raman_vector(T + shift, [1, 0, 0], [0, 1, 0])
    assert np.abs(v1 - v0).max() < 1e-12


def snte_lanczos():
    from tdscha_amd import DynamicalLanczos as DL
    d = load("projection_snte.npz")
    lz = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    lz.set_ensemble(d["X"], d["Y"], d["w"], d["rho"], d["T"], d["pols"], d["m"])
    return lz, d


def test_prepare_ir_closed_form():
    lz, d = snte_lanczos()
    rng = np.random.default_rng(4)
    Z = rng.standard_normal((2, 3, 3))
    pol = np.array([0.3, -0.5, 0.8])
    lz.prepare_ir(effective_charges=Z, pol_vec=pol)
    zc = np.array([[sum(Z[a, b, c] * pol[b] for b in range(3)) for c in range(3)]
                   for a in range(2)]).ravel()
    full = np.tile(zc, 8)
    expect = d["pols"].T @ (full / np.sqrt(d["m"]))
    assert np.abs(lz.psi[:45] - expect).max() < 1e-14 * np.abs(expect).max()
    assert np.all(lz.psi[45:] == 0)
    assert np.isclose(lz.perturbation_modulus, expect @ expect)
    with pytest.raises(ValueError):
        lz.prepare_ir(effective_charges=Z[:, :2])


def test_prepare_raman_unpolarized_is_the_sum_of_its_parts():
    lz, d = snte_lanczos()
    rng = np.random.default_rng(6)
    T = rng.standard_normal((3, 3, 6))
    T = 0.5 * (T + T.transpose(1, 0, 2))
    lz.prepare_raman(unpolarized=0, raman_tensor=T)
    got = lz.psi[:45].copy()
    parts = []
    for p in np.eye(3):
        lz.prepare_raman(p, p, raman_tensor=T)
        parts.append(lz.psi[:45].copy())
    assert np.abs(got - sum(parts) / 3).max() < 1e-13 * np.abs(got).max()
    lz.prepare_unpolarized_raman(0, raman_tensor=T)
    assert np.abs(lz.psi[:45] - sum(parts)).max() < 1e-13 * np.abs(got).max()
    lz.prepare_raman([1, 0, 0], [1, 0, 0], mixed=True, pol_in_2=[0, 1, 0], pol_out_2=[0, 1, 0],
                     raman_tensor=T)
    assert np.abs(lz.psi[:45] - parts[0] - parts[1]).max() < 1e-13 * np.abs(got).max()
    assert lz.get_prefactors_unpolarized_raman(4) == 21


def stein_ensemble(rng, n3, N):
    """u with mean 0 and sample covariance exactly Upsilon^-1 (rho = 1)."""
    M = rng.standard_normal((n3, n3))
    ups = M @ M.T + n3 * np.eye(n3)
    z = rng.standard_normal((N, n3))
    z -= z.mean(axis=0)
    C = z.T @ z / N
    z = z @ np.linalg.inv(np.linalg.cholesky(C)).T          # sample covariance = I
    Lc = np.linalg.cholesky(np.linalg.inv(ups))
    return z @ Lc.T, ups


def test_second_order_averages_obey_steins_identity():
    rng = np.random.default_rng(8)
    nat, N = 3, 40
    n3 = 3 * nat
    u, ups = stein_ensemble(rng, n3, N)
    rho = np.ones(N)
    A = rng.standard_normal((nat, 3, 3, n3))              # dZ[a, field, cart] / du_d
    Z0 = rng.standard_normal((nat, 3, 3))
    Z = Z0[None] + np.einsum("afcd,id->iafc", A, u)
    d2M = orc.d2M_dR_av(u, rho, Z, ups)
    # d2M[d, 3a + c, field] = dZ[a, field, c] / du_d
    expect = np.einsum("afcd->dacf", A).reshape(n3, n3, 3)
    assert np.abs(d2M - expect).max() < 1e-10 * np.abs(expect).max()
    d1M = orc.d1M_dR_av(rho, Z)
    assert np.abs(d1M - np.einsum("afc->acf", Z0).reshape(n3, 3)).max() < 1e-10
    B = rng.standard_normal((3, 3, n3, n3))
    R = np.einsum("bcxd,id->ibcx", B, u) + rng.standard_normal((3, 3, n3))[None]
    d2a = orc.d2alpha_dR_av(u, rho, R, ups)
    assert np.abs(d2a - B).max() < 1e-10 * np.abs(B).max()


# ---------------------------------------------------------------- symmetrization, two phonons
def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rock_salt_2():
    """The 2 x 2 x 2 rock-salt crystal of the synthetic model: 16 atoms, 45 kept modes in
    degenerate blocks, 384 space-group operations (Oh x 8 translations)."""
    from tdscha_amd import synthetic as sy
    cr = sy.make_crystal(2)
    w, E, blocks = sy.modes(cr, sy.force_constants(cr))
    rots, irt = sy.space_group(cr)
    starts = [b0 for b0, _ in blocks] + [blocks[-1][1]]
    S = orc.symmetries_on_modes(rots, irt, E, starts)
    return cr, w, E, starts, S, rots, irt


def _blockdiag(S, starts, g):
    n = starts[-1]
    M = np.zeros((n, n))
    for b, (b0, b1) in enumerate(zip(starts[:-1], starts[1:])):
        M[b0:b1, b0:b1] = S[b][g]
    return M


def test_group_average_matches_the_mode_blocks():
    """The oracle's space-group average acts on the atom-Cartesian indices as the symmetry
    operations do on the polarization vectors: for T built from mode-space data,
    (1/G) sum_g g.T equals the same average taken through the mode blocks
    S_g = E^T P_g E (symmetries_on_modes, pinned on the reference's SnTe blocks)."""
    _, w, E, starts, S, rots, irt = rock_salt_2()
    rng = np.random.default_rng(4)
    n = E.shape[1]
    x = rng.standard_normal(n)
    X = rng.standard_normal((n, n))
    X = X + X.T
    f = rng.standard_normal(3)
    W = rng.standard_normal((3, 3))
    W = W + W.T
    G = len(rots)
    e1 = np.zeros((E.shape[0], 3))
    e2 = np.zeros((E.shape[0], E.shape[0], 3))
    e3 = np.zeros((3, 3, E.shape[0]))
    e4 = np.zeros((3, 3, E.shape[0], E.shape[0]))
    for g in range(G):
        Sg = _blockdiag(S, starts, g)
        v = E @ (Sg @ x)
        M = E @ Sg @ X @ Sg.T @ E.T
        Rf = rots[g] @ f
        RW = rots[g] @ W @ rots[g].T
        e1 += np.outer(v, Rf)
        e2 += M[:, :, None] * Rf[None, None, :]
        e3 += RW[:, :, None] * v[None, None, :]
        e4 += RW[:, :, None, None] * M[None, None]
    C1 = np.outer(E @ x, f)
    C2 = (E @ X @ E.T)[:, :, None] * f
    C3 = W[:, :, None] * (E @ x)[None, None]
    C4 = W[:, :, None, None] * (E @ X @ E.T)[None, None]
    for kind, C, e in (("d1M", C1, e1), ("d2M", C2, e2), ("d1alpha", C3, e3),
                       ("d2alpha", C4, e4)):
        got = orc.symmetrize_average(kind, C, rots, irt, asr=False)
        scale = max(np.abs(e / G).max(), np.abs(C).max())
        assert np.abs(got - e / G).max() < 1e-12 * scale, kind
        if kind == "d1alpha":
            # rock salt is Raman inactive (inversion centres): the average vanishes
            assert np.abs(got).max() < 1e-14 * np.abs(C).max()
        # a wrong pairing of rotations and permutations fails (for d1M the Oh site symmetry
        # leaves only isotropic per-atom averages, which any pairing reproduces)
        if kind in ("d2M", "d2alpha"):
            for r_bad, i_bad in ((np.transpose(rots, (0, 2, 1)), irt),
                                 (rots, np.argsort(irt, axis=1))):
                bad = orc.symmetrize_average(kind, C, r_bad, i_bad, asr=False)
                assert rel(bad, e / G) > 1e-2, kind


def test_symmetrized_averages_obey_the_sum_rule_and_are_idempotent():
    _, _, E, _, _, rots, irt = rock_salt_2()
    rng = np.random.default_rng(5)
    n3 = E.shape[0]
    nat = n3 // 3
    shapes = {"d1M": (n3, 3), "d2M": (n3, n3, 3), "d1alpha": (3, 3, n3),
              "d2alpha": (3, 3, n3, n3)}
    for kind, shp in shapes.items():
        C = rng.standard_normal(shp)
        tol = 1e-12 * np.abs(C).max()       # (the first-order Raman average of rock salt is 0)
        S = orc.symmetrize_average(kind, C, rots, irt)
        assert np.abs(orc.symmetrize_average(kind, S, rots, irt) - S).max() < tol, kind
        if kind == "d1M":
            assert np.abs(S.reshape(nat, 3, 3).sum(0)).max() < tol
        elif kind == "d2M":
            T = S.reshape(nat, 3, nat, 3, 3)
            assert np.abs(T.sum(0)).max() < tol and np.abs(T.sum(2)).max() < tol
            assert np.abs(T - T.transpose(2, 3, 0, 1, 4)).max() < tol
        elif kind == "d1alpha":
            assert np.abs(S.reshape(3, 3, nat, 3).sum(2)).max() < tol
            assert np.abs(S - S.transpose(1, 0, 2)).max() < tol
        else:
            T = S.reshape(3, 3, nat, 3, nat, 3)
            assert np.abs(T.sum(2)).max() < tol and np.abs(T.sum(4)).max() < tol
            assert np.abs(S - S.transpose(1, 0, 2, 3)).max() < tol
            assert np.abs(S - S.transpose(0, 1, 3, 2)).max() < tol


def test_isotropic_charges_are_unchanged():
    """Rock salt: equal and opposite isotropic charges on the two species are already symmetric
    and obey the sum rule."""
    cr, _, E, _, _, rots, irt = rock_salt_2()
    q = np.where(cr.species == 0, 1.7, -1.7)
    Z = np.einsum("a, ij -> aij", q, np.eye(3)).reshape(-1, 3)      # [3 a + i, E]
    assert rel(orc.symmetrize_average("d1M", Z, rots, irt), Z) < 1e-14


def test_two_phonon_sectors_closed_forms():
    from tdscha_amd import perturbations as pt
    rng = np.random.default_rng(6)
    n = 7
    w = np.sort(rng.uniform(1e-3, 5e-3, n))
    d2 = rng.standard_normal((n, n))
    D = 0.5 * (d2 + d2.T)
    # T = 0: no occupations, chi^- = 0, chi^+ = (w_m + w_n) / (2 w_m w_n)
    a0, b0 = orc.two_phonon_sectors(d2, w, 0.0)
    assert np.all(a0 == 0)
    assert rel(b0, np.sqrt((w[:, None] + w[None]) / (4 * w[:, None] * w[None])) * D) < 1e-15
    # T > 0, one entry by hand
    T = 300.0
    nb = 1 / (np.exp(w * orc.RY_TO_K / T) - 1)
    a, b = orc.two_phonon_sectors(d2, w, T)
    i, j = 1, 5
    cm = (w[i] - w[j]) * (nb[i] - nb[j]) / (2 * w[i] * w[j])
    cp = (w[i] + w[j]) * (1 + nb[i] + nb[j]) / (2 * w[i] * w[j])
    assert abs(a[i, j] + np.sqrt(-0.5 * cm) * D[i, j]) < 1e-15 * abs(D[i, j]) + 1e-300
    assert abs(b[i, j] - np.sqrt(0.5 * cp) * D[i, j]) < 1e-14 * abs(D[i, j])
    # the library's host form is the oracle's
    pa, pb = pt.wigner_two_phonon(d2, w, T)
    assert rel(pa, a) < 1e-15 and rel(pb, b) < 1e-15
    # packing: upper triangles row by row, a'(1) then b'(1)
    psi = orc.pack_two_phonon(np.zeros(n + n * (n + 1)), a, b)
    iu = np.triu_indices(n)
    nw2 = len(iu[0])
    assert np.all(psi[:n] == 0)
    assert np.all(psi[n:n + nw2] == a[iu]) and np.all(psi[n + nw2:] == b[iu])
    assert np.all(pt.pack_wigner_two_phonon(np.zeros_like(psi), a, b) == psi)
