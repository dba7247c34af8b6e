"""Zone-centre reduction of the image sum (tdscha_amd.gamma): for a translation-invariant
perturbation the operator over all N_sym images equals the translation average of the
operator over one symmetry per rotation. Pinned here on the oracle (the C restatement of
the reference's kernel, CModules/LanczosFunctions.c:2082-2754, and the numpy operator layer of
Modules/DynamicalLanczos.py:2891-3570) on the 2 x 2 x 2 rock-salt crystal (384 operations =
48 rotations x 8 translations)."""
import numpy as np
import pytest

from oracle import oracle as orc


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def rock_salt():
    from tdscha_amd import symmetry, synthetic as sy
    ens = sy.make_ensemble(2, 300, T=250.0, weights="lognormal")
    rots, irt = sy.space_group(ens.crystal)
    reps, tr = symmetry.translation_cosets(rots, irt)
    sym = orc.SymBlocks(ens.symmetries, ens.N_degeneracy, ens.degenerate_space, ens.block_id)
    rsym = orc.SymBlocks([s[reps] for s in ens.symmetries], ens.N_degeneracy,
                         ens.degenerate_space, ens.block_id)
    full = orc.Ensemble(ens.X, ens.Y, ens.w, ens.rho, ens.T, sym)
    rep = orc.Ensemble(ens.X, ens.Y, ens.w, ens.rho, ens.T, rsym)
    return ens, rots, irt, reps, tr, full, rep


def test_cosets_of_the_translations(rock_salt):
    from tdscha_amd import symmetry, synthetic as sy
    ens, rots, irt, reps, tr, _, _ = rock_salt
    assert len(reps) == 48 and len(tr) == 8
    assert np.all(np.abs(rots[tr] - np.eye(3)).max(axis=(1, 2)) == 0)
    assert len({tuple(np.round(r.ravel(), 8)) for r in rots[reps]}) == 48
    assert symmetry.translation_cosets(rots[:100], irt[:100]) is None     # not a group
    bad = irt.copy()
    bad[reps[1]] = bad[reps[1]][::-1]                                       # broken coset
    assert symmetry.translation_cosets(rots, bad) is None
    cr6 = sy.make_crystal(6)
    r6, i6 = sy.space_group(cr6)
    reps6, tr6 = symmetry.translation_cosets(r6, i6)
    assert len(reps6) == 48 and len(tr6) == 216


@pytest.mark.parametrize("wigner", [False, True])
@pytest.mark.parametrize("transpose", [False, True])
def test_full_operator_is_the_translation_average_over_one_symmetry_per_rotation(
        rock_salt, wigner, transpose):
    ens, rots, irt, reps, tr, full, rep = rock_salt
    n = len(ens.w)
    adjoint = transpose and not wigner
    rng = np.random.default_rng(1)
    psi = orc.translation_average_psi(rng.standard_normal(n + n * (n + 1)), ens.pols, irt[tr],
                                      adjoint)
    assert rel(orc.translation_average_psi(psi, ens.pols, irt[tr], adjoint), psi) < 1e-13
    a = orc.apply_full_L(full, psi, transpose=transpose, use_wigner=wigner)
    b = orc.translation_average_psi(
        orc.apply_full_L(rep, psi, transpose=transpose, use_wigner=wigner), ens.pols, irt[tr],
        adjoint)
    assert rel(b, a) < 1e-12
    # the operator keeps the sector
    assert rel(orc.translation_average_psi(a, ens.pols, irt[tr], adjoint), a) < 1e-12


def test_reduction_needs_the_invariant_sector(rock_salt):
    """A perturbation with components outside the sector is not reproduced, and the wrong
    projector for the transposed standard form fails."""
    ens, rots, irt, reps, tr, full, rep = rock_salt
    n = len(ens.w)
    rng = np.random.default_rng(2)
    psi = rng.standard_normal(n + n * (n + 1))
    a = orc.apply_full_L(full, psi)
    b = orc.translation_average_psi(orc.apply_full_L(rep, psi), ens.pols, irt[tr])
    assert rel(b, a) > 0.1
    p = orc.translation_average_psi(psi, ens.pols, irt[tr], adjoint=True)
    a = orc.apply_full_L(full, p, transpose=True)
    b = orc.translation_average_psi(orc.apply_full_L(rep, p, transpose=True), ens.pols,
                                    irt[tr])
    assert rel(b, a) > 1e-6


def test_orbit_fold_is_the_translation_average(rock_salt):
    """The folded average (orbit x orbit x relative translation, the projector's fast form)
    equals the average over the translations' atom permutations, done here by brute force
    on the host with numpy; a non-free action has no fold."""
    from tdscha_amd.gamma import _orbit_fold
    ens, rots, irt, reps, tr, _, _ = rock_salt
    it = irt[tr]
    nt, nat = it.shape
    s, idx, nb = _orbit_fold(it)
    assert nb * nt == nat and np.all(np.bincount(s) == nt)
    rng = np.random.default_rng(7)
    C = rng.standard_normal((nat, nat, 3, 3))
    brute = np.mean([C[np.ix_(p, p)] for p in it], axis=0)
    G = np.zeros((nb * nb * nt, 3, 3))
    np.add.at(G, idx.ravel(), C.reshape(nat * nat, 3, 3))
    fold = (G[idx.ravel()] / nt).reshape(nat, nat, 3, 3)
    assert np.abs(fold - brute).max() < 1e-14
    fixed = it.copy()
    fixed[1] = np.arange(nat)          # a second identity: not a free action
    assert _orbit_fold(fixed) is None


def test_gamma_sector_needs_the_space_group_and_the_polarization_vectors():
    """Refused before any device work: no space group, or no polarization vectors."""
    from tdscha_amd import DynamicalLanczos as DL
    rng = np.random.default_rng(0)
    n, N = 4, 10
    lz = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    lz.set_ensemble(rng.standard_normal((N, n)), rng.standard_normal((N, n)),
                    np.linspace(1e-3, 2e-3, n), np.ones(N), 100.0)
    lz.set_symmetries([np.ones((1, 1, 1))] * n, np.ones(n, dtype=np.intc),
                      [np.array([i], dtype=np.intc) for i in range(n)],
                      np.arange(n, dtype=np.intc))
    lz.gamma_sector = True
    with pytest.raises(ValueError, match="space group"):
        lz.init()
    lz.sc_symmetries = (np.eye(3)[None], np.arange(2)[None])
    with pytest.raises(ValueError, match="polarization vectors"):
        lz.init()
