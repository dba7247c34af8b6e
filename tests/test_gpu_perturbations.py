"""Device ensemble averages of effective charges and Raman tensors
(tdscha_weighted_tn_product behind tdscha_amd.perturbations, Modules/Perturbations.py:195-375)
against the oracle's restatement."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("N,nat", [(1, 1), (777, 7), (130, 22)])
def test_averages_match_oracle(N, nat):
    from tdscha_amd import perturbations as pt
    print("case: N=%d nat=%d" % (N, nat), flush=True)
    rng = np.random.default_rng(N + nat)
    n3 = 3 * nat
    rho = np.exp(0.3 * rng.standard_normal(N))
    u = rng.standard_normal((N, nat, 3)) * 0.1
    Z = rng.standard_normal((N, nat, 3, 3))
    R = rng.standard_normal((N, 3, 3, n3))
    M = rng.standard_normal((n3, n3))
    ups = M @ M.T
    ub = u.reshape(N, -1) * pt.A_TO_BOHR
    assert rel(pt.get_d1M_dR_av(rho, Z, device=0), orc.d1M_dR_av(rho, Z)) < 1e-13
    assert rel(pt.get_d1alpha_dR_av(rho, R, device=0), orc.d1alpha_dR_av(rho, R)) < 1e-13
    assert rel(pt.get_d2M_dR_av(u, rho, Z, ups, device=0), orc.d2M_dR_av(ub, rho, Z, ups)) < 1e-12
    assert rel(pt.get_d2alpha_dR_av(u, rho, R, ups, device=0),
               orc.d2alpha_dR_av(ub, rho, R, ups)) < 1e-12


def test_stein_identity_on_device():
    import test_perturbations as tp
    from tdscha_amd import perturbations as pt
    print("case: Stein identity, nat=4 N=64", flush=True)
    rng = np.random.default_rng(9)
    nat, N = 4, 64
    n3 = 3 * nat
    u, ups = tp.stein_ensemble(rng, n3, N)
    A = rng.standard_normal((nat, 3, 3, n3))
    Z = rng.standard_normal((nat, 3, 3))[None] + np.einsum("afcd,id->iafc", A, u)
    d2M = pt.get_d2M_dR_av(u, np.ones(N), Z, ups, device=0, a_to_bohr=1.0)
    expect = np.einsum("afcd->dacf", A).reshape(n3, n3, 3)
    assert rel(d2M, expect) < 1e-10


@pytest.mark.parametrize("K,M,N", [(1, 1, 1), (5, 70, 3), (1000, 64, 64), (333, 129, 200)])
def test_weighted_tn_product_shapes(K, M, N):
    import torch
    from tdscha_amd import _lib as L
    print("case: K=%d M=%d N=%d" % (K, M, N), flush=True)
    rng = np.random.default_rng(K * 7 + M + N)
    dev = torch.device("cuda", 0)
    A = rng.standard_normal((K, M + 3))          # leading dimension M + 3
    B = rng.standard_normal((K, N))
    w = rng.standard_normal(K)
    At, Bt, wt = (torch.as_tensor(x).to(dev) for x in (A, B, w))
    C = torch.full((M, N + 2), 7.0, dtype=torch.float64, device=dev)
    L.check(L.lib().tdscha_weighted_tn_product(
        K, M, N, At.data_ptr(), M + 3, wt.data_ptr(), Bt.data_ptr(), N, 0.5, C.data_ptr(), N + 2,
        torch.cuda.current_stream(dev).cuda_stream), "tn")
    torch.cuda.synchronize()
    got = C.cpu().numpy()
    expect = 0.5 * A[:, :M].T @ (w[:, None] * B)
    assert rel(got[:, :N], expect) < 1e-13
    assert np.all(got[:, N:] == 7.0)


# ---------------------------------------------------------------- anharmonic IR / Raman builders
def _rock_salt_lanczos(N=150, T=250.0, seed=11):
    """A Wigner Lanczos on the 2 x 2 x 2 rock-salt crystal of the synthetic model (16 atoms,
    45 modes), with Gaussian displacements u (Angstrom), X = E^T sqrt(m) u (Bohr) and random
    forces; plus the crystal's 384 space-group operations."""
    import types
    import test_perturbations as tp
    from tdscha_amd import DynamicalLanczos as DL
    from tdscha_amd import perturbations as pt
    cr, w, E, starts, S, rots, irt = tp.rock_salt_2()
    rng = np.random.default_rng(seed)
    m3 = np.repeat(cr.masses, 3)
    nat = len(cr.masses)
    u = 0.05 * rng.standard_normal((N, nat, 3))
    X = (u.reshape(N, -1) * pt.A_TO_BOHR * np.sqrt(m3)) @ E
    Y = rng.standard_normal(X.shape)
    rho = np.exp(0.2 * rng.standard_normal(N))
    lz = DL.Lanczos(mode=DL.MODE_GPU_FP64, use_wigner=True)
    lz.verbose = False
    lz.set_ensemble(X, Y, w, rho, T, pols=E, m=m3)
    lz.sc_symmetries = (rots, irt)
    ens = types.SimpleNamespace(u_disps=u, rho=rho)
    return lz, ens, cr, rots, irt, rng


def _oracle_psi(lz, one_ph, d2_cart_proj, pt):
    """psi of the oracle: one-phonon vector v (3 nat) / sqrt(m) on the modes, two-phonon
    sectors from the projected second-order matrix (3 nat x 3 nat, already / sqrt(m))."""
    n = lz.n_modes
    psi = np.zeros(n + n * (n + 1))
    psi[:n] = (one_ph / np.sqrt(lz.m)) @ lz.pols
    if d2_cart_proj is not None:
        d2mu = lz.pols.T @ d2_cart_proj @ lz.pols
        a, b = orc.two_phonon_sectors(d2mu, lz.w, lz.T)
        orc.pack_two_phonon(psi, a, b)
    return psi


def _ups(lz):
    sm = np.sqrt(lz.m)
    return (sm[:, None] * lz.pols * orc.f_ups(lz.w, lz.T)[None]) @ (lz.pols.T * sm[None])


@pytest.mark.parametrize("kind", ["d1M", "d1alpha", "d2M", "d2alpha"])
@pytest.mark.parametrize("asr", [True, False])
def test_symmetrize_average_matches_oracle(kind, asr):
    import test_perturbations as tp
    from tdscha_amd import perturbations as pt
    print("case: symmetrize %s asr=%s" % (kind, asr), flush=True)
    _, _, E, _, _, rots, irt = tp.rock_salt_2()
    n3 = E.shape[0]
    shp = {"d1M": (n3, 3), "d2M": (n3, n3, 3), "d1alpha": (3, 3, n3),
           "d2alpha": (3, 3, n3, n3)}[kind]
    C = np.random.default_rng(3).standard_normal(shp)
    got = pt.symmetrize_average(kind, C, (rots, irt), device=0, asr=asr).cpu().numpy()
    expect = orc.symmetrize_average(kind, C, rots, irt, asr=asr)
    assert np.abs(got - expect).max() < 1e-13 * np.abs(C).max()


@pytest.mark.parametrize("symmetrize", [False, True])
def test_anharmonic_ir_FT_matches_oracle(symmetrize):
    from tdscha_amd import perturbations as pt
    print("case: prepare_anharmonic_ir_FT symmetrize=%s" % symmetrize, flush=True)
    lz, ens, cr, rots, irt, rng = _rock_salt_lanczos()
    N, nat = ens.u_disps.shape[:2]
    q = np.where(cr.species == 0, 1.3, -1.3)
    ec = np.einsum("a, ij -> aij", q, np.eye(3))[None] + 0.2 * rng.standard_normal((N, nat, 3, 3))
    ec_eq = np.stack([1.3 * np.eye(3), -1.3 * np.eye(3)])
    pol = np.array([0.3, -0.5, 0.8])
    lz.prepare_anharmonic_ir_FT(ec, ec_eq, pol, add_two_ph=True, symmetrize=symmetrize,
                                ensemble=ens)
    d1 = orc.d1M_dR_av(ens.rho, ec)
    ub = ens.u_disps.reshape(N, -1) * pt.A_TO_BOHR
    d2 = orc.d2M_dR_av(ub, ens.rho, ec - np.tile(ec_eq, (8, 1, 1))[None], _ups(lz))
    if symmetrize:
        d1 = orc.symmetrize_average("d1M", d1, rots, irt)
        d2 = orc.symmetrize_average("d2M", d2, rots, irt)
    sm = 1 / np.sqrt(lz.m)
    psi = _oracle_psi(lz, d1 @ pol, sm[:, None] * (d2 @ pol) * sm[None], pt)
    assert rel(lz.psi, psi) < 1e-11
    assert abs(lz.perturbation_modulus - psi @ (psi * lz.mask_dot_wigner())) < \
        1e-11 * lz.perturbation_modulus
    n = lz.n_modes
    assert np.abs(lz.psi[n:]).max() > 0
    if not symmetrize:          # (the symmetrized noise of the random charges is far smaller)
        assert np.abs(lz.psi[n:]).max() > 1e-3 * np.abs(lz.psi[:n]).max()


def test_anharmonic_ir_old_form_equals_new_form():
    """The earlier form (Py:1827, mode-basis X and f_ups) and the new one (Py:1693, Cartesian u
    with Upsilon) are the same average when X = E^T sqrt(m) u: the two device paths agree."""
    print("case: prepare_anharmonic_ir vs _FT", flush=True)
    lz, ens, cr, rots, irt, rng = _rock_salt_lanczos(seed=12)
    N, nat = ens.u_disps.shape[:2]
    ec = rng.standard_normal((N, nat, 3, 3))
    ec_eq = rng.standard_normal((2, 3, 3))
    pol = np.array([0.0, 1.0, 0.0])
    lz.prepare_anharmonic_ir_FT(ec, ec_eq, pol, add_two_ph=True, ensemble=ens)
    new, mod_new = lz.psi.copy(), lz.perturbation_modulus
    lz.prepare_anharmonic_ir(ec, ec_eq, pol, add_two_ph=True)
    assert rel(lz.psi, new) < 1e-10
    assert abs(lz.perturbation_modulus - mod_new) < 1e-10 * mod_new


def test_anharmonic_raman_FT_matches_oracle(tmp_path):
    from tdscha_amd import perturbations as pt
    print("case: prepare_anharmonic_raman_FT", flush=True)
    lz, ens, cr, rots, irt, rng = _rock_salt_lanczos(seed=13)
    N, nat = ens.u_disps.shape[:2]
    raman = rng.standard_normal((N, 3, 3, 3 * nat))
    raman_eq = rng.standard_normal((3, 3, 6))
    pairs = [(np.array([1.0, 0, 0]), np.array([0, 1.0, 0])),
             (np.array([0, 0, 1.0]), np.array([0, 0, 1.0])),
             (np.array([0.6, 0.8, 0]), np.array([0, 0.6, 0.8]))]
    kw = dict(pol_in=pairs[0][0], pol_out=pairs[0][1], mixed=True, pol_in_2=pairs[1][0],
              pol_out_2=pairs[1][1], pol_in_3=pairs[2][0], pol_out_3=pairs[2][1])
    lz.prepare_anharmonic_raman_FT(raman, raman_eq, add_two_ph=True, ensemble=ens, **kw)
    d1 = orc.d1alpha_dR_av(ens.rho, raman)
    v = sum(pt.raman_vector(d1, a, b) for a, b in pairs)
    ub = ens.u_disps.reshape(N, -1) * pt.A_TO_BOHR
    d2 = orc.d2alpha_dR_av(ub, ens.rho, raman - np.tile(raman_eq, 8)[None], _ups(lz))
    sm = 1 / np.sqrt(lz.m)
    d2 = d2 * sm[None, None, :, None] * sm[None, None, None, :]
    M = sum(np.einsum("abcd, a, b -> cd", d2, a, b) for a, b in pairs)
    psi = _oracle_psi(lz, v, M, pt)
    assert rel(lz.psi, psi) < 1e-11
    full = lz.psi.copy()
    # the saved second-order tensor gives the same two-phonon sectors through the _2ph form
    f = str(tmp_path / "raman2")
    lz.prepare_anharmonic_raman_FT(raman, raman_eq, add_two_ph=True, ensemble=ens,
                                   save_raman_tensor2=True, file_raman_tensor2=f, **kw)
    assert rel(np.load(f + ".npy"), d2) < 1e-11
    n = lz.n_modes
    lz.psi[n:] = 0
    lz.prepare_anharmonic_raman_FT_2ph(np.load(f + ".npy"), pol_in=pairs[0][0],
                                       pol_out=pairs[0][1], mixed=True, pol_in_2=pairs[1][0],
                                       pol_out_2=pairs[1][1])
    M2 = sum(np.einsum("abcd, a, b -> cd", d2, a, b) for a, b in pairs[:2])
    a2, b2 = orc.two_phonon_sectors(lz.pols.T @ M2 @ lz.pols, lz.w, lz.T)
    expect = orc.pack_two_phonon(np.zeros_like(full), a2, b2)
    assert rel(lz.psi[n:], expect[n:]) < 1e-11
    assert rel(lz.psi[:n], full[:n]) < 1e-15          # the one-phonon sector is left alone


def test_unpolarized_raman_FT_components():
    """The seven components of the unpolarized signal with configuration-dependent tensors
    (Py:1222-1336): (xx + yy + zz), (xx - yy), (xx - zz), (yy - zz), xy, xz, yz, each against
    the oracle's averages composed for that component's polarization pairs."""
    from tdscha_amd import perturbations as pt
    print("case: prepare_unpolarized_raman_FT", flush=True)
    lz, ens, cr, rots, irt, rng = _rock_salt_lanczos(seed=15)
    N, nat = ens.u_disps.shape[:2]
    raman = rng.standard_normal((N, 3, 3, 3 * nat))
    raman_eq = rng.standard_normal((3, 3, 6))
    d1 = orc.d1alpha_dR_av(ens.rho, raman)
    ub = ens.u_disps.reshape(N, -1) * pt.A_TO_BOHR
    d2 = orc.d2alpha_dR_av(ub, ens.rho, raman - np.tile(raman_eq, 8)[None], _ups(lz))
    sm = 1 / np.sqrt(lz.m)
    d2 = d2 * sm[None, None, :, None] * sm[None, None, None, :]
    x, y, z = np.eye(3)
    comps = {0: [(x, x), (y, y), (z, z)], 1: [(x, x), (-y, y)], 2: [(x, x), (-z, z)],
             3: [(y, y), (-z, z)], 4: [(x, y)], 5: [(x, z)], 6: [(y, z)]}
    for index, pairs in comps.items():
        lz.prepare_unpolarized_raman_FT(index, eq_raman_tns=raman_eq, use_symm=False,
                                        ens_av_raman=ens, raman_tns_ens=raman, add_2ph=True)
        v = sum(pt.raman_vector(d1, a, b) for a, b in pairs)
        M = sum(np.einsum("abcd, a, b -> cd", d2, a, b) for a, b in pairs)
        assert rel(lz.psi, _oracle_psi(lz, v, M, pt)) < 1e-11, index
    with pytest.raises(ValueError):
        lz.prepare_unpolarized_raman_FT(7, raman_tns_ens=raman, ens_av_raman=ens)


def test_anharmonic_raman_FT_symmetrized_is_raman_inactive():
    """Rock salt has no first-order Raman activity: with the space-group average the
    one-phonon sector vanishes while the two-phonon sectors do not."""
    print("case: prepare_anharmonic_raman_FT symmetrize", flush=True)
    lz, ens, cr, rots, irt, rng = _rock_salt_lanczos(seed=14)
    N, nat = ens.u_disps.shape[:2]
    raman = rng.standard_normal((N, 3, 3, 3 * nat))
    lz.prepare_anharmonic_raman_FT(raman, None, add_two_ph=True, symmetrize=True, ensemble=ens)
    n = lz.n_modes
    assert np.abs(lz.psi[:n]).max() < 1e-12 * np.abs(lz.psi[n:]).max()
    assert np.abs(lz.psi[n:]).max() > 0


def test_symmetrize_average_of_a_non_group_list_matches_oracle():
    """A list of operations that is not a group (no coset structure) is averaged operation by
    operation."""
    import test_perturbations as tp
    from tdscha_amd import perturbations as pt
    print("case: symmetrize d2M over 100 of the 384 operations", flush=True)
    _, _, E, _, _, rots, irt = tp.rock_salt_2()
    n3 = E.shape[0]
    from tdscha_amd.symmetry import translation_cosets
    assert translation_cosets(rots[:100], irt[:100]) is None
    assert translation_cosets(rots, irt) is not None
    C = np.random.default_rng(8).standard_normal((n3, n3, 3))
    got = pt.symmetrize_average("d2M", C, (rots[:100], irt[:100]), device=0).cpu().numpy()
    expect = orc.symmetrize_average("d2M", C, rots[:100], irt[:100])
    assert np.abs(got - expect).max() < 1e-13 * np.abs(C).max()
