"""Gradients of J = sum_k Im G(w_k) with respect to the perturbation vector and the
configuration forces (tdscha_amd.gradients), against central finite differences of the
same Lanczos value (h = 1e-4 relative), and the BF16x3 forward against the FP64 one."""
import numpy as np
import pytest

import fixtures

pytestmark = pytest.mark.gpu


def _problem(ens, omegas, smearing, n_steps, mode=None):
    from tdscha_amd import DynamicalLanczos as DL
    from tdscha_amd.gradients import SpectralProblem
    return SpectralProblem(ens.X, ens.w, ens.rho, ens.T, ens.sym, omegas, smearing, n_steps,
                           mode=DL.MODE_GPU_FP64 if mode is None else mode)


def _fd_check(problem, psi0, Y, seed, rel=1e-4, tol=1e-5, Dp=None):
    from tdscha_amd.gradients import spectral_gradients
    J, gp, gy = spectral_gradients(problem, psi0, Y)
    rng = np.random.default_rng(seed)
    if Dp is None:
        Dp = rng.standard_normal(psi0.shape)
    Dp = Dp * rel * np.linalg.norm(psi0) / np.linalg.norm(Dp)
    Dy = rng.standard_normal(Y.shape) * rel * np.abs(Y).mean()
    val = lambda p, y: problem.value(problem.lanczos(p, y))
    fdp = (val(psi0 + Dp, Y) - val(psi0 - Dp, Y)) / 2
    fdy = (val(psi0, Y + Dy) - val(psi0, Y - Dy)) / 2
    ep = abs(fdp - gp @ Dp) / np.abs(gp * Dp).sum()
    ey = abs(fdy - np.sum(gy * Dy)) / np.abs(gy * Dy).sum()
    print("J %.6e  psi0: fd %.6e grad %.6e (rel %.1e)  Y: fd %.6e grad %.6e (rel %.1e)" % (
        J, fdp, gp @ Dp, ep, fdy, np.sum(gy * Dy), ey))
    assert ep < tol and ey < tol
    return J, gp, gy


def test_complete_krylov_space_random_ensemble():
    """n = 6 (48-dimensional psi): the recurrence terminates on the whole space, so the
    Lanczos value is the exact resolvent and the gradient formulas are exact."""
    ens = fixtures.random_ensemble(91, 6, 20, 3, max_block=3)
    psi0 = np.random.default_rng(92).standard_normal(48)
    omegas = np.linspace(1e-4, 4e-3, 64)
    _fd_check(_problem(ens, omegas, 3e-4, 60), psi0, ens.Y, 93)


@pytest.fixture(scope="module")
def rocksalt2_small():
    import torch
    from tdscha_amd import synthetic as sy
    return sy.make_ensemble(2, 300, T=300.0, device=torch.device("cuda"), sigma3=2.0, sigma4=1.0)


def _rs_problem(ens, mode=None):
    from tdscha_amd.gradients import SpectralProblem
    from tdscha_amd.synthetic import RY_TO_CM
    from tdscha_amd import DynamicalLanczos as DL
    from oracle import oracle as orc
    sym = orc.SymBlocks(ens.symmetries, ens.N_degeneracy, ens.degenerate_space, ens.block_id)
    omegas = np.linspace(50.0, 250.0, 64) / RY_TO_CM
    return SpectralProblem(ens.X, ens.w, ens.rho, ens.T, sym, omegas, 1.0 / RY_TO_CM, 20,
                           mode=DL.MODE_GPU_FP64 if mode is None else mode)


def _ir_psi(ens):
    n = len(ens.w)
    psi0 = np.zeros(n + n * (n + 1))
    psi0[:n] = ens.ir_perturbation(0)
    return psi0


def test_ir_spectrum_rocksalt(rocksalt2_small):
    """2x2x2 rock salt, 300 configs, 384 images, IR perturbation, 64 frequencies in
    50-250 cm^-1 with 1 cm^-1 smearing. The IR vector spans an invariant Krylov space (the
    recurrence terminates after ~10 steps), so the Lanczos value is exact for starting
    vectors inside that space and for any force perturbation (which keeps the ensemble
    symmetric): the psi0 direction is drawn inside the space."""
    import torch
    ens = rocksalt2_small
    prob = _rs_problem(ens)
    psi0 = _ir_psi(ens)
    lz = prob.lanczos(psi0, ens.Y)
    Q = torch.stack(list(lz._device_basis[0])).cpu().numpy()
    D = np.random.default_rng(95).standard_normal(len(Q)) @ Q
    _fd_check(prob, psi0, ens.Y, 94, tol=1e-7, Dp=D)


def test_ir_spectrum_random_direction(rocksalt2_small):
    """A psi0 direction outside the invariant space: the perturbed recurrence resolves the new
    directions only partly in 20 steps, so the Lanczos value moves at first order in h away
    from the resolvent; the central difference approaches the gradient as h shrinks."""
    ens = rocksalt2_small
    prob = _rs_problem(ens)
    psi0 = _ir_psi(ens)
    from tdscha_amd.gradients import spectral_gradients
    J, gp, _ = spectral_gradients(prob, psi0, ens.Y)
    D = np.random.default_rng(96).standard_normal(psi0.shape)
    D /= np.linalg.norm(D)
    val = lambda p: prob.value(prob.lanczos(p, ens.Y))
    errs = []
    for rel in (1e-4, 1e-6):
        h = rel * np.linalg.norm(psi0)
        fd = (val(psi0 + h * D) - val(psi0 - h * D)) / (2 * h)
        errs.append(abs(fd - gp @ D) / np.abs(gp * D).sum())
    print("random psi0 direction: rel err at h = 1e-4, 1e-6: %.1e %.1e" % tuple(errs))
    assert errs[0] < 1e-3 and errs[1] < errs[0]


def test_bf16x3_forward_against_fp64(rocksalt2_small):
    from tdscha_amd import DynamicalLanczos as DL
    from tdscha_amd.gradients import spectral_gradients
    ens = rocksalt2_small
    psi0 = _ir_psi(ens)
    J0, gp0, gy0 = spectral_gradients(_rs_problem(ens), psi0, ens.Y)
    J1, gp1, gy1 = spectral_gradients(_rs_problem(ens, DL.MODE_GPU_BF16X3), psi0, ens.Y)
    ej = abs(J1 - J0) / abs(J0)
    ep = np.linalg.norm(gp1 - gp0) / np.linalg.norm(gp0)
    ey = np.linalg.norm(gy1 - gy0) / np.linalg.norm(gy0)
    print("BF16x3 forward vs FP64: J %.1e  dJ/dpsi0 %.1e  dJ/dY %.1e" % (ej, ep, ey))
    # the BF16x3 recurrence does not see the IR space close (its residual stays at the
    # operator's rounding level), and the extra Krylov vectors carry into dJ/dpsi0
    assert ej < 1e-6 and ep < 1e-3 and ey < 1e-5


def test_autograd_function_matches(rocksalt2_small):
    import torch
    from tdscha_amd.gradients import spectral_function_sum, spectral_gradients
    ens = rocksalt2_small
    prob = _rs_problem(ens)
    psi0 = torch.tensor(_ir_psi(ens), requires_grad=True)
    Y = torch.tensor(ens.Y, requires_grad=True)
    J = spectral_function_sum(prob, psi0, Y)
    J.backward()
    J2, gp, gy = spectral_gradients(prob, _ir_psi(ens), ens.Y)
    assert abs(float(J.detach()) - J2) <= 1e-12 * abs(J2)
    assert np.allclose(psi0.grad.numpy(), gp, rtol=1e-12, atol=0)
    assert np.allclose(Y.grad.numpy(), gy, rtol=1e-12, atol=0)


def test_gradients_on_the_zone_centre_operator(rocksalt2_small):
    """The IR perturbation is translation-invariant: the Lanczos and the force adjoint on the
    zone-centre operator (48 symmetries + translation average, the adjoint's seeds projected
    by the transposed projector) give J and both gradients of the full 384-image run."""
    from tdscha_amd import synthetic as sy
    from tdscha_amd.gradients import spectral_gradients
    print("case: gradients, full vs zone-centre operator", flush=True)
    ens = rocksalt2_small
    rots, irt = sy.space_group(ens.crystal)
    full = _rs_problem(ens)
    gam = _rs_problem(ens)
    gam.gamma_sector = (rots, irt, ens.pols)
    psi0 = _ir_psi(ens)
    J0, gp0, gy0 = spectral_gradients(full, psi0, ens.Y)
    J1, gp1, gy1 = spectral_gradients(gam, psi0, ens.Y)
    r = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    print("J %.6e %.6e  dpsi %.1e  dY %.1e" % (J0, J1, r(gp1, gp0), r(gy1, gy0)))
    assert abs(J1 - J0) < 1e-10 * abs(J0)
    assert r(gp1, gp0) < 1e-9 and r(gy1, gy0) < 1e-9
