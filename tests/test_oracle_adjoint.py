"""The oracle's force adjoint (d/dY of a linear functional of the kernel outputs), pinned by
finite differences of the oracle's own forward kernel and operator, and by Euler's
identity: every output is linear in Y, so J(Y) = <dJ/dY, Y> exactly."""
import numpy as np
import pytest

import fixtures
from oracle import oracle as orc


def _J_kernel(ens, Y, R1, Y1, uR, UD, d4):
    f, d = orc.perturb_average_sym(ens.X, Y, ens.w, ens.rho, R1, Y1, ens.T, d4, ens.sym)
    return uR.dot(f) + np.sum(UD * d)


@pytest.mark.parametrize("d4", [1, 0])
def test_force_grad_matches_finite_differences_of_the_kernel(d4):
    ens = fixtures.random_ensemble(71, 9, 5, 3)
    rng = np.random.default_rng(72)
    R1, Y1 = fixtures.random_perturbation(73, 9)
    uR = rng.standard_normal(9)
    A = rng.standard_normal((9, 9))
    UD = A + A.T
    g = orc.force_grad(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, ens.sym, uR, UD,
                       kf=1.0, kr=1.0, ky=float(d4))
    h = 1e-4 * np.abs(ens.Y).max()
    fd = np.zeros_like(ens.Y)
    for i in range(ens.Y.shape[0]):
        for m in range(ens.Y.shape[1]):
            Yp, Ym = ens.Y.copy(), ens.Y.copy()
            Yp[i, m] += h
            Ym[i, m] -= h
            fd[i, m] = (_J_kernel(ens, Yp, R1, Y1, uR, UD, d4) -
                        _J_kernel(ens, Ym, R1, Y1, uR, UD, d4)) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-7 * np.max(np.abs(g))
    # linear in Y: J(Y) = <dJ/dY, Y>
    J = _J_kernel(ens, ens.Y, R1, Y1, uR, UD, d4)
    assert abs(J - np.sum(g * ens.Y)) < 1e-12 * np.sum(np.abs(g * ens.Y))


def test_force_grad_ignores_the_antisymmetric_part_of_the_seed():
    ens = fixtures.random_ensemble(74, 7, 4, 2)
    R1, Y1 = fixtures.random_perturbation(75, 7)
    rng = np.random.default_rng(76)
    uR = rng.standard_normal(7)
    A = rng.standard_normal((7, 7))
    g1 = orc.force_grad(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, ens.sym, uR, A)
    g2 = orc.force_grad(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, ens.sym, uR, A.T)
    assert np.allclose(g1, g2, rtol=1e-13, atol=0)


@pytest.mark.parametrize("transpose,wigner,reverse", [
    (False, False, False), (True, False, False), (False, True, False), (False, False, True)])
def test_post_seeds_represent_the_post_transform(transpose, wigner, reverse):
    rng = np.random.default_rng(77)
    n = 6
    w = rng.uniform(2e-4, 2e-3, n)
    u = rng.standard_normal(orc.psi_len(n))
    uR, UD = orc.post_seeds(u, w, 250.0, transpose, wigner, reverse)
    f = rng.standard_normal(n)
    A = rng.standard_normal((n, n))
    d = A + A.T
    sign = -1.0 if reverse else 1.0
    lhs = sign * u.dot(orc.anharmonic_post(f, d, w, 250.0, transpose, wigner))
    assert abs(lhs - (uR.dot(f) + np.sum(UD * d))) < 1e-13 * (np.abs(lhs) + 1)
    assert np.allclose(UD, UD.T)


@pytest.mark.parametrize("transpose,wigner,ignore_v4", [
    (False, False, False), (True, False, False), (False, True, False), (False, False, True)])
def test_operator_level_force_gradient(transpose, wigner, ignore_v4):
    """u . apply_anharmonic_FT(v) through the pre-transform, the kernel and the post-transform:
    Euler's identity and central differences on a few entries."""
    ens = fixtures.random_ensemble(78, 6, 4, 3)
    rng = np.random.default_rng(79)
    n = 6
    u = rng.standard_normal(orc.psi_len(n))
    v = rng.standard_normal(orc.psi_len(n))

    def J(Y):
        e = orc.Ensemble(ens.X, Y, ens.w, ens.rho, ens.T, ens.sym)
        return u.dot(orc.apply_anharmonic_FT(e, v, transpose, wigner, ignore_v4))

    R1, Y1 = orc.anharmonic_pre(v, ens.w, ens.T, transpose, wigner)
    uR, UD = orc.post_seeds(u, ens.w, ens.T, transpose, wigner)
    g = orc.force_grad(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, ens.sym, uR, UD,
                       ky=0.0 if ignore_v4 else 1.0)
    assert abs(J(ens.Y) - np.sum(g * ens.Y)) < 1e-12 * np.sum(np.abs(g * ens.Y))
    h = 1e-4 * np.abs(ens.Y).max()
    for (i, m) in [(0, 0), (1, 3), (3, 5)]:
        Yp, Ym = ens.Y.copy(), ens.Y.copy()
        Yp[i, m] += h
        Ym[i, m] -= h
        fd = (J(Yp) - J(Ym)) / (2 * h)
        assert abs(fd - g[i, m]) < 1e-7 * np.max(np.abs(g))
