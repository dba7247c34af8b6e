"""The Krylov representation behind tdscha_amd.gradients, on a dense non-symmetric matrix
(CPU): the duals p~_j = p_j s_norm_j / c_(j-1) of the reference recurrence are biorthogonal
to the q_j, the Lanczos resolvent vectors equal the exact solves once the Krylov space is
complete, and the resulting gradients match central differences of J."""
import numpy as np

from oracle import oracle as orc
from tdscha_amd.gradients import resolvent_coefficients


def _matrix(n=8, seed=5):
    rng = np.random.default_rng(seed)
    A = np.diag(rng.uniform(1.0, 4.0, n)) + 0.3 * rng.standard_normal((n, n))
    return A


def _lanczos(A, psi0):
    st = orc.run_FT(lambda q: A @ q, lambda p: A.T @ p, psi0, A.shape[0])
    k = len(st.a)
    a, b, c = (np.asarray(x) for x in (st.a, st.b, st.c))
    s = np.asarray(st.s_norm)
    c_prev = np.concatenate([[1.0], c[:k - 1]])
    Q = np.array(st.Q[:k])
    Pd = np.array(st.P[:k]) * (s[:k] / c_prev)[:, None]
    return a, b, c, Q, Pd


def _J(A, psi0, omegas, eta):
    out = 0.0
    for w in omegas:
        z = w * w - 2j * w * eta
        out += np.imag(psi0 @ np.linalg.solve(A - z * np.eye(len(A)), psi0))
    return out


def test_duals_are_biorthogonal_and_resolvent_is_exact():
    A = _matrix()
    psi0 = np.random.default_rng(6).standard_normal(8)
    a, b, c, Q, Pd = _lanczos(A, psi0)
    k = len(a)
    assert k == 8
    assert np.allclose(Pd @ Q.T, np.eye(k), atol=1e-9)
    omegas = np.array([0.9, 1.3, 1.7])
    xi, et = resolvent_coefficients(a, b[:k - 1], c[:k - 1], omegas, 0.05)
    nrm = np.linalg.norm(psi0)
    for m, w in enumerate(omegas):
        z = w * w - 2j * w * 0.05
        x = np.linalg.solve(A - z * np.eye(8), psi0)
        y = np.linalg.solve(A.T - z * np.eye(8), psi0)
        assert np.allclose(nrm * (xi[:, m] @ Q), x, rtol=1e-8, atol=1e-10)
        assert np.allclose(nrm * (et[:, m] @ Pd), y, rtol=1e-8, atol=1e-10)


def test_gradient_formulas_match_central_differences():
    """dJ/dpsi0 = sum Im(x + y) and dJ/dA = -sum Im(y x^T) (the Y-gradient is this with
    dA/dY, done by the device adjoint), built from the Krylov coefficients as in
    spectral_gradients."""
    A = _matrix(seed=9)
    rng = np.random.default_rng(10)
    psi0 = rng.standard_normal(8)
    omegas = np.linspace(0.8, 2.0, 7)
    eta = 0.07
    a, b, c, Q, Pd = _lanczos(A, psi0)
    k = len(a)
    xi, et = resolvent_coefficients(a, b[:k - 1], c[:k - 1], omegas, eta)
    nrm = np.linalg.norm(psi0)
    g_psi = nrm * (np.imag(xi.sum(axis=1)) @ Q + np.imag(et.sum(axis=1)) @ Pd)
    W = np.imag(et @ xi.T)
    U = nrm * nrm * (W.T @ Pd)                 # u_j
    g_A = -(U.T @ Q)                           # -sum_j u_j q_j^T
    h = 1e-4
    D = rng.standard_normal(8)
    fd = (_J(A, psi0 + h * D, omegas, eta) - _J(A, psi0 - h * D, omegas, eta)) / (2 * h)
    assert abs(fd - g_psi @ D) < 1e-6 * np.abs(g_psi).sum()
    E = rng.standard_normal((8, 8))
    fd = (_J(A + h * E, psi0, omegas, eta) - _J(A - h * E, psi0, omegas, eta)) / (2 * h)
    assert abs(fd - np.sum(g_A * E)) < 1e-6 * np.abs(g_A).sum()
