"""Gradients of the Lanczos spectral function (no counterpart in the reference).

    J = sum_k Im G(w_k),   G(w) = psi0^T (L - z)^-1 psi0,   z = w^2 - 2 i w eta

with G evaluated as Lanczos.get_green_function_continued_fraction does (Py:4951-5047,
no terminator, smearing eta): G = |psi0|^2 [(T - z)^-1]_00 with T the tridiagonal matrix
of the biconjugate recurrence (L q_j = a_j q_j + b_j q_(j+1) + c_(j-1) q_(j-1)).

With x = (L - z)^-1 psi0 ~ |psi0| Q xi and y = (L - z)^-T psi0 ~ |psi0| P~ eta, where
xi = (T - z)^-1 e_0, eta = (T^T - z)^-1 e_0 and P~ are the dual vectors
p~_j = p_j s_norm_j / c_(j-1) (P~^T Q = 1):

    dJ/dpsi0 = sum_k Im (x + y)
    dJ/dY    = - sum_k Im y^T (dL/dY) x = - d/dY sum_j u_j . A(q_j; Y),
               u_j = |psi0|^2 sum_l W_lj p~_l,   W_lj = Im sum_k eta_l(w_k) xi_j(w_k)

The second line is the device force adjoint (tdscha_ensemble_force_grad) on the pairs
(u_j, q_j): A is the anharmonic part of L, the only part that depends on the forces Y.
The formulas are those of the exact resolvent; they are the gradient of the Lanczos value
once the recurrence has converged (or terminated on an invariant subspace).

The Lanczos is run by the forward pass and run again by the backward pass (it is
deterministic, so the second run rebuilds the same basis) instead of keeping the Krylov
basis alive between the two.
"""
import numpy as np

from .DynamicalLanczos import Lanczos, MODE_GPU_FP32, MODE_GPU_FP64


class SpectralProblem:
    """What stays fixed while psi0 and the forces Y vary: the rest of the ensemble, the
    frequencies w_k (Ry), the smearing (Ry), the number of Lanczos steps and the mode."""

    def __init__(self, X, w, rho, T, sym, omegas, smearing, n_steps, mode=MODE_GPU_FP64,
                 full_gs=True, gamma_sector=None, verbose=False, adjoint_precision=None):
        """gamma_sector: (rotations, irt, pols) of the supercell's space group to run the
        Lanczos and the adjoint on the zone-centre operator (tdscha_amd.gamma), for
        translation-invariant psi0 such as IR perturbations. adjoint_precision: "fp64" or
        "fp32" for the force adjoint's contractions (default: "fp32" with MODE_GPU_FP32,
        else "fp64")."""
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.w = np.ascontiguousarray(w, dtype=np.float64)
        self.rho = np.ascontiguousarray(rho, dtype=np.float64)
        self.T = float(T)
        self.sym = sym
        self.omegas = np.ascontiguousarray(omegas, dtype=np.float64)
        self.smearing = float(smearing)
        self.n_steps = int(n_steps)
        self.mode = mode
        self.full_gs = bool(full_gs)
        self.gamma_sector = gamma_sector
        self.verbose = bool(verbose)
        if adjoint_precision is None:
            adjoint_precision = "fp32" if mode == MODE_GPU_FP32 else "fp64"
        if adjoint_precision == "fp32" and mode != MODE_GPU_FP32:
            raise ValueError("the FP32 adjoint needs the FP32 operand copies: mode=MODE_GPU_FP32")
        self.adjoint_precision = adjoint_precision

    def lanczos(self, psi0, Y):
        lz = Lanczos(mode=self.mode)
        pols = None if self.gamma_sector is None else self.gamma_sector[2]
        lz.set_ensemble(self.X, Y, self.w, self.rho, self.T, pols=pols)
        s = self.sym
        lz.set_symmetries(s.symmetries, s.N_degeneracy, s.degenerate_space, s.block_id)
        if self.gamma_sector is not None:
            lz.sc_symmetries = tuple(self.gamma_sector[:2])
            lz.gamma_sector = True
        lz.init()
        lz.psi[:] = psi0
        lz.perturbation_modulus = float(np.dot(psi0, psi0))
        lz.run_FT(self.n_steps, verbose=self.verbose, full_gs=self.full_gs)
        return lz

    def value(self, lz):
        g = lz.get_green_function_continued_fraction(self.omegas, use_terminator=False,
                                                     smearing=self.smearing)
        return float(np.sum(np.imag(g)))


def _tridiagonal(a, b, c):
    k = len(a)
    T = np.diag(np.asarray(a, dtype=np.float64))
    for j in range(k - 1):
        T[j + 1, j] = b[j]
        T[j, j + 1] = c[j]
    return T


def resolvent_coefficients(a, b, c, omegas, smearing):
    """xi[:, k] = (T - z_k)^-1 e_0 and eta[:, k] = (T^T - z_k)^-1 e_0."""
    T = _tridiagonal(a, b, c)
    k = T.shape[0]
    e0 = np.zeros(k, dtype=np.complex128)
    e0[0] = 1.0
    xi = np.empty((k, len(omegas)), dtype=np.complex128)
    eta = np.empty_like(xi)
    eye = np.eye(k)
    for m, w in enumerate(omegas):
        z = w * w - 2j * w * smearing
        xi[:, m] = np.linalg.solve(T - z * eye, e0)
        eta[:, m] = np.linalg.solve(T.T - z * eye, e0)
    return xi, eta


def spectral_gradients(problem, psi0, Y, pairs_per_call=32):
    """(J, dJ/dpsi0, dJ/dY) as numpy arrays, after one (re-)run of the Lanczos."""
    import torch
    psi0 = np.ascontiguousarray(psi0, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    lz = problem.lanczos(psi0, Y)
    J = problem.value(lz)
    a = np.asarray(lz.a_coeffs, dtype=np.float64)
    b = np.asarray(lz.b_coeffs, dtype=np.float64)
    c = np.asarray(lz.c_coeffs, dtype=np.float64)
    s_norm = np.asarray(lz.s_norm, dtype=np.float64)
    k = len(a)
    ens = lz.device_ensemble
    dev = ens.device
    basis = getattr(lz, "_device_basis", None)
    if basis is not None and len(basis[0]) >= k:
        Q = torch.stack(list(basis[0][:k]))
        P = torch.stack(list(basis[1][:k]))
    else:
        Q = torch.as_tensor(np.asarray(lz.basis_Q[:k]), device=dev)
        P = torch.as_tensor(np.asarray(lz.basis_P[:k]), device=dev)
    c_prev = np.concatenate([[1.0], c[:k - 1]])
    pn = s_norm[:k] / c_prev
    Pd = P * torch.as_tensor(pn, device=dev)[:, None]
    xi, eta = resolvent_coefficients(a, b[:k - 1], c[:k - 1], problem.omegas, problem.smearing)
    nrm = float(np.linalg.norm(psi0))
    mod = nrm * nrm
    cq = torch.as_tensor(np.imag(xi.sum(axis=1)), device=dev)
    cp = torch.as_tensor(np.imag(eta.sum(axis=1)), device=dev)
    g_psi = (nrm * (cq @ Q + cp @ Pd)).cpu().numpy()
    W = np.imag(eta @ xi.T)                   # W[l, j] = Im sum_k eta_l xi_j
    U = mod * (torch.as_tensor(W.T, device=dev) @ Pd)   # row j: u_j
    gY = torch.zeros((ens.N, ens.n), dtype=torch.float64, device=dev)
    for j0 in range(0, k, pairs_per_call):
        j1 = min(k, j0 + pairs_per_call)
        ens.force_grad([U[j] for j in range(j0, j1)], [Q[j] for j in range(j0, j1)],
                       out=gY, accumulate=True, precision=problem.adjoint_precision)
        if problem.verbose:
            torch.cuda.synchronize()
            print("force adjoint: pairs %d-%d of %d done" % (j0, j1 - 1, k), flush=True)
    return J, g_psi, -gY.cpu().numpy()


def spectral_function_sum(problem, psi0, Y):
    """J = sum_k Im G(w_k) as a torch.autograd.Function of (psi0, Y) (CPU float64 tensors or
    anything torch.as_tensor takes); the backward pass is spectral_gradients."""
    import torch

    class _SpectralSum(torch.autograd.Function):
        @staticmethod
        def forward(ctx, p, y):
            ctx.save_for_backward(p, y)
            lz = problem.lanczos(p.detach().cpu().numpy(), y.detach().cpu().numpy())
            return torch.tensor(problem.value(lz), dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            p, y = ctx.saved_tensors
            _, gp, gy = spectral_gradients(problem, p.detach().cpu().numpy(),
                                           y.detach().cpu().numpy())
            g = gJ.detach().cpu()
            return (g * torch.as_tensor(gp)).to(p.device), (g * torch.as_tensor(gy)).to(y.device)

    return _SpectralSum.apply(torch.as_tensor(psi0, dtype=torch.float64),
                              torch.as_tensor(Y, dtype=torch.float64))


def _main():
    """Time J = sum_k Im G(w_k) and its gradients on the synthetic rock-salt ensemble:
    python -m tdscha_amd.gradients [--L 4] [--configs 10000] [--steps 150] [--anharmonic 20]
    IR perturbation along x (Z* = +-2 I), 64 frequencies over 50-250 cm^-1 with 1 cm^-1
    smearing, FP64; the full operator and the zone-centre one (tdscha_amd.gamma), and the
    largest differences between the two."""
    import argparse
    import json
    import time

    import torch

    from . import synthetic as sy
    from .ensemble import SymmetryBlocks
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=4)
    ap.add_argument("--configs", type=int, default=10000)
    ap.add_argument("--steps", type=int, default=150)
    ap.add_argument("--anharmonic", type=float, default=20.0)
    ap.add_argument("--skip-full", action="store_true")
    a = ap.parse_args()
    ens = sy.make_ensemble(a.L, a.configs, T=300.0, weights="lognormal",
                           device=torch.device("cuda", 0), sigma3=0.1 * a.anharmonic,
                           sigma4=0.05 * a.anharmonic)
    rots, irt = sy.space_group(ens.crystal)
    sym = SymmetryBlocks(ens.symmetries, ens.N_degeneracy, ens.degenerate_space, ens.block_id)
    omegas = np.linspace(50.0, 250.0, 64) / sy.RY_TO_CM
    n = len(ens.w)
    psi0 = np.zeros(n + n * (n + 1))
    psi0[:n] = ens.ir_perturbation(0)
    res = {"workload": "rock-salt %dx%dx%d, %d modes, %d configs, %d symmetries, IR, %d steps"
                       % (a.L, a.L, a.L, n, a.configs, ens.n_syms, a.steps)}
    out = {}
    for name, gam in (("full", None), ("zone_centre", (rots, irt, ens.pols))):
        if name == "full" and a.skip_full:
            continue
        prob = SpectralProblem(ens.X, ens.w, ens.rho, ens.T, sym, omegas, 1.0 / sy.RY_TO_CM,
                               a.steps, gamma_sector=gam, verbose=True)
        t0 = time.time()
        J, gp, gy = spectral_gradients(prob, psi0, ens.Y)
        torch.cuda.synchronize()
        out[name] = (J, gp, gy)
        res[name] = {"seconds": time.time() - t0, "J": J}
        print("%s: forward and backward done" % name, flush=True)
        print("%s: %.1f s" % (name, res[name]["seconds"]), flush=True)
    if "full" in out:
        J0, gp0, gy0 = out["full"]
        J1, gp1, gy1 = out["zone_centre"]
        res["rel_diff"] = {"J": abs(J1 - J0) / abs(J0),
                           "dJ_dpsi0": float(np.abs(gp1 - gp0).max() / np.abs(gp0).max()),
                           "dJ_dY": float(np.abs(gy1 - gy0).max() / np.abs(gy0).max())}
        res["speedup"] = res["full"]["seconds"] / res["zone_centre"]["seconds"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    _main()
