"""Perturbation builders: IR and Raman vectors and the ensemble averages of effective charges
and Raman tensors.

Reference: ``Lanczos.prepare_ir`` / ``prepare_raman`` / ``prepare_unpolarized_raman``
(Modules/DynamicalLanczos.py:1003-1221, 1647-1691) and Modules/Perturbations.py
(get_d1M_dR_av :195, get_d1alpha_dR_av :240, get_d2M_dR_av :275, get_d2alpha_dR_av :330).
The reference reads the tensors from CellConstructor objects; here they are arguments.

The vectors are host arithmetic on 3 nat numbers. The ensemble averages are contractions over
the configurations, C = A^T diag(rho) B / N_eff, and run as FP64 MFMA products on the device
(``tdscha_weighted_tn_product``).
"""
import numpy as np

from . import _lib as L
from .DynamicalLanczos import f_ups

# Angstrom -> Bohr as CellConstructor converts (CC.Units.A_TO_BOHR, Perturbations.py:308);
# its Bohr length, 0.529177249 Angstrom, is also what its dynamical-matrix files use.
A_TO_BOHR = 1.0 / 0.529177249

# Py:1139-1147, in index order 0 .. 6
UNPOLARIZED_RAMAN_PREFACTORS = (45 / 9, 7 / 2, 7 / 2, 7 / 2, 7 * 3, 7 * 3, 7 * 3)
UNPOLARIZED_RAMAN_LABELS = ("(xx+yy+zz)^2", "(xx-yy)^2", "(xx-zz)^2", "(yy-zz)^2", "(xy)^2",
                            "(xz)^2", "(yz)^2")

_E = np.eye(3)


def raman_vector(raman_tensor, pol_in, pol_out, asr=True):
    """sum_ab pol_in_a T[a, b, :] pol_out_b (the role of CellConstructor's GetRamanVector).

    raman_tensor (3, 3, 3 nat). asr: enforce the acoustic sum rule (the components of every
    Cartesian direction sum to zero over the atoms) by removing their mean; a tensor that
    already obeys it is unchanged."""
    v = np.einsum("abc, a, b -> c", np.asarray(raman_tensor, dtype=np.float64),
                  np.asarray(pol_in, dtype=np.float64), np.asarray(pol_out, dtype=np.float64))
    if asr:
        r = v.reshape(-1, 3)
        v = (r - r.mean(axis=0)[None]).ravel()
    return v


def unpolarized_raman_vector(raman_tensor, index, asr=True):
    """The seven components of the unpolarized Raman signal (Py:1152-1221):
    0 xx+yy+zz, 1 xx-yy, 2 xx-zz, 3 yy-zz, 4 xy, 5 xz, 6 yz."""
    x, y, z = _E
    combos = {0: ((x, x, 1), (y, y, 1), (z, z, 1)), 1: ((x, x, 1), (y, y, -1)),
              2: ((x, x, 1), (z, z, -1)), 3: ((y, y, 1), (z, z, -1)), 4: ((x, y, 1),),
              5: ((x, z, 1),), 6: ((y, z, 1),)}
    if index not in combos:
        raise ValueError("{} should be in {}".format(index, list(range(7))))
    v = 0
    for p_in, p_out, sign in combos[index]:
        v = v + sign * raman_vector(raman_tensor, p_in, p_out, asr)
    return v


def ir_vector(effective_charges, pol_vec):
    """sum_b Z[a, b, c] pol_b, flattened over (atom, Cartesian) (Py:1684-1687).
    effective_charges (nat, 3 field, 3 Cartesian)."""
    return np.einsum("abc, b", np.asarray(effective_charges, dtype=np.float64),
                     np.asarray(pol_vec, dtype=np.float64)).ravel()


def upsilon_matrix(w, pols, masses, T):
    """Upsilon = sqrt(m) E diag(2 w / (1 + 2 n(w))) E^T sqrt(m), 1/Bohr^2: the inverse of the
    SCHA displacement covariance (the role of CellConstructor's GetUpsilonMatrix). Modes of
    zero frequency (translations) contribute nothing."""
    w = np.asarray(w, dtype=np.float64)
    E = np.asarray(pols, dtype=np.float64)
    m = np.asarray(masses, dtype=np.float64).ravel()
    if len(m) * 3 == E.shape[0]:
        m = np.repeat(m, 3)
    f = np.zeros_like(w)
    ok = w > 0
    f[ok] = f_ups(w[ok], T)
    sm = np.sqrt(m)
    return (sm[:, None] * E * f[None]) @ (E.T * sm[None])


def _dev(device):
    import torch
    if isinstance(device, torch.device):
        return device
    return torch.device("cuda", torch.cuda.current_device() if device is None else int(device))


def _tn(A, w, B, scale, dev):
    """scale * A^T diag(w) B on the device (A: K x M, B: K x N, FP64 torch tensors)."""
    import torch
    K, M = A.shape
    N = B.shape[1]
    C = torch.empty((M, N), dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream(dev)
    L.check(L.lib().tdscha_weighted_tn_product(
        K, M, N, A.data_ptr(), A.stride(0), None if w is None else w.data_ptr(), B.data_ptr(),
        B.stride(0), float(scale), C.data_ptr(), C.stride(0), st.cuda_stream),
        "tdscha_weighted_tn_product")
    return C


def _put(a, dev):
    import torch
    return torch.as_tensor(np.asarray(a) if not hasattr(a, "data_ptr") else a,
                           dtype=torch.float64).to(dev).contiguous()


def _charges_rows(Z, dev):
    """(N, nat, 3 field, 3 cart) -> (N, 3 nat * 3): row [3 a + x] * 3 + field (Perturbations.py
    :219-221, :311-313)."""
    N = Z.shape[0]
    return Z.permute(0, 1, 3, 2).reshape(N, -1).contiguous()


def get_d1M_dR_av(rho, effective_charges, device=None):
    """<Z*> over the ensemble, (3 nat, 3 field) (Perturbations.py:195-238)."""
    dev = _dev(device)
    r = _put(rho, dev)
    Z = _put(effective_charges, dev)
    if Z.shape[0] != r.shape[0]:
        raise ValueError("the number of effective charges (%d) does not match the ensemble "
                         "size (%d)" % (Z.shape[0], r.shape[0]))
    B = _charges_rows(Z, dev)
    C = _tn(r.reshape(-1, 1), None, B, 1.0 / float(r.sum()), dev)
    return C.reshape(-1, 3).cpu().numpy()


def get_d1alpha_dR_av(rho, raman_tensors, device=None):
    """<dalpha/dR> over the ensemble, (3, 3, 3 nat) (Perturbations.py:240-273)."""
    dev = _dev(device)
    r = _put(rho, dev)
    R = _put(raman_tensors, dev)
    N = R.shape[0]
    C = _tn(r.reshape(-1, 1), None, R.reshape(N, -1), 1.0 / float(r.sum()), dev)
    return C.reshape(R.shape[1:]).cpu().numpy()


def get_d2M_dR_av(u_disps, rho, effective_charges, ups_mat, device=None, a_to_bohr=A_TO_BOHR):
    """<Z* (Upsilon u)> = Upsilon <u (x) Z*> over the ensemble, (3 nat, 3 nat, 3 field)
    (Perturbations.py:275-328). u_disps (N, nat, 3) or (N, 3 nat) in Angstrom; ups_mat in
    1/Bohr^2 (upsilon_matrix)."""
    dev = _dev(device)
    r = _put(rho, dev)
    Z = _put(effective_charges, dev)
    N = Z.shape[0]
    u = _put(u_disps, dev).reshape(N, -1)
    U = _put(ups_mat, dev)
    B = _charges_rows(Z, dev)
    C1 = _tn(u, r, B, a_to_bohr / float(r.sum()), dev)             # u^T diag(rho) Z / N_eff
    C = _tn(U.t().contiguous(), None, C1, 1.0, dev)                  # Upsilon . C1
    n3 = U.shape[0]
    return C.reshape(n3, n3, 3).cpu().numpy()


def get_d2alpha_dR_av(u_disps, rho, raman_tensors, ups_mat, device=None, a_to_bohr=A_TO_BOHR):
    """<dalpha/dR (Upsilon u)>, (3, 3, 3 nat, 3 nat) with the last index the displacement's
    (Perturbations.py:330-375)."""
    dev = _dev(device)
    r = _put(rho, dev)
    R = _put(raman_tensors, dev)
    N = R.shape[0]
    u = _put(u_disps, dev).reshape(N, -1)
    U = _put(ups_mat, dev)
    C1 = _tn(u, r, R.reshape(N, -1), a_to_bohr / float(r.sum()), dev)
    C = _tn(U.t().contiguous(), None, C1, 1.0, dev)                  # [a, (b c d)]
    n3 = U.shape[0]
    return C.reshape(n3, *R.shape[1:]).permute(1, 2, 3, 0).contiguous().cpu().numpy()


def _main():
    """Time the second-order effective-charge average at a given size:
    python -m tdscha_amd.perturbations [--configs N] [--atoms nat] [--reps K]. Synthetic
    displacements, charges and Upsilon; FP64."""
    import argparse
    import json

    import torch

    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", type=int, default=20000)
    ap.add_argument("--atoms", type=int, default=432)
    ap.add_argument("--reps", type=int, default=10)
    a = ap.parse_args()
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(5)
    N, n3 = a.configs, 3 * a.atoms
    u = torch.randn(N, n3, device=dev, dtype=torch.float64, generator=g)
    rho = torch.rand(N, device=dev, dtype=torch.float64, generator=g) + 0.5
    Zr = torch.randn(N, 3 * n3, device=dev, dtype=torch.float64, generator=g)
    U = torch.randn(n3, n3, device=dev, dtype=torch.float64, generator=g)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream(dev)
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]

    def run():
        C1 = _tn(u, rho, Zr, 1.0, dev)
        e[1].record(st)
        _tn(U, None, C1, 1.0, dev)

    run()
    torch.cuda.synchronize()
    t_ens = t_ups = 0.0
    for _ in range(a.reps):
        e[0].record(st)
        run()
        e[2].record(st)
        e[2].synchronize()
        t_ens += e[0].elapsed_time(e[1])
        t_ups += e[1].elapsed_time(e[2])
    t_ens /= a.reps
    t_ups /= a.reps
    f_ens = 2.0 * N * n3 * 3 * n3
    print(json.dumps({"kernel": "k_weighted_tn", "configs": N, "atoms": a.atoms,
                      "ensemble_product_ms": t_ens, "ensemble_tflops": f_ens / t_ens / 1e9,
                      "upsilon_product_ms": t_ups,
                      "fp64_mfma_peak_tflops": 78.6,
                      "frac": f_ens / t_ens / 1e9 / 78.6}))


if __name__ == "__main__":
    _main()
