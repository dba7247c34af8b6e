"""Perturbation builders: IR and Raman vectors and the ensemble averages of effective charges
and Raman tensors.

Reference: ``Lanczos.prepare_ir`` / ``prepare_raman`` / ``prepare_unpolarized_raman``
(Modules/DynamicalLanczos.py:1003-1221, 1647-1691) and Modules/Perturbations.py
(get_d1M_dR_av :195, get_d1alpha_dR_av :240, get_d2M_dR_av :275, get_d2alpha_dR_av :330).
The reference reads the tensors from CellConstructor objects; here they are arguments.

The vectors are host arithmetic on 3 nat numbers. The ensemble averages are contractions over
the configurations, C = A^T diag(rho) B / N_eff, and run as FP64 MFMA products on the device
(``tdscha_weighted_tn_product``).

Two-phonon perturbations (Wigner representation, Py:1340-1962) fill the a'(1), b'(1) sectors
of psi from a symmetric mode-space matrix weighted by sqrt(-chi^-/2) and sqrt(chi^+/2)
(``chi_minus_plus``, ``wigner_two_phonon``). The averages can be symmetrized over the
supercell's space group (Perturbations.py:15-193, CellConstructor's QE_Symmetry there): here
the group is the (rotations, irt) list the symmetry blocks are built from, and the average
over it runs on the device (``space_group_average``).
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
    A = A if A.stride(1) == 1 else A.contiguous()    # rows must be contiguous (any layout in)
    B = B if B.stride(1) == 1 else B.contiguous()
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


def _symmetrized(kind, C, symmetrize, symmetries, dev):
    if not symmetrize:
        return C
    if symmetries is None:
        raise ValueError("symmetrize=True needs symmetries=(rotations, irt) of the supercell "
                         "(the reference asks spglib for them, which is not available here)")
    return symmetrize_average(kind, C, symmetries, device=dev)


def get_d1M_dR_av(rho, effective_charges, device=None, symmetrize=False, symmetries=None):
    """<Z*> over the ensemble, (3 nat, 3 field) (Perturbations.py:195-238). symmetrize:
    averaged over the space group `symmetries` = (rotations, irt), with the acoustic sum rule
    (symmetrize_average)."""
    dev = _dev(device)
    r = _put(rho, dev)
    Z = _put(effective_charges, dev)
    if Z.shape[0] != r.shape[0]:
        raise ValueError("the number of effective charges (%d) does not match the ensemble "
                         "size (%d)" % (Z.shape[0], r.shape[0]))
    B = _charges_rows(Z, dev)
    C = _tn(r.reshape(-1, 1), None, B, 1.0 / float(r.sum()), dev).reshape(-1, 3)
    return _symmetrized("d1M", C, symmetrize, symmetries, dev).cpu().numpy()


def get_d1alpha_dR_av(rho, raman_tensors, device=None, symmetrize=False, symmetries=None):
    """<dalpha/dR> over the ensemble, (3, 3, 3 nat) (Perturbations.py:240-273)."""
    dev = _dev(device)
    r = _put(rho, dev)
    R = _put(raman_tensors, dev)
    N = R.shape[0]
    C = _tn(r.reshape(-1, 1), None, R.reshape(N, -1), 1.0 / float(r.sum()), dev)
    C = C.reshape(R.shape[1:])
    return _symmetrized("d1alpha", C, symmetrize, symmetries, dev).cpu().numpy()


def get_d2M_dR_av(u_disps, rho, effective_charges, ups_mat, device=None, a_to_bohr=A_TO_BOHR,
                  symmetrize=False, symmetries=None):
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
    C = C.reshape(n3, n3, 3)
    return _symmetrized("d2M", C, symmetrize, symmetries, dev).cpu().numpy()


def get_d2alpha_dR_av(u_disps, rho, raman_tensors, ups_mat, device=None, a_to_bohr=A_TO_BOHR,
                      symmetrize=False, symmetries=None):
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
    C = C.reshape(n3, *R.shape[1:]).permute(1, 2, 3, 0).contiguous()
    return _symmetrized("d2alpha", C, symmetrize, symmetries, dev).cpu().numpy()


# ------------------------------------------------------------------ space-group averages
def _ops(symmetries, nat, dev):
    import torch
    rot, irt = symmetries
    rot = torch.as_tensor(np.asarray(rot, dtype=np.float64), device=dev)
    irt = np.asarray(irt, dtype=np.int64)
    if rot.ndim != 3 or rot.shape[1:] != (3, 3) or irt.shape != (rot.shape[0], nat):
        raise ValueError("symmetries: rotations (n_sym, 3, 3) and irt (n_sym, %d) expected"
                         % nat)
    if np.any(np.sort(irt, axis=1) != np.arange(nat)[None]):
        raise ValueError("every row of irt must be a permutation of the atoms")
    inv = np.empty_like(irt)
    np.put_along_axis(inv, irt, np.arange(nat)[None].repeat(len(irt), 0), axis=1)
    return rot, torch.as_tensor(inv, device=dev)


def _act(T, R, inv, n_atom_axes, cart):
    """g.T for one operation (canonical layout): atoms gathered through the inverse permutation,
    then every Cartesian axis contracted with R."""
    import torch
    X = T
    for k in range(n_atom_axes):
        X = X.index_select(2 * k, inv)              # X[b] = T[irt^-1(b)]
    for ax in cart:
        X = torch.movedim(torch.tensordot(R, X, dims=([1], [ax])), 0, ax)
    return X


def space_group_average(Tc, n_atom_axes, symmetries, device=None):
    """(1 / n_sym) sum_g g.T for a tensor in canonical layout: n_atom_axes pairs (nat, 3)
    first, then field axes of length 3. g = (R, irt) acts as
    (g.T)[irt(a)][i] ... = R[i][p] ... T[a][p] ..., every Cartesian axis rotated by R (the
    convention of the symmetry blocks: (P_g v)[irt(a)] = R v[a]). For a space group the
    operations are cosets of the pure translations: the translations are averaged first
    (permutations only), then one operation per rotation is applied."""
    import torch
    dev = _dev(device)
    T = torch.as_tensor(Tc, dtype=torch.float64, device=dev)
    nat = T.shape[0]
    rot, inv = _ops(symmetries, nat, dev)
    irt = np.asarray(symmetries[1], dtype=np.int64)
    cart = [2 * k + 1 for k in range(n_atom_axes)] + list(range(2 * n_atom_axes, T.dim()))
    from .symmetry import translation_cosets
    cos = translation_cosets(symmetries[0], irt)
    if cos is None:
        acc = torch.zeros_like(T)
        for g in range(rot.shape[0]):
            acc += _act(T, rot[g], inv[g], n_atom_axes, cart)
        return acc / rot.shape[0]
    reps, tr = cos
    Tt = torch.zeros_like(T)
    for t in tr:
        X = T
        for k in range(n_atom_axes):
            X = X.index_select(2 * k, inv[t])
        Tt += X
    Tt /= len(tr)
    acc = torch.zeros_like(T)
    for g in reps:
        acc += _act(Tt, rot[g], inv[g], n_atom_axes, cart)
    return acc / len(reps)


def _asr(T, axes):
    """Remove the mean over the atoms of each listed atom axis (canonical layout)."""
    for ax in axes:
        T = T - T.mean(dim=ax, keepdim=True)
    return T


def symmetrize_average(kind, C, symmetries, device=None, asr=True):
    """Symmetrized ensemble average (the role of Perturbations.py:15-193 and CellConstructor's
    ApplySymmetryTo{EffCharge, RamanTensor, SecondOrderEffCharge, SecondOrderRamanTensor}):
    the space-group average, the acoustic sum rule (zero sum over the atoms of every atom
    index) and, for the second-order tensors, the permutation symmetry of the two
    displacements (and of the two field indices of the Raman one).

    kind: "d1M" (3 nat, 3), "d1alpha" (3, 3, 3 nat), "d2M" (3 nat, 3 nat, 3) or "d2alpha"
    (3, 3, 3 nat, 3 nat), the layouts of the get_*_av functions. asr=False skips the sum
    rule."""
    import torch
    dev = _dev(device)
    ax0, ax02 = ([0], [0, 2]) if asr else ([], [])
    C = torch.as_tensor(C if not isinstance(C, np.ndarray) else np.asarray(C),
                        dtype=torch.float64, device=dev)
    if kind == "d1M":
        nat = C.shape[0] // 3
        T = C.reshape(nat, 3, 3)                                   # [a, i, E]
        T = space_group_average(_asr(T, ax0), 1, symmetries, dev)
        return T.reshape(3 * nat, 3)
    if kind == "d1alpha":
        nat = C.shape[2] // 3
        T = C.reshape(3, 3, nat, 3).permute(2, 3, 0, 1)             # [a, i, E, F]
        T = 0.5 * (T + T.transpose(2, 3))
        T = space_group_average(_asr(T, ax0), 1, symmetries, dev)
        return T.permute(2, 3, 0, 1).reshape(3, 3, 3 * nat)
    if kind == "d2M":
        nat = C.shape[0] // 3
        T = C.reshape(nat, 3, nat, 3, 3)                           # [a, i, b, j, E]
        T = 0.5 * (T + T.permute(2, 3, 0, 1, 4))
        T = space_group_average(_asr(T, ax02), 2, symmetries, dev)
        return T.reshape(3 * nat, 3 * nat, 3)
    if kind == "d2alpha":
        nat = C.shape[2] // 3
        T = C.reshape(3, 3, nat, 3, nat, 3).permute(2, 3, 4, 5, 0, 1)   # [a, i, b, j, E, F]
        T = 0.5 * (T + T.permute(2, 3, 0, 1, 4, 5))
        T = 0.5 * (T + T.transpose(4, 5))
        T = space_group_average(_asr(T, ax02), 2, symmetries, dev)
        return T.permute(4, 5, 0, 1, 2, 3).reshape(3, 3, 3 * nat, 3 * nat)
    raise ValueError("kind must be d1M, d1alpha, d2M or d2alpha")


# ------------------------------------------------------------------ two-phonon sectors
def chi_minus_plus(w, T, epsilon=1e-12):
    """chi^-_{mn} = (w_m - w_n)(n_m - n_n) / (2 w_m w_n) and
    chi^+_{mn} = (w_m + w_n)(1 + n_m + n_n) / (2 w_m w_n), n the Bose occupations at T (Ry
    frequencies, T in K; n = 0 for T <= epsilon) (Py:2572-2627)."""
    from .DynamicalLanczos import RY_TO_K
    w = np.asarray(w, dtype=np.float64)
    occ = 1.0 / (np.exp(w * RY_TO_K / T) - 1.0) if T > epsilon else np.zeros_like(w)
    ww = 2.0 * w[:, None] * w[None, :]
    chi_m = (w[:, None] - w[None, :]) * (occ[:, None] - occ[None, :]) / ww
    chi_p = (w[:, None] + w[None, :]) * (1.0 + occ[:, None] + occ[None, :]) / ww
    return chi_m, chi_p


def wigner_two_phonon(d2, w, T):
    """a'(1) = -sqrt(-chi^-/2) o D and b'(1) = sqrt(chi^+/2) o D for the symmetric part D of
    the mode-space matrix d2 (Py:1488-1502, 1776-1790, 1918-1931)."""
    d2 = np.asarray(d2, dtype=np.float64)
    D = 0.5 * (d2 + d2.T)
    chi_m, chi_p = chi_minus_plus(w, T)
    # chi^- <= 0 (the occupations fall with the frequency); rounding can leave -0 or +tiny
    pert_a = -np.sqrt(np.maximum(-0.5 * chi_m, 0.0)) * D
    pert_b = np.sqrt(0.5 * chi_p) * D
    return pert_a, pert_b


def pack_wigner_two_phonon(psi, pert_a, pert_b):
    """Write the upper triangles (row-major, i <= j) of a'(1) and b'(1) after the n one-phonon
    entries of psi, as the reference's loops do (Py:1507-1516)."""
    n = pert_a.shape[0]
    iu = np.triu_indices(n)
    nw2 = len(iu[0])
    psi[n:n + nw2] = pert_a[iu]
    psi[n + nw2:n + 2 * nw2] = pert_b[iu]
    return psi


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
