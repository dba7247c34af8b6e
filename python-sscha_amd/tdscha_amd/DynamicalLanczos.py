"""Lanczos with the ensemble L-application on an MI355X.

Mirrors the interface of the reference class ``tdscha.DynamicalLanczos.Lanczos``
(Modules/DynamicalLanczos.py, "Py" below) for the path that every Lanczos step takes:
apply_full_L -> apply_anharmonic_FT -> GetPerturbAverageSym. Three new modes select the
device path (next to MODE_FAST_* at Py:166-169):

    MODE_GPU_FP64    plain FP64 contractions, reproduces the reference to <= 1e-10
    MODE_GPU_FP32    FP32 MFMA contractions, FP64 everything else
    MODE_GPU_BF16X3  BF16 x3 split contractions, FP64 everything else

Everything on the Krylov side stays FP64 on the device: the harmonic part L1, the
psi <-> (R1, Y1) transforms, the recurrence scalars a, b, c, the optional Gram-Schmidt
and the continued fraction.

Deviations from the reference, on purpose:
  * The ensemble is given either already projected on the modes (X, Y, w, rho, T and the
    symmetry blocks), as Lanczos.__init__ leaves it (Py:380-442) and save_status stores it
    (Py:3617-3646), or as raw configurations with the dynamical matrix's w and pols
    (set_ensemble_from_configs, projected on the device); CellConstructor / sscha objects
    are not required.
  * Errors raise ValueError / RuntimeError instead of exiting the interpreter.
  * run_FT(full_gs=True) adds a two-sided classical Gram-Schmidt (two passes) against every
    stored vector with the biorthogonal duals p_j * s_norm_j / c_(j-1), on the Python-driven
    and on the device-resident recurrence; n_rep_orth keeps the reference's own loop
    (Py:6245-6317) verbatim, quirks included.
"""
import os
import sys
import time

import numpy as np

from . import _lib as L
from .ensemble import DeviceEnsemble, SymmetryBlocks

TYPE_DP = np.double
__EPSILON__ = 1e-12          # Py:121
RY_TO_K = 157887.32400374097  # Py:134

MODE_SLOW_SERIAL = 0
MODE_FAST_SERIAL = 1
MODE_FAST_MPI = 2
MODE_FAST_JULIA = 3
MODE_GPU_FP64 = 4
MODE_GPU_FP32 = 5
MODE_GPU_BF16X3 = 6
_GPU_PRECISION = {MODE_GPU_FP64: "fp64", MODE_GPU_FP32: "fp32", MODE_GPU_BF16X3: "bf16x3"}


def f_ups(w, T):
    """Py:144-155."""
    w = np.asarray(w, dtype=np.float64)
    n_w = 1 / (np.exp(w * RY_TO_K / T) - 1) if T > 0 else 0
    return 2 * w / (1 + 2 * n_w)


class Lanczos(object):
    """Biconjugate / symmetric Lanczos on the TD-SCHA L operator (Py:182)."""

    def __init__(self, ensemble=None, mode=MODE_GPU_FP64, use_wigner=False, device=0,
                 s_batch=0):
        if mode not in _GPU_PRECISION:
            raise ValueError("mode %r: this package implements the device modes MODE_GPU_FP64, "
                             "MODE_GPU_FP32 and MODE_GPU_BF16X3" % (mode,))
        self.mode = mode
        self.device = device
        self.s_batch = s_batch
        self.verbose = True
        self.T = 0.0
        self.nat = 0
        self.m = np.zeros(0)
        self.w = np.zeros(0)
        self.pols = np.zeros((0, 0))
        self.n_modes = 0
        self.ignore_harmonic = False
        self.ignore_v3 = False
        self.ignore_v4 = False
        self.N = 0
        self.rho = np.zeros(0)
        self.N_eff = 0
        self.X = np.zeros((0, 0))
        self.Y = np.zeros((0, 0))
        self.psi = np.zeros(0)
        self.a_coeffs = []
        self.b_coeffs = []
        self.c_coeffs = []
        self.basis_P = []
        self.basis_Q = []
        self.s_norm = []
        self.reverse_L = False
        self.shift_value = 0
        self.symmetries = None
        self.degenerate_space = None
        self.N_degeneracy = None
        self.sym_block_id = None
        self.n_syms = 1
        self.initialized = False
        self.perturbation_modulus = 1
        self.use_wigner = use_wigner
        self.ignore_small_w = False
        self.effective_charges = None   # (nat_uc, 3, 3), read by prepare_ir
        self.raman_tensor = None        # (3, 3, 3 nat_uc), read by prepare_raman
        self.sc_symmetries = None       # (rotations, irt) given to prepare_symmetrization
        # translation-invariant perturbations (IR, Raman at the zone centre): L over one
        # symmetry per rotation and the translation average (tdscha_amd.gamma); set before init
        self.gamma_sector = False
        self.L_linop = None
        self._dev_ens = None
        if ensemble is not None and hasattr(ensemble, "u_disps"):
            self.set_ensemble_from_configs(
                ensemble.u_disps, ensemble.forces, ensemble.rho, ensemble.T, ensemble.m,
                ensemble.w, ensemble.pols, getattr(ensemble, "sscha_forces", None),
                getattr(ensemble, "f_mean", None), ignore_small_w=getattr(
                    ensemble, "ignore_small_w", False))
        elif ensemble is not None:
            self.set_ensemble(ensemble.X, ensemble.Y, ensemble.w, ensemble.rho, ensemble.T,
                              getattr(ensemble, "pols", None), getattr(ensemble, "m", None))
            sym = getattr(ensemble, "sym", None)
            if sym is not None:
                self.set_symmetries(sym.symmetries, sym.N_degeneracy, sym.degenerate_space,
                                    sym.block_id)

    # ------------------------------------------------------------------ setup
    def set_ensemble(self, X, Y, w, rho, T, pols=None, m=None):
        """The projected ensemble (the arrays Py:380-442 builds)."""
        self.X = np.ascontiguousarray(X, dtype=TYPE_DP)
        self.Y = np.ascontiguousarray(Y, dtype=TYPE_DP)
        self.w = np.ascontiguousarray(w, dtype=TYPE_DP)
        self.rho = np.ascontiguousarray(rho, dtype=TYPE_DP)
        self.T = float(T)
        self.N, self.n_modes = self.X.shape
        self.N_eff = np.sum(self.rho)
        if pols is not None:
            self.pols = np.asarray(pols, dtype=TYPE_DP)
        if m is not None:
            self.m = np.asarray(m, dtype=TYPE_DP)
            self.nat = len(self.m) // 3
        n = self.n_modes
        self.psi = np.zeros(n + n * (n + 1), dtype=TYPE_DP)
        self._dev_ens = None
        self.initialized = False

    def set_ensemble_from_configs(self, u_disps, forces, rho, T, masses, w, pols,
                                  sscha_forces=None, f_mean=None, select_modes=None,
                                  ignore_small_w=False, eps_w=1e-8, bohr=None):
        """Project raw configurations on the SCHA modes on the device (Py:330-442).

        u_disps, forces, sscha_forces: (N, nat, 3) or (N, 3 nat), Angstrom / Ry/Angstrom.
        masses: (nat,) or (3 nat,); w, pols: frequencies [Ry] and mass-scaled polarization
        vectors (3 nat, 3 nat) of the supercell dynamical matrix, translations included (they
        are removed here, Py:333-359). f_mean: symmetrized average force over the supercell,
        Ry/Angstrom (the reference symmetrizes it with the crystal's space group, Py:391-397;
        pass None when it vanishes by symmetry).
        """
        import torch
        from . import projection as pj

        pols = np.asarray(pols, dtype=TYPE_DP)
        good = pj.good_modes(w, pols, masses, ignore_small_w, select_modes, eps_w)
        self.ignore_small_w = bool(ignore_small_w)
        w_sel = np.asarray(w, dtype=TYPE_DP)[good]
        pols_sel = pols[:, good]
        if len(w_sel) == 1:
            pols_sel = pols_sel.reshape((pols.shape[0], 1))
        m = pj._masses_per_component(masses, pols.shape[0])
        X, Y = pj.project_ensemble(u_disps, forces, m, pols_sel, sscha_forces, f_mean,
                                   pj.BOHR if bohr is None else bohr,
                                   device=torch.device("cuda", self.device))
        self.set_ensemble(X.cpu().numpy(), Y.cpu().numpy(), w_sel, rho, T, pols_sel, m)

    def prepare_symmetrization(self, no_sym=False, verbose=True, symmetries=None,
                               deg_tol=1e-8):
        """Mode-basis symmetry blocks built on the device (Py:686-778).

        symmetries: (rotations, irt) of the supercell's space group, Cartesian 3 x 3 rotations
        (n_sym, 3, 3) and the atom each operation carries every atom to (n_sym, nat) (see
        symmetry.atom_permutations); the reference asks spglib for them, which is not
        available here. no_sym: one 1 x 1 identity block per mode (Py:717-722). Blocks are the
        runs of consecutive modes whose frequencies agree to deg_tol (relative)."""
        from . import symmetry as sy

        n = self.n_modes
        if no_sym:
            self.set_symmetries([np.ones((1, 1, 1), dtype=TYPE_DP)] * n,
                                np.ones(n, dtype=np.intc),
                                [np.array([i], dtype=np.intc) for i in range(n)],
                                np.arange(n).astype(np.intc))
            return
        if symmetries is None:
            raise ValueError("pass symmetries=(rotations, irt): the space group is not searched "
                             "here (the reference uses spglib)")
        pols = np.asarray(self.pols)
        if pols.ndim != 2 or pols.shape[1] != n or pols.shape[0] % 3:
            raise ValueError("the polarization vectors of the kept modes are not set")
        rot, irt = symmetries
        self.sc_symmetries = (np.asarray(rot, dtype=TYPE_DP), np.asarray(irt))
        t0 = time.time()
        starts = sy.degenerate_blocks(self.w, deg_tol)
        blocks = sy.symmetries_on_modes(rot, irt, pols, starts, device=self.device)
        ndeg, degs, bid = sy.block_arrays(starts, n)
        self.set_symmetries(blocks, ndeg, degs, bid)
        if verbose:
            print("Time to convert symmetries in the polarization space: {} s".format(
                time.time() - t0))

    def set_symmetries(self, symmetries, N_degeneracy, degenerate_space, block_id):
        """Mode-basis symmetry blocks (what prepare_symmetrization stores, Py:741-778)."""
        self.symmetries = [np.ascontiguousarray(s, dtype=TYPE_DP) for s in symmetries]
        self.N_degeneracy = np.ascontiguousarray(N_degeneracy, dtype=np.intc)
        self.degenerate_space = [np.ascontiguousarray(d, dtype=np.intc) for d in degenerate_space]
        self.sym_block_id = np.ascontiguousarray(block_id, dtype=np.intc)
        self.n_syms = self.symmetries[0].shape[0]
        self._dev_ens = None
        self.initialized = False

    def init(self, use_symmetries=True):
        """Upload the ensemble to the device (Py:556-575 + prepare_symmetrization)."""
        if self.symmetries is None or not use_symmetries:
            if use_symmetries:
                raise ValueError("no symmetry blocks were given: call set_symmetries() or "
                                 "init(use_symmetries=False)")
            sym = SymmetryBlocks.identity(self.n_modes)
            self.set_symmetries(sym.symmetries, sym.N_degeneracy, sym.degenerate_space,
                                sym.block_id)
        sym = SymmetryBlocks(self.symmetries, self.N_degeneracy, self.degenerate_space,
                             self.sym_block_id)
        precs = ("fp64", _GPU_PRECISION[self.mode]) if self.mode != MODE_GPU_FP64 else ("fp64",)
        if self.gamma_sector:
            from .gamma import GammaSectorOperator
            if self.sc_symmetries is None or not use_symmetries:
                raise ValueError("gamma_sector needs the space group: call "
                                 "prepare_symmetrization(symmetries=(rotations, irt)) first")
            pols = np.asarray(self.pols)
            if pols.ndim != 2 or pols.shape[1] != self.n_modes or pols.shape[0] % 3 or \
                    pols.shape[0] == 0:
                raise ValueError("gamma_sector needs the polarization vectors of the kept modes, "
                                 "(3 nat, n_modes): set_ensemble(..., pols=...)")
            rot, irt = self.sc_symmetries
            self._dev_ens = GammaSectorOperator(self.X, self.Y, self.w, self.rho, self.T, sym,
                                                rot, irt, self.pols, precisions=precs,
                                                device=self.device, s_batch=self.s_batch)
        else:
            self._dev_ens = DeviceEnsemble(self.X, self.Y, self.w, self.rho, self.T, sym,
                                           precisions=precs, device=self.device,
                                           s_batch=self.s_batch)
        self.initialized = True
        import scipy.sparse.linalg
        n_psi = len(self.psi)
        self.L_linop = scipy.sparse.linalg.LinearOperator(
            shape=(n_psi, n_psi), matvec=self.apply_full_L,
            rmatvec=lambda p: self.apply_full_L(p, transpose=True), dtype=TYPE_DP)

    def reset(self):
        """Py:1000-1002 equivalent: forget the Lanczos coefficients and bases."""
        self.a_coeffs = []
        self.b_coeffs = []
        self.c_coeffs = []
        self.basis_P = []
        self.basis_Q = []
        self.s_norm = []

    @property
    def precision(self):
        return _GPU_PRECISION[self.mode]

    @property
    def device_ensemble(self):
        if not self.initialized:
            raise ValueError("Error, this class must be initialized before lunching the "
                             "computation:\n Use the .init()")
        return self._dev_ens

    # ------------------------------------------------------------------ perturbations
    def prepare_mode(self, index):
        """Py:2017-2031."""
        self.reset()
        self.psi[:] = 0
        self.psi[index] = 1
        self.perturbation_modulus = 1

    def prepare_perturbation(self, vector, masses_exp=1, add=False):
        """Py:1963-2015 (R sector only)."""
        if not add:
            self.reset()
            self.psi = np.zeros(self.psi.shape, dtype=TYPE_DP)
        m_on = np.sqrt(self.m) ** masses_exp
        new_v = np.einsum("a, a, ab->b", m_on, vector, self.pols)
        self.psi[:self.n_modes] += new_v
        self.perturbation_modulus = new_v.dot(new_v)

    def _n_supercell(self, n_cart_uc):
        n_cart = len(self.m)
        if n_cart == 0 or n_cart % n_cart_uc:
            raise ValueError("the masses (3 nat of the supercell) are not set or do not match "
                             "the tensor's unit cell")
        return n_cart // n_cart_uc

    def prepare_ir(self, effective_charges=None, pol_vec=np.array([1, 0, 0])):
        """IR perturbation (Py:1647-1691): Z* contracted with the light polarization, tiled
        over the supercell, divided by sqrt(m) and projected on the modes. effective_charges
        (nat_uc, 3 field, 3 Cartesian); defaults to self.effective_charges."""
        from . import perturbations as pt
        ec = self.effective_charges if effective_charges is None else effective_charges
        if ec is None:
            raise ValueError("Error, no effective charge found. Cannot initialize IR responce")
        ec = np.asarray(ec, dtype=TYPE_DP)
        if ec.ndim != 3 or ec.shape[1] != 3 or ec.shape[2] != 3:
            raise ValueError("Error, effective charges of the wrong shape: {}".format(ec.shape))
        n_sc = self._n_supercell(3 * ec.shape[0])
        if not self.ignore_small_w and ec.shape[0] * ec.shape[2] * n_sc != self.n_modes + 3:
            raise ValueError("effective charges do not match the modes (n_modes + 3 = %d)"
                             % (self.n_modes + 3))
        self.prepare_perturbation(np.tile(pt.ir_vector(ec, pol_vec), n_sc), masses_exp=-1)

    def prepare_raman(self, pol_vec_in=np.array([1, 0, 0]), pol_vec_out=np.array([1, 0, 0]),
                      mixed=False, pol_in_2=None, pol_out_2=None, unpolarized=None,
                      raman_tensor=None):
        """Raman perturbation (Py:1003-1124). raman_tensor (3, 3, 3 nat_uc); defaults to
        self.raman_tensor. mixed adds the (pol_in_2, pol_out_2) component. unpolarized = 0..6
        selects the components of the unpolarized signal (alpha^2 from xx + yy + zz over 3;
        (xx - yy) / sqrt 2, (xx - zz) / sqrt 2, (yy - zz) / sqrt 2; sqrt 3 xy, sqrt 3 yz,
        sqrt 3 xz).

        Deliberate fixes: the reference computes the mixed vector and then overwrites it
        (Py:1037-1049), and after the unpolarized branches it prepares the last component
        alone again (Py:1123), discarding the sum built with add=True; here the mixed vector
        and the whole unpolarized sum are what is prepared."""
        from . import perturbations as pt
        T = self.raman_tensor if raman_tensor is None else raman_tensor
        if T is None:
            raise ValueError("Error, no Raman tensor found. Cannot initialize the Raman responce")
        T = np.asarray(T, dtype=TYPE_DP)
        n_sc = self._n_supercell(T.shape[2])
        x, y, z = np.eye(3)
        if unpolarized is None:
            v = pt.raman_vector(T, pol_vec_in, pol_vec_out)
            if mixed:
                v = v + pt.raman_vector(T, pol_in_2, pol_out_2)
            self.prepare_perturbation(np.tile(v, n_sc), masses_exp=-1)
            return
        r2, r3 = 1 / np.sqrt(2), np.sqrt(3)
        parts = {0: ((x, x, 1 / 3), (y, y, 1 / 3), (z, z, 1 / 3)),
                 1: ((x, x, r2), (y, y, -r2)), 2: ((x, x, r2), (z, z, -r2)),
                 3: ((y, y, r2), (z, z, -r2)), 4: ((x, y, r3),), 5: ((y, z, r3),),
                 6: ((x, z, r3),)}
        if unpolarized not in parts:
            raise ValueError("Error, unpolarized must be between [0, ... ,6] got invalid {}."
                             .format(unpolarized))
        for k, (p_in, p_out, c) in enumerate(parts[unpolarized]):
            v = np.tile(pt.raman_vector(T, p_in, p_out), n_sc) * c
            self.prepare_perturbation(v, masses_exp=-1, add=k > 0)

    def get_prefactors_unpolarized_raman(self, index):
        """Py:1126-1150."""
        from . import perturbations as pt
        if index not in range(7):
            raise ValueError("{} should be in {}".format(index, list(range(7))))
        return pt.UNPOLARIZED_RAMAN_PREFACTORS[index]

    def prepare_unpolarized_raman(self, index=0, debug=False, raman_tensor=None):
        """Py:1152-1221: component `index` of the unpolarized Raman signal (see
        perturbations.unpolarized_raman_vector), tiled over the supercell, / sqrt(m)."""
        from . import perturbations as pt
        T = self.raman_tensor if raman_tensor is None else raman_tensor
        if T is None:
            raise ValueError("Error, no Raman tensor found. Cannot initialize the Raman responce")
        T = np.asarray(T, dtype=TYPE_DP)
        v = pt.unpolarized_raman_vector(T, index)
        if debug:
            np.save("raman_v_{}".format(index), v)
        self.prepare_perturbation(np.tile(v, self._n_supercell(T.shape[2])), masses_exp=-1)

    # ------------------------------------------------------------------ anharmonic IR / Raman
    def get_chi_minus(self):
        """chi^- of the Wigner formalism (Py:2572-2598)."""
        from . import perturbations as pt
        return pt.chi_minus_plus(self.w, self.T, __EPSILON__)[0]

    def get_chi_plus(self):
        """chi^+ of the Wigner formalism (Py:2601-2627)."""
        from . import perturbations as pt
        return pt.chi_minus_plus(self.w, self.T, __EPSILON__)[1]

    def _set_two_phonon(self, d2_mu, label):
        """Fill psi's a'(1), b'(1) sectors from the mode-space matrix d2_mu and take the
        perturbation modulus with the Wigner mask (Py:1488-1526)."""
        from . import perturbations as pt
        pert_a, pert_b = pt.wigner_two_phonon(d2_mu, self.w, self.T)
        pt.pack_wigner_two_phonon(self.psi, pert_a, pert_b)
        self.perturbation_modulus = self.psi.dot(self.psi * self.mask_dot_wigner())
        if self.verbose:
            print("Perturbation modulus after adding two ph contributions{} = {}".format(
                label, self.perturbation_modulus))

    def _to_mode_space(self, M):
        """E^T diag(1/sqrt m) M diag(1/sqrt m) E for a (3 nat, 3 nat) matrix."""
        sm = 1.0 / np.sqrt(self.m)
        P = self.pols * sm[:, None]
        return P.T @ M @ P

    def _ensemble_arrays(self, ensemble, n_expected):
        """u_disps (N, 3 nat) in Angstrom and rho of a duck-typed ensemble (attributes u_disps,
        rho; the reference's sscha.Ensemble has both)."""
        if ensemble is None or not hasattr(ensemble, "u_disps"):
            raise ValueError("pass ensemble= (an object with u_disps (N, nat, 3) in Angstrom and "
                             "rho (N)) for the ensemble averages")
        u = np.asarray(ensemble.u_disps, dtype=TYPE_DP).reshape(len(ensemble.rho), -1)
        rho = np.asarray(ensemble.rho, dtype=TYPE_DP)
        if len(rho) != n_expected:
            raise ValueError("the tensors are given for {} configurations, the ensemble has {}"
                             .format(n_expected, len(rho)))
        return u, rho

    def _upsilon(self):
        from . import perturbations as pt
        return pt.upsilon_matrix(self.w, self.pols, self.m, self.T)

    def _sym_ops(self, symmetrize):
        if symmetrize and self.sc_symmetries is None:
            raise ValueError("symmetrize=True uses the space group given to "
                             "prepare_symmetrization(symmetries=(rotations, irt))")
        return self.sc_symmetries if symmetrize else None

    def prepare_anharmonic_ir_FT(self, ec=None, ec_eq=None, pol_vec_light=np.array([1., 0., 0.]),
                                 add_two_ph=False, symmetrize=False, ensemble=None):
        """IR perturbation with configuration-dependent effective charges (Py:1693-1824).

        ec (N, nat_sc, 3 field, 3 Cartesian); ec_eq (nat_uc, 3, 3), subtracted (tiled over the
        supercell) from ec in the two-phonon average only; ensemble: u_disps (Angstrom) and
        rho of the configurations. One-phonon sector: <Z*> . pol / sqrt(m) on the modes;
        two-phonon sectors (Wigner only): E^T m^-1/2 (Upsilon <u (x) Z*> . pol) m^-1/2 E,
        symmetrized, weighted by sqrt(-chi^-/2), sqrt(chi^+/2). Upsilon is built from this
        object's frequencies, polarization vectors, masses and temperature (the reference takes
        the ensemble's current dynamical matrix, the same one in a converged run).
        symmetrize: the averages are symmetrized over the space group given to
        prepare_symmetrization (perturbations.symmetrize_average)."""
        from . import perturbations as pt
        if not self.use_wigner and add_two_ph:
            raise NotImplementedError("The two phonon processes are implemented only in Wigner")
        if ec is None:
            raise ValueError("Must specify the effective charges for all configurations!")
        ec = np.asarray(ec, dtype=TYPE_DP)
        if ec.ndim != 4 or ec.shape[1] != self.nat or ec.shape[2:] != (3, 3):
            raise ValueError("The effective charges in input have the wrong shape. The required "
                             "is N_conf N_at_sc E_field cart")
        u, rho = self._ensemble_arrays(ensemble, ec.shape[0])
        ops = self._sym_ops(symmetrize)
        pol = np.asarray(pol_vec_light, dtype=TYPE_DP)
        d1 = pt.get_d1M_dR_av(rho, ec, self.device, symmetrize, ops)
        self.prepare_perturbation(d1 @ pol, masses_exp=-1)
        if self.verbose:
            print("Pertubation modulus with one ph effects only = {}".format(
                self.perturbation_modulus))
        if not add_two_ph:
            return
        if ec_eq is not None:
            ec_eq = np.asarray(ec_eq, dtype=TYPE_DP)
            if ec_eq.ndim != 3 or ec_eq.shape[1:] != (3, 3):
                raise ValueError("Error, effective charges of the wrong shape: {}".format(
                    ec_eq.shape))
            n_sc = self._n_supercell(3 * ec_eq.shape[0])
            if not self.ignore_small_w and ec_eq.shape[0] * 3 * n_sc != self.n_modes + 3:
                raise ValueError("effective charges do not match the modes")
            ec = ec - np.tile(ec_eq, (n_sc, 1, 1))[None]
        d2 = pt.get_d2M_dR_av(u, rho, ec, self._upsilon(), self.device,
                              symmetrize=symmetrize, symmetries=ops)
        self._set_two_phonon(self._to_mode_space(d2 @ pol), "")

    def prepare_anharmonic_ir(self, ec=None, ec_eq=None, pol_vec_light=np.array([1., 0., 0.]),
                              add_two_ph=False):
        """IR perturbation with configuration-dependent effective charges, earlier form
        (Py:1827-1960): the averages are taken on this object's ensemble in the mode basis,
        <X_n f_ups(w_n) (E^T m^-1/2 Z* . pol)_m>, both as FP64 products on the device."""
        import torch
        from . import perturbations as pt
        if not self.use_wigner and add_two_ph:
            raise NotImplementedError("The two phonon processes are implemented only in Wigner")
        if ec is None:
            raise ValueError("Must specify the effective charges for all configurations!")
        ec = np.asarray(ec, dtype=TYPE_DP)
        if ec.shape != (self.N, self.nat, 3, 3):
            raise ValueError("The effective charges in input have the wrong shape. The required "
                             "is {}".format((self.N, self.nat, 3, 3)))
        pol = np.asarray(pol_vec_light, dtype=TYPE_DP)
        z = np.einsum("iabc, b -> iac", ec, pol).reshape(self.N, -1)     # (N, 3 nat)
        d1 = self.rho @ z / np.sum(self.rho)
        self.prepare_perturbation(d1, masses_exp=-1)
        if self.verbose:
            print("Pertubation modulus with one ph effects only = {}".format(
                self.perturbation_modulus))
        if not add_two_ph:
            return
        if ec_eq is not None:
            ec_eq = np.asarray(ec_eq, dtype=TYPE_DP)
            if ec_eq.ndim != 3 or ec_eq.shape[1:] != (3, 3):
                raise ValueError("Error, effective charges of the wrong shape: {}".format(
                    ec_eq.shape))
            n_sc = self._n_supercell(3 * ec_eq.shape[0])
            if not self.ignore_small_w and ec_eq.shape[0] * 3 * n_sc != self.n_modes + 3:
                raise ValueError("effective charges do not match the modes")
            z = z - np.tile(np.einsum("abc, b -> ac", ec_eq, pol).ravel(), n_sc)[None]
        dev = pt._dev(self.device)
        P = torch.as_tensor(self.pols / np.sqrt(self.m)[:, None], device=dev)    # (3 nat, n)
        zt = torch.as_tensor(z, device=dev)
        r = torch.as_tensor(self.rho, device=dev)
        X = torch.as_tensor(np.ascontiguousarray(self.X), device=dev)
        C1 = pt._tn(X, r, zt, 1.0 / float(self.rho.sum()), dev)              # <X (x) z>
        d2 = pt._tn(C1.t().contiguous(), None, P, 1.0, dev).cpu().numpy()    # (n, n)
        d2 = f_ups(self.w, self.T)[:, None] * d2
        self._set_two_phonon(d2, "")

    def prepare_anharmonic_raman_FT(self, raman=None, raman_eq=None,
                                    pol_in=np.array([1., 0., 0.]), pol_out=np.array([1., 0., 0.]),
                                    mixed=False, pol_in_2=None, pol_out_2=None, pol_in_3=None,
                                    pol_out_3=None, add_two_ph=False, symmetrize=False,
                                    ensemble=None, save_raman_tensor2=False,
                                    file_raman_tensor2=None):
        """Raman perturbation with configuration-dependent Raman tensors (Py:1340-1532).

        raman (N, 3, 3, 3 nat_sc); raman_eq (3, 3, 3 nat_uc), subtracted (tiled) in the
        two-phonon average only. One-phonon sector: the Raman vector of <dalpha/dR> for
        (pol_in, pol_out) (plus the _2 and _3 pairs when mixed), with the acoustic sum rule of
        raman_vector; two-phonon sectors as prepare_anharmonic_ir_FT with the second-order
        Raman tensor. save_raman_tensor2: np.save the mass-scaled second-order tensor to
        file_raman_tensor2 and return before the two-phonon sectors, as the reference does."""
        from . import perturbations as pt
        if not self.use_wigner and add_two_ph:
            raise NotImplementedError("The two phonon processes are implemented only in Wigner")
        if raman is None:
            raise ValueError("Must specify the raman tensors for all configurations!")
        pairs = [(pol_in, pol_out)]
        if mixed:
            for a, b, k in ((pol_in_2, pol_out_2, 2), (pol_in_3, pol_out_3, 3)):
                if a is None or b is None:
                    raise ValueError("Must specify pol_in_{0} pol_out_{0} if mixed = True!"
                                     .format(k))
                if len(a) != 3 or len(b) != 3:
                    raise ValueError("pol_in_{0} pol_out_{0} must be array of len 3".format(k))
                pairs.append((a, b))
        raman = np.asarray(raman, dtype=TYPE_DP)
        if raman.ndim != 4 or raman.shape[1:3] != (3, 3) or raman.shape[3] != 3 * self.nat:
            raise ValueError("The raman tensor in input have the wrong shape. The required is "
                             "N_conf - E_field - E_field - 3 * N_at_sc")
        u, rho = self._ensemble_arrays(ensemble, raman.shape[0])
        ops = self._sym_ops(symmetrize)
        d1 = pt.get_d1alpha_dR_av(rho, raman, self.device, symmetrize, ops)
        v = sum(pt.raman_vector(d1, a, b) for a, b in pairs)
        self.prepare_perturbation(v, masses_exp=-1)
        if self.verbose:
            print("Pertubation modulus with one ph effects only = {}".format(
                self.perturbation_modulus))
        if not add_two_ph:
            return
        if raman_eq is not None:
            raman_eq = np.asarray(raman_eq, dtype=TYPE_DP)
            if raman_eq.ndim != 3 or raman_eq.shape[:2] != (3, 3):
                raise ValueError("Error, raman tns of the wrong shape: {}".format(raman_eq.shape))
            n_sc = self._n_supercell(raman_eq.shape[2])
            if not self.ignore_small_w and raman_eq.shape[2] * n_sc != self.nat * 3:
                raise ValueError("the equilibrium Raman tensor does not match the supercell")
            raman = raman - np.tile(raman_eq, n_sc)[None]
        d2 = pt.get_d2alpha_dR_av(u, rho, raman, self._upsilon(), self.device,
                                  symmetrize=symmetrize, symmetries=ops)
        sm = 1.0 / np.sqrt(self.m)
        d2 = d2 * sm[None, None, :, None] * sm[None, None, None, :]
        if save_raman_tensor2:
            np.save("{}".format(file_raman_tensor2), d2)
            return
        M = sum(np.einsum("abcd, a, b -> cd", d2, np.asarray(a, dtype=TYPE_DP),
                          np.asarray(b, dtype=TYPE_DP)) for a, b in pairs)
        self._set_two_phonon(self.pols.T @ M @ self.pols, " RAMAN")

    def prepare_anharmonic_raman_FT_2ph(self, d2alpha_dR=None, pol_in=np.array([1., 0., 0.]),
                                        pol_out=np.array([1., 0., 0.]), mixed=False,
                                        pol_in_2=None, pol_out_2=None):
        """Two-phonon sectors from a given second-order Raman tensor (3, 3, 3 nat_sc,
        3 nat_sc), already divided by the masses (Py:1535-1644); the one-phonon sector of psi
        is left as it is."""
        if not self.use_wigner:
            raise NotImplementedError("The two phonon processes are implemented only in Wigner")
        if d2alpha_dR is None:
            raise ValueError("Must specify the 2nd order Raman tensor!")
        d2 = np.asarray(d2alpha_dR, dtype=TYPE_DP)
        exp_shape = (3, 3, self.nat * 3, self.nat * 3)
        if d2.shape != exp_shape:
            raise ValueError("The shape of the 2nd order Raman tensor is not correct, expected "
                             "{}".format(exp_shape))
        pairs = [(pol_in, pol_out)]
        if mixed:
            if pol_in_2 is None or pol_out_2 is None:
                raise ValueError("Must specify pol_in_2 pol_out_2 if mixed = True!")
            if len(pol_in_2) != 3 or len(pol_out_2) != 3:
                raise ValueError("pol_in_2 pol_out_2 must be array of len 3")
            pairs.append((pol_in_2, pol_out_2))
        M = sum(np.einsum("abcd, a, b -> cd", d2, np.asarray(a, dtype=TYPE_DP),
                          np.asarray(b, dtype=TYPE_DP)) for a, b in pairs)
        self._set_two_phonon(self.pols.T @ M @ self.pols, " RAMAN")

    def prepare_unpolarized_raman_FT(self, index=0, debug=False, eq_raman_tns=None,
                                     use_symm=True, ens_av_raman=None, raman_tns_ens=None,
                                     add_2ph=True):
        """Component `index` (0-6) of the unpolarized Raman signal with configuration-dependent
        Raman tensors (Py:1222-1336): prepare_anharmonic_raman_FT on (xx + yy + zz), (xx - yy),
        (xx - zz), (yy - zz), xy, xz, yz; debug saves the second-order tensor under the
        component's name, as the reference does. The tensors are arguments (the reference also
        requires the dynamical matrix to carry an equilibrium Raman tensor, which it does not
        use here)."""
        if index not in range(7):
            raise ValueError('{} should be in {}'.format(index, list(range(7))))
        e = np.eye(3)
        z = np.zeros(3)
        # (pol_in, pol_out, mixed, pol_in_2, pol_out_2, pol_in_3, pol_out_3, file name)
        table = {0: (e[0], e[0], True, e[1], e[1], e[2], e[2], "xx_plus_yy_plus_zz"),
                 1: (e[0], e[0], True, -e[1], e[1], z, z, "xx_minus_yy"),
                 2: (e[0], e[0], True, -e[2], e[2], z, z, "xx_minus_zz"),
                 3: (e[1], e[1], True, -e[2], e[2], z, z, "yy_minus_zz"),
                 4: (e[0], e[1], False, None, None, None, None, "xy_square"),
                 5: (e[0], e[2], False, None, None, None, None, "xz_square"),
                 6: (e[1], e[2], False, None, None, None, None, "yz_square")}
        pi, po, mixed, pi2, po2, pi3, po3, name = table[index]
        self.prepare_anharmonic_raman_FT(raman=raman_tns_ens, raman_eq=eq_raman_tns,
                                         pol_in=pi, pol_out=po, mixed=mixed, pol_in_2=pi2,
                                         pol_out_2=po2, pol_in_3=pi3, pol_out_3=po3,
                                         add_two_ph=add_2ph, symmetrize=use_symm,
                                         ensemble=ens_av_raman, save_raman_tensor2=debug,
                                         file_raman_tensor2=name)

    def prepare_two_ph(self, a, b):
        """Two-phonon perturbation on the mode pair (a, b) (Py:2035-2087, Wigner only): psi
        is reset and holds a'(1) = -sqrt(-chi^-/2) o D, b'(1) = sqrt(chi^+/2) o D with
        D_ab = D_ba = 1/2 (D_aa = 1); the modulus is taken with the Wigner mask.
        Deviation: an index equal to n_modes is refused (the reference's check lets it through
        to an IndexError)."""
        if not self.use_wigner:
            raise NotImplementedError('The two phonon response is available only in Wigner')
        if not (0 <= a < self.n_modes and 0 <= b < self.n_modes):
            raise ValueError('The a-b indices must be smaller than {}'.format(self.n_modes))
        if self.verbose:
            print('PREPARE THE TWO PHONON PERTUBATION')
            print('Indices selected a = {} b = {}'.format(a, b))
        self.reset()
        self.psi = np.zeros(self.n_modes + self.n_modes * (self.n_modes + 1), dtype=TYPE_DP)
        D = np.zeros((self.n_modes, self.n_modes))
        D[a, b] = 0.5
        D[b, a] += 0.5
        self._set_two_phonon(D, "")

    def set_max_frequency(self, freq):
        """Py:2167-2192: run the Lanczos on 4 freq^2 - L (reverse_L with shift 4 freq^2), whose
        largest eigenvalues are the lowest of L."""
        self.shift_value = 4 * freq * freq
        self.reverse_L = True
        if self.verbose:
            print("Shift value:", self.shift_value)

    def build_lanczos_matrix_from_coeffs(self, use_arnoldi=False):
        """The tridiagonal matrix of the run (Py:4571-4612): a on the diagonal, b below and c
        above (b above too for a symmetric run without c), mapped back through reverse_L and
        shift_value, so that it represents L itself. This package keeps no Arnoldi matrix:
        use_arnoldi=True raises."""
        if use_arnoldi:
            raise NotImplementedError("no Arnoldi matrix is kept by this Lanczos")
        N = len(self.a_coeffs)
        M = np.zeros((N, N), dtype=TYPE_DP)
        for i in range(N):
            M[i, i] = self.a_coeffs[i]
            if i >= 1:
                c = self.c_coeffs[i - 1] if len(self.c_coeffs) == len(self.b_coeffs) else \
                    self.b_coeffs[i - 1]
                M[i - 1, i] = c
                M[i, i - 1] = self.b_coeffs[i - 1]
        sign = -1 if self.reverse_L else 1
        return sign * M - sign * np.eye(N) * self.shift_value

    # ------------------------------------------------------------------ psi helpers
    def get_Y1(self, half_off_diagonal=False):
        """Py:2827-2858."""
        n = self.n_modes
        nw2 = n * (n + 1) // 2
        Y1 = np.zeros((n, n))
        Y1[np.triu_indices(n)] = self.psi[n:n + nw2]
        Y1 = Y1 + np.triu(Y1, 1).T
        if half_off_diagonal:
            Y1 *= 0.5
            Y1[np.diag_indices(n)] *= 2
        return Y1

    def get_ReA1(self, half_off_diagonal=False):
        """Py:2861-2888."""
        n = self.n_modes
        nw2 = n * (n + 1) // 2
        R = np.zeros((n, n))
        R[np.triu_indices(n)] = self.psi[n + nw2:]
        R = R + np.triu(R, 1).T
        if half_off_diagonal:
            R *= 0.5
            R[np.diag_indices(n)] *= 2
        return R

    def _wigner_pair_factors(self):
        """Per independent pair (row i, column j >= i, in psi's packing order): w_i w_j, X =
        (1 + 2 n_i)(1 + 2 n_j) / 8, sqrt(-chi^-/2) and sqrt(chi^+/2) (Py:2747-2770)."""
        n = self.n_modes
        iu, ju = np.triu_indices(n)
        w = np.asarray(self.w, dtype=TYPE_DP)
        occ = 1 / (np.exp(w * RY_TO_K / self.T) - 1) if self.T > 0 else np.zeros(n)
        wa, wb, na, nb = w[ju], w[iu], occ[ju], occ[iu]
        X = (1 + 2 * na) * (1 + 2 * nb) / 8
        chi_m = (wa - wb) * (na - nb) / (2 * wa * wb)
        chi_p = (wa + wb) * (1 + na + nb) / (2 * wa * wb)
        return wa * wb, X, np.sqrt(np.maximum(-0.5 * chi_m, 0.0)), np.sqrt(0.5 * chi_p)

    def _unpack_symmetric(self, v):
        n = self.n_modes
        M = np.zeros((n, n), dtype=TYPE_DP)
        M[np.triu_indices(n)] = v
        return M + np.triu(M, 1).T

    def get_alpha1_beta1_wigner(self, get_alpha=True):
        """Upsilon^(1) = alpha^(1) (get_alpha) or beta^(1) as a symmetric (n, n) matrix from
        the a'(1), b'(1) sectors of psi (Py:2713-2822):
        alpha = w_i w_j / X (sqrt(-chi^-/2) a' - sqrt(chi^+/2) b'),
        beta = 1 / X (sqrt(-chi^-/2) a' + sqrt(chi^+/2) b')."""
        n = self.n_modes
        nw2 = n * (n + 1) // 2
        psi = np.asarray(self.psi, dtype=TYPE_DP)
        a_all, b_all = psi[n:n + nw2], psi[n + nw2:n + 2 * nw2]
        ww, X, sm, sp = self._wigner_pair_factors()
        if get_alpha:
            return self._unpack_symmetric((ww / X) * (sm * a_all - sp * b_all))
        return self._unpack_symmetric((sm * a_all + sp * b_all) / X)

    def get_a1_b1_wigner(self, get_a1=True):
        """The inverse change of variables (Py:2630-2710), from get_alpha1_beta1_wigner's
        matrices back to packed vectors: sqrt(-chi^-/2) a'(1) = X / 2 (alpha / (w_i w_j) +
        beta) (get_a1), or b'(1) = -X / 2 (alpha / (w_i w_j) - beta) / sqrt(chi^+/2)."""
        iu = np.triu_indices(self.n_modes)
        alpha = self.get_alpha1_beta1_wigner(True)[iu]
        beta = self.get_alpha1_beta1_wigner(False)[iu]
        ww, X, _sm, sp = self._wigner_pair_factors()
        if get_a1:
            return 0.5 * X * (alpha / ww + beta)
        return -0.5 * X * (alpha / ww - beta) / sp

    def mask_dot_wigner(self, debug=False):
        """Py:5790-5839: weight 2 on off-diagonal pair entries."""
        n = self.n_modes
        mask = np.ones(n + n * (n + 1))
        iu = np.triu_indices(n)
        off = iu[0] < iu[1]
        nw2 = n * (n + 1) // 2
        mask[n:n + nw2][off] = 2
        mask[n + nw2:][off] = 2
        return mask

    # ------------------------------------------------------------------ L application
    def _flags(self, transpose, extra=0):
        fl = extra
        if transpose:
            fl |= L.OP_TRANSPOSE
        if self.use_wigner:
            fl |= L.OP_WIGNER
        if self.ignore_v3:
            fl |= L.OP_IGNORE_V3
        if self.ignore_v4:
            fl |= L.OP_IGNORE_V4
        if self.ignore_harmonic:
            fl |= L.OP_IGNORE_HARMONIC
        if self.reverse_L:
            fl |= L.OP_REVERSE
        return fl

    def apply_full_L_device(self, psi, transpose=False, out=None):
        """L psi with psi already a device tensor; returns a device tensor."""
        ens = self.device_ensemble
        return ens.apply_full_L(psi, flags=self._flags(transpose), shift=float(self.shift_value),
                                precision=self.precision, out=out)

    def apply_full_L(self, target=None, force_t_0=False, force_FT=True, transpose=False,
                     fast_lanczos=True):
        """Py:3467-3570 (finite-temperature branch): overwrites and returns self.psi."""
        if force_t_0 and force_FT:
            raise ValueError("Error, only one between force_t_0 and force_FT can be True")
        if target is not None:
            self.psi = np.asarray(target, dtype=TYPE_DP)
        t1 = time.time()
        out = self.apply_full_L_device(self.psi, transpose)
        self.psi = out.cpu().numpy()
        if self.verbose:
            print("Time to apply the full L: {}".format(time.time() - t1))
        return self.psi

    def apply_L1_FT(self, transpose=False):
        """Harmonic part only (Py:2227-2385), evaluated on the device."""
        ens = self.device_ensemble
        fl = self._flags(transpose) & ~(L.OP_REVERSE | L.OP_IGNORE_HARMONIC)
        if self.ignore_harmonic:
            return np.zeros_like(self.psi)
        return ens.apply_full_L(self.psi, flags=fl | L.OP_IGNORE_V3, shift=0.0,
                                precision=self.precision).cpu().numpy()

    def apply_anharmonic_FT(self, transpose=False, test_weights=False, use_old_version=False):
        """Anharmonic part only (Py:2891-3281), evaluated on the device."""
        ens = self.device_ensemble
        fl = (self._flags(transpose) & ~(L.OP_REVERSE | L.OP_IGNORE_V3)) | L.OP_IGNORE_HARMONIC
        return ens.apply_full_L(self.psi, flags=fl, shift=0.0,
                                precision=self.precision).cpu().numpy()

    # ------------------------------------------------------------------ Lanczos
    def run_FT(self, n_iter, save_dir=None, save_each=5, verbose=True, n_rep_orth=0, n_ortho=10,
               flush_output=True, debug=False, prefix="LANCZOS", run_simm=False, optimized=False,
               full_gs=False, device_loop=False, graph=True):
        """Biconjugate (or Wigner symmetric) Lanczos, Py:5843-6372, Krylov vectors on device.

        device_loop=True runs the whole recurrence on the GPU (tdscha_lanczos_*, one hipGraph
        per step when graph=True), with or without full_gs; it needs a fresh start and
        n_rep_orth = 0.
        """
        import torch
        self.verbose = verbose
        ens = self.device_ensemble
        if device_loop:
            if len(self.a_coeffs) or n_rep_orth:
                raise ValueError("device_loop needs a fresh start and n_rep_orth=0")
            return self._run_device_loop(n_iter, run_simm, graph, save_dir, prefix, full_gs)
        dev = ens.device
        psi_host = np.asarray(self.psi, dtype=TYPE_DP)
        psi_norm = np.sum(psi_host ** 2)
        if np.isnan(psi_norm) or psi_norm == 0:
            raise ValueError("Error, you must initialize a perturbation to start the Lanczos.")
        if run_simm and not self.use_wigner:
            raise NotImplementedError("The symmetric Lanczos works only with Wigner.")
        if full_gs and optimized:
            raise ValueError("full_gs needs every stored vector; optimized=True drops them")
        if save_dir is not None and not os.path.exists(save_dir):
            os.makedirs(save_dir)

        mask = torch.as_tensor(self.mask_dot_wigner(), device=dev) if run_simm else None
        if run_simm:
            dot = lambda u, v: torch.dot(u, v * mask)
        else:
            dot = lambda u, v: torch.dot(u, v)
        to_dev = lambda a: torch.as_tensor(np.asarray(a, dtype=TYPE_DP), device=dev).clone()

        i_step = len(self.a_coeffs)
        if i_step == 0:
            first = to_dev(psi_host)
            first = first / torch.sqrt(dot(first, first))
            if self.gamma_sector:
                for adj in ((False,) if run_simm or self.use_wigner else (False, True)):
                    r = ens.proj.residual(first, adjoint=adj)
                    if r > 1e-8:
                        raise ValueError("gamma_sector: the perturbation is not "
                                         "translation-invariant (|psi - P0 psi| / |psi| = %g)"
                                         % r)
            Q, P, s_norm_l = [first], [first.clone()], [1.0]
        else:
            Q = [to_dev(q) for q in self.basis_Q]
            P = [to_dev(p) for p in self.basis_P]
            s_norm_l = [float(s) for s in self.s_norm]
        a_l = [float(a) for a in self.a_coeffs]
        b_l = [float(b) for b in self.b_coeffs]
        c_l = [float(c) for c in self.c_coeffs]
        if not (len(Q) == len(P) == len(s_norm_l)) or len(b_l) != len(c_l):
            raise ValueError("Something wrong when restoring the Lanczos.")

        if self.gamma_sector:
            # the starting vector was checked above and L keeps the sector
            with ens.unchecked():
                return self._run_FT_steps(Q, P, s_norm_l, a_l, b_l, c_l, i_step, n_iter,
                                          save_dir, save_each, verbose, n_rep_orth, n_ortho,
                                          flush_output, prefix, run_simm, optimized, full_gs,
                                          dot, mask, dev)
        return self._run_FT_steps(Q, P, s_norm_l, a_l, b_l, c_l, i_step, n_iter, save_dir,
                                  save_each, verbose, n_rep_orth, n_ortho, flush_output, prefix,
                                  run_simm, optimized, full_gs, dot, mask, dev)

    def _run_FT_steps(self, Q, P, s_norm_l, a_l, b_l, c_l, i_step, n_iter, save_dir, save_each,
                      verbose, n_rep_orth, n_ortho, flush_output, prefix, run_simm, optimized,
                      full_gs, dot, mask, dev):
        """The step loop of run_FT (Py:6007-6365) on device vectors."""
        import torch
        ens = self.device_ensemble
        psi_q, psi_p = Q[-1], P[-1]
        work_q = torch.empty_like(psi_q)
        work_p = torch.empty_like(psi_q)
        next_converged = False
        for i in range(i_step, i_step + n_iter):
            if verbose:
                print("\n ===== NEW STEP %d =====\n" % (i + 1))
                if flush_output:
                    sys.stdout.flush()
            if run_simm:
                L_q = self.apply_full_L_device(psi_q, transpose=False, out=work_q).clone()
                p_L = L_q.clone()
            else:
                # L q and L^T p share one pass over the symmetries
                L_q, p_L = ens.apply_full_L_batched(
                    [psi_q, psi_p], flags=[self._flags(False), self._flags(True)],
                    shifts=[float(self.shift_value)] * 2, precision=self.precision,
                    outs=[work_q, work_p])
                L_q, p_L = L_q.clone(), p_L.clone()
            c_old = c_l[-1] if len(c_l) > 0 else 1.0
            p_norm = s_norm_l[-1] / c_old
            a_coeff = float(dot(psi_p, L_q)) * p_norm
            if np.isnan(a_coeff):
                raise ValueError("Invalid value encountered during the Lanczos.\n"
                                 "Check if you have correctly initialized the algorithm.")
            rk = L_q - a_coeff * psi_q
            if len(Q) > 1:
                rk -= c_l[-1] * Q[-2]
            sk = p_L - a_coeff * psi_p
            if len(P) > 1:
                old_p_norm = s_norm_l[-2] if len(c_l) < 2 else s_norm_l[-2] / c_l[-2]
                sk -= b_l[-1] * P[-2] * (old_p_norm / p_norm)
            s_norm = float(torch.sqrt(dot(sk, sk)))
            sk_tilde = sk / s_norm
            s_norm *= p_norm
            b_coeff = float(torch.sqrt(dot(rk, rk)))
            c_coeff = float(dot(sk_tilde, rk / b_coeff)) * s_norm
            a_l.append(a_coeff)
            if np.abs(b_coeff) < __EPSILON__ or next_converged:
                if verbose:
                    print("Converged (b coefficient is {}, |b| < {})".format(b_coeff, __EPSILON__))
                break
            if np.abs(c_coeff) < __EPSILON__:
                if verbose:
                    print("Converged (c coefficient is {}, |c| < {})".format(c_coeff, __EPSILON__))
                break
            new_q = rk / b_coeff
            new_p = sk_tilde.clone()

            if full_gs:
                # two-sided classical Gram-Schmidt against every stored pair, two passes:
                # q -= sum_j (pt_j . q) Q_j, p -= sum_j (Q_j . p) pt_j with the duals
                # pt_j = P_j s_norm_j / c_(j-1); every coefficient of a pass comes from the
                # same vectors (the form the device-resident recurrence uses)
                Qm = torch.stack(Q)
                pns = torch.as_tensor([s_norm_l[j] / (c_l[j - 1] if j >= 1 else 1.0)
                                       for j in range(len(P))], dtype=Qm.dtype, device=dev)
                Pt = torch.stack(P) * pns[:, None]
                wq = mask if run_simm else None
                for _rep in range(2):
                    hq = Pt @ (new_q * wq if wq is not None else new_q)
                    hp = Qm @ (new_p * wq if wq is not None else new_p)
                    new_q -= hq @ Qm
                    new_p -= hp @ Pt
                del Qm, Pt
                normq = float(torch.sqrt(dot(new_q, new_q)))
                normp = float(torch.sqrt(dot(new_p, new_p)))
                if normq < __EPSILON__ or normp < __EPSILON__:
                    next_converged = True
                new_q /= normq
                new_p /= normp
                s_norm = c_coeff / float(dot(new_p, new_q))
            for _k in range(n_rep_orth):
                # the reference's loop, Py:6245-6317, as written
                ortho_q = 0.0
                ortho_p = 0.0
                start = 0 if n_ortho is None else max(0, len(P) - n_ortho)
                for j in range(start, len(P)):
                    coeff1 = dot(P[j], new_q)
                    coeff2 = dot(Q[j], new_p)
                    new_q -= coeff1 * P[j]
                    new_p -= coeff2 * Q[j]
                    ortho_q += abs(float(coeff1))
                    ortho_p += abs(ortho_p)
                normq = float(torch.sqrt(dot(new_q, new_q)))
                if normq < __EPSILON__:
                    next_converged = True
                new_q /= normq
                normp = float(dot(new_p, new_p))
                if np.abs(normp) < __EPSILON__:
                    next_converged = True
                new_p /= normp
                s_norm = c_coeff / float(dot(new_p, new_q))
                if ortho_p < __EPSILON__ and ortho_q < __EPSILON__:
                    break

            Q.append(new_q)
            P.append(new_p)
            psi_q, psi_p = new_q, new_p
            b_l.append(b_coeff)
            c_l.append(c_coeff)
            s_norm_l.append(s_norm)
            if optimized and len(P) > 3:
                P.pop(-4)
                Q.pop(-4)
                s_norm_l.pop(-4)
            if verbose:
                print("a_%d = %.8e" % (i, a_l[-1]))
                print("b_%d = %.8e" % (i, b_l[-1]))
                print("c_%d = %.8e" % (i, c_l[-1]))
            self.a_coeffs, self.b_coeffs, self.c_coeffs = a_l, b_l, c_l
            self.s_norm = s_norm_l
            self._device_basis = (Q, P)
            if save_dir is not None and (i + 1) % save_each == 0:
                self.basis_Q = [q.cpu().numpy() for q in Q]
                self.basis_P = [p.cpu().numpy() for p in P]
                self.save_status("%s/%s_STEP%d" % (save_dir, prefix, i + 1))
        self.a_coeffs, self.b_coeffs, self.c_coeffs = a_l, b_l, c_l
        self.s_norm = s_norm_l
        self.basis_Q = [q.cpu().numpy() for q in Q]
        self.basis_P = [p.cpu().numpy() for p in P]
        return self

    def _run_device_loop(self, n_iter, run_simm, graph, save_dir, prefix, full_gs=False):
        from .lanczos_device import DeviceLanczos
        ens = self.device_ensemble     # a GammaSectorOperator with gamma_sector: the
        #                               translation average runs inside the captured step
        if run_simm and not self.use_wigner:
            raise NotImplementedError("The symmetric Lanczos works only with Wigner.")
        psi = np.asarray(self.psi, dtype=TYPE_DP)
        if np.isnan(np.sum(psi ** 2)) or np.sum(psi ** 2) == 0:
            raise ValueError("Error, you must initialize a perturbation to start the Lanczos.")
        fl = self._flags(False) & ~L.OP_TRANSPOSE
        dl = DeviceLanczos(ens, 1, self.precision, [fl], [float(self.shift_value)],
                           symmetric=run_simm, max_steps=n_iter, keep_basis=True,
                           mask=self.mask_dot_wigner() if run_simm else None, full_gs=full_gs)
        dl.start([psi])
        dl.run(n_iter, graph=graph)
        a, b, c, s = dl.coefficients(0)
        Q, P = dl.basis(0, len(b))
        self.a_coeffs, self.b_coeffs, self.c_coeffs = list(a), list(b), list(c)
        self.s_norm = [1.0] + list(s)
        self.basis_Q = [q for q in Q.cpu().numpy()]
        self.basis_P = [p for p in P.cpu().numpy()]
        dl.close()
        if save_dir is not None:
            if not os.path.exists(save_dir):
                os.makedirs(save_dir)
            self.save_status("%s/%s_STEP%d" % (save_dir, prefix, len(a)))
        return self

    # ------------------------------------------------------------------ Green function
    def get_green_function_continued_fraction(self, w_array, use_terminator=True, last_average=1,
                                              smearing=0):
        """Py:4951-5047, evaluated in complex128 on the device."""
        import torch
        dev = self.device_ensemble.device if self.initialized else torch.device("cpu")
        a_c = np.asarray(self.a_coeffs, dtype=np.float64)
        b_c = np.asarray(self.b_coeffs, dtype=np.float64)
        c_c = np.asarray(self.c_coeffs, dtype=np.float64)
        return continued_fraction(a_c, b_c, c_c, w_array, use_terminator, last_average, smearing,
                                  self.use_wigner, self.reverse_L, self.shift_value,
                                  self.perturbation_modulus, device=dev)

    # ------------------------------------------------------------------ status files
    def save_abc(self, file):
        """Py:3572-3585."""
        total_len = len(self.c_coeffs)
        abc = np.zeros((total_len, 3), dtype=np.double)
        abc[:, 0] = self.a_coeffs[:total_len]
        abc[:, 1] = self.b_coeffs
        abc[:, 2] = self.c_coeffs
        np.savetxt(file, abc, header="a; b; c")

    def load_abc(self, file):
        """Py:3587-3595."""
        abc = np.loadtxt(file)
        self.a_coeffs = abc[:, 0]
        self.b_coeffs = abc[:, 1]
        self.c_coeffs = abc[:, 2]

    def save_status(self, file):
        """Py:3597-3646 (same keys; sym_julia / deg_julia hold the padded blocks)."""
        if not ".npz" in file.lower():
            file += ".npz"
        sym_julia, deg_julia = self._padded_symmetries()
        np.savez_compressed(
            file, T=self.T, nat=self.nat, m=self.m, w=self.w, pols=self.pols, n_modes=self.n_modes,
            ignore_v3=self.ignore_v3, ignore_v4=self.ignore_v4, N=self.N, rho=self.rho, X=self.X,
            Y=self.Y, psi=self.psi, a_coeffs=self.a_coeffs, b_coeffs=self.b_coeffs,
            c_coeffs=self.c_coeffs, basis_Q=self.basis_Q, basis_P=self.basis_P, s_norm=self.s_norm,
            krilov_basis=[], arnoldi_matrix=[], reverse=self.reverse_L, shift=self.shift_value,
            perturbation_modulus=self.perturbation_modulus, use_wigner=self.use_wigner,
            ignore_small_w=self.ignore_small_w, sym_julia=sym_julia, deg_julia=deg_julia,
            n_syms=self.n_syms, N_degeneracy=self.N_degeneracy
            if self.N_degeneracy is not None else np.zeros(0, dtype=np.intc))

    def prepare_input_files(self, root_name="tdscha", n_steps=100, start_from_scratch=True,
                            directory=".", run_symm=False):
        """Text inputs of the reference's standalone executable (Py:862-919), same files and
        formats: XXX.json, .ndegs, .blockid, .masses, .freqs, .rho, .X.dat, .Y.dat, .psi,
        .block<i>, .symsb<i>."""
        import json
        if self.symmetries is None:
            raise ValueError("symmetries are not set")
        n_syms = self.symmetries[0].shape[0]
        json_data = {"T": self.T, "n_steps": n_steps, "ignore_v2": self.ignore_harmonic,
                     "ignore_v3": self.ignore_v3, "ignore_v4": self.ignore_v4,
                     "use_wigner": self.use_wigner, "run_sym": run_symm,
                     "data": {"n_configs": int(self.N), "n_modes": int(self.n_modes),
                              "n_syms": int(n_syms), "n_blocks": len(self.symmetries),
                              "perturbation_modulus": self.perturbation_modulus,
                              "reverse": self.reverse_L, "shift": self.shift_value}}
        if not start_from_scratch:
            raise NotImplementedError("Error, restarting from a previous calculation is not "
                                      "yet implemented.")
        root = os.path.join(directory, root_name)
        os.makedirs(directory, exist_ok=True)
        with open(root + ".json", "w") as fp:
            json.dump(json_data, fp)
        np.savetxt(root + ".ndegs", self.N_degeneracy, fmt="%d")
        np.savetxt(root + ".blockid", self.sym_block_id, fmt="%d")
        np.savetxt(root + ".masses", self.m)
        np.savetxt(root + ".freqs", self.w)
        np.savetxt(root + ".rho", self.rho)
        np.savetxt(root + ".X.dat", self.X)
        np.savetxt(root + ".Y.dat", self.Y)
        np.savetxt(root + ".psi", self.psi)
        for i, S in enumerate(self.symmetries):
            np.savetxt(root + ".block{:d}".format(i), self.degenerate_space[i], fmt="%d")
            ns, b1, b2 = S.shape
            with open(root + ".symsb{:d}".format(i), "w") as fp:
                for isym in range(ns):
                    for k1 in range(b1):
                        fp.write("".join(" {:22.16f}".format(v) for v in S[isym, k1]))
                        fp.write("\n")
                    fp.write("\n")

    def read_input_files(self, root_name="tdscha", directory="."):
        """Load what prepare_input_files wrote (the inputs the standalone executable reads):
        ensemble, symmetry blocks, psi and the json flags."""
        import json
        root = os.path.join(directory, root_name)
        with open(root + ".json") as fp:
            meta = json.load(fp)
        data = meta["data"]
        n, N = int(data["n_modes"]), int(data["n_configs"])
        X = np.loadtxt(root + ".X.dat", ndmin=2).reshape(N, n)
        Y = np.loadtxt(root + ".Y.dat", ndmin=2).reshape(N, n)
        w = np.loadtxt(root + ".freqs", ndmin=1)
        rho = np.loadtxt(root + ".rho", ndmin=1)
        m = np.loadtxt(root + ".masses", ndmin=1) if os.path.exists(root + ".masses") else None
        self.set_ensemble(X, Y, w, rho, float(meta["T"]), None, m)
        ndegs = np.loadtxt(root + ".ndegs", dtype=np.intc, ndmin=1)
        bid = np.loadtxt(root + ".blockid", dtype=np.intc, ndmin=1)
        ns = int(data["n_syms"])
        syms, degs = [], []
        for i in range(int(data["n_blocks"])):
            d = int(ndegs[i])
            degs.append(np.loadtxt(root + ".block{:d}".format(i), dtype=np.intc, ndmin=1))
            syms.append(np.loadtxt(root + ".symsb{:d}".format(i), ndmin=1).reshape(ns, d, d))
        self.set_symmetries(syms, ndegs, degs, bid)
        self.psi = np.loadtxt(root + ".psi", ndmin=1)
        self.ignore_harmonic = bool(meta["ignore_v2"])
        self.ignore_v3 = bool(meta["ignore_v3"])
        self.ignore_v4 = bool(meta["ignore_v4"])
        self.use_wigner = bool(meta["use_wigner"])
        self.perturbation_modulus = data["perturbation_modulus"]
        self.reverse_L = bool(data["reverse"])
        self.shift_value = data["shift"]
        return meta

    def load_from_input_files(self, root_name="tdscha", directory="."):
        """Results of the standalone executable (Py:920-998): XXX.abc, and the Krylov bases
        and s norms when present (XXX.qbasis.out, .pbasis.out, .snorm.out), then the json."""
        import json
        import warnings
        if not os.path.exists(directory):
            raise IOError("Error, the specified directory '{}' does not exist".format(directory))
        abc_file = os.path.join(directory, "{}.abc".format(root_name))
        if not os.path.exists(abc_file):
            raise IOError("Error, the file '{}' does not exist.".format(abc_file))
        abc = np.loadtxt(abc_file, dtype=np.double, ndmin=2)
        self.a_coeffs = abc[:, 0]
        self.b_coeffs = abc[:, 1]
        self.c_coeffs = abc[:, 2]
        for attr, ext in (("basis_Q", "qbasis.out"), ("basis_P", "pbasis.out"),
                          ("s_norm", "snorm.out")):
            fn = os.path.join(directory, "{}.{}".format(root_name, ext))
            if os.path.exists(fn):
                v = np.loadtxt(fn, dtype=np.double)
                setattr(self, attr, v.ravel() if attr == "s_norm" else v)
            else:
                warnings.warn("File {} not found. {} not loaded.".format(fn, attr))
        with open(os.path.join(directory, root_name + ".json")) as fp:
            meta = json.load(fp)
        self.perturbation_modulus = meta["data"]["perturbation_modulus"]
        self.T = meta["T"]
        self.ignore_harmonic = meta["ignore_v2"]
        self.ignore_v3 = meta["ignore_v3"]
        self.ignore_v4 = meta["ignore_v4"]
        self.N = meta["data"]["n_configs"]
        self.n_modes = meta["data"]["n_modes"]
        self.reverse_L = meta["data"]["reverse"]
        self.shift_value = meta["data"]["shift"]

    def _padded_symmetries(self):
        if self.symmetries is None:
            return np.zeros(0), np.zeros(0, dtype=np.int32)
        nb = len(self.symmetries)
        dmax = max(s.shape[1] for s in self.symmetries)
        sj = np.zeros((nb, self.n_syms, dmax, dmax))
        dj = np.zeros((nb, dmax), dtype=np.int32)
        for i, s in enumerate(self.symmetries):
            c = s.shape[1]
            sj[i, :, :c, :c] = s
            dj[i, :c] = self.degenerate_space[i]
        return sj, dj

    def load_status(self, file, is_file_instance=False):
        """Py:3649-3758. Restores the ensemble, the symmetry blocks and the Lanczos state."""
        if not is_file_instance and not os.path.exists(file):
            if os.path.exists(file + ".npz"):
                file += ".npz"
        with np.load(file, allow_pickle=False) as d:
            data = {k: d[k] for k in d.files}
        self.set_ensemble(data["X"], data["Y"], data["w"], data["rho"], float(data["T"]),
                          data.get("pols"), data.get("m"))
        self.psi = np.array(data["psi"], dtype=TYPE_DP)
        self.a_coeffs = list(data["a_coeffs"])
        self.b_coeffs = list(data["b_coeffs"])
        self.c_coeffs = list(data["c_coeffs"])
        self.basis_Q = list(data["basis_Q"])
        self.basis_P = list(data["basis_P"])
        self.s_norm = list(data["s_norm"])
        self.reverse_L = bool(data["reverse"])
        self.shift_value = float(data["shift"])
        self.perturbation_modulus = float(data["perturbation_modulus"])
        self.use_wigner = bool(data["use_wigner"])
        self.ignore_v3 = bool(data["ignore_v3"])
        self.ignore_v4 = bool(data["ignore_v4"])
        sj, dj = data.get("sym_julia"), data.get("deg_julia")
        if sj is not None and sj.ndim == 4 and sj.size:
            self.set_symmetries(*blocks_from_padded(sj, dj, self.n_modes,
                                                    data.get("N_degeneracy")))


def blocks_from_padded(sym_julia, deg_julia, n_modes, N_degeneracy=None):
    """sym_julia (n_blocks, N_sym, d_max, d_max) + deg_julia -> block lists (Py:753-778).

    The padded form does not store the block sizes; without N_degeneracy they are recovered
    as the largest index whose symmetry rows are not all zero.
    """
    nb = sym_julia.shape[0]
    syms, degs, ndeg = [], [], []
    for b in range(nb):
        if N_degeneracy is not None and len(N_degeneracy) == nb:
            d = int(N_degeneracy[b])
        else:
            nz = np.nonzero(np.any(sym_julia[b] != 0, axis=(0, 2)))[0]
            d = int(nz[-1]) + 1
        syms.append(np.ascontiguousarray(sym_julia[b, :, :d, :d]))
        degs.append(np.ascontiguousarray(deg_julia[b, :d].astype(np.intc)))
        ndeg.append(d)
    bid = np.full(n_modes, -1, dtype=np.intc)
    for b, dg in enumerate(degs):
        bid[dg] = b
    if (bid < 0).any():
        raise ValueError("the padded symmetry blocks do not cover every mode")
    return syms, np.array(ndeg, dtype=np.intc), degs, bid


def continued_fraction(a, b, c, w_array, use_terminator=True, last_average=1, smearing=0.0,
                       use_wigner=False, reverse_L=False, shift_value=0.0, perturbation_modulus=1.0,
                       device=None):
    """Continued-fraction Green function of Py:4951-5047 in complex128 (torch, any device)."""
    import torch
    dev = device if device is not None else torch.device("cpu")
    w = torch.as_tensor(np.asarray(w_array, dtype=np.float64), device=dev)
    w2 = w * w
    cplx = torch.complex128
    n_iters = len(a)
    sign = -1.0 if reverse_L else 1.0
    sym = len(c) == len(b)
    if use_terminator:
        a_av = float(np.mean(a[-last_average:]))
        b_av = float(np.mean(b[-last_average:]))
        c_av = float(np.mean(c[-last_average:])) if sym else b_av
        aa = a_av * sign - sign * shift_value
        bc = (b_av * sign) * (c_av * sign)
        zz = (aa + w2) if use_wigner else (aa - w2)
        disc = torch.sqrt((zz * zz - 4 * bc).to(cplx))
        gf = ((zz + disc) if use_wigner else (zz - disc)) / (2 * bc)
    else:
        aa = float(a[-1]) * sign - sign * shift_value
        zz = (aa + w2) if use_wigner else (aa - w2)
        gf = 1.0 / (zz.to(cplx) + 2j * w * smearing)
    for i in range(n_iters - 2, -1, -1):
        aa = float(a[i]) * sign - sign * shift_value
        bb = float(b[i]) * sign
        cc = float(c[i]) * sign if sym else bb
        zz = (aa + w2) if use_wigner else (aa - w2)
        gf = 1.0 / (zz.to(cplx) + 2j * w * smearing - bb * cc * gf)
    if not use_wigner:
        out = gf * perturbation_modulus
    else:
        out = (-gf.real + 1j * gf.imag) * perturbation_modulus
    return out.cpu().numpy()
