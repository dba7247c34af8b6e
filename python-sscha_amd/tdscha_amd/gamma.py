"""L on translation-invariant (Gamma-point) perturbations over one symmetry per rotation.

The symmetrized ensemble average (the sum over the N_sym images of every configuration that
Modules/DynamicalLanczos.py:3051-3106 hands to CModules/LanczosFunctions.c:2082-2754) runs over
a space group: N_sym = |P| |T| operations, the cosets g = t o p of the pure translations T
(rotation = identity) by one operation p per rotation. When the perturbation is
translation-invariant, every image of a coset sees the same rotated operands
(S_g^T Y1 S_g = S_p^T Y1 S_p because S_t^T Y1 S_t = Y1), and the coset's outputs are the
translates of the representative's. So

    L psi = P0 (L_P psi)

with L_P the same operator over the |P| representatives only and P0 the average over the
translations (the projector on the translation-invariant sector). IR and Raman perturbations
at the zone centre are in that sector and L keeps them there, so their Lanczos chains need
|T| times fewer images (216 for the 6 x 6 x 6 rock-salt supercell: 48 instead of 10368).

P0 runs in Cartesian space, where a translation is a permutation of the atoms:
S_t M S_t^T = E^T P_t (E M E^T) P_t^T E for the polarization vectors E of the kept modes; the
average over the translations of a Cartesian matrix depends only on the orbits of its two
atoms and their relative translation (the translations commute and act freely), so it is a
sum over the pairs of each (orbit, orbit, relative translation) key, written back to every
pair of the key (the library's tdscha_projector_* kernels). psi
holds upper triangles of symmetric matrices, whose dot product counts an off-diagonal pair
once; in that metric the transpose of L is covariant with P0^T = W P0 W^-1 (W = 2 on the
off-diagonal entries), which is the projector for a transposed application in the standard
representation. In the Wigner representation both directions use P0.
"""
import ctypes
import warnings

import numpy as np

from . import _lib as L
from .ensemble import DeviceEnsemble, SymmetryBlocks
from .symmetry import translation_cosets


def _torch():
    import torch
    return torch


def _orbit_fold(irt):
    """For pure translations (an abelian group acting freely on the atoms): the orbit s(a) of
    every atom and, for every atom pair, the index (s(a) nb + s(b)) n_t + d(a, b) with d the
    translation tau_a^-1 tau_b (tau_x carries the orbit's first atom to x). None when the
    action is not free."""
    nt, nat = irt.shape
    if nat % nt:
        return None
    s = -np.ones(nat, dtype=np.int64)
    tau = -np.ones(nat, dtype=np.int64)
    reps = []
    for a in range(nat):
        if s[a] >= 0:
            continue
        img = irt[:, a]
        if len(np.unique(img)) != nt or np.any(s[img] >= 0):
            return None
        s[img] = len(reps)
        tau[img] = np.arange(nt)
        reps.append(a)
    r0 = reps[0]
    key = -np.ones(nat, dtype=np.int64)
    key[irt[:, r0]] = np.arange(nt)
    inv = np.argsort(irt, axis=1)
    img = inv[tau[:, None], irt[tau[None, :], r0]]       # (tau_a^-1 tau_b)(r0)
    d = key[img]
    if np.any(d < 0):
        return None
    nb = len(reps)
    return s, (s[:, None] * nb + s[None, :]) * nt + d, nb


class TranslationProjector:
    """P0 and its packed adjoint on device vectors psi = (R (n), two packed n x n upper
    triangles). pols: (3 nat, n) polarization vectors of the kept modes (orthonormal
    columns); irt_translations: (n_t, nat) the atom each pure translation carries every atom
    to.

    When the translations act freely on the atoms (as the pure translations of a supercell
    do) the projection runs in the library's kernels (tdscha_projector_*: the orbit-folded
    average, deterministic, four FP64 MFMA products per matrix), which the device Lanczos can
    also run inside its captured step (``handle``). Otherwise it falls back to the explicit
    average over every translation, on the device with torch."""

    def __init__(self, pols, irt_translations, device=0, chunk=16):
        torch = _torch()
        self.torch = torch
        self.dev = torch.device("cuda", device) if isinstance(device, int) else device
        E = np.ascontiguousarray(pols, dtype=np.float64)
        if E.ndim != 2 or E.shape[0] % 3:
            raise ValueError("pols must be (3 nat, n_modes)")
        irt = np.asarray(irt_translations, dtype=np.int64)
        nat = E.shape[0] // 3
        if irt.ndim != 2 or irt.shape[1] != nat:
            raise ValueError("irt_translations must be (n_t, %d)" % nat)
        if np.any(np.sort(irt, axis=1) != np.arange(nat)[None]):
            raise ValueError("every row of irt_translations must be a permutation of the atoms")
        self.n = E.shape[1]
        self.nat, self.nt = nat, len(irt)
        self._h = None
        if _orbit_fold(irt) is not None:
            irt32 = np.ascontiguousarray(irt, dtype=np.intc)
            h = ctypes.c_void_p()
            L.check(L.lib().tdscha_projector_create(
                ctypes.byref(h), self.dev.index or 0, self.n, nat, E.ctypes.data_as(L._dp),
                len(irt), irt32.ctypes.data_as(L._ip)), "tdscha_projector_create")
            self._h = h
            return
        warnings.warn("translation average: the translations do not act freely on the "
                      "atoms; using the explicit average over every translation")
        self.E = torch.as_tensor(E, device=self.dev)
        cart = (3 * irt[:, :, None] + np.arange(3)[None, None]).reshape(len(irt), -1)
        self.perm = torch.as_tensor(cart, device=self.dev)           # (n_t, 3 nat)
        self.chunk = chunk
        n = self.n
        iu = np.triu_indices(n)
        self.iu = (torch.as_tensor(iu[0], device=self.dev), torch.as_tensor(iu[1], device=self.dev))
        w = np.where(iu[0] < iu[1], 2.0, 1.0)
        self.w_off = torch.as_tensor(w, device=self.dev)

    @property
    def handle(self):
        """The library projector (None for the explicit-average fallback)."""
        return self._h

    def close(self):
        if self._h is not None and self._h.value:
            L.lib().tdscha_projector_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _avg_mat_gather(self, C):
        acc = self.torch.zeros_like(C)
        for t0 in range(0, self.perm.shape[0], self.chunk):
            p = self.perm[t0:t0 + self.chunk]
            acc += C[p[:, :, None], p[:, None, :]].sum(0)
        return acc / self.perm.shape[0]

    def project(self, psi, adjoint=False, out=None):
        """P0 psi, or P0^T psi (adjoint=True)."""
        torch = self.torch
        n = self.n
        nw2 = n * (n + 1) // 2
        if psi.shape[0] != n + 2 * nw2:
            raise ValueError("psi has %d entries, expected %d" % (psi.shape[0], n + 2 * nw2))
        if self._h is not None:
            psi = psi.to(device=self.dev, dtype=torch.float64).contiguous()
            if out is None:
                out = torch.empty_like(psi)
            elif (out.dtype != torch.float64 or out.device != psi.device or
                  not out.is_contiguous() or out.shape != psi.shape):
                raise ValueError("out must be a contiguous float64 tensor like psi")
            L.ops().projector_apply(self._h.value, self.n, bool(adjoint), psi, out)
            return out
        if out is None:
            out = torch.empty_like(psi)
        E = self.E
        out[:n] = E.T @ (E @ psi[:n])[self.perm].mean(0)
        for k in range(2):
            v = psi[n + k * nw2:n + (k + 1) * nw2]
            if adjoint:
                v = v / self.w_off
            M = torch.zeros((n, n), dtype=psi.dtype, device=psi.device)
            M[self.iu] = v
            M = M + M.triu(1).T
            r = (E.T @ self._avg_mat_gather(E @ M @ E.T) @ E)[self.iu]
            out[n + k * nw2:n + (k + 1) * nw2] = r * self.w_off if adjoint else r
        return out

    def residual(self, psi, adjoint=False):
        """|psi - P psi| / |psi|: how far psi is from the sector."""
        d = self.project(psi, adjoint) - psi
        return float(self.torch.linalg.norm(d) / self.torch.linalg.norm(psi))


class GammaSectorOperator:
    """L on translation-invariant perturbations: a DeviceEnsemble over one symmetry per
    rotation, followed by the translation average. Same apply_full_L / apply_full_L_batched
    interface as DeviceEnsemble; results equal the full operator's for psi in the sector
    (transposed standard-representation applications: psi in the sector of P0^T).

    sym: the symmetry blocks of all N_sym operations, in the order of (rotations, irt), the
    supercell's space group; pols: (3 nat, n) polarization vectors of the kept modes.

    check_tol: every apply refuses (ValueError) a psi farther than this from the sector,
    |psi - P0 psi| / |psi| (None: no check). A caller that has checked its vectors already
    (run_FT checks the starting vector, and L keeps the sector) may skip the per-call check
    inside ``with op.unchecked():``."""

    def __init__(self, X, Y, w, rho, T, sym: SymmetryBlocks, rotations, irt, pols,
                 precisions=("fp64",), device=0, s_batch=0, check_tol=1e-8):
        cos = translation_cosets(rotations, irt)
        if cos is None:
            raise ValueError("the operations are not cosets of the pure translations (a space "
                             "group with at least two translations is needed)")
        reps, trans = cos
        if sym.symmetries[0].shape[0] != len(rotations):
            raise ValueError("the symmetry blocks hold %d operations, the space group %d"
                             % (sym.symmetries[0].shape[0], len(rotations)))
        self.representatives = reps
        self.translations = trans
        sub = SymmetryBlocks([np.ascontiguousarray(s[reps]) for s in sym.symmetries],
                             sym.N_degeneracy, sym.degenerate_space, sym.block_id)
        self.ens = DeviceEnsemble(X, Y, w, rho, T, sub, precisions=precisions, device=device,
                                  s_batch=s_batch)
        self.proj = TranslationProjector(pols, np.asarray(irt)[trans], device=device)
        self.n = self.ens.n
        self.device = device
        self.torch = self.ens.torch
        self.check_tol = check_tol

    @staticmethod
    def _adjoint(flags):
        return bool(flags & L.OP_TRANSPOSE) and not bool(flags & L.OP_WIGNER)

    def unchecked(self):
        """Context in which the applies skip the per-call sector check."""
        import contextlib

        @contextlib.contextmanager
        def ctx():
            saved = self.check_tol
            self.check_tol = None
            try:
                yield self
            finally:
                self.check_tol = saved
        return ctx()

    def _check(self, psi, flags):
        if self.check_tol is not None:
            r = self.proj.residual(psi, self._adjoint(flags))
            if r > self.check_tol:
                raise ValueError("psi is not translation-invariant (|psi - P0 psi| / |psi| = %g):"
                                 " use the full operator" % r)

    def apply_full_L(self, psi, flags=0, shift=0.0, precision="fp64", out=None):
        psi = self.ens._dev(psi, (self.n + self.n * (self.n + 1),))
        self._check(psi, flags)
        tmp = self.ens.apply_full_L(psi, flags, shift, precision)
        return self.proj.project(tmp, self._adjoint(flags), out=out)

    def apply_full_L_batched(self, psis, flags=None, shifts=None, precision="fp64", outs=None):
        C = len(psis)
        flags = [0] * C if flags is None else list(flags)
        psis = [self.ens._dev(p, (self.n + self.n * (self.n + 1),)) for p in psis]
        for p, f in zip(psis, flags):
            self._check(p, f)
        tmp = self.ens.apply_full_L_batched(psis, flags, shifts, precision)
        if outs is None:
            outs = [None] * C
        return [self.proj.project(t, self._adjoint(f), out=o)
                for t, f, o in zip(tmp, flags, outs)]

    @property
    def N(self):
        return self.ens.N

    def force_grad(self, us, vs, flags=None, out=None, accumulate=False, precision="fp64"):
        """d/dY sum_p u_p . A_p(v_p) for v_p in the sector: A_p = P_f A_P, so
        u . A_p(v) = (P_f^T u) . A_P(v), the representatives' adjoint with u projected by the
        transposed projector (P0 <-> P0^T in the packed metric)."""
        C = len(us)
        flags = [0] * C if flags is None else list(flags)
        n_psi = self.n + self.n * (self.n + 1)
        us = [self.proj.project(self.ens._dev(u, (n_psi,)), adjoint=not self._adjoint(f))
              for u, f in zip(us, flags)]
        return self.ens.force_grad(us, vs, flags, out=out, accumulate=accumulate,
                                   precision=precision)

    def autotune(self, precision="bf16x3", n_chains=2):
        return self.ens.autotune(precision, n_chains)

    def set_tuning(self, bf16_tiles=0, pass_slots=0):
        return self.ens.set_tuning(bf16_tiles, pass_slots)

    def tuning(self):
        return self.ens.tuning()

    def close(self):
        self.ens.close()
        self.proj.close()


def _main():
    """Time the zone-centre operator on the synthetic rock-salt ensemble:
    python -m tdscha_amd.gamma [--L 6] [--configs 20000] [--precision bf16x3] [--reps 10]
    Two chains (L q and L^T p, q and p drawn in their sectors) per batched application; the
    image sum over the representatives and the projection are timed apart with events."""
    import argparse
    import json
    import time

    import torch

    from . import synthetic as sy
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=6)
    ap.add_argument("--configs", type=int, default=20000)
    ap.add_argument("--precision", default="bf16x3")
    ap.add_argument("--reps", type=int, default=10)
    a = ap.parse_args()
    dev = torch.device("cuda", 0)
    ens = sy.make_ensemble(a.L, a.configs, T=300.0, device=dev)
    rots, irt = sy.space_group(ens.crystal)
    sym = SymmetryBlocks(ens.symmetries, ens.N_degeneracy, ens.degenerate_space, ens.block_id)
    precs = ("fp64",) if a.precision == "fp64" else ("fp64", a.precision)
    op = GammaSectorOperator(ens.X, ens.Y, ens.w, ens.rho, ens.T, sym, rots, irt, ens.pols,
                             precisions=precs)
    if a.precision == "bf16x3":
        op.autotune(a.precision, 2)
    n = op.n
    g = torch.Generator().manual_seed(3)
    flags = [0, L.OP_TRANSPOSE]
    vs = []
    for f in flags:
        v = torch.randn(n + n * (n + 1), generator=g, dtype=torch.float64).to(dev)
        v = op.proj.project(v, adjoint=op._adjoint(f))
        vs.append(v / v.norm())
    tmp = [torch.empty_like(v) for v in vs]
    outs = [torch.empty_like(v) for v in vs]
    st = torch.cuda.current_stream(dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    t_sum = t_proj = 0.0
    for r in range(a.reps + 1):
        ev[0].record(st)
        op.ens.apply_full_L_batched(vs, flags, None, a.precision, outs=tmp)
        ev[1].record(st)
        for t, o, f in zip(tmp, outs, flags):
            op.proj.project(t, op._adjoint(f), out=o)
        ev[2].record(st)
        ev[2].synchronize()
        if r:                                  # the first application is the warm-up
            t_sum += ev[0].elapsed_time(ev[1])
            t_proj += ev[1].elapsed_time(ev[2])
    ms = (t_sum + t_proj) / a.reps
    print(json.dumps({"workload": "rock-salt %dx%dx%d, %d modes, %d configs, %d symmetries = "
                                  "%d rotations x %d translations"
                                  % (a.L, a.L, a.L, n, a.configs, ens.n_syms,
                                     len(op.representatives), len(op.translations)),
                      "precision": a.precision, "ms_per_batched_application": ms,
                      "L_applications_per_s": 2e3 / ms,
                      "image_sum_ms": t_sum / a.reps, "projection_ms": t_proj / a.reps,
                      "tuning": op.tuning()}), flush=True)


if __name__ == "__main__":
    _main()
