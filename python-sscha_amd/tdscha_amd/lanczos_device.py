"""Device-resident Lanczos over several chains (tdscha_lanczos_* in include/tdscha_hip.h).

Each step is one batched L application plus the FP64 recurrence of run_FT
(Modules/DynamicalLanczos.py:5843-6372) on the GPU, captured once as a hipGraph and
replayed; the host reads the coefficients when the run is over.
"""
import ctypes

import numpy as np

from . import _lib as L
from .ensemble import _stream_handle

EPSILON = 1e-12   # the reference's stopping threshold (Py:121, Py:6144-6153)


def truncate(a, b, c, s_norm=None, gs_norms=None):
    """Cut the coefficients where the reference's run_FT stops: a_k is kept, b_k / c_k of the
    first step with |b_k| or |c_k| below 1e-12 are not (Py:6143-6153). gs_norms (full_gs runs):
    min(|q|, |p|) of each step's new pair after the projection; below 1e-12 at step k the run
    keeps b_k, c_k and ends after a_(k+1) (run_FT's next_converged, Py:6283-6303)."""
    a, b, c = (np.asarray(x, dtype=np.float64) for x in (a, b, c))
    s = None if s_norm is None else np.asarray(s_norm)
    bad = np.nonzero((np.abs(b) < EPSILON) | (np.abs(c) < EPSILON) | ~np.isfinite(b) |
                     ~np.isfinite(c))[0]
    cut = len(b) if not len(bad) else int(bad[0])   # b, c kept below this index
    n_a = cut + 1 if len(bad) else len(a)
    if gs_norms is not None:
        g = np.asarray(gs_norms, dtype=np.float64)[:cut]
        dep = np.nonzero((g < EPSILON) | ~np.isfinite(g))[0]
        if len(dep) and dep[0] + 1 < cut:
            cut = int(dep[0]) + 1
            n_a = cut + 1
    n_a = min(n_a, len(a))
    return a[:n_a], b[:cut], c[:cut], (None if s is None else s[:cut])


class DeviceLanczos:
    """n_chains Lanczos chains on one DeviceEnsemble.

    flags[c], shifts[c]: TDSCHA_OP_* bits (WIGNER, REVERSE, IGNORE_*) and shift of chain c.
    symmetric: the Wigner symmetric form (run_simm), needs mask (mask_dot_wigner).
    dev_ens may be a gamma.GammaSectorOperator: the chains then run on the zone-centre
    operator, its translation average inside every captured step (their starting vectors
    must be translation-invariant; this is checked).
    """

    def __init__(self, dev_ens, n_chains=1, precision="fp64", flags=None, shifts=None,
                 symmetric=False, max_steps=100, keep_basis=False, mask=None, full_gs=False):
        self.gamma = None
        if hasattr(dev_ens, "proj") and hasattr(dev_ens, "ens"):
            if dev_ens.proj.handle is None:
                raise ValueError("the device Lanczos needs the library projector (translations "
                                 "acting freely on the atoms)")
            self.gamma = dev_ens
            dev_ens = dev_ens.ens
        self.ens = dev_ens
        self.torch = dev_ens.torch
        self.nc = int(n_chains)
        self.max_steps = int(max_steps)
        self.symmetric = bool(symmetric)
        self.full_gs = bool(full_gs)
        self.keep_basis = bool(keep_basis) or self.full_gs
        n = dev_ens.n
        self.n_psi = n + n * (n + 1)
        flags = [0] * self.nc if flags is None else list(flags)
        shifts = [0.0] * self.nc if shifts is None else list(shifts)
        fl = (ctypes.c_int * self.nc)(*[int(f) for f in flags])
        sh = (ctypes.c_double * self.nc)(*[float(s) for s in shifts])
        m = None
        if symmetric:
            if mask is None:
                raise ValueError("the symmetric Lanczos needs the Wigner mask")
            self._mask = np.ascontiguousarray(mask, dtype=np.float64)
            m = self._mask.ctypes.data_as(L._dp)
        h = ctypes.c_void_p()
        rc = L.lib().tdscha_lanczos_create(ctypes.byref(h), dev_ens._h, L.PRECISIONS[precision],
                                           self.nc, fl, sh, int(symmetric), self.max_steps,
                                           int(self.keep_basis), int(self.full_gs), m)
        L.check(rc, "tdscha_lanczos_create")
        self._h = h
        self._flags = flags
        if self.gamma is not None:
            L.check(L.lib().tdscha_lanczos_set_projector(self._h, self.gamma.proj.handle),
                    "tdscha_lanczos_set_projector")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().tdscha_lanczos_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def start(self, psis):
        t = self.torch
        if len(psis) != self.nc:
            raise ValueError("expected %d starting vectors" % self.nc)
        self._psi0 = [self.ens._dev(p, (self.n_psi,)) for p in psis]
        if self.gamma is not None:
            for p, f in zip(self._psi0, self._flags):
                for adj in ((False,) if self.symmetric or f & L.OP_WIGNER else (False, True)):
                    r = self.gamma.proj.residual(p, adjoint=adj)
                    if r > 1e-8:
                        raise ValueError("the zone-centre Lanczos needs translation-invariant "
                                         "starting vectors (|psi - P0 psi| / |psi| = %g)" % r)
        pp = (ctypes.c_void_p * self.nc)(*[p.data_ptr() for p in self._psi0])
        rc = L.lib().tdscha_lanczos_start(self._h, pp, _stream_handle(t, self.ens.device))
        L.check(rc, "tdscha_lanczos_start")

    def run(self, n_steps, graph=True):
        rc = L.lib().tdscha_lanczos_run(self._h, int(n_steps), int(bool(graph)),
                                        _stream_handle(self.torch, self.ens.device))
        L.check(rc, "tdscha_lanczos_run")

    def coefficients(self, chain=0, truncated=True):
        k = ctypes.c_int()
        a = np.zeros(self.max_steps)
        b = np.zeros(self.max_steps)
        c = np.zeros(self.max_steps)
        s = np.zeros(self.max_steps)
        rc = L.lib().tdscha_lanczos_read(self._h, int(chain), a.ctypes.data_as(L._dp),
                                         b.ctypes.data_as(L._dp), c.ctypes.data_as(L._dp),
                                         s.ctypes.data_as(L._dp), ctypes.byref(k))
        L.check(rc, "tdscha_lanczos_read")
        a, b, c, s = a[:k.value], b[:k.value], c[:k.value], s[:k.value]
        if truncated:
            return truncate(a, b, c, s, self.gs_norms(chain) if self.full_gs else None)
        return a, b, c, s

    def gs_norms(self, chain=0):
        """min(|q|, |p|) of every step's new pair after the projection (full_gs runs)."""
        k = ctypes.c_int()
        g = np.zeros(self.max_steps)
        L.check(L.lib().tdscha_lanczos_read_gs(self._h, int(chain), g.ctypes.data_as(L._dp),
                                               ctypes.byref(k)), "tdscha_lanczos_read_gs")
        return g[:k.value]

    def basis(self, chain, steps):
        """(Q, P) device views of the first steps + 1 stored vectors of a chain."""
        if not self.keep_basis:
            raise ValueError("created without keep_basis")
        q = ctypes.c_void_p()
        p = ctypes.c_void_p()
        L.check(L.lib().tdscha_lanczos_basis(self._h, int(chain), ctypes.byref(q), ctypes.byref(p)),
                "tdscha_lanczos_basis")
        count = (steps + 1) * self.n_psi
        shape = (steps + 1, self.n_psi)
        # views of library-owned device memory, valid while this object lives
        Q = _DeviceView(q.value, count, self.ens.device, self.torch).tensor().reshape(shape)
        P = _DeviceView(p.value, count, self.ens.device, self.torch).tensor().reshape(shape)
        return Q, P


class LanczosGroup:
    """DeviceLanczos runs on different ensembles advanced together, one graph launch per
    step (tdscha_lanczos_group_*). Start every member first."""

    def __init__(self, members):
        self.members = list(members)
        self.torch = self.members[0].torch
        arr = (ctypes.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        h = ctypes.c_void_p()
        L.check(L.lib().tdscha_lanczos_group_create(ctypes.byref(h), arr, len(self.members)),
                "tdscha_lanczos_group_create")
        self._h = h

    def run(self, n_steps):
        m = self.members[0]
        L.check(L.lib().tdscha_lanczos_group_run(self._h, int(n_steps),
                                                 _stream_handle(self.torch, m.ens.device)),
                "tdscha_lanczos_group_run")

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().tdscha_lanczos_group_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _DeviceView:
    """A torch tensor over library-owned device memory (via __cuda_array_interface__)."""

    def __init__(self, ptr, count, device, torch):
        self.__cuda_array_interface__ = {
            "shape": (count,), "typestr": "<f8", "data": (ptr, False), "version": 2,
            "strides": None}
        self._device = device
        self._torch = torch

    def tensor(self):
        return self._torch.as_tensor(self, device=self._device)
