"""Device-resident ensemble: the Python face of tdscha_ensemble_* (include/tdscha_hip.h).

PyTorch supplies device memory and the current HIP stream; all arithmetic happens in the
HIP kernels of libtdscha_hip.so. The per-step calls (the perturbed averages, L, the force
adjoint) go through the PyTorch-ROCm operators torch.ops.tdscha.*; set-up and teardown
through the C ABI with ctypes.
"""
import ctypes

import numpy as np

from . import _lib as L


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class SymmetryBlocks:
    """Mode-basis symmetry blocks, as Lanczos.prepare_symmetrization leaves them
    (Modules/DynamicalLanczos.py:741-778):

    symmetries[b]       (N_sym, d_b, d_b) float64
    N_degeneracy[b]     d_b
    degenerate_space[b] mode ids of block b (contiguous, sorted)
    block_id[mu]        block of mode mu
    """

    def __init__(self, symmetries, N_degeneracy, degenerate_space, block_id):
        self.symmetries = [_c(s) for s in symmetries]
        self.N_degeneracy = _c(N_degeneracy, np.intc)
        self.degenerate_space = [_c(d, np.intc) for d in degenerate_space]
        self.block_id = _c(block_id, np.intc)
        self.n_syms = int(self.symmetries[0].shape[0])

    @property
    def n_blocks(self):
        return len(self.symmetries)

    @staticmethod
    def identity(n):
        return SymmetryBlocks([np.ones((1, 1, 1))] * n, np.ones(n, dtype=np.intc),
                              [np.array([i], dtype=np.intc) for i in range(n)],
                              np.arange(n, dtype=np.intc))

    def pointer_arrays(self):
        """(double**, int**) views; the caller keeps `self` alive during the C call."""
        nb = self.n_blocks
        syms = (L._dp * nb)(*[s.ctypes.data_as(L._dp) for s in self.symmetries])
        degs = (L._ip * nb)(*[d.ctypes.data_as(L._ip) for d in self.degenerate_space])
        return syms, degs


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is visible: the ensemble L-application runs on the GPU only")
    return torch


def _stream_handle(torch, device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class DeviceEnsemble:
    """X, Y, rho, w and the symmetry blocks uploaded once to one GPU.

    precisions: iterable of "fp64", "fp32", "bf16x3" whose operand copies to build.
    """

    def __init__(self, X, Y, w, rho, T, sym: SymmetryBlocks, precisions=("fp64",), device=0,
                 s_batch=0):
        torch = _torch()
        self.torch = torch
        self.device = torch.device("cuda", device)
        X, Y, w, rho = _c(X), _c(Y), _c(w), _c(rho)
        N, n = X.shape
        if Y.shape != (N, n) or w.shape != (n,) or rho.shape != (N,):
            raise ValueError("inconsistent shapes: X %s, Y %s, w %s, rho %s"
                             % (X.shape, Y.shape, w.shape, rho.shape))
        self.N, self.n, self.T = N, n, float(T)
        self.sym = sym
        self.w = w
        mask = 0
        for p in precisions:
            mask |= 1 << L.PRECISIONS[p]
        syms, degs = sym.pointer_arrays()
        h = ctypes.c_void_p()
        rc = L.lib().tdscha_ensemble_create(
            ctypes.byref(h), int(device), X.ctypes.data_as(L._dp), Y.ctypes.data_as(L._dp),
            w.ctypes.data_as(L._dp), rho.ctypes.data_as(L._dp), float(T), n, N, sym.n_syms,
            sym.n_blocks, syms, sym.N_degeneracy.ctypes.data_as(L._ip), degs,
            sym.block_id.ctypes.data_as(L._ip), mask, int(s_batch))
        L.check(rc, "tdscha_ensemble_create")
        self._h = h
        dims = (ctypes.c_int * 6)()
        L.check(L.lib().tdscha_ensemble_dims(h, dims), "tdscha_ensemble_dims")
        self.n_syms, self.s_batch, self.n_pad, self.N_pad = dims[2], dims[3], dims[4], dims[5]

    @property
    def nbytes(self):
        return int(L.lib().tdscha_ensemble_bytes(self._h))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().tdscha_ensemble_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _dev(self, a, shape):
        t = self.torch
        if a is None:
            return None
        if not isinstance(a, t.Tensor):
            a = t.as_tensor(np.asarray(a, dtype=np.float64))
        a = a.to(device=self.device, dtype=t.float64).contiguous()
        if tuple(a.shape) != shape:
            raise ValueError("expected shape %s, got %s" % (shape, tuple(a.shape)))
        return a

    def _check_out(self, o, shape, what="out"):
        """A caller's output buffer: the kernels write its whole extent through a raw device
        pointer, so anything but a contiguous FP64 tensor of the expected shape on this
        ensemble's device is refused before a launch."""
        t = self.torch
        if not isinstance(o, t.Tensor) or o.dtype != t.float64 or o.device != self.device or \
                not o.is_contiguous() or tuple(o.shape) != tuple(shape):
            desc = ("%s %s %s%s" % (o.dtype, tuple(o.shape), o.device,
                                    "" if o.is_contiguous() else " non-contiguous")
                    if isinstance(o, t.Tensor) else type(o).__name__)
            raise ValueError("%s must be a contiguous float64 tensor of shape %s on %s, got %s"
                             % (what, tuple(shape), self.device, desc))
        return o

    def apply(self, R1, Y1, terms=L.TERM_F | L.TERM_D2V_R | L.TERM_D2V_Y, precision="fp64",
              f=None, d2v=None, d2v_accumulate=False):
        """Perturbed averages (f_av, d2v) as device tensors (GetPerturbAverageSym contract)."""
        t = self.torch
        n = self.n
        R1 = self._dev(R1, (n,))
        Y1 = self._dev(Y1, (n, n))
        if f is None:
            f = t.empty(n, dtype=t.float64, device=self.device)
        else:
            self._check_out(f, (n,), "f")
        if d2v is None:
            d2v = t.zeros((n, n), dtype=t.float64, device=self.device)
        else:
            self._check_out(d2v, (n, n), "d2v")
        L.ops().ensemble_apply(self._h.value, L.PRECISIONS[precision], int(terms), R1, Y1, f, d2v,
                               bool(d2v_accumulate))
        return f, d2v

    def apply_full_L(self, psi, flags=0, shift=0.0, precision="fp64", out=None):
        """L psi on the device (Lanczos.apply_full_L, Modules/DynamicalLanczos.py:3467)."""
        t = self.torch
        n = self.n
        psi = self._dev(psi, (n + n * (n + 1),))
        if out is None:
            out = t.empty_like(psi)
        else:
            self._check_out(out, psi.shape)
            if out.data_ptr() == psi.data_ptr():
                raise ValueError("out must not alias psi")
        L.ops().apply_full_L_batched(self._h.value, L.PRECISIONS[precision], [psi], [int(flags)],
                                     [float(shift)], [out])
        return out

    def apply_full_L_batched(self, psis, flags=None, shifts=None, precision="fp64", outs=None):
        """L on several device vectors in one pass over the symmetries
        (tdscha_apply_full_L_batched). flags / shifts per vector."""
        t = self.torch
        n = self.n
        C = len(psis)
        psis = [self._dev(p, (n + n * (n + 1),)) for p in psis]
        if outs is None:
            outs = [t.empty_like(p) for p in psis]
        else:
            if len(outs) != C:
                raise ValueError("expected %d output tensors, got %d" % (C, len(outs)))
            ins = {p.data_ptr() for p in psis}
            for o in outs:
                self._check_out(o, (n + n * (n + 1),), "each of outs")
                if o.data_ptr() in ins:
                    raise ValueError("an output must not alias an input")
        flags = [0] * C if flags is None else list(flags)
        shifts = [0.0] * C if shifts is None else list(shifts)
        L.ops().apply_full_L_batched(self._h.value, L.PRECISIONS[precision], list(psis),
                                     [int(f) for f in flags], [float(s) for s in shifts],
                                     list(outs))
        return outs

    def autotune(self, precision="bf16x3", n_chains=2):
        """Pick the slot count per pass by timing (tdscha_ensemble_autotune); returns
        (slots, microseconds per chain-configuration-symmetry image)."""
        slots = ctypes.c_int()
        per = ctypes.c_double()
        L.check(L.lib().tdscha_ensemble_autotune(self._h, L.PRECISIONS[precision], int(n_chains),
                                                 ctypes.byref(slots), ctypes.byref(per)),
                "tdscha_ensemble_autotune")
        return slots.value, per.value

    def set_tuning(self, bf16_tiles=0, pass_slots=0):
        """BF16x3 contraction kernels (0: 256/192 x 128 tiles, 1: 256 x 256, 2: 256 x 256 with
        direct global->LDS copies, 3: 256/192 x 192 with the A operands in registers, 4: the
        same with 96-column tiles, 5: GEMM1 of 4 with GEMM2 of 3, 6: GEMM1 of 4 with 256 x 96
        GEMM2 tiles, 7: GEMM1 of 4 with 512 x 96 GEMM2 tiles of eight row waves, 8: the same
        with 448-row GEMM2 tiles) and image slots per pass (0 = all), as the autotuner sets
        them (tdscha_ensemble_set_tuning)."""
        L.check(L.lib().tdscha_ensemble_set_tuning(self._h, int(bf16_tiles), int(pass_slots)),
                "tdscha_ensemble_set_tuning")

    def tuning(self):
        v, s = ctypes.c_int(), ctypes.c_int()
        L.check(L.lib().tdscha_ensemble_get_tuning(self._h, ctypes.byref(v), ctypes.byref(s)),
                "tdscha_ensemble_get_tuning")
        return v.value, s.value

    def force_grad(self, us, vs, flags=None, out=None, accumulate=False, precision="fp64"):
        """d/dY of sum_p u_p . A_p(v_p) (tdscha_ensemble_force_grad): A_p is the anharmonic
        part of the operator with flags[p]. Returns an (N, n) FP64 device tensor.
        precision "fp64", or "fp32" (FP32 MFMA contractions with FP32 accumulation; needs an
        ensemble created with the fp32 copies)."""
        t = self.torch
        n = self.n
        C = len(us)
        if len(vs) != C:
            raise ValueError("us and vs must have the same length")
        us = [self._dev(u, (n + n * (n + 1),)) for u in us]
        vs = [self._dev(v, (n + n * (n + 1),)) for v in vs]
        flags = [0] * C if flags is None else list(flags)
        if out is None:
            out = t.zeros((self.N, n), dtype=t.float64, device=self.device)
            accumulate = False
        else:
            self._check_out(out, (self.N, n))
        if precision not in ("fp64", "fp32"):
            raise ValueError("the force adjoint runs in fp64 or fp32, not %r" % (precision,))
        L.ops().force_grad(self._h.value, L.PRECISIONS[precision], us, vs,
                           [int(f) for f in flags], out, bool(accumulate))
        return out
