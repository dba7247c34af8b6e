"""Drop-in for the reference's extension module ``sscha_HP_odd`` on the Lanczos path.

``GetPerturbAverageSym`` keeps the argument list and in-place output contract of
CModules/odd_corr_module.c:491-634 (format "OOOOOOdiiOOOOOO"): force_out is overwritten,
d2v_out is overwritten with the D3 part and the D4 part is added when apply_D4 is 1.
It runs on the GPU through libtdscha_hip (tdscha_get_perturb_average_sym). Bad input raises
ValueError instead of terminating the interpreter.
"""
import numpy as np

from . import _lib as L


def _check_f64(name, a, shape):
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("%s must be a C-contiguous float64 numpy array" % name)
    if a.shape != shape:
        raise ValueError("%s has shape %s, expected %s" % (name, a.shape, shape))


def GetPerturbAverageSym(X, Y, w, rho, R1, Y1, T, apply_D4, N_syms, sym_list, n_deg,
                         sym_basis_list, blocks_ids, force_out, d2v_out):
    """Perturbed averages <f>^(1), <d2V/dR2>^(1) over the symmetry-expanded ensemble."""
    X = np.asarray(X)
    if X.ndim != 2:
        raise ValueError("X must be (N_configs, N_modes)")
    N, n = X.shape
    if len(w) != n:
        raise ValueError("N_modes from X is %d, while len(w) = %d" % (n, len(w)))
    if apply_D4 not in (0, 1):
        raise ValueError("the apply_D4 value passed must be either 0 or 1 (got %r)" % (apply_D4,))
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    R1 = np.ascontiguousarray(R1, dtype=np.float64)
    Y1 = np.ascontiguousarray(Y1, dtype=np.float64)
    _check_f64("force_out", force_out, (n,))
    _check_f64("d2v_out", d2v_out, (n, n))
    syms = [np.ascontiguousarray(s, dtype=np.float64) for s in sym_list]
    degs = [np.ascontiguousarray(d, dtype=np.intc) for d in sym_basis_list]
    n_deg = np.ascontiguousarray(n_deg, dtype=np.intc)
    bids = np.ascontiguousarray(blocks_ids, dtype=np.intc)
    nb = len(syms)
    for b in range(nb):
        d = int(n_deg[b])
        if syms[b].shape != (int(N_syms), d, d) or degs[b].shape != (d,):
            raise ValueError("symmetry block %d: expected (%d, %d, %d) and (%d,)"
                             % (b, N_syms, d, d, d))
    psyms = (L._dp * nb)(*[s.ctypes.data_as(L._dp) for s in syms])
    pdegs = (L._ip * nb)(*[d.ctypes.data_as(L._ip) for d in degs])
    rc = L.lib().tdscha_get_perturb_average_sym(
        X.ctypes.data_as(L._dp), Y.ctypes.data_as(L._dp), w.ctypes.data_as(L._dp),
        rho.ctypes.data_as(L._dp), R1.ctypes.data_as(L._dp), Y1.ctypes.data_as(L._dp), float(T),
        int(apply_D4), int(N_syms), nb, psyms, n_deg.ctypes.data_as(L._ip), pdegs,
        bids.ctypes.data_as(L._ip), N, n, force_out.ctypes.data_as(L._dp),
        d2v_out.ctypes.data_as(L._dp))
    L.check(rc, "GetPerturbAverageSym")
    return None
