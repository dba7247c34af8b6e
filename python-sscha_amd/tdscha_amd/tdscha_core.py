"""Range-sliced perturbed averages with the interface of the reference's Julia level
(Modules/tdscha_core.jl, "JL" below), evaluated on the GPU through libtdscha_hip
(tdscha_jl_get_perturb_*_averages_sym).

The three functions take the arguments in the order, and with the meaning, that
Lanczos.apply_anharmonic_FT passes them in MODE_FAST_JULIA (Modules/DynamicalLanczos.py:
3065-3103): X and Y as (n_modes, n_configs) matrices (self.X.T, self.Y.T), the padded
symmetry array sym_julia (n_blocks, n_syms, d_max, d_max), deg_julia (n_blocks, d_max),
and a 1-based inclusive range [start_index, end_index] of bigindex = i n_syms + s + 1 over
configurations i and symmetries s (JL:71-73). Each returns the sum over the images in the
range divided by n_syms sum(rho) over all configurations, so the results over the split
of ``split_indices`` add up to the full average.

Deviations, on purpose: a bad range raises ValueError (the reference's @assert raises a
Julia AssertionError, JL:171-173), and R1 / Y1 may be None (zeros).
"""
import numpy as np

from . import _lib as L


def _f64(a, name, shape=None):
    a = np.asarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError("%s has shape %s, expected %s" % (name, a.shape, shape))
    return np.ascontiguousarray(a)


def _args(X, Y, omega, rho, R1, Y1, temperature, apply_v4, symmetries, n_degeneracies,
          degenerate_space, blocks, start_index, end_index):
    X = np.asarray(X)
    if X.ndim != 2:
        raise ValueError("X must be the (n_modes, n_configs) matrix self.X.T")
    n, N = X.shape
    # X.T of the (n_modes, n_configs) view is the (n_configs, n_modes) row-major array the
    # library reads (no copy when X is self.X.T of a C-contiguous self.X)
    Xc = _f64(X.T, "X.T")
    Yc = _f64(np.asarray(Y).T, "Y.T", (N, n))
    omega = _f64(omega, "omega", (n,))
    rho = _f64(rho, "rho", (N,))
    R1 = np.zeros(n) if R1 is None else _f64(R1, "R1", (n,))
    Y1 = np.zeros((n, n)) if Y1 is None else _f64(Y1, "Y1", (n, n))
    sym = np.asarray(symmetries, dtype=np.float64)
    if sym.ndim != 4 or sym.shape[2] != sym.shape[3]:
        raise ValueError("symmetries must be the (n_blocks, n_syms, d_max, d_max) array")
    nb, nsym, dmax, _ = sym.shape
    sym = np.ascontiguousarray(sym)
    ndeg = np.ascontiguousarray(n_degeneracies, dtype=np.intc)
    deg = np.ascontiguousarray(degenerate_space, dtype=np.intc)
    bids = np.ascontiguousarray(blocks, dtype=np.intc)
    if ndeg.shape != (nb,) or deg.shape != (nb, dmax) or bids.shape != (n,):
        raise ValueError("n_degeneracies %s, degenerate_space %s, blocks %s do not match "
                         "%d blocks of at most %d modes over %d modes"
                         % (ndeg.shape, deg.shape, bids.shape, nb, dmax, n))
    keep = (Xc, Yc, omega, rho, R1, Y1, sym, ndeg, deg, bids)
    c = (Xc.ctypes.data_as(L._dp), Yc.ctypes.data_as(L._dp), omega.ctypes.data_as(L._dp),
         rho.ctypes.data_as(L._dp), R1.ctypes.data_as(L._dp), Y1.ctypes.data_as(L._dp),
         float(temperature), int(bool(apply_v4)), sym.ctypes.data_as(L._dp),
         ndeg.ctypes.data_as(L._ip), deg.ctypes.data_as(L._ip), bids.ctypes.data_as(L._ip),
         int(start_index), int(end_index), n, N, nsym, nb, dmax)
    return keep, c, n


def _call(fn, c):
    rc = fn(*c)
    if rc != 0:
        msg = L.lib().tdscha_last_error().decode()
        if rc in (1, 2, 6):
            raise ValueError(msg)
        raise RuntimeError("%s failed (code %d): %s" % (fn.__name__, rc, msg))


def get_perturb_averages_sym(X, Y, omega, rho, R1, Y1, temperature, apply_v4, symmetries,
                             n_degeneracies, degenerate_space, blocks, start_index, end_index):
    """(f_average, d2v_dr2) over the images in [start_index, end_index] (JL:202-229)."""
    keep, c, n = _args(X, Y, omega, rho, R1, Y1, temperature, apply_v4, symmetries,
                       n_degeneracies, degenerate_space, blocks, start_index, end_index)
    f = np.zeros(n)
    d2v = np.zeros((n, n))
    _call(L.lib().tdscha_jl_get_perturb_averages_sym,
          c + (f.ctypes.data_as(L._dp), d2v.ctypes.data_as(L._dp)))
    del keep
    return f, d2v


def get_perturb_d2v_averages_sym(X, Y, omega, rho, R1, Y1, temperature, apply_v4, symmetries,
                                 n_degeneracies, degenerate_space, blocks, start_index,
                                 end_index):
    """d2v_dr2 over the images in [start_index, end_index] (JL:232-259)."""
    keep, c, n = _args(X, Y, omega, rho, R1, Y1, temperature, apply_v4, symmetries,
                       n_degeneracies, degenerate_space, blocks, start_index, end_index)
    d2v = np.zeros((n, n))
    _call(L.lib().tdscha_jl_get_perturb_d2v_averages_sym, c + (d2v.ctypes.data_as(L._dp),))
    del keep
    return d2v


def get_perturb_f_averages_sym(X, Y, omega, rho, R1, Y1, temperature, apply_v4, symmetries,
                               n_degeneracies, degenerate_space, blocks, start_index, end_index):
    """f_average over the images in [start_index, end_index] (JL:262-284)."""
    keep, c, n = _args(X, Y, omega, rho, R1, Y1, temperature, apply_v4, symmetries,
                       n_degeneracies, degenerate_space, blocks, start_index, end_index)
    f = np.zeros(n)
    _call(L.lib().tdscha_jl_get_perturb_f_averages_sym, c + (f.ctypes.data_as(L._dp),))
    del keep
    return f


def split_indices(n_total, n_processors):
    """The [start, end] image range of every rank, as Lanczos.apply_anharmonic_FT divides
    n_syms * N over the processors (Modules/DynamicalLanczos.py:3083-3099): count =
    n_total // P each, the first n_total % P ranks one more."""
    if n_processors < 1:
        raise ValueError("n_processors must be positive")
    count, rem = divmod(int(n_total), int(n_processors))
    out = []
    for rank in range(int(n_processors)):
        start = rank * (count + 1) if rank < rem else rank * count + rem
        stop = start + count + (1 if rank < rem else 0)
        out.append([start + 1, stop])
    return out


def drop_in_cache(enable):
    """Keep (1) or release and stop keeping (0) the device ensemble of the last host-pointer
    call (tdscha_drop_in_cache); returns the previous setting."""
    return bool(L.lib().tdscha_drop_in_cache(int(bool(enable))))
