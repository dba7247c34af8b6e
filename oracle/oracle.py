"""CPU oracle for the TD-SCHA ensemble L-application (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker. The product
library never routes through it.

Contents
--------
* ctypes wrappers around ``liboracle.so`` (this project's C restatement of the
  per-image kernel, ``tdscha_oracle.c``) and, when it has been built, around
  ``_ref/libref_lanczos.so`` (the reference's own C trio compiled from its
  source by ``oracle/Makefile``).
* A numpy restatement of the Python operator layer that wraps the kernel in
  the reference ``Modules/DynamicalLanczos.py`` ("Py" below):

  - Bose occupation / f_ups                 Py:144-163
  - psi <-> symmetric matrices              Py:2827-2888 (get_Y1, get_ReA1)
  - Wigner change of variables              Py:2572-2627, Py:2713-2822
  - harmonic part apply_L1_FT               Py:2227-2385
  - anharmonic pre/post transforms          Py:2891-3281 (apply_anharmonic_FT)
  - apply_full_L (reverse / shift)          Py:3467-3570
  - Wigner dot-product mask                 Py:5790-5839
  - biconjugate / symmetric Lanczos run_FT  Py:5843-6372
  - continued-fraction Green function       Py:4951-5047
  - ensemble projection on the modes        Py:380-442 (``project_ensemble``)
  - symmetry operations on mode blocks      Py:735-746 (``symmetries_on_modes``)
  - ensemble averages of effective charges and Raman tensors, Modules/Perturbations.py
    :195-375 (``d1M_dR_av``, ``d1alpha_dR_av``, ``d2M_dR_av``, ``d2alpha_dR_av``)
* Two-phonon (Wigner) perturbation sectors, Py:1488-1526 (``two_phonon_sectors``,
  ``pack_two_phonon``), and the space-group symmetrization of the ensemble averages
  (``symmetrize_average``, the role of Modules/Perturbations.py:15-193). The reference does the
  latter with CellConstructor (absent here): its exact form (the acoustic sum rule in
  particular) is parity unpinned; the restatement is pinned on invariance under every
  operation, idempotence, the sum rule and group-invariant model tensors.
* The Julia level's range-sliced averages over the flattened (configuration, symmetry) index
  (``perturb_averages_range``, Modules/tdscha_core.jl:54-284) and the reference's split of
  that index over ranks (``julia_split``, Py:3083-3099). Pinned: the full range equals
  ``perturb_average_sym``; the split's pieces add up to it; a single image equals the
  reference's own C build on that one configuration and symmetry.
* The translation average of a psi vector (``translation_average_psi``): the projector on
  translation-invariant perturbations, restated with explicit loops over the translations'
  atom permutations; the checker of the zone-centre reduction of the image sum.
* The derivative of a linear functional of the kernel outputs with respect to the
  configuration forces Y (``force_grad``), image by image in the reference's own form
  (u = S x_i, f = S y_i, C:2158-2170), and the adjoint seeds of the post-transform
  (``post_seeds``). Pinned by finite differences of ``perturb_average_sym`` and of
  ``apply_anharmonic_FT``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

RY_TO_K = 157887.32400374097   # Py:134, CModules/LanczosFunctions.h:9
EPSILON = 1e-12                # Py:121

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE_SO = os.path.join(_HERE, "liboracle.so")
_REF_SO = os.path.join(_HERE, "_ref", "libref_lanczos.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_dpp = ctypes.POINTER(_dp)
_ipp = ctypes.POINTER(_ip)


# --------------------------------------------------------------------------
# Symmetry description (the reference binding's sym_list / n_deg /
# sym_basis_list / blocks_ids, odd_corr_module.c:519-601)
# --------------------------------------------------------------------------
@dataclass
class SymBlocks:
    """Mode-basis symmetry blocks.

    symmetries[b]       : (N_sym, d_b, d_b) float64, row-major
    N_degeneracy[b]     : d_b
    degenerate_space[b] : (d_b,) int32 mode ids, contiguous and sorted
    block_id[mu]        : block of mode mu
    """
    symmetries: list
    N_degeneracy: np.ndarray
    degenerate_space: list
    block_id: np.ndarray
    n_syms: int = field(init=False)

    def __post_init__(self):
        self.symmetries = [np.ascontiguousarray(s, dtype=np.float64) for s in self.symmetries]
        self.N_degeneracy = np.ascontiguousarray(self.N_degeneracy, dtype=np.intc)
        self.degenerate_space = [np.ascontiguousarray(d, dtype=np.intc) for d in self.degenerate_space]
        self.block_id = np.ascontiguousarray(self.block_id, dtype=np.intc)
        self.n_syms = int(self.symmetries[0].shape[0])

    @property
    def n_blocks(self):
        return len(self.symmetries)

    def subset(self, n_sym):
        """The first n_sym symmetries (a bounded sample of the image set)."""
        return SymBlocks([s[:n_sym] for s in self.symmetries], self.N_degeneracy,
                         self.degenerate_space, self.block_id)

    def dense(self, s):
        """Dense n x n matrix of symmetry s."""
        n = len(self.block_id)
        out = np.zeros((n, n))
        for b, sb in enumerate(self.symmetries):
            idx = self.degenerate_space[b]
            out[np.ix_(idx, idx)] = sb[s]
        return out

    @staticmethod
    def identity(n):
        """No symmetries: one 1x1 identity block per mode (Py:721-726)."""
        return SymBlocks([np.ones((1, 1, 1))] * n, np.ones(n, dtype=np.intc),
                         [np.array([i], dtype=np.intc) for i in range(n)],
                         np.arange(n, dtype=np.intc))


class _PtrArrays:
    """double** / int** views kept alive for the duration of a C call."""

    def __init__(self, sym: SymBlocks):
        self._syms = sym.symmetries
        self._degs = sym.degenerate_space
        nb = sym.n_blocks
        self.syms = (_dp * nb)(*[s.ctypes.data_as(_dp) for s in self._syms])
        self.degs = (_ip * nb)(*[d.ctypes.data_as(_ip) for d in self._degs])
        self.ndeg = sym.N_degeneracy
        self.bid = sym.block_id


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


_lib_oracle = None
_lib_ref = None


def _load_oracle():
    global _lib_oracle
    if _lib_oracle is None:
        if not os.path.exists(_ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        lib = ctypes.CDLL(_ORACLE_SO)
        lib.oracle_perturb_average_sym_range.restype = ctypes.c_int
        lib.oracle_perturb_average_sym_range.argtypes = [
            _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, _dpp, _ip, _ipp, _ip, ctypes.c_int, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
            ctypes.c_int, _dp, _dp, _dp, _dp]
        lib.oracle_f_ups.restype = ctypes.c_double
        lib.oracle_f_ups.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.oracle_f_psi.restype = ctypes.c_double
        lib.oracle_f_psi.argtypes = [ctypes.c_double, ctypes.c_double]
        _lib_oracle = lib
    return _lib_oracle


def ref_available():
    return os.path.exists(_REF_SO)


def _load_ref():
    global _lib_ref
    if _lib_ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref/libref_lanczos.so is not built")
        lib = ctypes.CDLL(_REF_SO)
        argt = [_dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dpp,
                ctypes.c_int, _ip, _ipp, _ip, _dp]
        for name in ("get_f_average_from_Y_pert_sym_fast", "get_d2v_dR2_from_R_pert_sym_fast",
                     "get_d2v_dR2_from_Y_pert_sym_fast"):
            fn = getattr(lib, name)
            fn.restype = None
            fn.argtypes = argt
        _lib_ref = lib
    return _lib_ref


F_TERM, D2V_R_TERM, D2V_Y_TERM = 1, 2, 4


def perturb_average_sym(X, Y, w, rho, R1, Y1, T, apply_d4, sym: SymBlocks, n_threads=None,
                        cfg_range=None, sym_range=None, what=None, norm=None, with_abs=False):
    """Oracle (f_av, d2v) with the output contract of GetPerturbAverageSym.

    cfg_range / sym_range restrict the images to a slice; `norm` overrides the
    (N_eff, N_sym) normalisation pair (default: full-ensemble values).
    with_abs=True also returns (f_abs, d2v_abs), the same sums taken over the
    absolute values of every term (the rounding-error scale).
    """
    lib = _load_oracle()
    X, Y, w, rho, R1, Y1 = (_c(a) for a in (X, Y, w, rho, R1, Y1))
    N, n = X.shape
    if n_threads is None:
        n_threads = max(1, min(8, os.cpu_count() or 1))
    c0, c1 = cfg_range if cfg_range is not None else (0, N)
    s0, s1 = sym_range if sym_range is not None else (0, sym.n_syms)
    if what is None:
        what = F_TERM | D2V_R_TERM | (D2V_Y_TERM if apply_d4 else 0)
    neff, nsym = norm if norm is not None else (float(np.sum(rho)), float(sym.n_syms))
    p = _PtrArrays(sym)
    f = np.zeros(n)
    d2v = np.zeros((n, n))
    fa = np.zeros(n) if with_abs else None
    da = np.zeros((n, n)) if with_abs else None
    rc = lib.oracle_perturb_average_sym_range(
        X.ctypes.data_as(_dp), Y.ctypes.data_as(_dp), w.ctypes.data_as(_dp),
        rho.ctypes.data_as(_dp), R1.ctypes.data_as(_dp), Y1.ctypes.data_as(_dp), float(T), n, N,
        sym.n_syms, sym.n_blocks, p.syms, p.ndeg.ctypes.data_as(_ip), p.degs,
        p.bid.ctypes.data_as(_ip), int(what), int(c0), int(c1), int(s0), int(s1), float(neff),
        float(nsym), int(n_threads), f.ctypes.data_as(_dp), d2v.ctypes.data_as(_dp),
        fa.ctypes.data_as(_dp) if with_abs else None, da.ctypes.data_as(_dp) if with_abs else None)
    if rc != 0:
        raise ValueError("oracle rejected its arguments (code %d)" % rc)
    if with_abs:
        return f, d2v, fa, da
    return f, d2v


def ref_perturb_average_sym(X, Y, w, rho, R1, Y1, T, apply_d4, sym: SymBlocks,
                            n_configs=None, n_syms=None, quiet=True):
    """The reference's C trio, called in the order of odd_corr_module.c:613-627.

    n_configs / n_syms take the first configs / symmetries only (the arrays are
    laid out config-major and symmetry-major, so a prefix is a valid input).
    With quiet=True the per-call printf of LanczosFunctions.c:2139 is sent to
    /dev/null.
    """
    lib = _load_ref()
    X, Y, w, rho = (_c(a) for a in (X, Y, w, rho))
    R1 = _c(R1).copy()
    Y1 = _c(Y1).copy()
    N, n = X.shape
    if n_configs is not None:
        N = int(n_configs)
    nsym = sym.n_syms if n_syms is None else int(n_syms)
    p = _PtrArrays(sym)
    f = np.zeros(n)
    d2v = np.zeros((n, n))
    args_common = (X.ctypes.data_as(_dp), Y.ctypes.data_as(_dp), w.ctypes.data_as(_dp))
    tail = (float(T), n, N, rho.ctypes.data_as(_dp), p.syms, nsym, p.ndeg.ctypes.data_as(_ip),
            p.degs, p.bid.ctypes.data_as(_ip))
    saved = None
    if quiet:
        import sys
        sys.stdout.flush()
        saved = os.dup(1)
        devnull = os.open(os.devnull, os.O_WRONLY)
        os.dup2(devnull, 1)
        os.close(devnull)
    try:
        lib.get_f_average_from_Y_pert_sym_fast(*args_common, Y1.ctypes.data_as(_dp), *tail,
                                               f.ctypes.data_as(_dp))
        lib.get_d2v_dR2_from_R_pert_sym_fast(*args_common, R1.ctypes.data_as(_dp), *tail,
                                             d2v.ctypes.data_as(_dp))
        if apply_d4:
            lib.get_d2v_dR2_from_Y_pert_sym_fast(*args_common, Y1.ctypes.data_as(_dp), *tail,
                                                 d2v.ctypes.data_as(_dp))
    finally:
        if saved is not None:
            import ctypes.util
            libc = ctypes.CDLL(ctypes.util.find_library("c"))
            libc.fflush(None)
            os.dup2(saved, 1)
            os.close(saved)
    return f, d2v


def image_range_pieces(start, end, n_syms):
    """The images bigindex in [start, end] (1-based, inclusive) as at most three rectangles
    ((c0, c1), (s0, s1)) of configurations x symmetries: bigindex = i n_syms + s + 1, i.e.
    configuration (bigindex - 1) // n_syms and symmetry (bigindex - 1) % n_syms
    (Modules/tdscha_core.jl:71-73, :124-126, :175-177)."""
    lo, hi = start - 1, end - 1
    i0, j0 = divmod(lo, n_syms)
    i1, j1 = divmod(hi, n_syms)
    if i0 == i1:
        return [((i0, i0 + 1), (j0, j1 + 1))]
    pieces = [((i0, i0 + 1), (j0, n_syms))]
    if i1 > i0 + 1:
        pieces.append(((i0 + 1, i1), (0, n_syms)))
    pieces.append(((i1, i1 + 1), (0, j1 + 1)))
    return pieces


def perturb_averages_range(X, Y, w, rho, R1, Y1, T, apply_v4, sym: SymBlocks, start, end,
                           what="both"):
    """The Julia level's range-sliced averages (Modules/tdscha_core.jl, "JL"):
    what = "both"  get_perturb_averages_sym      (JL:202-229) -> (f, d2v)
           "d2v"   get_perturb_d2v_averages_sym  (JL:232-259) -> d2v
           "f"     get_perturb_f_averages_sym    (JL:262-284) -> f
    over the images with bigindex in [start, end], each term as JL:54-198 (the same per-image
    arithmetic as the C trio), divided by n_syms sum(rho) over every configuration (JL:98,
    JL:150, JL:197). The force asserts 1 <= start <= end <= N n_syms (JL:171-173); the d2v
    loops are simply empty when start > end."""
    N = np.asarray(X).shape[0]
    n = np.asarray(X).shape[1]
    nsym = sym.n_syms
    want_f = what in ("both", "f")
    want_d = what in ("both", "d2v")
    if want_f and not (1 <= start <= end <= N * nsym):
        raise AssertionError("image range [%d, %d] outside [1, %d]" % (start, end, N * nsym))
    if start > end:
        return np.zeros((n, n))
    if not (1 <= start and end <= N * nsym):
        raise AssertionError("image range [%d, %d] outside [1, %d]" % (start, end, N * nsym))
    terms = (F_TERM if want_f else 0) | ((D2V_R_TERM | (D2V_Y_TERM if apply_v4 else 0))
                                        if want_d else 0)
    norm = (float(np.sum(rho)), float(nsym))
    f = np.zeros(n)
    d2v = np.zeros((n, n))
    for cr, sr in image_range_pieces(start, end, nsym):
        fp, dp = perturb_average_sym(X, Y, w, rho, R1, Y1, T, apply_v4, sym, cfg_range=cr,
                                     sym_range=sr, what=terms, norm=norm)
        f += fp
        d2v += dp
    if what == "both":
        return f, d2v
    return d2v if what == "d2v" else f


def julia_split(n_total, n_processors):
    """The [start, end] ranges of every rank (Modules/DynamicalLanczos.py:3083-3099): the
    first n_total % P ranks take one image more."""
    count, rem = divmod(n_total, n_processors)
    out = []
    for rank in range(n_processors):
        if rank < rem:
            start = rank * (count + 1)
            stop = start + count + 1
        else:
            start = rank * count + rem
            stop = start + count
        out.append((start + 1, stop))
    return out


# --------------------------------------------------------------------------
# Python operator layer (restated from Modules/DynamicalLanczos.py)
# --------------------------------------------------------------------------
def bose_occupation(w, T):
    """Py:157-163 (threshold T > 0)."""
    w = np.asarray(w, dtype=np.float64)
    if T > 0:
        return 1.0 / (np.exp(w * RY_TO_K / T) - 1.0)
    return np.zeros_like(w)


def f_ups(w, T):
    """Py:144-155: 2w / (1 + 2n)."""
    return 2 * np.asarray(w) / (1 + 2 * bose_occupation(w, T))


def pair_indices(n):
    """Independent (row b, col a >= b) pairs in psi order (Py:2280-2286).

    Returns (i_row, i_col): the upper triangle, row-major.
    """
    iu = np.triu_indices(n)
    return iu[0], iu[1]


def psi_len(n):
    return n + n * (n + 1)


def unpack_sym(vec, n, half_off_diagonal=False):
    """Upper-triangle row-major vector -> symmetric matrix (Py:2827-2858)."""
    M = np.zeros((n, n))
    iu = np.triu_indices(n)
    M[iu] = vec
    M = M + np.triu(M, 1).T
    if half_off_diagonal:
        M = M * 0.5
        M[np.diag_indices(n)] *= 2
    return M


def pack_upper(M):
    """Matrix -> upper-triangle row-major vector (Py:3250-3262)."""
    return np.asarray(M)[np.triu_indices(M.shape[0])].copy()


def _pair_freqs(w, T):
    """w_a, w_b, n_a, n_b per independent pair as apply_L1_FT builds them (Py:2280-2301)."""
    i_b, i_a = pair_indices(len(w))
    w_a, w_b = w[i_a], w[i_b]
    if T > 0:
        n_a = 1 / (np.exp(w_a / np.double(T / RY_TO_K)) - 1)
        n_b = 1 / (np.exp(w_b / np.double(T / RY_TO_K)) - 1)
    else:
        n_a = np.zeros_like(w_a)
        n_b = np.zeros_like(w_b)
    return w_a, w_b, n_a, n_b


def apply_L1_FT(psi, w, T, transpose=False, use_wigner=False, ignore_harmonic=False):
    """Harmonic propagator, Py:2227-2385."""
    n = len(w)
    out = np.zeros_like(psi)
    if ignore_harmonic:
        return out
    w_a, w_b, n_a, n_b = _pair_freqs(w, T)
    nw2 = len(w_a)
    sY, sA = n, n + nw2
    assert len(psi) == sA + nw2
    if not use_wigner:
        out[:n] = (psi[:n] * w) * w
        X_ab = -w_a**2 - w_b**2 - (2 * w_a * w_b) / ((2 * n_a + 1) * (2 * n_b + 1))
        out[sY:sA] = -X_ab * psi[sY:sA]
        Y_ab = -(8 * w_a * w_b) / ((2 * n_a + 1) * (2 * n_b + 1))
        if not transpose:
            out[sY:sA] += -Y_ab * psi[sA:]
        else:
            out[sA:] += -Y_ab * psi[sY:sA]
        X1_ab = -(2 * n_a * n_b + n_a + n_b) * (2 * n_a * n_b + n_a + n_b + 1) * (2 * w_a * w_b) / (
            (2 * n_a + 1) * (2 * n_b + 1))
        if not transpose:
            out[sA:] += -X1_ab * psi[sY:sA]
        else:
            out[sY:sA] += -X1_ab * psi[sA:]
        Y1_ab = -w_a**2 - w_b**2 + (2 * w_a * w_b) / ((2 * n_a + 1) * (2 * n_b + 1))
        out[sA:] += -Y1_ab * psi[sA:]
    else:
        out[:n] = -(psi[:n] * w) * w
        out[sY:sA] = -(w_a**2 + w_b**2 - 2.0 * w_a * w_b) * psi[sY:sA]
        out[sA:] = -(w_a**2 + w_b**2 + 2.0 * w_a * w_b) * psi[sA:]
    return out


def chi_minus_plus(w, T):
    """Py:2572-2627 (threshold T > EPSILON)."""
    n = len(w)
    W = np.tile(w, (n, 1))
    nn = np.zeros((n, n))
    if T > EPSILON:
        nn = 1.0 / (np.exp(W * RY_TO_K / T) - 1.0)
    chi_m = (W - W.T) * (nn - nn.T) / (2.0 * W * W.T)
    chi_p = (W + W.T) * (1 + nn + nn.T) / (2.0 * W * W.T)
    return chi_m, chi_p


def wigner_alpha1(psi, w, T):
    """Upsilon^(1) = alpha^(1) from (a', b'), Py:2713-2800 (get_alpha=True)."""
    n = len(w)
    nw2 = n * (n + 1) // 2
    a_all = psi[n:n + nw2]
    b_all = psi[n + nw2:]
    w_a, w_b, n_a, n_b = _pair_freqs(w, T)
    Xf = (1 + 2 * n_a) * (1 + 2 * n_b) / 8
    w2_on_X = (w_a * w_b) / Xf
    chi_m = ((w_a - w_b) * (n_a - n_b)) / (2 * w_a * w_b)
    chi_p = ((w_a + w_b) * (1 + n_a + n_b)) / (2 * w_a * w_b)
    new_a = w2_on_X * np.sqrt(-0.5 * chi_m) * a_all
    new_b = w2_on_X * np.sqrt(+0.5 * chi_p) * b_all
    return unpack_sym(new_a - new_b, n)


def anharmonic_pre(psi, w, T, transpose=False, use_wigner=False):
    """Py:2946-2993: the (R1, Y1) handed to the ensemble kernel."""
    n = len(w)
    nw2 = n * (n + 1) // 2
    R1 = psi[:n].copy()
    if use_wigner:
        return R1, wigner_alpha1(psi, w, T)
    Y1 = unpack_sym(psi[n:n + nw2], n, half_off_diagonal=transpose)
    if transpose:
        n_mu = bose_occupation(w, T) if T > EPSILON else np.zeros(n)
        Y_w = 2 * w / (2 * n_mu + 1)
        ReA_w = 2 * w * n_mu * (n_mu + 1) / (2 * n_mu + 1)
        ReA1 = unpack_sym(psi[n + nw2:], n, half_off_diagonal=True)
        coeff_Y = 2 * (np.einsum("a,b,b->ab", Y_w, Y_w, Y_w) + np.einsum("a,a,b->ab", Y_w, Y_w, Y_w))
        coeff_RA = 2 * (np.einsum("a,b,b->ab", Y_w, ReA_w, Y_w) + np.einsum("a,a,b->ab", Y_w, ReA_w, Y_w))
        Y1 = Y1 * coeff_Y + ReA1 * coeff_RA
    return R1, Y1


def anharmonic_post(f_av, d2v, w, T, transpose=False, use_wigner=False):
    """Py:3200-3281: kernel outputs -> psi' (sign included)."""
    n = len(w)
    if use_wigner:
        chi_m, chi_p = chi_minus_plus(w, T)
        pert_Y = np.sqrt(-0.5 * chi_m) * d2v
        pert_RA = -np.sqrt(0.5 * chi_p) * d2v
        sign = 1.0
    else:
        n_mu = 1.0 / (np.exp(w * RY_TO_K / T) - 1.0) if T > EPSILON else np.zeros(n)
        Y_w = 2 * w / (2 * n_mu + 1)
        ReA_w = 2 * w * n_mu * (n_mu + 1) / (2 * n_mu + 1)
        if not transpose:
            pert_Y = d2v * Y_w[:, None] + d2v * Y_w[None, :]
            pert_RA = d2v * ReA_w[:, None] + d2v * ReA_w[None, :]
        else:
            Y_inv = 1 / Y_w
            pert_Y = 0.5 * np.einsum("a,ab,b->ab", Y_inv, d2v, Y_inv)
            pert_RA = np.zeros_like(pert_Y)
            mask = np.full(pert_Y.shape, 2.0)
            np.fill_diagonal(mask, 1)
            pert_Y *= mask
            pert_RA *= mask
        sign = -1.0
    out = np.concatenate([f_av, pack_upper(pert_Y), pack_upper(pert_RA)])
    return sign * out


def f_psi(w, T):
    """C:24-30: (1 + 2n) / (2w)."""
    return (1 + 2 * bose_occupation(w, T)) / (2 * np.asarray(w))


def force_grad(X, Y, w, rho, R1, Y1, T, sym: SymBlocks, uR, UD, kf=1.0, kr=1.0, ky=1.0,
               norm=None):
    """dJ/dY (N x n) for J = uR . f_av + <UD, d2v>.

    f_av, d2v are the outputs of perturb_average_sym with the per-term switches of
    C:2082-2754 (kf: the f_av terms, kr: the R1 terms of d2v, ky: the Y1 terms of d2v).
    Per image, with u = S x, f = S y, ut = f_ups u, ft = f_psi f, q = Y1s u:
      dJ/df = [e1 uR + (-kf rho/3)(uR.ut) f_psi q + 2 c1 U ut
               + rho (ut^T U ut)(ky/4 f_psi q - kr/3 R1)] / (N_eff N_sym)
    and dJ/dy = S^T dJ/df; U is the symmetric part of UD (d2v is symmetric).
    """
    X, Y, R1 = (np.asarray(a, dtype=np.float64) for a in (X, Y, R1))
    N, n = X.shape
    fu, fp = f_ups(w, T), f_psi(w, T)
    Y1s = 0.5 * (np.asarray(Y1) + np.asarray(Y1).T)
    U = 0.5 * (np.asarray(UD) + np.asarray(UD).T)
    uR = np.asarray(uR, dtype=np.float64)
    rho = np.asarray(rho, dtype=np.float64)
    neff, nsym = norm if norm is not None else (float(np.sum(rho)), float(sym.n_syms))
    G = np.zeros((N, n))
    for s in range(sym.n_syms):
        S = sym.dense(s)
        u = X @ S.T
        ut = u * fu
        q = u @ Y1s
        alpha = np.einsum("ia,ia->i", u, q)
        gamma = ut @ R1
        e1 = -kf * rho * alpha / 6
        c1 = rho * (ky * alpha / 8 - kr * gamma / 3)
        nu = ut @ uR
        Uut = ut @ U
        mu = np.einsum("ia,ia->i", ut, Uut)
        g = (e1[:, None] * uR[None, :] + 2 * c1[:, None] * Uut
             + (rho * (-kf * nu / 3 + ky * mu / 4))[:, None] * (fp[None, :] * q)
             - (kr * rho * mu / 3)[:, None] * R1[None, :])
        G += g @ S
    return G / (neff * nsym)


def post_seeds(u, w, T, transpose=False, use_wigner=False, reverse_L=False):
    """(uR, UD) with u . anharmonic_post(f, d2v) = uR . f + <UD, d2v> for every f and
    symmetric d2v, found by applying anharmonic_post to unit inputs (n + n(n+1)/2 calls)."""
    n = len(w)
    sign = -1.0 if reverse_L else 1.0
    uR = np.zeros(n)
    UD = np.zeros((n, n))
    zero_d = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        uR[j] = sign * u.dot(anharmonic_post(e, zero_d, w, T, transpose, use_wigner))
    zero_f = np.zeros(n)
    for i in range(n):
        for j in range(i, n):
            D = np.zeros((n, n))
            D[i, j] = D[j, i] = 1.0
            v = sign * u.dot(anharmonic_post(zero_f, D, w, T, transpose, use_wigner))
            if i == j:
                UD[i, i] = v
            else:
                UD[i, j] = UD[j, i] = 0.5 * v
    return uR, UD


@dataclass
class Ensemble:
    """The projected ensemble the operator is built on (Py:380-442)."""
    X: np.ndarray
    Y: np.ndarray
    w: np.ndarray
    rho: np.ndarray
    T: float
    sym: SymBlocks


def apply_anharmonic_FT(ens: Ensemble, psi, transpose=False, use_wigner=False, ignore_v4=False,
                        kernel=None):
    """Py:2891-3281. `kernel(R1, Y1, apply_d4) -> (f, d2v)` defaults to the C oracle."""
    R1, Y1 = anharmonic_pre(psi, ens.w, ens.T, transpose, use_wigner)
    apply_d4 = 0 if ignore_v4 else 1
    if kernel is None:
        f_av, d2v = perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, apply_d4, ens.sym)
    else:
        f_av, d2v = kernel(R1, Y1, apply_d4)
    return anharmonic_post(f_av, d2v, ens.w, ens.T, transpose, use_wigner)


def apply_full_L(ens: Ensemble, psi, transpose=False, use_wigner=False, ignore_v3=False,
                 ignore_v4=False, ignore_harmonic=False, reverse_L=False, shift_value=0.0,
                 kernel=None):
    """Py:3467-3570 (finite-temperature branch, force_FT=True default)."""
    out = apply_L1_FT(psi, ens.w, ens.T, transpose, use_wigner, ignore_harmonic)
    if not ignore_v3:
        out = out + apply_anharmonic_FT(ens, psi, transpose, use_wigner, ignore_v4, kernel)
    if reverse_L:
        out *= -1
    out += shift_value * psi
    return out


def mask_dot_wigner(n):
    """Py:5790-5839: weight 2 on the off-diagonal pair entries."""
    mask = np.ones(psi_len(n))
    i_b, i_a = pair_indices(n)
    nw2 = len(i_a)
    off = i_b < i_a
    mask[n:n + nw2][off] = 2
    mask[n + nw2:][off] = 2
    return mask


@dataclass
class LanczosState:
    a: list = field(default_factory=list)
    b: list = field(default_factory=list)
    c: list = field(default_factory=list)
    Q: list = field(default_factory=list)
    P: list = field(default_factory=list)
    s_norm: list = field(default_factory=list)


def biorthogonal_gs(new_q, new_p, Q, P, s_norm, c, dot, passes=2):
    """Two-sided classical Gram-Schmidt, applied `passes` times, of a new Lanczos pair against
    every stored pair (Q_j, P_j), followed by unit normalisation.

    The duals of the stored left vectors are pt_j = P_j * s_norm_j / c_(j-1) (c_(-1) = 1):
    the recurrence of Py:6059-6113 keeps pt_j . q_k = delta_jk. Each pass removes from q its
    components along Q_j measured by the duals, and from p its components along the duals
    measured by Q_j, all coefficients of a pass taken from the same vectors:
        q <- q - sum_j (pt_j . q) Q_j,    p <- p - sum_j (Q_j . p) pt_j
    Returns (q / |q|, p / |p|, |q|, |p|); the caller sets the new s_norm = c_k / (p . q) so
    that the new dual is p * s_norm / c_k.

    No counterpart in the reference: its loop (Py:6245-6317, the n_rep_orth branch of
    run_FT below) subtracts (P_j . q) P_j from q and (Q_j . p) Q_j from p (the left vectors
    along the right ones and vice versa, without the dual scale), normalises p by p . p and
    never grows its ortho_p test; no fixture of the reference exercises it."""
    q = new_q.copy()
    p = new_p.copy()
    pts = [P[j] * (s_norm[j] / (c[j - 1] if j >= 1 else 1.0)) for j in range(len(P))]
    for _ in range(passes):
        hq = [dot(pt, q) for pt in pts]
        hp = [dot(Q[j], p) for j in range(len(Q))]
        for j in range(len(Q)):
            q = q - hq[j] * Q[j]
            p = p - hp[j] * pts[j]
    nq = np.sqrt(dot(q, q))
    nP = np.sqrt(dot(p, p))
    return q / nq, p / nP, nq, nP


def run_FT(matvec, rmatvec, psi0, n_iter, run_simm=False, mask=None, n_rep_orth=0, n_ortho=10,
           state: LanczosState | None = None, full_gs=False):
    """Biconjugate / symmetric Lanczos of Py:5843-6372 (verbatim recurrence).

    matvec(q) = L q, rmatvec(p) = L^T p. Returns the LanczosState.

    full_gs=True re-biorthogonalises every new pair against all stored pairs with
    ``biorthogonal_gs`` (two passes). Pinned by mathematics: the duals and the basis stay
    biorthogonal (P~^T Q = I), the tridiagonal matrix of a complete Krylov space has the
    eigenvalues of the operator, and the first steps equal the reference's own
    unreorthogonalised run (where that run is still exact).

    n_rep_orth > 0 follows the reference's Gram-Schmidt block Py:6245-6317 as written
    (including its normalisation of p by p.p and its ortho_p accumulator). That branch is a
    documented quirk of the reference and is PARITY UNPINNED: no reference fixture uses it,
    and its p normalisation breaks the recurrence's own biorthogonality.
    """
    st = state if state is not None else LanczosState()
    dot = (lambda u, v: u.dot(v * mask)) if run_simm else (lambda u, v: u.dot(v))
    if len(st.a) == 0:
        first = psi0 / np.sqrt(dot(psi0, psi0))
        st.Q, st.P, st.s_norm = [first], [first], [1.0]
    psi_q, psi_p = st.Q[-1], st.P[-1]
    next_converged = False
    for _ in range(n_iter):
        L_q = matvec(psi_q)
        p_L = L_q.copy() if run_simm else rmatvec(psi_p)
        c_old = st.c[-1] if len(st.c) > 0 else 1
        p_norm = st.s_norm[-1] / c_old
        a_coeff = dot(psi_p, L_q) * p_norm
        if np.isnan(a_coeff):
            raise ValueError("Invalid value encountered during the Lanczos.")
        rk = L_q - a_coeff * psi_q
        if len(st.Q) > 1:
            rk -= st.c[-1] * st.Q[-2]
        sk = p_L - a_coeff * psi_p
        if len(st.P) > 1:
            old_p_norm = st.s_norm[-2] if len(st.c) < 2 else st.s_norm[-2] / st.c[-2]
            sk -= st.b[-1] * st.P[-2] * (old_p_norm / p_norm)
        s_norm = np.sqrt(dot(sk, sk))
        sk_tilde = sk / s_norm
        s_norm *= p_norm
        b_coeff = np.sqrt(dot(rk, rk))
        c_coeff = dot(sk_tilde, rk / b_coeff) * s_norm
        st.a.append(a_coeff)
        if np.abs(b_coeff) < EPSILON or next_converged:
            break
        if np.abs(c_coeff) < EPSILON:
            break
        new_q = rk / b_coeff
        new_p = sk_tilde.copy()
        if full_gs:
            new_q, new_p, nq, nP = biorthogonal_gs(new_q, new_p, st.Q, st.P, st.s_norm, st.c,
                                                   dot)
            if nq < EPSILON or nP < EPSILON:
                next_converged = True
            s_norm = c_coeff / dot(new_p, new_q)
        for _k in range(n_rep_orth):
            ortho_q = 0.0
            ortho_p = 0.0
            start = 0 if n_ortho is None else max(0, len(st.P) - n_ortho)
            for j in range(start, len(st.P)):
                coeff1 = dot(st.P[j], new_q)
                coeff2 = dot(st.Q[j], new_p)
                new_q = new_q - coeff1 * st.P[j]
                new_p = new_p - coeff2 * st.Q[j]
                ortho_q += np.abs(coeff1)
                ortho_p += np.abs(ortho_p)
            normq = np.sqrt(dot(new_q, new_q))
            if normq < EPSILON:
                next_converged = True
            new_q = new_q / normq
            normp = dot(new_p, new_p)
            if np.abs(normp) < EPSILON:
                next_converged = True
            new_p = new_p / normp
            s_norm = c_coeff / dot(new_p, new_q)
            if ortho_p < EPSILON and ortho_q < EPSILON:
                break
        st.Q.append(new_q)
        st.P.append(new_p)
        psi_q, psi_p = new_q.copy(), new_p.copy()
        st.b.append(b_coeff)
        st.c.append(c_coeff)
        st.s_norm.append(s_norm)
    return st


def green_function_continued_fraction(a, b, c, w_array, use_terminator=True, last_average=1,
                                      smearing=0.0, use_wigner=False, reverse_L=False,
                                      shift_value=0.0, perturbation_modulus=1.0):
    """Py:4951-5047."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    c = np.asarray(c, dtype=np.float64)
    w_array = np.asarray(w_array, dtype=np.float64)
    n_iters = len(a)
    sign = -1 if reverse_L else 1
    gf = np.zeros(w_array.shape, dtype=np.complex128)
    if use_terminator:
        a_av = np.mean(a[-last_average:])
        b_av = np.mean(b[-last_average:])
        c_av = b_av
        if len(c) == len(b):
            c_av = np.mean(c[-last_average:])
        aa = a_av * sign - sign * shift_value
        bb = b_av * sign
        cc = c_av * sign
        if not use_wigner:
            gf[:] = (aa - w_array**2 - np.sqrt((aa - w_array**2)**2 - 4 * bb * cc + 0j)) / (2 * bb * cc)
        else:
            gf[:] = (aa + w_array**2 + np.sqrt((aa + w_array**2)**2 - 4 * bb * cc + 0j)) / (2 * bb * cc)
    else:
        aa = a[-1] * sign - sign * shift_value
        if not use_wigner:
            gf[:] = 1 / (aa - w_array**2 + 2j * w_array * smearing)
        else:
            gf[:] = 1 / (aa + w_array**2 + 2j * w_array * smearing)
    for i in range(n_iters - 2, -1, -1):
        aa = a[i] * sign - sign * shift_value
        bb = b[i] * sign
        cc = bb
        if len(c) == len(b):
            cc = c[i] * sign
        if not use_wigner:
            gf = 1.0 / (aa - w_array**2 + 2j * w_array * smearing - bb * cc * gf)
        else:
            gf = 1.0 / (aa + w_array**2 + 2j * w_array * smearing - bb * cc * gf)
    if not use_wigner:
        return gf * perturbation_modulus
    return (-np.real(gf) + 1j * np.imag(gf)) * perturbation_modulus


def project_ensemble(u_disps, forces, sscha_forces, f_mean, masses, pols, bohr):
    """X, Y of Lanczos.__init__, step by step as Py:384-439 does them (FP64 numpy).

    u_disps, forces, sscha_forces: (N, 3 nat) Angstrom / Ry/Angstrom; f_mean: (3 nat,) the
    symmetrized average force tiled over the supercell (Py:391-400); masses: (3 nat,);
    pols: (3 nat, n) kept modes. sscha_forces / f_mean may be None (zero).
    """
    N, nc = np.shape(u_disps)
    u = np.array(u_disps, dtype=np.float64) / bohr                      # Py:385
    f = np.array(forces, dtype=np.float64).reshape(N, nc).copy()        # Py:387
    if sscha_forces is not None:
        f -= np.asarray(sscha_forces, dtype=np.float64).reshape(N, nc)  # Py:388
    fm = np.zeros(nc) if f_mean is None else np.array(f_mean, dtype=np.float64).ravel()
    fm = fm * bohr                                                      # Py:400
    f *= bohr                                                           # Py:411
    f[:, :] -= np.tile(fm, (N, 1))                                      # Py:422
    m = np.asarray(masses, dtype=np.float64)
    u *= np.tile(np.sqrt(m), (N, 1))                                    # Py:425
    f /= np.tile(np.sqrt(m), (N, 1))                                    # Py:426
    X = u.dot(pols)                                                     # Py:438
    Y = f.dot(pols)                                                     # Py:439
    return X, Y


def symmetries_on_modes(rotations, irt, pols, block_start):
    """S_b[s] = E_b^T P_s E_b with (P_s E)[irt_s(a)] = R_s E[a], per block of consecutive
    modes [block_start[b], block_start[b+1]) (what prepare_symmetrization stores through
    CellConstructor's GetSymmetriesOnModesDeg, Py:735-746)."""
    rot = np.asarray(rotations, dtype=np.float64)
    irt = np.asarray(irt)
    nat = irt.shape[1]
    E = np.asarray(pols, dtype=np.float64).reshape(nat, 3, -1)
    out = []
    for b0, b1 in zip(block_start[:-1], block_start[1:]):
        Eb = E[:, :, b0:b1]
        S = np.empty((len(rot), b1 - b0, b1 - b0))
        for s in range(len(rot)):
            RE = np.einsum("xy,ayj->axj", rot[s], Eb)
            S[s] = np.einsum("axi,axj->ij", Eb[irt[s]], RE)
        out.append(S)
    return out


def d1M_dR_av(rho, Z):
    """Perturbations.py:195-238: Z (N, nat, 3 field, 3 cart) -> (3 nat, 3)."""
    N = len(rho)
    new = np.array([np.einsum("abc ->bac", Z[i]).reshape((3, -1)).T for i in range(N)])
    return np.einsum("iab, i", new, rho) / np.sum(rho)


def d1alpha_dR_av(rho, R):
    """Perturbations.py:240-273: R (N, 3, 3, 3 nat) -> (3, 3, 3 nat)."""
    return np.einsum("iabc, i", R, rho) / np.sum(rho)


def d2M_dR_av(u_bohr, rho, Z, ups):
    """Perturbations.py:275-328 with u already in Bohr: (3 nat, 3 nat, 3)."""
    N = len(rho)
    v = np.einsum("ab, ib -> ia", ups, u_bohr)
    new = np.array([np.einsum("abc ->bac", Z[i]).reshape((3, -1)).T for i in range(N)])
    return np.einsum("i, ia, ibc -> abc", rho, v, new) / np.sum(rho)


def d2alpha_dR_av(u_bohr, rho, R, ups):
    """Perturbations.py:330-375 with u already in Bohr: (3, 3, 3 nat, 3 nat)."""
    v = np.einsum("ab, ib -> ia", ups, u_bohr)
    return np.einsum("i, ia, ibcd -> bcda", rho, v, R) / np.sum(rho)


def two_phonon_sectors(d2_mu, w, T):
    """Py:1484-1502: D = (d2 + d2^T) / 2, a'(1) = -sqrt(-chi^-/2) D, b'(1) = +sqrt(chi^+/2) D
    (element by element)."""
    D = 0.5 * (d2_mu + d2_mu.T)
    chi_m, chi_p = chi_minus_plus(w, T)
    pert_a = -np.einsum("nm, nm -> nm", np.sqrt(-0.5 * chi_m), D)
    pert_b = +np.einsum("nm, nm -> nm", np.sqrt(+0.5 * chi_p), D)
    return pert_a, pert_b


def pack_two_phonon(psi, pert_a, pert_b):
    """Py:1507-1516: rows i of the upper triangles (columns j >= i) one after the other, a'(1)
    then b'(1), after the n one-phonon entries."""
    n = pert_a.shape[0]
    cur = n
    for i in range(n):
        psi[cur:cur + n - i] = pert_a[i, i:]
        cur += n - i
    for i in range(n):
        psi[cur:cur + n - i] = pert_b[i, i:]
        cur += n - i
    return psi


def _act(T, R, irt, n_atom):
    """g.T for one operation, canonical layout (nat, 3) per atom axis then field axes of 3:
    (g.T)[irt(a)][i] ... = R[i][p] ... T[a][p] ... (the convention of symmetries_on_modes)."""
    nd = T.ndim
    cart = [2 * k + 1 for k in range(n_atom)] + list(range(2 * n_atom, nd))
    Tr = T
    for ax in cart:                       # rotate every Cartesian axis
        Tr = np.moveaxis(np.tensordot(R, Tr, axes=([1], [ax])), 0, ax)
    out = np.empty_like(Tr)
    if n_atom == 1:
        out[irt] = Tr
    else:
        out[np.ix_(irt, np.arange(3), irt)] = Tr
    return out


def symmetrize_average(kind, C, rotations, irt, asr=True):
    """The symmetrized ensemble averages of Modules/Perturbations.py:15-193, restated as the
    space-group average (1/n_sym) sum_g g.T after the acoustic sum rule (the mean over the
    atoms of every atom index removed) and, for the second-order tensors, the symmetrization
    in the two displacements (and in the two field indices of the Raman tensors). Layouts as
    d1M_dR_av ("d1M"), d1alpha_dR_av ("d1alpha"), d2M_dR_av ("d2M"), d2alpha_dR_av
    ("d2alpha"). Loops over the operations: small systems only."""
    C = np.asarray(C, dtype=np.float64)
    rot = np.asarray(rotations, dtype=np.float64)
    irt = np.asarray(irt)
    if kind == "d1M":
        nat = C.shape[0] // 3
        T = C.reshape(nat, 3, 3)
        if asr:
            T = T - T.mean(axis=0, keepdims=True)
        na = 1
    elif kind == "d1alpha":
        nat = C.shape[2] // 3
        T = C.reshape(3, 3, nat, 3).transpose(2, 3, 0, 1)
        T = 0.5 * (T + T.transpose(0, 1, 3, 2))
        if asr:
            T = T - T.mean(axis=0, keepdims=True)
        na = 1
    elif kind == "d2M":
        nat = C.shape[0] // 3
        T = C.reshape(nat, 3, nat, 3, 3)
        T = 0.5 * (T + T.transpose(2, 3, 0, 1, 4))
        if asr:
            T = T - T.mean(axis=0, keepdims=True)
            T = T - T.mean(axis=2, keepdims=True)
        na = 2
    elif kind == "d2alpha":
        nat = C.shape[2] // 3
        T = C.reshape(3, 3, nat, 3, nat, 3).transpose(2, 3, 4, 5, 0, 1)
        T = 0.5 * (T + T.transpose(2, 3, 0, 1, 4, 5))
        T = 0.5 * (T + T.transpose(0, 1, 2, 3, 5, 4))
        if asr:
            T = T - T.mean(axis=0, keepdims=True)
            T = T - T.mean(axis=2, keepdims=True)
        na = 2
    else:
        raise ValueError(kind)
    S = np.zeros_like(T)
    for g in range(len(rot)):
        S += _act(T, rot[g], irt[g], na)
    S /= len(rot)
    if kind == "d1M":
        return S.reshape(3 * nat, 3)
    if kind == "d1alpha":
        return S.transpose(2, 3, 0, 1).reshape(3, 3, 3 * nat)
    if kind == "d2M":
        return S.reshape(3 * nat, 3 * nat, 3)
    return S.transpose(4, 5, 0, 1, 2, 3).reshape(3, 3, 3 * nat, 3 * nat)


def translation_average_psi(psi, pols, irt_translations, adjoint=False):
    """(1/n_t) sum_t Ad_t psi: R -> S_t R, each packed symmetric matrix M -> S_t M S_t^T with
    S_t = E^T P_t E (P_t the Cartesian permutation of translation t, E the polarization
    vectors of the kept modes). adjoint: the transpose in the packed metric, W P0 W^-1 with
    W = 2 on the off-diagonal entries."""
    E = np.asarray(pols, dtype=np.float64)
    n = E.shape[1]
    nw2 = n * (n + 1) // 2
    iu = np.triu_indices(n)
    w = np.ones(n + 2 * nw2)
    off = iu[0] < iu[1]
    w[n:n + nw2][off] = 2
    w[n + nw2:][off] = 2
    x = psi / w if adjoint else psi.copy()
    perms = [(3 * np.asarray(p)[:, None] + np.arange(3)[None]).ravel() for p in irt_translations]
    out = np.empty_like(x)
    c = E @ x[:n]
    acc = np.zeros_like(c)
    for p in perms:
        o = np.empty_like(c)
        o[p] = c
        acc += o
    out[:n] = E.T @ acc / len(perms)
    for k in range(2):
        M = np.zeros((n, n))
        M[iu] = x[n + k * nw2:n + (k + 1) * nw2]
        M = M + np.triu(M, 1).T
        C = E @ M @ E.T
        acc = np.zeros_like(C)
        for p in perms:
            o = np.empty_like(C)
            o[np.ix_(p, p)] = C
            acc += o
        out[n + k * nw2:n + (k + 1) * nw2] = (E.T @ acc @ E / len(perms))[iu]
    return out * w if adjoint else out
