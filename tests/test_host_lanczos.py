"""Host-side logic of the Lanczos mirror that needs no GPU: status files, psi helpers."""
import numpy as np
import pytest

import fixtures
from oracle import oracle as orc


def _lanczos(ens):
    from tdscha_amd import DynamicalLanczos as DL
    lz = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    lz.set_ensemble(ens.X, ens.Y, ens.w, ens.rho, ens.T)
    s = ens.sym
    lz.set_symmetries(s.symmetries, s.N_degeneracy, s.degenerate_space, s.block_id)
    return lz


def test_status_round_trip(tmp_path, mode10):
    st = fixtures.load_npz("lanczos_std.npz")
    lz = _lanczos(mode10)
    lz.psi = st["psi"].copy()
    lz.a_coeffs, lz.b_coeffs, lz.c_coeffs = list(st["a"]), list(st["b"]), list(st["c"])
    lz.basis_Q, lz.basis_P, lz.s_norm = list(st["Q"]), list(st["P"]), list(st["s_norm"])
    f = str(tmp_path / "status")
    lz.save_status(f)
    from tdscha_amd import DynamicalLanczos as DL
    lz2 = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    lz2.load_status(f + ".npz")
    assert np.array_equal(lz2.X, mode10.X) and np.array_equal(lz2.w, mode10.w)
    assert np.array_equal(np.asarray(lz2.basis_Q), st["Q"])
    assert np.array_equal(np.asarray(lz2.c_coeffs), st["c"])
    for b in range(mode10.sym.n_blocks):
        assert np.array_equal(lz2.symmetries[b], mode10.sym.symmetries[b])
        assert np.array_equal(lz2.degenerate_space[b], mode10.sym.degenerate_space[b])
    assert np.array_equal(lz2.sym_block_id, mode10.sym.block_id)


def test_padded_blocks_without_sizes(mode10):
    """Reference status files store sym_julia / deg_julia only (Py:753-768)."""
    from tdscha_amd.DynamicalLanczos import blocks_from_padded
    lz = _lanczos(mode10)
    sj, dj = lz._padded_symmetries()
    syms, ndeg, degs, bid = blocks_from_padded(sj, dj, 45)
    assert np.array_equal(ndeg, mode10.sym.N_degeneracy)
    assert np.array_equal(bid, mode10.sym.block_id)


def test_psi_helpers_match_oracle(mode10):
    st = fixtures.load_npz("lanczos_std.npz")
    lz = _lanczos(mode10)
    lz.psi = st["Q"][4].copy()
    n = 45
    nw2 = n * (n + 1) // 2
    for half in (False, True):
        assert np.array_equal(lz.get_Y1(half), orc.unpack_sym(lz.psi[n:n + nw2], n, half))
        assert np.array_equal(lz.get_ReA1(half), orc.unpack_sym(lz.psi[n + nw2:], n, half))
    assert np.array_equal(lz.mask_dot_wigner(), orc.mask_dot_wigner(n))


def test_lanczos_requires_init_and_rejects_cpu_modes(mode10):
    from tdscha_amd import DynamicalLanczos as DL
    with pytest.raises(ValueError):
        DL.Lanczos(mode=DL.MODE_FAST_SERIAL)
    lz = _lanczos(mode10)
    with pytest.raises(ValueError):
        lz.apply_full_L(np.zeros(len(lz.psi)))


def test_continued_fraction_host_matches_oracle():
    from tdscha_amd.DynamicalLanczos import continued_fraction
    st = fixtures.load_npz("lanczos_std.npz")
    w = np.linspace(0, 3e-3, 101)
    for term in (True, False):
        for wig in (False, True):
            g = continued_fraction(st["a"], st["b"], st["c"], w, term, 1, 1e-5, wig)
            g_o = orc.green_function_continued_fraction(st["a"], st["b"], st["c"], w, term, 1,
                                                        1e-5, wig)
            assert np.allclose(g, g_o, rtol=1e-12, atol=0)
