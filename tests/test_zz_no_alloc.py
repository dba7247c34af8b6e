"""The no-allocation entry point tdscha_get_perturb_average_sym_noalloc (device buffers carved
from a static per-GPU scratch array) against tdscha_get_perturb_average_sym, through the C
ABI: two calls in a row, the second on a smaller ensemble, and an ensemble too large for the
scratch, rejected before any upload. Both entry points run the same FP64 kernels and the
result does not depend on how the images are split into passes, so they agree bit for bit."""
import ctypes

import numpy as np
import pytest

import fixtures

pytestmark = pytest.mark.gpu


def _call(fn, ens, R1, Y1, d4):
    from tdscha_amd import _lib as L
    s = ens.sym
    syms = (L._dp * s.n_blocks)(*[a.ctypes.data_as(L._dp) for a in s.symmetries])
    degs = (L._ip * s.n_blocks)(*[d.ctypes.data_as(L._ip) for d in s.degenerate_space])
    X, Y, w, rho = (np.ascontiguousarray(a, dtype=np.float64) for a in (ens.X, ens.Y, ens.w,
                                                                          ens.rho))
    R1c, Y1c = np.ascontiguousarray(R1), np.ascontiguousarray(Y1)
    N, n = X.shape
    f = np.full(n, 7.0)
    d = np.full((n, n), 7.0)
    ndeg = np.ascontiguousarray(s.N_degeneracy, dtype=np.intc)
    bid = np.ascontiguousarray(s.block_id, dtype=np.intc)
    rc = getattr(L.lib(), fn)(
        X.ctypes.data_as(L._dp), Y.ctypes.data_as(L._dp), w.ctypes.data_as(L._dp),
        rho.ctypes.data_as(L._dp), R1c.ctypes.data_as(L._dp), Y1c.ctypes.data_as(L._dp), ens.T,
        d4, int(s.symmetries[0].shape[0]), s.n_blocks, syms, ndeg.ctypes.data_as(L._ip), degs,
        bid.ctypes.data_as(L._ip), N, n, f.ctypes.data_as(L._dp), d.ctypes.data_as(L._dp))
    return rc, f, d


def test_noalloc_equals_the_allocating_entry_point(mode10):
    from tdscha_amd import _lib as L
    assert L.lib().tdscha_noalloc_scratch_bytes() == 256 << 20
    cases = [("mode10 (45 modes, 10 configs, 384 symmetries)", mode10, 1),
             ("random ensemble (20 modes, 6 configs, 5 symmetries)",
              fixtures.random_ensemble(81, 20, 6, 5), 0)]
    for k, (name, ens, d4) in enumerate(cases):
        print("case %d: %s, apply_D4=%d" % (k, name, d4), flush=True)
        R1, Y1 = fixtures.random_perturbation(82 + k, ens.X.shape[1])
        rc0, f0, d0 = _call("tdscha_get_perturb_average_sym", ens, R1, Y1, d4)
        rc1, f1, d1 = _call("tdscha_get_perturb_average_sym_noalloc", ens, R1, Y1, d4)
        assert rc0 == 0 and rc1 == 0, L.lib().tdscha_last_error()
        assert np.array_equal(f0, f1) and np.array_equal(d0, d1)
        assert np.abs(f1).max() > 0 and np.abs(d1).max() > 0


def test_noalloc_rejects_an_ensemble_beyond_the_scratch():
    from tdscha_amd import _lib as L
    print("case: ensemble larger than the scratch (64 modes, 140000 configs)", flush=True)
    ens = fixtures.random_ensemble(83, 64, 140000, 2, max_block=4)
    R1, Y1 = fixtures.random_perturbation(84, 64)
    rc, f, d = _call("tdscha_get_perturb_average_sym_noalloc", ens, R1, Y1, 1)
    assert rc == 5
    assert "no-allocation scratch" in L.lib().tdscha_last_error().decode()
    assert np.all(f == 7.0) and np.all(d == 7.0)     # outputs untouched
