"""The Julia level's range-sliced entries on the GPU (tdscha_jl_get_perturb_*_averages_sym,
Modules/tdscha_core.jl:202-284) against the oracle's range restatement, and the reuse of the
drop-ins' device ensemble between calls.

Tolerance: plain relative error ||out - ref|| / ||ref|| <= 1e-10 (FP64 path).
"""
import numpy as np
import pytest

import fixtures
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _padded(sym):
    """sym_julia / deg_julia as Modules/DynamicalLanczos.py:753-768 builds them."""
    nb = sym.n_blocks
    dmax = max(s.shape[1] for s in sym.symmetries)
    sj = np.zeros((nb, sym.n_syms, dmax, dmax))
    dj = np.zeros((nb, dmax), dtype=np.int32)
    for b, s in enumerate(sym.symmetries):
        c = s.shape[1]
        sj[b, :, :c, :c] = s
        dj[b, :c] = sym.degenerate_space[b]
    return sj, dj


def _args(ens, R1, Y1, v4=True):
    sj, dj = _padded(ens.sym)
    return (ens.X.T, ens.Y.T, ens.w, ens.rho, R1, Y1, ens.T, v4, sj, ens.sym.N_degeneracy, dj,
            ens.sym.block_id)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _ranges(N, nsym):
    total = N * nsym
    return [(1, total),                        # everything
            (1, 1), (total, total),            # one image, first and last
            (nsym + 2, nsym + 2),              # one image inside a configuration
            (3, nsym - 1),                     # inside the first configuration
            (nsym + 1, 2 * nsym),              # exactly one configuration
            (2, 3 * nsym + 1),                 # from inside one configuration to inside another
            (nsym, total - nsym + 1)]          # last image of the first to first of the last


@pytest.mark.parametrize("seed", [1, 2])
def test_ranges_against_the_oracle(seed):
    from tdscha_amd import tdscha_core as tc
    ens = fixtures.random_ensemble(60 + seed, 30, 40, 6, max_block=5)
    R1, Y1 = fixtures.random_perturbation(70 + seed, 30)
    args = _args(ens, R1, Y1)
    N, nsym = ens.X.shape[0], ens.sym.n_syms
    for a, b in _ranges(N, nsym):
        f_o, d_o = orc.perturb_averages_range(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 1,
                                              ens.sym, a, b)
        f, d = tc.get_perturb_averages_sym(*args, a, b)
        d2 = tc.get_perturb_d2v_averages_sym(*args, a, b)
        f2 = tc.get_perturb_f_averages_sym(*args, a, b)
        print("[%d, %d]: f %.2e  d2v %.2e" % (a, b, _rel(f, f_o), _rel(d, d_o)), flush=True)
        assert _rel(f, f_o) <= TOL and _rel(d, d_o) <= TOL, (a, b)
        assert np.array_equal(d2, d) and np.array_equal(f2, f), (a, b)


def test_d3_only_and_split_adds_up():
    """apply_v4 = False leaves the D4 part out; the split of Py:3083-3103 over 5 ranks adds
    up to the full average of GetPerturbAverageSym's contract."""
    from tdscha_amd import tdscha_core as tc
    from tdscha_amd.sscha_HP_odd import GetPerturbAverageSym
    ens = fixtures.random_ensemble(65, 24, 30, 4, max_block=4)
    R1, Y1 = fixtures.random_perturbation(66, 24)
    total = 30 * 4
    d_o = orc.perturb_averages_range(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 0, ens.sym,
                                     7, 77, what="d2v")
    d = tc.get_perturb_d2v_averages_sym(*_args(ens, R1, Y1, v4=False), 7, 77)
    assert _rel(d, d_o) <= TOL
    args = _args(ens, R1, Y1)
    f = np.zeros(24)
    d = np.zeros((24, 24))
    for a, b in tc.split_indices(total, 5):
        f += tc.get_perturb_f_averages_sym(*args, a, b)
        d += tc.get_perturb_d2v_averages_sym(*args, a, b)
    s = ens.sym
    f0, d0 = np.zeros(24), np.zeros((24, 24))
    GetPerturbAverageSym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 1, s.n_syms, s.symmetries,
                         s.N_degeneracy, s.degenerate_space, s.block_id, f0, d0)
    assert _rel(f, f0) <= TOL and _rel(d, d0) <= TOL


def test_unequal_blocks_image_path():
    """A block whose modes differ in frequency runs on the exact image path: the range
    window applies there too."""
    from tdscha_amd import tdscha_core as tc
    ens = fixtures.split_frequencies(fixtures.random_ensemble(67, 20, 25, 4, max_block=4), 1e-3)
    R1, Y1 = fixtures.random_perturbation(68, 20)
    for a, b in ((3, 58), (41, 41), (1, 100)):
        f_o, d_o = orc.perturb_averages_range(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 1,
                                              ens.sym, a, b)
        f, d = tc.get_perturb_averages_sym(*_args(ens, R1, Y1), a, b)
        assert _rel(f, f_o) <= TOL and _rel(d, d_o) <= TOL, (a, b)


def test_mode10_range_against_the_reference_golden_vectors(mode10):
    """The reference ensemble (45 modes, 10 configurations, 384 symmetries), split over 3
    ranks: the pieces add up to the reference C trio's own outputs (an output that the
    symmetry makes vanish must vanish to 1e-12 of the sum of |every term|)."""
    from tdscha_amd import tdscha_core as tc
    kv = fixtures.load_npz("ref_kernel_vectors.npz")
    for k in range(int(kv["n_cases"])):
        R1, Y1, d4 = kv["R1_%d" % k], kv["Y1_%d" % k], int(kv["d4_%d" % k])
        _, _, fa, da = orc.perturb_average_sym(mode10.X, mode10.Y, mode10.w, mode10.rho, R1, Y1,
                                               mode10.T, d4, mode10.sym, with_abs=True)
        args = _args(mode10, R1, Y1, v4=bool(d4))
        f = np.zeros(45)
        d = np.zeros((45, 45))
        for a, b in tc.split_indices(10 * 384, 3):
            fp, dp = tc.get_perturb_averages_sym(*args, a, b)
            f += fp
            d += dp
        for out, ref, scale in ((f, kv["f_%d" % k], fa), (d, kv["d2v_%d" % k], da)):
            if np.linalg.norm(ref) >= 1e-6 * np.linalg.norm(scale):
                assert _rel(out, ref) <= TOL, k
            else:
                assert np.linalg.norm(out) <= 1e-12 * np.linalg.norm(scale), k


def test_drop_in_reuses_its_ensemble_only_for_equal_inputs():
    """Two calls with the same arrays give the same bits (the second on the kept ensemble);
    a change inside X, or in rho, is seen; with the reuse turned off the results agree."""
    from tdscha_amd import tdscha_core as tc
    from tdscha_amd.sscha_HP_odd import GetPerturbAverageSym
    ens = fixtures.random_ensemble(69, 18, 22, 3, max_block=3)
    R1, Y1 = fixtures.random_perturbation(70, 18)
    s = ens.sym
    X = ens.X.copy()
    rho = ens.rho.copy()

    def call():
        f, d = np.zeros(18), np.zeros((18, 18))
        GetPerturbAverageSym(X, ens.Y, ens.w, rho, R1, Y1, ens.T, 1, s.n_syms, s.symmetries,
                             s.N_degeneracy, s.degenerate_space, s.block_id, f, d)
        return f, d

    def oracle():
        return orc.perturb_average_sym(X, ens.Y, ens.w, rho, R1, Y1, ens.T, 1, s)

    tc.drop_in_cache(True)
    f1, d1 = call()
    f2, d2 = call()
    assert np.array_equal(f1, f2) and np.array_equal(d1, d2)
    X[11, 5] *= 1.5
    f3, d3 = call()
    f_o, d_o = oracle()
    assert _rel(f3, f_o) <= TOL and _rel(d3, d_o) <= TOL
    rho[3] *= 2.0
    f4, d4 = call()
    f_o, d_o = oracle()
    assert _rel(f4, f_o) <= TOL and _rel(d4, d_o) <= TOL
    was = tc.drop_in_cache(False)
    assert was
    f5, d5 = call()
    assert _rel(f5, f_o) <= TOL and _rel(d5, d_o) <= TOL
    tc.drop_in_cache(True)
