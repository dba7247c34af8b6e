"""Parity of the HIP path (through the C ABI) with the CPU oracle and the golden fixtures.

Tolerance (FP64 path): plain relative error ||out - ref|| / ||ref|| <= 1e-10 for every output
that does not cancel (||ref|| at least 1e-6 of the oracle's absolute-value sums, the sum of
|every term|). An output that cancels to rounding (f_av of a perturbation that the symmetry
makes vanish) must itself vanish: ||out|| <= 1e-12 of those sums. Reduced-precision forms are
measured against the absolute-value sums, the scale their rounding error follows.
Full L: plain relative 1e-10 on the whole vector and on its anharmonic part alone.
"""
import numpy as np
import pytest

import fixtures
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _dev_ens(ens, precisions=("fp64",), s_batch=0):
    from tdscha_amd.ensemble import DeviceEnsemble, SymmetryBlocks
    s = ens.sym
    sym = SymmetryBlocks(s.symmetries, s.N_degeneracy, s.degenerate_space, s.block_id)
    return DeviceEnsemble(ens.X, ens.Y, ens.w, ens.rho, ens.T, sym, precisions=precisions,
                          s_batch=s_batch)


def _scaled_err(out, ref, scale):
    return np.linalg.norm(out - ref) / max(np.linalg.norm(scale), 1e-300)


CANCEL = 1e-6     # ||ref|| below this fraction of the abs-sum scale: the output cancels
VANISH = 1e-12    # a cancelling output must vanish to this fraction of the scale


def _assert_fp64(out, ref, scale, what):
    """Plain relative error for a non-cancelling output, vanishing for a cancelling one."""
    nref, nsc = np.linalg.norm(ref), np.linalg.norm(scale)
    if nref >= CANCEL * nsc:
        err = np.linalg.norm(out - ref) / nref
        print("%s: relative error %.3e (||ref|| / scale %.1e)" % (what, err, nref / nsc))
        assert err <= TOL64, (what, err)
    else:
        v = np.linalg.norm(out) / nsc
        print("%s: cancels (||ref|| / scale %.1e), ||out|| / scale %.1e" % (what, nref / nsc, v))
        assert v <= VANISH, (what, v)


def _check(ens, R1, Y1, d4, dev=None, tol=TOL64, precision="fp64"):
    f_o, d_o, fa, da = orc.perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, d4,
                                               ens.sym, with_abs=True)
    dev = dev or _dev_ens(ens)
    terms = 1 | 2 | (4 if d4 else 0)
    f, d = dev.apply(R1, Y1, terms=terms, precision=precision)
    f, d = f.cpu().numpy(), d.cpu().numpy()
    ef, ed = _scaled_err(f, f_o, fa), _scaled_err(d, d_o, da)
    print("f err %.3e  d2v err %.3e (of the abs-sum scale)" % (ef, ed))
    if precision == "fp64" and tol == TOL64:
        _assert_fp64(f, f_o, fa, "f_av")
        _assert_fp64(d, d_o, da, "d2v")
    else:
        assert ef < tol and ed < tol, (ef, ed)
    return f, d


def test_mode10_golden_reference_vectors(mode10):
    """The reference C trio's own outputs (compiled from CModules/LanczosFunctions.c)."""
    kv = fixtures.load_npz("ref_kernel_vectors.npz")
    dev = _dev_ens(mode10)
    for k in range(int(kv["n_cases"])):
        R1, Y1, d4 = kv["R1_%d" % k], kv["Y1_%d" % k], int(kv["d4_%d" % k])
        _, _, fa, da = orc.perturb_average_sym(mode10.X, mode10.Y, mode10.w, mode10.rho, R1, Y1,
                                               mode10.T, d4, mode10.sym, with_abs=True)
        f, d = dev.apply(R1, Y1, terms=1 | 2 | (4 if d4 else 0))
        print("case", k)
        _assert_fp64(f.cpu().numpy(), kv["f_%d" % k], fa, "f_av")
        _assert_fp64(d.cpu().numpy(), kv["d2v_%d" % k], da, "d2v")


@pytest.mark.parametrize("seed,n,N,n_sym,d4", [
    (1, 45, 37, 12, 1),     # ragged sizes, several blocks
    (2, 70, 130, 5, 0),     # D4 off
    (3, 1, 3, 2, 1),        # a single mode
    (4, 64, 64, 1, 1),      # exactly one tile, one symmetry
    (5, 129, 65, 3, 1),     # one past the tile sizes
])
def test_random_ensembles(seed, n, N, n_sym, d4):
    ens = fixtures.random_ensemble(seed, n, N, n_sym)
    R1, Y1 = fixtures.random_perturbation(seed + 100, n, symmetric=(seed % 2 == 0))
    _check(ens, R1, Y1, d4)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_blocks_wider_than_rotation_tile(precision):
    """Blocks of more than 64 modes take the unfused rotation kernels."""
    ens = fixtures.random_ensemble(20, 160, 50, 3, max_block=90)
    assert max(ens.sym.N_degeneracy) > 64
    R1, Y1 = fixtures.random_perturbation(24, 160)
    tol = {"fp64": TOL64, "fp32": 2e-5, "bf16x3": 2e-4}[precision]
    dev = _dev_ens(ens, precisions=("fp64", precision))
    for v in ((0, 3, 5, 7, 8) if precision == "bf16x3" else (0,)):
        # tile forms 3-8 keep their register-operand GEMM1 here; GEMM2 falls back to the form
        # that writes the FP64 B the one-sided rotations read
        dev.set_tuning(v, 0)
        _check(ens, R1, Y1, 1, dev=dev, tol=tol, precision=precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("max_block,seed", [(32, 40), (50, 41)])
def test_group_tiled_rotation_widths(precision, max_block, seed):
    """Blocks of up to 32 modes take the 32-wide group tiles, up to 64 the 64-wide ones."""
    ens = fixtures.random_ensemble(seed, 170, 70, 4, max_block=max_block)
    assert 16 < max(ens.sym.N_degeneracy) <= max_block
    R1, Y1 = fixtures.random_perturbation(seed + 5, 170)
    tol = {"fp64": TOL64, "fp32": 2e-5, "bf16x3": 2e-4}[precision]
    dev = _dev_ens(ens, precisions=("fp64", precision))
    _check(ens, R1, Y1, 1, dev=dev, tol=tol, precision=precision)


@pytest.mark.parametrize("n,N,seed", [(150, 300, 31), (300, 520, 32), (45, 256, 33)])
def test_bf16x3_tile_shapes(n, N, seed):
    """The nine contraction kernel forms against the oracle and against each other. Forms
    0-2 share their arithmetic and differ only in the order of GEMM1's FP64 row sums (1e-12);
    so do forms 3-8 among themselves. Forms 3-8 add each lane's six alpha / beta products in
    FP32 and take gamma / delta from an FP64 product, so they agree with forms 0-2 to the
    FP32 rounding of those partial sums (1e-6 relative). Forms 6, 7 and 8 run form 4's GEMM1
    and, in their taller GEMM2 tiles (256 rows of four waves, 512 rows of eight, 448 rows of
    eight waves of four and of three row blocks), the same MFMA sequence per output element:
    form 4 bit for bit. The sizes put row and column tiles across the edges of the modes (a
    448- or 512-row GEMM2 tile over 150 or 300 modes)."""
    ens = fixtures.random_ensemble(seed, n, N, 4, max_block=8)
    R1, Y1 = fixtures.random_perturbation(seed + 1, n)
    dev = _dev_ens(ens, precisions=("fp64", "bf16x3"))
    out = []
    for v in (0, 1, 2, 3, 4, 5, 6, 7, 8):
        dev.set_tuning(v, 0)
        assert dev.tuning()[0] == v
        out.append(_check(ens, R1, Y1, 1, dev=dev, tol=2e-4, precision="bf16x3"))
    for k in (6, 7, 8):
        for a, b in zip(out[k], out[4]):
            assert np.array_equal(a, b), k
    for ref, group, tol in ((0, (1, 2), 1e-12), (3, (4, 5, 6, 7, 8), 1e-12), (0, (3,), 1e-6)):
        for k in group:
            for a, b in zip(out[ref], out[k]):
                assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(a)), (ref, k)


@pytest.mark.parametrize("precision,tol", [("fp32", 2e-5), ("bf16x3", 2e-4)])
def test_precision_forms(precision, tol):
    """FP32 / BF16x3 contractions against the FP64 oracle (relative to the abs-sum scale)."""
    ens = fixtures.random_ensemble(17, 150, 300, 6)
    R1, Y1 = fixtures.random_perturbation(18, 150)
    dev = _dev_ens(ens, precisions=("fp64", precision))
    _check(ens, R1, Y1, 1, dev=dev, tol=tol, precision=precision)


@pytest.mark.parametrize("rel,d4", [(1e-6, 1), (0.05, 1), (0.05, 0)])
def test_blocks_with_unequal_frequencies(rel, d4):
    """Blocks whose modes differ in frequency (by 1e-6 and 5 % relative) take the exact image
    path: the perturbed averages and the full L against the oracle (which the host tests pin
    on the reference's own C build for such blocks), plain relative 1e-10."""
    from tdscha_amd import _lib as L
    ens = fixtures.split_frequencies(fixtures.random_ensemble(95, 23, 40, 5, max_block=6), rel)
    print("case: split blocks, rel %g, d4 %d" % (rel, d4), flush=True)
    R1, Y1 = fixtures.random_perturbation(96, 23)
    dev = _dev_ens(ens)
    _check(ens, R1, Y1, d4, dev=dev)
    psi = np.random.default_rng(97).standard_normal(orc.psi_len(23))
    for fl, kw in ((0, {}), (L.OP_TRANSPOSE, dict(transpose=True)),
                   (L.OP_WIGNER, dict(use_wigner=True))):
        out = dev.apply_full_L(psi, flags=fl).cpu().numpy()
        ref = orc.apply_full_L(ens, psi, **kw)
        err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
        print(kw, "full L relative error %.2e" % err)
        assert err < TOL64


def test_unequal_blocks_refuse_reduced_precision_and_the_adjoint():
    ens = fixtures.split_frequencies(fixtures.random_ensemble(98, 12, 20, 3, max_block=4), 0.01)
    dev = _dev_ens(ens, precisions=("fp64", "bf16x3"))
    R1, Y1 = fixtures.random_perturbation(99, 12)
    with pytest.raises(ValueError, match="FP64 only"):
        dev.apply(R1, Y1, precision="bf16x3")
    u = np.ones(orc.psi_len(12))
    with pytest.raises(ValueError, match="degenerate"):
        dev.force_grad([u], [u], [0])


def test_zero_temperature_and_uniform_weights():
    ens = fixtures.random_ensemble(11, 30, 40, 4, T=0.0, weights="uniform")
    R1, Y1 = fixtures.random_perturbation(12, 30)
    _check(ens, R1, Y1, 1)


def test_s_batch_independent(mode10):
    """Results do not depend on how the symmetries are split into passes."""
    R1, Y1 = fixtures.random_perturbation(7, 45)
    out = []
    for sb in (1, 7, 384):
        f, d = _dev_ens(mode10, s_batch=sb).apply(R1, Y1)
        out.append((f.cpu().numpy(), d.cpu().numpy()))
    for f, d in out[1:]:
        assert np.allclose(f, out[0][0], rtol=0, atol=1e-12 * np.abs(out[0][0]).max() + 1e-300)
        assert np.allclose(d, out[0][1], rtol=0, atol=1e-12 * np.abs(out[0][1]).max())


def test_trio_entry_points(mode10):
    """The three drop-in C symbols keep the reference's overwrite / accumulate contract."""
    import ctypes
    from tdscha_amd import _lib as L
    ens = mode10
    R1, Y1 = fixtures.random_perturbation(21, 45)
    f_o, d_o = orc.perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 1, ens.sym)
    _, _, fa, da = orc.perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 1, ens.sym,
                                           with_abs=True)
    s = ens.sym
    psyms = (L._dp * s.n_blocks)(*[a.ctypes.data_as(L._dp) for a in s.symmetries])
    pdegs = (L._ip * s.n_blocks)(*[a.ctypes.data_as(L._ip) for a in s.degenerate_space])
    X, Y, w, rho = (np.ascontiguousarray(a) for a in (ens.X, ens.Y, ens.w, ens.rho))
    R1c, Y1c = np.ascontiguousarray(R1), np.ascontiguousarray(Y1)
    common = lambda pert: (X.ctypes.data_as(L._dp), Y.ctypes.data_as(L._dp),
                           w.ctypes.data_as(L._dp), pert.ctypes.data_as(L._dp), ens.T, 45, 10,
                           rho.ctypes.data_as(L._dp), psyms, 384,
                           s.N_degeneracy.ctypes.data_as(L._ip), pdegs,
                           s.block_id.ctypes.data_as(L._ip))
    f = np.full(45, 7.0)
    d = np.full((45, 45), 7.0)   # garbage: must be overwritten, then accumulated into
    lib = L.lib()
    lib.get_f_average_from_Y_pert_sym_fast(*common(Y1c), f.ctypes.data_as(L._dp))
    lib.get_d2v_dR2_from_R_pert_sym_fast(*common(R1c), d.ctypes.data_as(L._dp))
    lib.get_d2v_dR2_from_Y_pert_sym_fast(*common(Y1c), d.ctypes.data_as(L._dp))
    _assert_fp64(f, f_o, fa, "f_av")
    _assert_fp64(d, d_o, da, "d2v")


def test_get_perturb_average_sym_shim(mode10):
    from tdscha_amd.sscha_HP_odd import GetPerturbAverageSym
    ens = mode10
    R1, Y1 = fixtures.random_perturbation(31, 45)
    f_o, d_o, fa, da = orc.perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 0,
                                               ens.sym, with_abs=True)
    f = np.zeros(45)
    d = np.zeros((45, 45))
    GetPerturbAverageSym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 0, 384, ens.sym.symmetries,
                         ens.sym.N_degeneracy, ens.sym.degenerate_space, ens.sym.block_id, f, d)
    _assert_fp64(f, f_o, fa, "f_av")
    _assert_fp64(d, d_o, da, "d2v")
    with pytest.raises(ValueError):
        GetPerturbAverageSym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 2, 384,
                             ens.sym.symmetries, ens.sym.N_degeneracy, ens.sym.degenerate_space,
                             ens.sym.block_id, f, d)


def test_linearity_and_symmetry(mode10):
    dev = _dev_ens(mode10)
    R1a, Y1a = fixtures.random_perturbation(41, 45)
    R1b, Y1b = fixtures.random_perturbation(42, 45)
    fa, da = (x.cpu().numpy() for x in dev.apply(R1a, Y1a))
    fb, db = (x.cpu().numpy() for x in dev.apply(R1b, Y1b))
    fc, dc = (x.cpu().numpy() for x in dev.apply(2 * R1a - 3 * R1b, 2 * Y1a - 3 * Y1b))
    assert np.linalg.norm(fc - (2 * fa - 3 * fb)) <= 1e-12 * np.linalg.norm(fc) + 1e-300
    assert np.linalg.norm(dc - (2 * da - 3 * db)) <= 1e-12 * np.linalg.norm(dc)
    assert np.array_equal(dc, dc.T) or np.abs(dc - dc.T).max() <= 1e-14 * np.abs(dc).max()
    z, dz = (x.cpu().numpy() for x in dev.apply(np.zeros(45), np.zeros((45, 45))))
    assert not z.any() and not dz.any()


# ---- full L (harmonic + anharmonic) against the reference's own Krylov recurrences ----

def _lanczos_from(ens, mode_kw=None):
    from tdscha_amd import DynamicalLanczos as DL
    lz = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    lz.set_ensemble(ens.X, ens.Y, ens.w, ens.rho, ens.T)
    s = ens.sym
    lz.set_symmetries(s.symmetries, s.N_degeneracy, s.degenerate_space, s.block_id)
    lz.init()
    lz.verbose = False
    return lz


def test_full_L_recurrence_standard(mode10):
    """L q_k = a_k q_k + b_k q_k+1 + c_k-1 q_k-1 on the stored basis (Py:6084-6160)."""
    st = fixtures.load_npz("lanczos_std.npz")
    lz = _lanczos_from(mode10)
    Q, P, a, b, c, s = st["Q"], st["P"], st["a"], st["b"], st["c"], st["s_norm"]
    for k in range(10):
        Lq = lz.apply_full_L(Q[k].copy())
        rhs = a[k] * Q[k] + b[k] * Q[k + 1] + (c[k - 1] * Q[k - 1] if k else 0)
        err = np.linalg.norm(Lq - rhs) / np.linalg.norm(Lq)
        Lq_o = orc.apply_full_L(mode10, Q[k])
        err_o = np.linalg.norm(Lq - Lq_o) / np.linalg.norm(Lq_o)
        print(k, err, err_o)
        assert err < 1e-10 and err_o < TOL64


def test_anharmonic_part_alone_matches_oracle(mode10):
    """The anharmonic part of L (L - L1) on its own at 1e-10 relative: on q_0 it is only
    ~1e-4 of ||L q_0||, so the whole-vector check above does not bound it. Standard and
    transposed on the stored bases, Wigner on its own basis."""
    from tdscha_amd import _lib as L
    st = fixtures.load_npz("lanczos_std.npz")
    sw = fixtures.load_npz("lanczos_wigner.npz")
    dev = _dev_ens(mode10)
    cases = [(st["Q"][k], 0, dict()) for k in range(10)]
    cases += [(st["P"][k], L.OP_TRANSPOSE, dict(transpose=True)) for k in (0, 4, 9)]
    cases += [(sw["Q"][k], L.OP_WIGNER, dict(use_wigner=True)) for k in (0, 5, 9)]
    for v, fl, kw in cases:
        out = dev.apply_full_L(v, flags=fl | L.OP_IGNORE_HARMONIC).cpu().numpy()
        ref = orc.apply_anharmonic_FT(mode10, v, **kw)
        err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
        ratio = np.linalg.norm(ref) / np.linalg.norm(orc.apply_full_L(mode10, v, **kw))
        print(kw, "anharmonic / full %.1e, relative error %.2e" % (ratio, err))
        assert err < TOL64, (kw, err)


def test_full_L_transpose_matches_oracle(mode10):
    st = fixtures.load_npz("lanczos_std.npz")
    lz = _lanczos_from(mode10)
    for k in (0, 3, 9):
        p = st["P"][k]
        out = lz.apply_full_L(p.copy(), transpose=True)
        ref = orc.apply_full_L(mode10, p, transpose=True)
        assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < TOL64


def test_full_L_wigner_and_flags(mode10):
    st = fixtures.load_npz("lanczos_wigner.npz")
    lz = _lanczos_from(mode10)
    lz.use_wigner = True
    for k in (0, 4):
        q = st["Q"][k]
        out = lz.apply_full_L(q.copy())
        ref = orc.apply_full_L(mode10, q, use_wigner=True)
        assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < TOL64
    lz.use_wigner = False
    q = st["Q"][3]
    for kw in (dict(ignore_v4=True), dict(ignore_v3=True), dict(ignore_harmonic=True),
               dict(reverse_L=True, shift_value=1.5e-6)):
        for key, val in kw.items():
            setattr(lz, key, val)
        out = lz.apply_full_L(q.copy())
        ref = orc.apply_full_L(mode10, q, **kw)
        assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < TOL64, kw
        for key in kw:
            setattr(lz, key, type(getattr(lz, key))(0))


@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("bf16x3", TOL64)])
def test_batched_chains_equal_single_applications(mode10, precision, tol):
    """Several chains with different flags in one pass give what separate calls give."""
    import torch
    from tdscha_amd import _lib as L
    ens = fixtures.random_ensemble(33, 60, 90, 7)
    dev = _dev_ens(ens, precisions=("fp64", precision), s_batch=3)   # fewer slots than chains
    rng = np.random.default_rng(34)
    n_psi = orc.psi_len(60)
    psis = [torch.as_tensor(rng.standard_normal(n_psi), device=dev.device) for _ in range(5)]
    flags = [0, L.OP_TRANSPOSE, L.OP_WIGNER, L.OP_REVERSE, L.OP_TRANSPOSE | L.OP_REVERSE]
    shifts = [0.0, 1e-6, 0.0, 2e-6, 0.0]
    outs = dev.apply_full_L_batched(psis, flags, shifts, precision=precision)
    for p, f, s, o in zip(psis, flags, shifts, outs):
        ref = dev.apply_full_L(p, flags=f, shift=s, precision=precision)
        err = (torch.linalg.norm(o - ref) / torch.linalg.norm(ref)).item()
        assert err < tol, (f, err)
    # and against the oracle for the fp64 chains
    if precision == "fp64":
        for p, f, s, o in zip(psis, flags, shifts, outs):
            kw = dict(transpose=bool(f & L.OP_TRANSPOSE), use_wigner=bool(f & L.OP_WIGNER),
                      reverse_L=bool(f & L.OP_REVERSE), shift_value=s)
            ref = orc.apply_full_L(ens, p.cpu().numpy(), **kw)
            assert np.linalg.norm(o.cpu().numpy() - ref) / np.linalg.norm(ref) < TOL64


def test_run_FT_coefficients_and_kat(mode10):
    """a/b/c of the reference's 10-step run and the KAT c_4 = 6.59744344e-07
    (tests/test_julia/test_lanczos.py:60-62)."""
    st = fixtures.load_npz("lanczos_std.npz")
    lz = _lanczos_from(mode10)
    lz.prepare_mode(10)
    lz.run_FT(10, verbose=False)
    a, b, c = map(np.asarray, (lz.a_coeffs, lz.b_coeffs, lz.c_coeffs))
    assert abs(c[4] - 6.59744344e-07) / 6.59744344e-07 < 1e-6
    # beyond step 7 the unreorthogonalised recurrence amplifies rounding (the fixture's own
    # b_9 moves by 1e-2 under a change of summation order), so compare the first 8 steps
    assert np.max(np.abs(a - st["a"]) / np.abs(st["a"])) < 1e-8
    assert np.max(np.abs(b[:8] - st["b"][:8]) / np.abs(st["b"][:8])) < 1e-6
    assert np.max(np.abs(c[:8] - st["c"][:8]) / np.abs(st["c"][:8])) < 1e-6


def test_run_FT_wigner_symmetric(mode10):
    st = fixtures.load_npz("lanczos_wigner.npz")
    lz = _lanczos_from(mode10)
    lz.use_wigner = True
    lz.psi = st["Q"][0].copy()
    lz.run_FT(10, verbose=False, run_simm=True)
    a, b = np.asarray(lz.a_coeffs), np.asarray(lz.b_coeffs)
    assert np.max(np.abs(a - st["a"]) / np.abs(st["a"])) < 1e-8
    assert np.max(np.abs(b[:8] - st["b"][:8]) / np.abs(st["b"][:8])) < 1e-6


def test_continued_fraction_device_matches_oracle(mode10):
    st = fixtures.load_npz("lanczos_std.npz")
    from tdscha_amd.DynamicalLanczos import continued_fraction
    import torch
    w = np.linspace(0, 3e-3, 257)
    for term in (True, False):
        g = continued_fraction(st["a"], st["b"], st["c"], w, use_terminator=term, smearing=1e-5,
                               device=torch.device("cuda"))
        g_o = orc.green_function_continued_fraction(st["a"], st["b"], st["c"], w,
                                                    use_terminator=term, smearing=1e-5)
        assert np.allclose(g, g_o, rtol=1e-12, atol=0)
