"""GPU tests at the sizes of the benchmark configurations (synthetic rock-salt ensembles,
tdscha_amd.synthetic: the crystal, force model and sampling of the workload definitions).

  C2  4x4x4 (128 atoms, 381 modes), 10000 configurations, log-normal weights (sigma 0.2),
      all 3072 symmetry images, FP64:
        * kernel parity against the oracle on a slice of 6 configurations x all 3072 images
          (the oracle's cost is linear in the images), through the device ensemble and
          through the drop-in GetPerturbAverageSym entry point: plain relative 1e-10;
        * 3 IR polarizations, 200 Lanczos steps each with full re-biorthogonalisation, on the
          zone-centre operator: by the cubic symmetry of the symmetrised ensemble the three
          chains have the same a_k and b_k c_k (the biconjugate recurrence matches moments
          p_0 . L^j q_0, which the rotation between the polarizations leaves unchanged; b_k
          alone depends on the packed metric), so their spectra agree; the first 10 steps
          equal the full operator's.
  C3  6x6x6 (432 atoms, 1293 modes), 20000 configurations, 10368 images: IR Lanczos, 300
      steps with full re-biorthogonalisation on the zone-centre operator, in BF16x3 against
      FP64: peak positions of -Im G over 0-300 cm^-1 (2000 points, 1 cm^-1 smearing) within
      0.1 cm^-1, max |Im G - Im G_FP64| below the stochastic spread (the same FP64 run on
      half of the configurations against the whole).
  C4  12 chains (3 IR, 6 Raman-like, 3 random) on each of 2x2x2 / 4x4x4 / 5x5x5 with
      4000 / 10000 / 15000 configurations, BF16x3, autotuned, one graph per step for the
      three systems: two steps of the group equal two steps of each member run alone, bit
      for bit.
  C5  4x4x4, 381 modes, 10000 configurations: gradient of sum_k Im G(w_k) over 64 points in
      50-250 cm^-1 after 150 steps (zone-centre operator) against central differences, in
      FP64 and with FP32 contractions and FP32 accumulation (forward and adjoint) against the
      FP64 gradient.
The Lanczos tests scale the model's cubic and quartic terms by 20: with the benchmark model
the IR vector couples to the two-phonon sector below the reference's absolute stop
threshold |b| < 1e-12 (Py:6144-6153) and the recurrence stops after one step.
"""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ANH = 20.0
W_CM = np.linspace(0.0, 300.0, 2000)


def _cuda():
    import torch
    return torch.device("cuda", 0)


def _sym(ens):
    from tdscha_amd.ensemble import SymmetryBlocks
    return SymmetryBlocks(ens.symmetries, ens.N_degeneracy, ens.degenerate_space, ens.block_id)


def _lanczos(ens, mode, gamma=True, rows=None):
    from tdscha_amd import DynamicalLanczos as DL
    from tdscha_amd import synthetic as sy
    lz = DL.Lanczos(mode=mode)
    lz.verbose = False
    r = slice(None) if rows is None else rows
    lz.set_ensemble(ens.X[r], ens.Y[r], ens.w, ens.rho[r], ens.T, pols=ens.pols, m=ens.m)
    lz.set_symmetries(ens.symmetries, ens.N_degeneracy, ens.degenerate_space, ens.block_id)
    if gamma:
        lz.sc_symmetries = sy.space_group(ens.crystal)
        lz.gamma_sector = True
    lz.init()
    return lz


def _ir(lz, direction):
    pol = np.zeros(3)
    pol[direction] = 1.0
    lz.prepare_ir(effective_charges=np.stack([2 * np.eye(3), -2 * np.eye(3)]), pol_vec=pol)


def _spectrum(lz):
    from tdscha_amd.synthetic import RY_TO_CM
    g = lz.get_green_function_continued_fraction(W_CM / RY_TO_CM, use_terminator=False,
                                                 smearing=1.0 / RY_TO_CM)
    return -np.imag(np.asarray(g))


def _peaks(s):
    idx = [i for i in range(1, len(s) - 1) if s[i] > s[i - 1] and s[i] >= s[i + 1]
           and s[i] > 0.05 * s.max()]
    out = []
    for i in idx:
        y0, y1, y2 = s[i - 1], s[i], s[i + 1]
        out.append(W_CM[i] + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2) * (W_CM[1] - W_CM[0]))
    return np.array(out)


def _rel(x, y, k=None):
    x, y = np.asarray(x)[:k], np.asarray(y)[:k]
    return float(np.abs(x - y).max() / np.abs(y).max())


# ---------------------------------------------------------------------------- C2

@pytest.fixture(scope="module")
def c2():
    from tdscha_amd import synthetic as sy
    print("case: C2 ensemble, 4x4x4, 10000 configurations, log-normal weights", flush=True)
    return sy.make_ensemble(4, 10000, T=300.0, weights="lognormal", device=_cuda())


@pytest.fixture(scope="module")
def c2_anh():
    from tdscha_amd import synthetic as sy
    print("case: C2 ensemble with anharmonic terms x20", flush=True)
    return sy.make_ensemble(4, 10000, T=300.0, weights="lognormal", device=_cuda(),
                            sigma3=0.1 * ANH, sigma4=0.05 * ANH)


def test_c2_kernel_parity_on_a_slice_of_all_images(c2):
    import fixtures
    from tdscha_amd.ensemble import DeviceEnsemble
    from tdscha_amd.sscha_HP_odd import GetPerturbAverageSym
    print("case: C2 kernel parity, 6 configurations x 3072 images", flush=True)
    e = c2
    n = len(e.w)
    assert n == 381 and e.n_syms == 3072
    rows = slice(0, 6)
    X, Y, rho = e.X[rows], e.Y[rows], e.rho[rows]
    sym = orc.SymBlocks(e.symmetries, e.N_degeneracy, e.degenerate_space, e.block_id)
    R1, Y1 = fixtures.random_perturbation(91, n)
    f_o, d_o = orc.perturb_average_sym(X, Y, e.w, rho, R1, Y1, e.T, 1, sym)
    dev = DeviceEnsemble(X, Y, e.w, rho, e.T, _sym(e))
    f, d = (t.cpu().numpy() for t in dev.apply(R1, Y1))
    ef = np.linalg.norm(f - f_o) / np.linalg.norm(f_o)
    ed = np.linalg.norm(d - d_o) / np.linalg.norm(d_o)
    print("device ensemble: f %.2e d2v %.2e" % (ef, ed))
    assert ef < 1e-10 and ed < 1e-10
    f2 = np.zeros(n)
    d2 = np.zeros((n, n))
    GetPerturbAverageSym(X, Y, e.w, rho, R1, Y1, e.T, 1, e.n_syms, e.symmetries,
                         e.N_degeneracy, e.degenerate_space, e.block_id, f2, d2)
    assert np.linalg.norm(f2 - f_o) / np.linalg.norm(f_o) < 1e-10
    assert np.linalg.norm(d2 - d_o) / np.linalg.norm(d_o) < 1e-10
    # the whole operator on a random vector
    oens = orc.Ensemble(X, Y, e.w, rho, e.T, sym)
    psi = np.random.default_rng(92).standard_normal(orc.psi_len(n))
    for transpose in (False, True):
        from tdscha_amd import _lib as L
        out = dev.apply_full_L(psi, flags=L.OP_TRANSPOSE if transpose else 0).cpu().numpy()
        ref = orc.apply_full_L(oens, psi, transpose=transpose)
        err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
        print("full L (transpose=%s): %.2e" % (transpose, err))
        assert err < 1e-10
    dev.close()


@pytest.fixture(scope="module")
def c2_runs(c2_anh):
    from tdscha_amd import DynamicalLanczos as DL
    runs = []
    for d in range(3):
        print("case: C2 IR polarization %d, 200 steps, FP64, zone-centre operator" % d,
              flush=True)
        lz = _lanczos(c2_anh, DL.MODE_GPU_FP64)
        _ir(lz, d)
        lz.run_FT(200, verbose=False, full_gs=True)
        runs.append((np.asarray(lz.a_coeffs), np.asarray(lz.b_coeffs), np.asarray(lz.c_coeffs),
                     _spectrum(lz)))
        lz.device_ensemble.close()
    return runs


def test_c2_three_ir_polarizations_agree(c2_runs):
    (a0, b0, c0, s0) = c2_runs[0]
    print("steps per polarization: %s" % [len(r[0]) for r in c2_runs])
    assert len(a0) > 20
    for d in (1, 2):
        a, b, c, s = c2_runs[d]
        k = min(20, len(a), len(a0))
        ea = _rel(a, a0, k)
        ebc = _rel(b[:k - 1] * c[:k - 1], b0[:k - 1] * c0[:k - 1])
        es = float(np.abs(s - s0).max() / np.abs(s0).max())
        print("polarization %d against x: a %.1e, b c %.1e (first %d steps), Im G %.1e"
              % (d, ea, ebc, k, es))
        assert ea < 1e-8 and ebc < 1e-8 and es < 1e-5


def test_c2_zone_centre_equals_the_full_operator(c2_anh):
    from tdscha_amd import DynamicalLanczos as DL
    print("case: C2 IR x, 10 steps on the full operator against the zone-centre one",
          flush=True)
    res = []
    for gamma in (False, True):
        lz = _lanczos(c2_anh, DL.MODE_GPU_FP64, gamma=gamma)
        _ir(lz, 0)
        lz.run_FT(10, verbose=False, full_gs=True)
        res.append((np.asarray(lz.a_coeffs), np.asarray(lz.b_coeffs), np.asarray(lz.c_coeffs)))
        lz.device_ensemble.close()
    for x, y in zip(res[1], res[0]):
        assert len(x) == len(y)
        print("relative difference %.2e" % _rel(x, y))
        assert _rel(x, y) < 1e-10


# ---------------------------------------------------------------------------- C3

@pytest.fixture(scope="module")
def c3():
    from tdscha_amd import synthetic as sy
    print("case: C3 ensemble, 6x6x6, 20000 configurations, anharmonic terms x20", flush=True)
    return sy.make_ensemble(6, 20000, T=300.0, device=_cuda(), sigma3=0.1 * ANH,
                            sigma4=0.05 * ANH)


@pytest.fixture(scope="module")
def c3_fp64(c3):
    from tdscha_amd import DynamicalLanczos as DL
    print("case: C3 IR, 300 steps, FP64", flush=True)
    lz = _lanczos(c3, DL.MODE_GPU_FP64)
    _ir(lz, 0)
    lz.run_FT(300, verbose=False, full_gs=True)
    out = (np.asarray(lz.a_coeffs), _spectrum(lz))
    lz.device_ensemble.close()
    return out


def test_c3_fp64_reference_run(c3, c3_fp64):
    n = len(c3.w)
    assert n == 1293 and c3.n_syms == 10368
    a, s = c3_fp64
    print("FP64: %d steps, peaks %s cm-1" % (len(a), np.round(_peaks(s), 3)))
    assert len(a) > 50 and len(_peaks(s)) >= 1


def test_c3_bf16x3_spectrum_against_fp64(c3, c3_fp64):
    from tdscha_amd import DynamicalLanczos as DL
    print("case: C3 IR, 300 steps, BF16x3", flush=True)
    lz = _lanczos(c3, DL.MODE_GPU_BF16X3)
    _ir(lz, 0)
    lz.run_FT(300, verbose=False, full_gs=True)
    s1 = _spectrum(lz)
    lz.device_ensemble.close()
    a0, s0 = c3_fp64
    print("case: C3 IR, 300 steps, FP64 on the first 10000 configurations", flush=True)
    lh = _lanczos(c3, DL.MODE_GPU_FP64, rows=slice(0, 10000))
    _ir(lh, 0)
    lh.run_FT(300, verbose=False, full_gs=True)
    sh = _spectrum(lh)
    lh.device_ensemble.close()
    p0, p1, ph = _peaks(s0), _peaks(s1), _peaks(sh)
    shift = float(max(np.abs(p0[:, None] - p1[None]).min(axis=1)))
    g_err = float(np.abs(s1 - s0).max() / s0.max())
    g_spread = float(np.abs(sh - s0).max() / s0.max())
    p_spread = float(max(np.abs(p0[:, None] - ph[None]).min(axis=1))) if len(ph) else None
    print("BF16x3 against FP64: peak shift %.2e cm-1, Im G %.2e; half ensemble: peak shift "
          "%s cm-1, Im G %.2e" % (shift, g_err, p_spread, g_spread))
    assert len(p1) == len(p0)
    assert shift < 0.1
    assert g_err < g_spread


# ---------------------------------------------------------------------------- C4

def test_c4_group_of_uneven_systems_equals_members_alone():
    import torch
    from tdscha_amd import batched
    from tdscha_amd.lanczos_device import DeviceLanczos
    print("case: C4 group, 2x2x2 / 4x4x4 / 5x5x5 with 4000 / 10000 / 15000 configurations, "
          "12 chains each, BF16x3, autotuned", flush=True)
    group, members, ensembles, info = batched.build(precision="bf16x3", max_steps=2,
                                                    log=lambda *a: print(*a, flush=True))
    assert [(i["n_modes"], i["n_configs"], i["n_syms"], i["chains"]) for i in info] == [
        (45, 4000, 384, 12), (381, 10000, 3072, 12), (747, 15000, 6000, 12)]
    group.run(2)
    torch.cuda.synchronize()
    print("group: 2 steps done", flush=True)
    for m, d in zip(members, ensembles):
        alone = DeviceLanczos(d, m.nc, "bf16x3", max_steps=2)
        alone.start(m._psi0)
        alone.run(2)
        for ch in range(m.nc):
            for x, y in zip(m.coefficients(ch, truncated=False),
                            alone.coefficients(ch, truncated=False)):
                assert len(x) == 2 and np.all(np.isfinite(x))
                assert np.array_equal(x, y)
        alone.close()
        print("member n=%d: equal alone" % d.n, flush=True)
    group.close()
    for m in members:
        m.close()
    for d in ensembles:
        d.close()


def test_c4_zone_centre_group_equals_members_alone():
    """The same three systems with the nine IR and Raman chains of each on the zone-centre
    operator (translation average inside the captured step) and the three random chains on
    the full image set: six group members, two steps, equal to each member alone."""
    import torch
    from tdscha_amd import batched
    from tdscha_amd.lanczos_device import DeviceLanczos
    print("case: C4 group with zone-centre members, BF16x3, autotuned", flush=True)
    group, members, ensembles, info = batched.build(precision="bf16x3", max_steps=2,
                                                    log=lambda *a: print(*a, flush=True),
                                                    zone_centre=True)
    assert [(i["n_syms"], i["chains"], i["operator"]) for i in info] == [
        (48, 9, "zone_centre"), (384, 3, "full"), (48, 9, "zone_centre"), (3072, 3, "full"),
        (48, 9, "zone_centre"), (6000, 3, "full")]
    group.run(2)
    torch.cuda.synchronize()
    print("group: 2 steps done", flush=True)
    for m, d in zip(members, ensembles):
        alone = DeviceLanczos(d, m.nc, "bf16x3", max_steps=2)
        alone.start(m._psi0)
        alone.run(2)
        for ch in range(m.nc):
            for x, y in zip(m.coefficients(ch, truncated=False),
                            alone.coefficients(ch, truncated=False)):
                assert len(x) == 2 and np.all(np.isfinite(x))
                assert np.array_equal(x, y)
        alone.close()
    group.close()
    for m in members:
        m.close()
    for d in ensembles:
        d.close()


# ---------------------------------------------------------------------------- C5

def test_c5_gradients_against_central_differences(c2_anh):
    """J = sum_k Im G(w_k), 64 points in 50-250 cm^-1, 1 cm^-1 smearing, 150 steps, FP64, on
    the zone-centre operator. Directions: a translation-invariant change of the effective
    charges (psi0), and the forces (Y) along the adjoint gradient itself. Step h = 1e-4
    relative. After 150 steps J itself carries rounding noise of ~7.6e-7 relative (measured
    between two summation orders), ~0.15 here: a random direction in the 3.8 million force
    entries changes J by only ~0.02 at h = 1e-4 (measured: adjoint -94.9, quotient -51.1,
    all noise), so the forces are checked along the gradient, where the change is ~2000x
    larger. The quotients then carry <= 1 % noise: both are checked to 3 %."""
    from tdscha_amd import synthetic as sy
    from tdscha_amd.gradients import SpectralProblem, spectral_gradients
    print("case: C5 gradients, 150 steps, zone-centre operator", flush=True)
    e = c2_anh
    n = len(e.w)
    rots, irt = sy.space_group(e.crystal)
    omegas = np.linspace(50.0, 250.0, 64) / sy.RY_TO_CM
    prob = SpectralProblem(e.X, e.w, e.rho, e.T, _sym(e), omegas, 1.0 / sy.RY_TO_CM, 150,
                           gamma_sector=(rots, irt, e.pols))
    psi0 = np.zeros(orc.psi_len(n))
    psi0[:n] = e.ir_perturbation(0)
    J, gp, gy = spectral_gradients(prob, psi0, e.Y)
    print("J = %.6e" % J, flush=True)
    rng = np.random.default_rng(93)
    # psi0 direction: effective charges changed by a per-species, translation-invariant amount
    nat = e.crystal.nat
    zeta = rng.standard_normal((2, 3))
    v = zeta[e.crystal.species]                        # (nat, 3), the same in every cell
    dpsi = np.zeros_like(psi0)
    dpsi[:n] = np.einsum("a,a,ab->b", 1 / np.sqrt(e.m), v.ravel(), e.pols)
    dpsi *= np.linalg.norm(psi0) / np.linalg.norm(dpsi)
    dY = gy * (np.linalg.norm(e.Y) / np.linalg.norm(gy))
    h = 1e-4
    for name, fwd, grad_dot in (
            ("psi0", lambda t: spectral_value(prob, psi0 + t * dpsi, e.Y), float(gp @ dpsi)),
            ("Y", lambda t: spectral_value(prob, psi0, e.Y + t * dY), float(np.sum(gy * dY)))):
        fd = (fwd(h) - fwd(-h)) / (2 * h)
        err = abs(fd - grad_dot) / abs(fd)
        print("%s direction: adjoint %.6e, central difference %.6e, relative %.2e"
              % (name, grad_dot, fd, err), flush=True)
        assert err < 3e-2, name


def spectral_value(prob, psi0, Y):
    lz = prob.lanczos(psi0, Y)
    J = prob.value(lz)
    lz.device_ensemble.close()
    return J


def test_c5_fp32_gradients_against_fp64(c2_anh):
    """The C5 gradients with FP32 MFMA contractions and FP32 accumulation in the forward
    Lanczos and in the force adjoint (MODE_GPU_FP32), against the FP64 run. The recurrence
    amplifies any operator change over 150 steps (between two FP64 summation orders: J
    7.6e-7, dJ/dY 0.02 %, dJ/dpsi0 3.8 % of the largest entry), so the FP32 run is held to
    J within 1e-4, and to directional derivatives along the FP64 gradients (psi0 and Y)
    within 5 %, which its own central differences must also match."""
    from tdscha_amd import DynamicalLanczos as DL
    from tdscha_amd import synthetic as sy
    from tdscha_amd.gradients import SpectralProblem, spectral_gradients
    print("case: C5 gradients, FP32 forward and adjoint against FP64", flush=True)
    e = c2_anh
    n = len(e.w)
    rots, irt = sy.space_group(e.crystal)
    omegas = np.linspace(50.0, 250.0, 64) / sy.RY_TO_CM
    psi0 = np.zeros(orc.psi_len(n))
    psi0[:n] = e.ir_perturbation(0)
    res = {}
    probs = {}
    for name, mode in (("fp64", DL.MODE_GPU_FP64), ("fp32", DL.MODE_GPU_FP32)):
        probs[name] = SpectralProblem(e.X, e.w, e.rho, e.T, _sym(e), omegas, 1.0 / sy.RY_TO_CM,
                                      150, mode=mode, gamma_sector=(rots, irt, e.pols))
        res[name] = spectral_gradients(probs[name], psi0, e.Y)
        print("%s: J = %.8e" % (name, res[name][0]), flush=True)
    J0, gp0, gy0 = res["fp64"]
    J1, gp1, gy1 = res["fp32"]
    dJ = abs(J1 - J0) / abs(J0)
    up = gp0 / np.linalg.norm(gp0)
    uy = gy0 / np.linalg.norm(gy0)
    dp = abs(gp1 @ up - gp0 @ up) / abs(gp0 @ up)
    dy = abs(np.sum(gy1 * uy) - np.sum(gy0 * uy)) / abs(np.sum(gy0 * uy))
    ey = np.abs(gy1 - gy0).max() / np.abs(gy0).max()
    print("FP32 against FP64: J %.2e, d/dpsi0 along the FP64 gradient %.2e, d/dY along it "
          "%.2e, max |dJ/dY| difference %.2e of the largest entry" % (dJ, dp, dy, ey))
    assert dJ < 1e-4 and dp < 5e-2 and dy < 5e-2
    # the FP32 gradient against the FP32 run's own central differences along uY
    h = 1e-4
    dY = uy * np.linalg.norm(e.Y)
    fd = (spectral_value(probs["fp32"], psi0, e.Y + h * dY) -
          spectral_value(probs["fp32"], psi0, e.Y - h * dY)) / (2 * h)
    adj = float(np.sum(gy1 * dY))
    print("FP32 Y direction: adjoint %.6e, central difference %.6e, relative %.2e"
          % (adj, fd, abs(fd - adj) / abs(fd)))
    assert abs(fd - adj) / abs(fd) < 5e-2
