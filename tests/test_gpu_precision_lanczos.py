"""Lanczos with the reduced-precision contractions against the FP64 device path.

Workload: synthetic rock-salt 2x2x2 supercell (16 atoms, 45 modes), 2000 configs, the full
384-image symmetry set, T = 300 K, IR perturbation along x with Z* = +-2. The bond
anharmonicity is 20x the benchmark model's (sigma 2 eV/A^3 and 1 eV/A^4): with the weaker
model the first Lanczos residual (3e-13) is below the reference's absolute convergence
threshold of 1e-12 (Py:121, Py:6144-6153) and the recurrence stops at once. Criteria: Lanczos
a / b / c within 1e-5 relative and the peak positions of -Im G(w) (0-300 cm^-1, 1 cm^-1
smearing) within 0.1 cm^-1 of FP64, each beside the ensemble's own stochastic spread
(the same Lanczos on two disjoint halves of the configurations).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_STEPS = 50


@pytest.fixture(scope="module")
def rocksalt2():
    import torch
    from tdscha_amd import synthetic as sy
    return sy.make_ensemble(2, 2000, T=300.0, device=torch.device("cuda"), sigma3=2.0, sigma4=1.0)


def _run(ens, mode, cfg=slice(None), steps=N_STEPS, full_gs=True, start="ir"):
    from tdscha_amd import DynamicalLanczos as DL
    lz = DL.Lanczos(mode=mode)
    lz.set_ensemble(ens.X[cfg], ens.Y[cfg], ens.w, ens.rho[cfg], ens.T, ens.pols, ens.m)
    lz.set_symmetries(ens.symmetries, ens.N_degeneracy, ens.degenerate_space, ens.block_id)
    lz.init()
    lz.verbose = False
    lz.psi[:] = 0
    if start == "ir":
        lz.psi[:len(ens.w)] = ens.ir_perturbation(0)
    else:
        lz.psi[:] = np.random.default_rng(3).standard_normal(len(lz.psi))
    lz.perturbation_modulus = float(lz.psi.dot(lz.psi))
    lz.run_FT(steps, verbose=False, full_gs=full_gs)
    return lz


def _peaks(lz):
    from tdscha_amd.synthetic import RY_TO_CM
    w_cm = np.linspace(0.0, 300.0, 2000)
    g = lz.get_green_function_continued_fraction(w_cm / RY_TO_CM, use_terminator=False,
                                                 smearing=1.0 / RY_TO_CM)
    s = -np.imag(g)
    idx = [i for i in range(1, len(s) - 1) if s[i] > s[i - 1] and s[i] >= s[i + 1]
           and s[i] > 0.05 * s.max()]
    # parabolic refinement of each maximum
    out = []
    for i in idx:
        y0, y1, y2 = s[i - 1], s[i], s[i + 1]
        d = 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
        out.append(w_cm[i] + d * (w_cm[1] - w_cm[0]))
    return np.array(out), s


@pytest.mark.parametrize("start", ["ir", "random"])
@pytest.mark.parametrize("mode_name", ["MODE_GPU_FP32", "MODE_GPU_BF16X3"])
def test_reduced_precision_lanczos(rocksalt2, mode_name, start):
    """The IR vector spans a small symmetric Krylov space (the recurrence terminates after a
    few steps); a random vector exercises all N_STEPS."""
    from tdscha_amd import DynamicalLanczos as DL
    ref = _run(rocksalt2, DL.MODE_GPU_FP64, start=start)
    lp = _run(rocksalt2, getattr(DL, mode_name), start=start)
    assert len(ref.b_coeffs) >= 5, "the FP64 Lanczos stopped after %d steps" % len(ref.b_coeffs)
    a0, b0, c0 = map(np.asarray, (ref.a_coeffs, ref.b_coeffs, ref.c_coeffs))
    a1, b1, c1 = map(np.asarray, (lp.a_coeffs, lp.b_coeffs, lp.c_coeffs))
    k = min(len(a0), len(a1))
    ea = np.max(np.abs(a1[:k] - a0[:k]) / np.abs(a0[:k]))
    kb = min(len(b0), len(b1))
    eb = np.max(np.abs(b1[:kb] - b0[:kb]) / np.abs(b0[:kb]))
    ec = np.max(np.abs(c1[:kb] - c0[:kb]) / np.abs(c0[:kb]))
    p0, _ = _peaks(ref)
    p1, _ = _peaks(lp)
    # stochastic spread: two halves of the ensemble in FP64
    h1 = _run(rocksalt2, DL.MODE_GPU_FP64, slice(0, 1000), start=start)
    ph, _ = _peaks(h1)
    m = min(len(p0), len(p1))
    dpk = np.max(np.abs(p1[:m] - p0[:m])) if m else 0.0
    spread = np.max(np.abs(ph[:min(len(ph), len(p0))] - p0[:min(len(ph), len(p0))])) \
        if len(ph) else float("nan")
    print("%s %s (%d steps): max rel err a %.2e b %.2e c %.2e | peaks fp64 %s | max peak shift %.4f cm-1 "
          "| half-ensemble peak spread %.3f cm-1" % (mode_name, start, k, ea, eb, ec,
                                                    np.round(p0, 3),
                                                    dpk, spread))
    assert len(p1) == len(p0)
    assert ea < 1e-5 and eb < 1e-5 and ec < 1e-5
    assert dpk < 0.1
