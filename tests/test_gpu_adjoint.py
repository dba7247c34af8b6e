"""The device force adjoint (tdscha_ensemble_force_grad) against the oracle's adjoint, which
is itself pinned by finite differences (tests/test_oracle_adjoint.py)."""
import numpy as np
import pytest

import fixtures
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _dev_ens(ens, s_batch=0):
    from tdscha_amd.ensemble import DeviceEnsemble, SymmetryBlocks
    s = ens.sym
    return DeviceEnsemble(ens.X, ens.Y, ens.w, ens.rho, ens.T,
                          SymmetryBlocks(s.symmetries, s.N_degeneracy, s.degenerate_space,
                                         s.block_id), s_batch=s_batch)


def _oracle_grad(ens, u, v, flags):
    from tdscha_amd import _lib as L
    tr, wig = bool(flags & L.OP_TRANSPOSE), bool(flags & L.OP_WIGNER)
    if flags & L.OP_IGNORE_V3:
        return np.zeros_like(ens.Y)
    R1, Y1 = orc.anharmonic_pre(v, ens.w, ens.T, tr, wig)
    uR, UD = orc.post_seeds(u, ens.w, ens.T, tr, wig, bool(flags & L.OP_REVERSE))
    return orc.force_grad(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, ens.sym, uR, UD,
                          ky=0.0 if flags & L.OP_IGNORE_V4 else 1.0)


def _flags(name):
    from tdscha_amd import _lib as L
    return {"std": 0, "transpose": L.OP_TRANSPOSE, "wigner": L.OP_WIGNER,
            "reverse": L.OP_REVERSE, "no_v4": L.OP_IGNORE_V4, "no_v3": L.OP_IGNORE_V3}[name]


@pytest.mark.parametrize("name", ["std", "transpose", "wigner", "reverse", "no_v4", "no_v3"])
def test_force_grad_matches_oracle(name):
    ens = fixtures.random_ensemble(81, 20, 30, 4)
    rng = np.random.default_rng(82)
    u, v = (rng.standard_normal(orc.psi_len(20)) for _ in range(2))
    fl = _flags(name)
    ref = _oracle_grad(ens, u, v, fl)
    g = _dev_ens(ens).force_grad([u], [v], [fl]).cpu().numpy()
    scale = max(np.max(np.abs(ref)), 1e-300)
    err = np.max(np.abs(g - ref)) / scale
    print(name, err)
    if name == "no_v3":
        assert np.all(g == 0)
    else:
        assert err < TOL


@pytest.mark.parametrize("max_block,seed", [(6, 83), (32, 84), (50, 85), (90, 86)])
def test_force_grad_block_widths(max_block, seed):
    """32- and 64-wide group tiles and the one-sided rotations of blocks wider than 64."""
    ens = fixtures.random_ensemble(seed, 100, 40, 3, max_block=max_block)
    rng = np.random.default_rng(seed + 1)
    u, v = (rng.standard_normal(orc.psi_len(100)) for _ in range(2))
    ref = _oracle_grad(ens, u, v, 0)
    g = _dev_ens(ens).force_grad([u], [v], [0]).cpu().numpy()
    assert np.max(np.abs(g - ref)) / np.max(np.abs(ref)) < TOL


def test_pairs_add_and_accumulate():
    """Several pairs in one call (split over passes with a small s_batch) give the sum of
    single-pair calls; accumulate adds onto the output."""
    ens = fixtures.random_ensemble(87, 24, 35, 5)
    dev = _dev_ens(ens, s_batch=3)
    rng = np.random.default_rng(88)
    us = [rng.standard_normal(orc.psi_len(24)) for _ in range(3)]
    vs = [rng.standard_normal(orc.psi_len(24)) for _ in range(3)]
    fl = [0, _flags("transpose"), _flags("wigner")]
    both = dev.force_grad(us, vs, fl)
    single = sum(dev.force_grad([u], [v], [f]) for u, v, f in zip(us, vs, fl))
    assert float((both - single).abs().max() / single.abs().max()) < 1e-12
    again = dev.force_grad(us, vs, fl, out=both.clone(), accumulate=True)
    assert float((again - 2 * both).abs().max() / both.abs().max()) < 1e-14


def test_euler_identity_on_the_device_operator():
    """The anharmonic part of L is linear in Y: u . A(v; Y) = <dJ/dY, Y>."""
    import torch
    from tdscha_amd import _lib as L
    ens = fixtures.random_ensemble(89, 30, 50, 4)
    dev = _dev_ens(ens)
    rng = np.random.default_rng(90)
    u, v = (rng.standard_normal(orc.psi_len(30)) for _ in range(2))
    Av = dev.apply_full_L(torch.as_tensor(v, device=dev.device),
                          flags=L.OP_IGNORE_HARMONIC).cpu().numpy()
    g = dev.force_grad([u], [v], [0]).cpu().numpy()
    J = u.dot(Av)
    assert abs(J - np.sum(g * ens.Y)) < 1e-11 * np.sum(np.abs(g * ens.Y))


def _dev_ens32(ens, s_batch=0):
    from tdscha_amd.ensemble import DeviceEnsemble, SymmetryBlocks
    s = ens.sym
    return DeviceEnsemble(ens.X, ens.Y, ens.w, ens.rho, ens.T,
                          SymmetryBlocks(s.symmetries, s.N_degeneracy, s.degenerate_space,
                                         s.block_id), precisions=("fp64", "fp32"),
                          s_batch=s_batch)


# FP32 adjoint: the three contractions of a pass in FP32 MFMA with FP32 accumulation (the
# operands carry 2^-24 relative rounding; X, the rotated seeds and the rotated perturbation
# enter products of ~n terms): measured against the FP64 oracle adjoint, relative to its
# largest entry.
TOL32 = 2e-5


@pytest.mark.parametrize("name", ["std", "transpose", "wigner", "reverse", "no_v4"])
def test_fp32_force_grad_matches_oracle(name):
    ens = fixtures.random_ensemble(81, 20, 30, 4)
    rng = np.random.default_rng(82)
    u, v = (rng.standard_normal(orc.psi_len(20)) for _ in range(2))
    fl = _flags(name)
    ref = _oracle_grad(ens, u, v, fl)
    dev = _dev_ens32(ens)
    g = dev.force_grad([u], [v], [fl], precision="fp32").cpu().numpy()
    g64 = dev.force_grad([u], [v], [fl]).cpu().numpy()
    err = np.max(np.abs(g - ref)) / np.max(np.abs(ref))
    err64 = np.max(np.abs(g64 - ref)) / np.max(np.abs(ref))
    print("%s: fp32 %.2e, fp64 %.2e" % (name, err, err64))
    assert err < TOL32 and err64 < TOL


@pytest.mark.parametrize("max_block,seed", [(32, 84), (90, 86)])
def test_fp32_force_grad_block_widths(max_block, seed):
    ens = fixtures.random_ensemble(seed, 100, 40, 3, max_block=max_block)
    rng = np.random.default_rng(seed + 1)
    u, v = (rng.standard_normal(orc.psi_len(100)) for _ in range(2))
    ref = _oracle_grad(ens, u, v, 0)
    g = _dev_ens32(ens).force_grad([u], [v], [0], precision="fp32").cpu().numpy()
    err = np.max(np.abs(g - ref)) / np.max(np.abs(ref))
    print("max_block %d: fp32 %.2e" % (max_block, err))
    assert err < TOL32


def test_fp32_force_grad_needs_the_fp32_copies():
    ens = fixtures.random_ensemble(87, 12, 10, 2)
    u = np.ones(orc.psi_len(12))
    with pytest.raises(ValueError):
        _dev_ens(ens).force_grad([u], [u], [0], precision="fp32")
