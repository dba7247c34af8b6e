"""Device construction of the mode-basis symmetry blocks (tdscha_symmetries_on_modes, the
blocks of Lanczos.prepare_symmetrization, Modules/DynamicalLanczos.py:686-778) against the
oracle, and a Lanczos L-application with generated blocks against one with the reference's."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def test_device_blocks_match_oracle_and_reference_on_snte():
    import test_symmetry as ts
    from tdscha_amd import symmetry as sy
    print("case: SnTe 2x2x2, 384 operations", flush=True)
    rots, trans, cen, cell, species, d = ts.snte_space_group()
    irt = sy.atom_permutations(cen, cell, rots, trans, species)
    starts = sy.degenerate_blocks(d["w"])
    dev = sy.symmetries_on_modes(rots, irt, d["pols"], starts, device=0)
    ora = orc.symmetries_on_modes(rots, irt, d["pols"], starts)
    g = ts.load("mode10_ensemble.npz")
    for b, (S, O) in enumerate(zip(dev, ora)):
        assert np.abs(S - O).max() < 1e-13, b
        assert np.abs(ts.canonical(S) - ts.canonical(g["sym_%d" % b])).max() < 1e-12, b


def test_device_blocks_random_shapes():
    """Non-group operations and blocks of 1 to 100 modes (one pass and several passes)."""
    rng = np.random.default_rng(11)
    nat, n_sym = 40, 37
    sizes = [1, 2, 3, 5, 17, 48, 34]
    n = sum(sizes)
    print("case: nat=%d n_sym=%d blocks=%s" % (nat, n_sym, sizes), flush=True)
    from tdscha_amd import symmetry as sy
    rots = np.array([np.linalg.qr(rng.standard_normal((3, 3)))[0] for _ in range(n_sym)])
    irt = np.array([rng.permutation(nat) for _ in range(n_sym)], dtype=np.intc)
    pols = rng.standard_normal((3 * nat, n))
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.intc)
    dev = sy.symmetries_on_modes(rots, irt, pols, starts, device=0)
    ora = orc.symmetries_on_modes(rots, irt, pols, starts)
    for S, O in zip(dev, ora):
        assert np.abs(S - O).max() < 1e-12 * max(1.0, np.abs(O).max())
    wide = [100]
    starts = np.array([0, 100], dtype=np.intc)
    pols = rng.standard_normal((3 * nat, 100))
    print("case: one block of 100 modes", flush=True)
    S = sy.symmetries_on_modes(rots[:5], irt[:5], pols, starts, device=0)[0]
    O = orc.symmetries_on_modes(rots[:5], irt[:5], pols, starts)[0]
    assert np.abs(S - O).max() < 1e-12 * np.abs(O).max()
    with pytest.raises(ValueError):
        bad = irt[:2].copy()
        bad[0, 0] = bad[0, 1]
        sy.symmetries_on_modes(rots[:2], bad, pols, starts, device=0)


def test_lanczos_with_generated_symmetries_matches_reference_blocks():
    import test_symmetry as ts
    import fixtures
    from tdscha_amd import DynamicalLanczos as DL
    from tdscha_amd import symmetry as sy
    print("case: SnTe L application, generated vs reference blocks", flush=True)
    rots, trans, cen, cell, species, d = ts.snte_space_group()
    irt = sy.atom_permutations(cen, cell, rots, trans, species)
    ens = fixtures.load_mode10()
    psi = ts.load("mode10_ensemble.npz")["psi"]
    out = []
    for generated in (False, True):
        lz = DL.Lanczos(mode=DL.MODE_GPU_FP64)
        lz.verbose = False
        lz.set_ensemble(ens.X, ens.Y, ens.w, ens.rho, ens.T, d["pols"], d["m"])
        if generated:
            lz.prepare_symmetrization(symmetries=(rots, irt), verbose=False)
        else:
            s = ens.sym
            lz.set_symmetries(s.symmetries, s.N_degeneracy, s.degenerate_space, s.block_id)
        lz.init()
        out.append(lz.apply_full_L(psi.copy()))
    assert np.abs(out[0] - out[1]).max() < 1e-12 * np.abs(out[0]).max()
