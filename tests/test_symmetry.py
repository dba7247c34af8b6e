"""Mode-basis symmetry blocks (Lanczos.prepare_symmetrization, Modules/DynamicalLanczos.py:
686-778): CPU checks of the oracle and the host logic.

Pin: the reference's own blocks for its SnTe 2x2x2 test ensemble (384 operations, 9 blocks,
tests/golden/mode10_ensemble.npz) are reproduced as sets of matrices from the Oh operations
about a Te site times the 8 supercell translations, the centroid positions and the saved
polarization vectors (tests/golden/projection_snte.npz). Sets, because the reference orders
the operations as spglib lists them; the symmetrized average does not depend on the order.
"""
import itertools
import os

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def oh():
    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3))
            for r, c in enumerate(perm):
                R[r, c] = signs[r]
            ops.append(R)
    return np.array(ops)


def snte_space_group():
    """Cartesian rotations, translations, supercell and species of the SnTe 2x2x2 cell."""
    d = load("projection_snte.npz")
    cen = d["centroid"]
    nat = len(cen)
    species = np.arange(nat) % 2
    alat = 12.4 * 0.529177249
    prim = np.array([[-0.5, 0, 0.5], [0, 0.5, 0.5], [-0.5, 0.5, 0]]) * alat
    cell = 2 * prim
    origin = cen[0]
    rots, trans = [], []
    for R in oh():
        for a in range(0, nat, 2):           # the 8 lattice translations Te_0 -> Te_a
            rots.append(R)
            trans.append(origin - R @ origin + (cen[a] - cen[0]))
    return np.array(rots), np.array(trans), cen, cell, species, d


def canonical(mats):
    key = lambda M: tuple(np.round(M.ravel(), 7))
    return np.array(sorted(mats, key=key))


def test_oracle_blocks_match_the_reference_set():
    from tdscha_amd import symmetry as sy
    rots, trans, cen, cell, species, d = snte_space_group()
    irt = sy.atom_permutations(cen, cell, rots, trans, species)
    assert irt.shape == (384, 16)
    g = load("mode10_ensemble.npz")
    starts = sy.degenerate_blocks(d["w"])
    ndeg, degs, bid = sy.block_arrays(starts, len(d["w"]))
    assert np.array_equal(ndeg, g["N_degeneracy"])
    assert np.array_equal(bid, g["block_id"])
    blocks = orc.symmetries_on_modes(rots, irt, d["pols"], starts)
    for b, S in enumerate(blocks):
        ref = g["sym_%d" % b]
        assert S.shape == ref.shape
        assert np.abs(canonical(S) - canonical(ref)).max() < 1e-12, b
    # a transposed rotation (the inverse operation on the Cartesian part only) breaks it
    bad = orc.symmetries_on_modes(np.transpose(rots, (0, 2, 1)), irt, d["pols"], starts)
    assert max(np.abs(canonical(S) - canonical(g["sym_%d" % b])).max()
               for b, S in enumerate(bad)) > 1e-3


def test_blocks_are_orthogonal_representations():
    """Each block is an invariant subspace: S_b[s] orthogonal, and S_b[s] S_b[t] is again
    in the set (closure), on the reference ensemble's largest block."""
    from tdscha_amd import symmetry as sy
    rots, trans, cen, cell, species, d = snte_space_group()
    irt = sy.atom_permutations(cen, cell, rots, trans, species)
    starts = sy.degenerate_blocks(d["w"])
    S = orc.symmetries_on_modes(rots, irt, d["pols"], starts)[5]
    dim = S.shape[1]
    for s in range(0, 384, 37):
        assert np.abs(S[s] @ S[s].T - np.eye(dim)).max() < 1e-12
    flat = S.reshape(384, -1)
    for s, t in ((3, 77), (100, 250), (383, 5)):
        prod = (S[s] @ S[t]).ravel()
        assert np.abs(flat - prod[None]).max(axis=1).min() < 1e-12


def test_atom_permutations_rejects_a_non_symmetry():
    from tdscha_amd import symmetry as sy
    rots, trans, cen, cell, species, d = snte_space_group()
    with pytest.raises(ValueError):
        sy.atom_permutations(cen, cell, rots[:2], trans[:2] + 0.3, species)
    shift = np.array([cen[1] - cen[0]])          # Te -> Sn: a translation of the lattice only
    with pytest.raises(ValueError):
        sy.atom_permutations(cen, cell, rots[:1], shift, species)


def test_degenerate_blocks_and_no_sym():
    from tdscha_amd import symmetry as sy
    from tdscha_amd import DynamicalLanczos as DL
    w = np.array([1.0, 1.0 + 1e-12, 2.0, 3.0, 3.0, 3.0])
    assert list(sy.degenerate_blocks(w)) == [0, 2, 3, 6]
    lz = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    lz.set_ensemble(np.zeros((2, 6)), np.zeros((2, 6)), w, np.ones(2), 100.0)
    lz.prepare_symmetrization(no_sym=True)
    assert len(lz.symmetries) == 6 and lz.n_syms == 1
    assert np.array_equal(lz.sym_block_id, np.arange(6))
    with pytest.raises(ValueError):
        lz.prepare_symmetrization()


def test_atom_permutations_match_the_synthetic_crystal():
    """Position matching agrees with the synthetic generator's integer-lattice permutations
    (rock salt 3x3x3, 48 x 27 operations)."""
    from tdscha_amd import symmetry as sy
    from tdscha_amd import synthetic as syn
    cr = syn.make_crystal(3)
    rots, perms = syn.space_group(cr)
    trans = cr.cell[cr.species == 0] @ syn.PRIM
    T = np.array([t for _ in syn.oh_operations() for t in trans], dtype=float)
    irt = sy.atom_permutations(cr.pos.astype(float), 3 * syn.PRIM.astype(float), rots, T,
                               cr.species)
    assert np.array_equal(irt, perms)
