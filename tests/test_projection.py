"""Ensemble projection on the modes (Modules/DynamicalLanczos.py:330-442) and the text input
files of the standalone executable (Py:862-998): CPU checks.

The oracle's projection is pinned against the reference's own saved X, Y for its test
ensemble (SnTe 2x2x2, 10 configurations: tests/test_lanczos_fast/data + normal/
LANCZOS_STEP5.npz, in tests/golden/projection_snte.npz), and against a closed form. Two
inputs of that fixture are rebuilt rather than read, which sets the tolerance: the
centroid positions (from the dynamical-matrix file, CellConstructor's alat conversion) and
the SSCHA forces, -Phi u with Phi = sqrt(m) P diag(w^2) P^T sqrt(m) from the saved modes.
The symmetrized average force vanishes: every atom of rock salt sits on an Oh site.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
BOHR = 0.52917721067


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def snte_inputs():
    d = load("projection_snte.npz")
    N = d["xats"].shape[0]
    u = (d["xats"] - d["centroid"][None]).reshape(N, -1)          # Angstrom
    m, P, w = d["m"], d["pols"], d["w"]
    phi = np.sqrt(m)[:, None] * (P * w ** 2) @ P.T * np.sqrt(m)[None, :]   # Ry / Bohr^2
    fs = -(u / BOHR) @ phi.T / BOHR                                 # Ry / Angstrom
    return d, u, d["forces"].reshape(N, -1), fs


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_oracle_projection_matches_reference_fixture():
    d, u, f, fs = snte_inputs()
    X, Y = orc.project_ensemble(u, f, fs, None, d["m"], d["pols"], BOHR)
    assert rel(X, d["X"]) < 3e-8
    assert rel(Y, d["Y"]) < 1e-6
    # a dropped unit or mass factor is far outside these bounds
    X2, Y2 = orc.project_ensemble(u * BOHR, f, fs, None, d["m"], d["pols"], BOHR)
    assert rel(X2, d["X"]) > 0.1
    X3, Y3 = orc.project_ensemble(u, f, None, None, d["m"], d["pols"], BOHR)
    assert rel(Y3, d["Y"]) > 1e-3


def test_oracle_projection_closed_form():
    """With an orthonormal mode basis P, displacements B P x / sqrt(m) project to x and
    forces (y sqrt(m) P^T rows) / B + sscha + mean project back to y."""
    rng = np.random.default_rng(5)
    nat, N = 7, 11
    nc = 3 * nat
    P, _ = np.linalg.qr(rng.standard_normal((nc, nc)))
    m = np.repeat(rng.uniform(1e4, 2e5, nat), 3)
    x = rng.standard_normal((N, nc))
    y = rng.standard_normal((N, nc))
    fs = rng.standard_normal((N, nc))
    fm = rng.standard_normal(nc)
    u = BOHR * (x @ P.T) / np.sqrt(m)
    f = (y @ P.T) * np.sqrt(m) / BOHR + fs + fm
    X, Y = orc.project_ensemble(u, f, fs, fm, m, P, BOHR)
    assert np.abs(X - x).max() < 1e-12
    assert np.abs(Y - y).max() < 1e-9


def test_translation_mask_finds_the_rigid_shifts():
    from tdscha_amd import projection as pj
    d = load("projection_snte.npz")
    m, P = d["m"], d["pols"]
    nc = len(m)
    T = np.zeros((nc, 3))
    for a in range(3):
        T[a::3, a] = np.sqrt(m[a::3])
    T /= np.linalg.norm(T, axis=0)
    full = np.concatenate([P, T], axis=1)
    order = np.random.default_rng(1).permutation(full.shape[1])
    full = full[:, order]
    mask = pj.translation_mask(full, m[::3])
    assert mask.sum() == 3
    assert set(np.nonzero(mask)[0]) == set(np.nonzero(order >= P.shape[1])[0])
    w = np.concatenate([d["w"], np.zeros(3)])[order]
    good = pj.good_modes(w, full, m)
    assert np.array_equal(good, ~mask)
    assert np.array_equal(pj.good_modes(w, full, m, ignore_small_w=True), ~mask)
    sel = np.zeros(len(w), dtype=bool)
    sel[:5] = True
    assert np.array_equal(pj.good_modes(w, full, m, select_modes=sel), ~mask & sel)
    with pytest.raises(ValueError):
        pj.good_modes(w, full, m, select_modes=sel[:-1])


def mode10_lanczos(mode10):
    from tdscha_amd import DynamicalLanczos as DL
    lz = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    lz.set_ensemble(mode10.X, mode10.Y, mode10.w, mode10.rho, mode10.T, None,
                    load("input_files_mode10.npz")["masses"])
    s = mode10.sym
    lz.set_symmetries(s.symmetries, s.N_degeneracy, s.degenerate_space, s.block_id)
    lz.psi = load("mode10_ensemble.npz")["psi"].copy()
    return lz


def test_prepare_input_files_writes_the_reference_bytes(mode10, tmp_path):
    """Every mode10.* text file the reference wrote for its executable is reproduced byte
    for byte (the json is compared as data: the reference kept T as an integer)."""
    ref = json.load(open(os.path.join(GOLDEN, "input_files_mode10.json")))
    lz = mode10_lanczos(mode10)
    lz.prepare_input_files(root_name="mode10", n_steps=5, directory=str(tmp_path))
    written = sorted(os.listdir(tmp_path))
    assert written == sorted(ref["sha256"])
    for name, digest in ref["sha256"].items():
        path = os.path.join(tmp_path, name)
        if name.endswith(".json"):
            assert json.load(open(path)) == ref["json"]
        else:
            assert hashlib.sha256(open(path, "rb").read()).hexdigest() == digest, name
    with pytest.raises(NotImplementedError):
        lz.prepare_input_files("x", directory=str(tmp_path), start_from_scratch=False)


def test_read_input_files_round_trip(mode10, tmp_path):
    from tdscha_amd import DynamicalLanczos as DL
    lz = mode10_lanczos(mode10)
    lz.reverse_L = True
    lz.shift_value = 0.25
    lz.prepare_input_files(root_name="rt", n_steps=7, directory=str(tmp_path))
    back = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    meta = back.read_input_files("rt", str(tmp_path))
    assert meta["n_steps"] == 7
    assert np.array_equal(back.X, lz.X) and np.array_equal(back.Y, lz.Y)
    assert np.array_equal(back.w, lz.w) and np.array_equal(back.rho, lz.rho)
    assert np.array_equal(back.m, lz.m) and np.array_equal(back.psi, lz.psi)
    assert back.T == lz.T and back.reverse_L and back.shift_value == 0.25
    assert np.array_equal(back.sym_block_id, lz.sym_block_id)
    for a, b in zip(back.symmetries, lz.symmetries):
        assert np.abs(a - b).max() < 1e-16 * 10
    for a, b in zip(back.degenerate_space, lz.degenerate_space):
        assert np.array_equal(a, b)


def test_load_from_input_files_reads_results(mode10, tmp_path):
    from tdscha_amd import DynamicalLanczos as DL
    lz = mode10_lanczos(mode10)
    lz.prepare_input_files(root_name="res", n_steps=3, directory=str(tmp_path))
    abc = np.arange(12, dtype=float).reshape(4, 3) + 0.5
    np.savetxt(os.path.join(tmp_path, "res.abc"), abc)
    np.savetxt(os.path.join(tmp_path, "res.snorm.out"), np.array([1.0, 2.0, 3.0]))
    back = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    with pytest.warns(UserWarning):
        back.load_from_input_files("res", str(tmp_path))
    assert np.array_equal(back.a_coeffs, abc[:, 0]) and np.array_equal(back.c_coeffs, abc[:, 2])
    assert np.array_equal(back.s_norm, [1.0, 2.0, 3.0])
    assert back.N == mode10.X.shape[0] and back.n_modes == mode10.X.shape[1]
    with pytest.raises(IOError):
        back.load_from_input_files("missing", str(tmp_path))


def test_project_entry_point_rejects_bad_arguments():
    from tdscha_amd import _lib as L
    lib = L.lib()
    buf = (ctypes.c_double * 4)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    assert lib.tdscha_project_ensemble(0, 3, 1, p, p, None, None, p, p, BOHR, p, p, None) == 2
    assert lib.tdscha_project_ensemble(1, 3, 1, None, p, None, None, p, p, BOHR, p, p, None) == 1
    assert lib.tdscha_project_ensemble(1, 3, 1, p, p, None, None, p, p, 0.0, p, p, None) == 2
    assert b"bohr" in lib.tdscha_last_error()


def test_load_status_reads_a_reference_status_file(mode10):
    """The npz the reference's save_status wrote (its 5-step run on the SnTe ensemble) loads
    without pickle: ensemble, padded symmetry blocks (block sizes recovered from the padding),
    Krylov coefficients and bases."""
    from tdscha_amd import DynamicalLanczos as DL
    lz = DL.Lanczos(mode=DL.MODE_GPU_FP64)
    lz.load_status(os.path.join(GOLDEN, "lanczos_status_step5.npz"))
    assert np.array_equal(lz.X, mode10.X) and np.array_equal(lz.Y, mode10.Y)
    assert np.array_equal(lz.w, mode10.w) and lz.T == mode10.T
    assert lz.n_syms == 384 and len(lz.symmetries) == 9
    s = mode10.sym
    assert np.array_equal(lz.N_degeneracy, s.N_degeneracy)
    assert np.array_equal(lz.sym_block_id, s.block_id)
    for a, b in zip(lz.symmetries, s.symmetries):
        assert np.array_equal(a, b)
    std = load("lanczos_std.npz")        # the 10-step run of the same calculation
    assert len(lz.a_coeffs) == 5
    assert np.allclose(lz.a_coeffs, std["a"][:5], rtol=1e-12, atol=0)
    assert np.allclose(lz.b_coeffs, std["b"][:5], rtol=1e-12, atol=0)
    assert np.array_equal(np.asarray(lz.basis_Q)[:6], std["Q"][:6])
