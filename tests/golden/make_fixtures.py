"""Write the committed golden fixtures from the reference tree's own test data.

Run where the reference tree is present (it reads /root/reference; the GPU tests only
read the fixtures it writes):  python tests/golden/make_fixtures.py

Inputs (all data files shipped in the reference's tests, loaded with
allow_pickle=False / text readers):
  tests/test_lanczos_fast/normal/LANCZOS_STEP10.npz      standard run_FT, 10 steps
  tests/test_julia/wigner_nowigner_julia/lanczos_wigner_STEP10.npz   Wigner run_simm
  tests/test_julia/wigner_nowigner_julia/lanczos_nowigner_STEP10.npz standard (second run)
  tests/test_lanczos_fast/test_c_code_lanczos/mode10.*   kernel inputs as text

Outputs (tests/golden/):
  mode10_ensemble.npz   X, Y, w, rho, T, symmetry blocks, psi (the 2x2x2 SnTe ensemble,
                        45 modes, 10 configs, 384 symmetries, 9 blocks)
  lanczos_std.npz       Krylov bases Q/P, a/b/c, s_norm of the standard run
  lanczos_nowigner.npz  the same for the second standard run
  lanczos_wigner.npz    the same for the Wigner symmetric run
  ref_kernel_vectors.npz (f_av, d2v) of the reference C trio (oracle/_ref, compiled from
                        CModules/LanczosFunctions.c) on mode10 for several (R1, Y1, D4) inputs
  projection_snte.npz   raw ensemble of tests/test_lanczos_fast/data (xats_pop1.npy,
                        forces_pop1.npy), the centroid positions rebuilt from dyn_gen_pop1_1,
                        and the reference's projected X, Y with its m, pols, w, rho
                        (normal/LANCZOS_STEP5.npz): the inputs and outputs of Py:380-442
  input_files_mode10.npz / .json  masses of mode10 and the SHA-256 of every mode10.* text
                        file that prepare_input_files (Py:862-919) wrote
  raman_unpol.npz       tests/test_raman_unpol: the Raman tensor and the seven unpolarized
                        Raman vectors its test checks (raman_tensor_eq.npy, raman_v_*.npy)
  lanczos_status_step5.npz  tests/test_lanczos_fast/normal/LANCZOS_STEP5.npz as the
                        reference's save_status wrote it (copied byte for byte)
"""
import os
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

from oracle import oracle as orc  # noqa: E402


def load_npz(path):
    with np.load(path, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def blocks_from_status(st):
    """sym_julia (n_blocks, N_sym, d_max, d_max) + deg_julia -> SymBlocks (Py:753-778)."""
    sj, dj = st["sym_julia"], st["deg_julia"]
    n = int(st["n_modes"])
    ndeg = []
    syms, degs = [], []
    text_dir = os.path.join(REF, "tests/test_lanczos_fast/test_c_code_lanczos")
    nd_text = np.loadtxt(os.path.join(text_dir, "mode10.ndegs"), dtype=int)
    for b in range(sj.shape[0]):
        d = int(nd_text[b])
        ndeg.append(d)
        syms.append(np.ascontiguousarray(sj[b, :, :d, :d]))
        degs.append(np.ascontiguousarray(dj[b, :d].astype(np.intc)))
    bid = np.full(n, -1, dtype=np.intc)
    for b, dg in enumerate(degs):
        bid[dg] = b
    assert (bid >= 0).all()
    return orc.SymBlocks(syms, np.array(ndeg), degs, bid)


def check_against_text(st, sym):
    """The binary status and the mode10 text inputs describe the same ensemble."""
    d = os.path.join(REF, "tests/test_lanczos_fast/test_c_code_lanczos")
    X = np.loadtxt(os.path.join(d, "mode10.X.dat"))
    Y = np.loadtxt(os.path.join(d, "mode10.Y.dat"))
    w = np.loadtxt(os.path.join(d, "mode10.freqs"))
    rho = np.loadtxt(os.path.join(d, "mode10.rho"))
    bid = np.loadtxt(os.path.join(d, "mode10.blockid"), dtype=int)
    assert np.array_equal(X, st["X"]) and np.array_equal(Y, st["Y"])
    assert np.array_equal(w, st["w"]) and np.array_equal(rho, st["rho"])
    assert np.array_equal(bid, sym.block_id)
    for b in range(sym.n_blocks):
        blk = np.loadtxt(os.path.join(d, "mode10.block%d" % b), dtype=int).reshape(-1)
        assert np.array_equal(blk, sym.degenerate_space[b])
        s_txt = np.loadtxt(os.path.join(d, "mode10.symsb%d" % b)).reshape(sym.symmetries[b].shape)
        assert np.max(np.abs(s_txt - sym.symmetries[b])) < 1e-15


def main():
    std = load_npz(os.path.join(REF, "tests/test_lanczos_fast/normal/LANCZOS_STEP10.npz"))
    sym = blocks_from_status(std)
    check_against_text(std, sym)
    psi_txt = np.loadtxt(os.path.join(REF, "tests/test_lanczos_fast/test_c_code_lanczos/mode10.psi"))

    ens = dict(X=std["X"], Y=std["Y"], w=std["w"], rho=std["rho"], T=np.float64(std["T"]),
               N_degeneracy=sym.N_degeneracy, block_id=sym.block_id, psi=psi_txt,
               n_syms=np.int64(sym.n_syms))
    for b in range(sym.n_blocks):
        ens["sym_%d" % b] = sym.symmetries[b]
        ens["deg_%d" % b] = sym.degenerate_space[b]
    np.savez_compressed(os.path.join(HERE, "mode10_ensemble.npz"), **ens)

    def krylov(src, dst):
        st = load_npz(src)
        assert np.array_equal(st["X"], std["X"]) and np.array_equal(st["w"], std["w"])
        keep = dict(Q=st["basis_Q"], P=st["basis_P"], a=st["a_coeffs"], b=st["b_coeffs"],
                    c=st["c_coeffs"], s_norm=st["s_norm"], psi=st["psi"],
                    use_wigner=st["use_wigner"], T=np.float64(st["T"]),
                    perturbation_modulus=np.float64(st["perturbation_modulus"]))
        np.savez_compressed(os.path.join(HERE, dst), **keep)

    krylov(os.path.join(REF, "tests/test_lanczos_fast/normal/LANCZOS_STEP10.npz"), "lanczos_std.npz")
    krylov(os.path.join(REF, "tests/test_julia/wigner_nowigner_julia/lanczos_nowigner_STEP10.npz"),
           "lanczos_nowigner.npz")
    krylov(os.path.join(REF, "tests/test_julia/wigner_nowigner_julia/lanczos_wigner_STEP10.npz"),
           "lanczos_wigner.npz")

    # Reference kernel outputs on mode10 (the C trio built from source by oracle/Makefile).
    if not orc.ref_available():
        raise SystemExit("build oracle/_ref first: make -C oracle")
    X, Y, w, rho, T = std["X"], std["Y"], std["w"], std["rho"], float(std["T"])
    n = len(w)
    rng = np.random.default_rng(20240601)
    cases = []
    Q = std["basis_Q"]
    P = std["basis_P"]
    # (1) standard pre-transform of a Krylov vector, D4 on
    R1, Y1 = orc.anharmonic_pre(Q[3], w, T, transpose=False)
    cases.append((R1, Y1, 1))
    # (2) transposed pre-transform of a P vector, D4 on
    R1, Y1 = orc.anharmonic_pre(P[2], w, T, transpose=True)
    cases.append((R1, Y1, 1))
    # (3) random symmetric Y1 and random R1, D4 off
    A = rng.standard_normal((n, n))
    cases.append((rng.standard_normal(n), (A + A.T) * 1e-3, 0))
    # (4) random non-symmetric Y1, D4 on (the kernel symmetrises implicitly)
    cases.append((rng.standard_normal(n), rng.standard_normal((n, n)) * 1e-3, 1))
    out = {}
    for k, (R1, Y1, d4) in enumerate(cases):
        f, d2v = orc.ref_perturb_average_sym(X, Y, w, rho, R1, Y1, T, d4, sym)
        out["R1_%d" % k] = R1
        out["Y1_%d" % k] = Y1
        out["d4_%d" % k] = np.int64(d4)
        out["f_%d" % k] = f
        out["d2v_%d" % k] = d2v
    out["n_cases"] = np.int64(len(cases))
    np.savez_compressed(os.path.join(HERE, "ref_kernel_vectors.npz"), **out)
    projection_fixture()
    input_files_fixture()
    raman_fixture()
    import shutil
    shutil.copyfile(os.path.join(REF, "tests/test_lanczos_fast/normal/LANCZOS_STEP5.npz"),
                    os.path.join(HERE, "lanczos_status_step5.npz"))
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))


def centroid_from_dyn(path, supercell, bohr_to_angstrom):
    """Supercell centroid positions (Angstrom) from a Quantum ESPRESSO dynamical-matrix file
    with ibrav = 0: unit cell and atoms in units of alat (celldm(1), Bohr); supercell atoms
    ordered cell by cell (third lattice index fastest), unit-cell atoms fastest."""
    import itertools
    lines = open(path).read().split("\n")
    head = lines[2].split()
    ntyp, nat, ibrav, alat = int(head[0]), int(head[1]), int(head[2]), float(head[3])
    assert ibrav == 0
    cell = np.array([[float(x) for x in lines[4 + i].split()] for i in range(3)])
    pos = np.array([[float(x) for x in lines[7 + ntyp + a].split()[2:5]] for a in range(nat)])
    cell = cell * alat * bohr_to_angstrom
    pos = pos * alat * bohr_to_angstrom
    out = [pos[a] + np.array(ijk, dtype=float) @ cell
           for ijk in itertools.product(*[range(s) for s in supercell]) for a in range(nat)]
    return np.array(out)


def projection_fixture():
    d = os.path.join(REF, "tests/test_lanczos_fast")
    st = load_npz(os.path.join(d, "normal/LANCZOS_STEP5.npz"))
    xats = np.load(os.path.join(d, "data/xats_pop1.npy"), allow_pickle=False)
    forces = np.load(os.path.join(d, "data/forces_pop1.npy"), allow_pickle=False)
    # CellConstructor's Bohr -> Angstrom constant for alat (the reference's X is matched to
    # 8e-9 with it, 1.5e-6 with the CODATA value).
    centroid = centroid_from_dyn(os.path.join(d, "data/dyn_gen_pop1_1"), (2, 2, 2), 0.529177249)
    assert np.abs(xats - centroid[None]).max() < 0.5
    np.savez_compressed(os.path.join(HERE, "projection_snte.npz"), xats=xats, forces=forces,
                        centroid=centroid, m=st["m"], pols=st["pols"], w=st["w"], X=st["X"],
                        Y=st["Y"], rho=st["rho"], T=np.float64(st["T"]))


def input_files_fixture():
    import hashlib
    import json
    d = os.path.join(REF, "tests/test_lanczos_fast/test_c_code_lanczos")
    hashes = {}
    for f in sorted(os.listdir(d)):
        if f.startswith("mode10."):
            hashes[f] = hashlib.sha256(open(os.path.join(d, f), "rb").read()).hexdigest()
    with open(os.path.join(d, "mode10.json")) as fp:
        meta = json.load(fp)
    with open(os.path.join(HERE, "input_files_mode10.json"), "w") as fp:
        json.dump({"sha256": hashes, "json": meta}, fp, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "input_files_mode10.npz"),
                        masses=np.loadtxt(os.path.join(d, "mode10.masses")))




def raman_fixture():
    d = os.path.join(REF, "tests/test_raman_unpol")
    out = {"raman_tensor": np.load(os.path.join(d, "raman_tensor_eq.npy"), allow_pickle=False)}
    for i in range(7):
        out["raman_v_%d" % i] = np.load(os.path.join(d, "raman_v_%d.npy" % i), allow_pickle=False)
    np.savez_compressed(os.path.join(HERE, "raman_unpol.npz"), **out)


if __name__ == "__main__":
    main()
