"""Ensemble -> (X, Y) projection on the SCHA modes, on the device.

The reference does this inside ``Lanczos.__init__`` (Modules/DynamicalLanczos.py:330-442,
"Py" below) from a ``sscha.Ensemble`` and its ``CellConstructor`` dynamical matrix. Neither
package is needed here: the caller hands over the arrays those objects hold.

    translations / small-frequency modes removed         Py:333-362
    u  = u_disps / bohr * sqrt(m)                        Py:385, 425
    f  = ((forces - sscha_forces) * bohr - f_mean * bohr) / sqrt(m)
                                                         Py:387-388, 404-422, 426
    X  = u . pols,   Y = f . pols                        Py:438-439

The rescaling and both products run in one FP64 MFMA kernel (``tdscha_project_ensemble``);
the mode selection is host logic on the (3 nat x 3 nat) polarization matrix.
"""
import numpy as np

from . import _lib as L

# Angstrom per Bohr used by the reference's ensemble rescaling (sscha's Ensemble.Bohr,
# CODATA 2014). CellConstructor converts the dynamical matrix's alat to Angstrom with its own
# older constant, 0.529177249; that one only matters when centroid positions are rebuilt
# from a dynamical-matrix file.
BOHR = 0.52917721067


def translation_mask(pols, masses, tol=1e-6):
    """Modes that rigidly translate the crystal (the role of CC.Methods.get_translations,
    Py:336): the displacement pattern pols[:, mu] / sqrt(m) is the same vector on every atom.

    pols: (3 nat, n) mass-scaled polarization vectors; masses: (nat,) or (3 nat,).
    tol: largest spread of the per-atom vectors allowed, relative to their size.
    """
    pols = np.asarray(pols, dtype=np.float64)
    nc, n = pols.shape
    if nc % 3:
        raise ValueError("pols must have 3 nat rows")
    m = _masses_per_component(masses, nc)
    disp = (pols / np.sqrt(m)[:, None]).reshape(nc // 3, 3, n)
    mean = disp.mean(axis=0)
    spread = np.abs(disp - mean[None]).max(axis=(0, 1))
    size = np.abs(disp).max(axis=(0, 1))
    return (spread <= tol * size) & (size > 0)


def good_modes(w, pols, masses, ignore_small_w=False, select_modes=None, eps_w=1e-8):
    """Boolean mask of the modes the Lanczos keeps (Py:333-359).

    ignore_small_w: drop |w| < eps_w instead of detecting translations (the reference takes
    eps_w from CellConstructor's Phonons module; here it is an argument).
    select_modes: optional boolean array of the same length, ANDed with the mask.
    """
    w = np.asarray(w, dtype=np.float64)
    if ignore_small_w:
        trans = np.abs(w) < eps_w
    else:
        trans = translation_mask(pols, masses)
    good = ~trans
    if select_modes is not None:
        select_modes = np.asarray(select_modes, dtype=bool)
        if len(select_modes) != len(trans):
            raise ValueError("'select_modes' should be an array of the same length of the "
                             "number of modes. n_modes = %d | len(select_modes) = %d"
                             % (len(w), len(select_modes)))
        good = good & select_modes
    return good


def _masses_per_component(masses, n_cart):
    m = np.asarray(masses, dtype=np.float64).ravel()
    if len(m) * 3 == n_cart:
        m = np.repeat(m, 3)          # np.tile(m, (3, 1)).T.ravel(), Py:331
    if len(m) != n_cart:
        raise ValueError("masses must have nat or 3 nat entries")
    return m


def project_ensemble(u_disps, forces, masses, pols, sscha_forces=None, f_mean=None, bohr=BOHR,
                     device=None, stream=None):
    """X, Y (torch FP64 tensors on the device) of Py:438-439.

    u_disps, forces, sscha_forces: (N, nat, 3) or (N, 3 nat); Angstrom and Ry/Angstrom.
    masses: (nat,) or (3 nat,). pols: (3 nat, n) of the kept modes. f_mean: the symmetrized
    average force over the supercell, (3 nat,) or (nat, 3), Ry/Angstrom (None = zero).
    Inputs may be numpy arrays or torch tensors; they are copied to the device as FP64.
    """
    import torch

    if isinstance(device, torch.device):
        dev = device
    else:
        dev = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))

    def put(a, shape):
        t = torch.as_tensor(a, dtype=torch.float64)
        return t.reshape(shape).to(dev).contiguous()

    pols_t = torch.as_tensor(pols, dtype=torch.float64)
    if pols_t.dim() == 1:
        pols_t = pols_t.reshape(-1, 1)
    nc, n = pols_t.shape
    N = int(np.prod(np.shape(u_disps))) // nc
    if N * nc != int(np.prod(np.shape(u_disps))) or N * nc != int(np.prod(np.shape(forces))):
        raise ValueError("u_disps and forces must hold N x %d values" % nc)
    u = put(u_disps, (N, nc))
    f = put(forces, (N, nc))
    fs = put(sscha_forces, (N, nc)) if sscha_forces is not None else None
    fm = put(f_mean, (nc,)) if f_mean is not None else None
    m = put(_masses_per_component(masses, nc), (nc,))
    P = pols_t.to(dev).contiguous()
    X = torch.empty((N, n), dtype=torch.float64, device=dev)
    Y = torch.empty((N, n), dtype=torch.float64, device=dev)
    st = stream if stream is not None else torch.cuda.current_stream(dev)
    ptr = lambda t: None if t is None else t.data_ptr()
    L.check(L.lib().tdscha_project_ensemble(
        N, nc, n, ptr(u), ptr(f), ptr(fs), ptr(fm), ptr(m), ptr(P), float(bohr), ptr(X), ptr(Y),
        st.cuda_stream), "tdscha_project_ensemble")
    return X, Y


def _main():
    """Time the projection at a given size: python -m tdscha_amd.projection [--configs N]
    [--atoms nat] [--reps K] [--cpu]. Synthetic displacements, forces and an orthonormal
    mode basis with the three translations removed; FP64."""
    import argparse
    import json
    import time

    import torch

    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", type=int, default=20000)
    ap.add_argument("--atoms", type=int, default=432)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--cpu", action="store_true",
                    help="also time the reference's host path (numpy u.dot(pols), f.dot(pols))")
    a = ap.parse_args()
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(7)
    N, nc = a.configs, 3 * a.atoms
    n = nc - 3
    u = torch.randn(N, nc, device=dev, dtype=torch.float64, generator=g) * 0.1
    f = torch.randn(N, nc, device=dev, dtype=torch.float64, generator=g) * 1e-2
    fs = torch.randn(N, nc, device=dev, dtype=torch.float64, generator=g) * 1e-2
    m = torch.rand(nc, device=dev, dtype=torch.float64, generator=g) * 1e5 + 1e4
    P = torch.linalg.qr(torch.randn(nc, nc, device=dev, dtype=torch.float64,
                                    generator=g))[0][:, :n].contiguous()
    X = torch.empty(N, n, device=dev, dtype=torch.float64)
    Y = torch.empty_like(X)
    st = torch.cuda.current_stream(dev)

    def launch():
        L.check(L.lib().tdscha_project_ensemble(
            N, nc, n, u.data_ptr(), f.data_ptr(), fs.data_ptr(), None, m.data_ptr(),
            P.data_ptr(), BOHR, X.data_ptr(), Y.data_ptr(), st.cuda_stream), "project")

    launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(a.reps):
        launch()
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / a.reps
    flop = 2.0 * 2.0 * N * nc * n
    out = {"kernel": "k_project", "configs": N, "atoms": a.atoms, "n_modes": n,
           "ms_per_projection": ms, "tflops": flop / (ms * 1e-3) / 1e12,
           "fp64_mfma_peak_tflops": 78.6,
           "frac": flop / (ms * 1e-3) / 1e12 / 78.6,
           "hbm_bytes": 8.0 * (3 * N * nc + 2 * N * n + nc * n)}
    if a.cpu:
        uh, fh, Ph = u.cpu().numpy(), (f - fs).cpu().numpy(), P.cpu().numpy()
        t0 = time.perf_counter()
        uh.dot(Ph)
        fh.dot(Ph)
        out["cpu_numpy_ms"] = (time.perf_counter() - t0) * 1e3
        try:
            from threadpoolctl import threadpool_info
            out["cpu_threads"] = max(i["num_threads"] for i in threadpool_info())
        except Exception:
            out["cpu_threads"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    _main()
