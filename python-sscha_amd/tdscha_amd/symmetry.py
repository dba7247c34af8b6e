"""Symmetry operations of a supercell in the SCHA mode basis, on the device.

The reference's ``Lanczos.prepare_symmetrization`` (Modules/DynamicalLanczos.py:686-778,
"Py" below) asks spglib for the supercell's space group and CellConstructor's
GetSymmetriesOnModesDeg for the blocks. Neither package is needed here: the caller gives the
operations as Cartesian rotations and the atom each one moves every atom to (or the
rotations and translations with the supercell, from which ``atom_permutations`` finds the
atoms), and ``symmetries_on_modes`` builds, for every block of degenerate modes,

    S_b[s] = E_b^T P_s E_b,   (P_s E)[irt_s(a)] = R_s E[a]

in one HIP kernel launch per block (``tdscha_symmetries_on_modes``).
"""
import numpy as np

from . import _lib as L


def degenerate_blocks(w, tol=1e-8):
    """Blocks of consecutive modes whose frequencies agree to `tol` (relative): the block
    structure the device path needs (contiguous, sorted; Py:741-746, LanczosFunctions.c:2162).
    Returns the n_blocks + 1 block offsets."""
    w = np.asarray(w, dtype=np.float64)
    starts = [0]
    for k in range(1, len(w)):
        if abs(w[k] - w[k - 1]) > tol * max(abs(w[k]), abs(w[k - 1])):
            starts.append(k)
    starts.append(len(w))
    return np.array(starts, dtype=np.intc)


def atom_permutations(positions, cell, rotations, translations, species=None, tol=1e-4):
    """irt[s][a]: the atom that R_s x_a + t_s lands on, modulo the supercell.

    positions (nat, 3) and translations (n_sym, 3) Cartesian (any length unit, the same as
    cell, whose rows are the supercell vectors); rotations (n_sym, 3, 3) Cartesian. Raises if
    an image lands on no atom (or on one of another species) within `tol` (fractional).
    Images are matched by their fractional coordinates rounded to a grid of spacing ~tol/2
    (sorted keys, one search per image); an image next to a grid edge falls back to the
    nearest atom."""
    pos = np.asarray(positions, dtype=np.float64)
    cell = np.asarray(cell, dtype=np.float64)
    inv = np.linalg.inv(cell)
    frac = pos @ inv
    nat = len(pos)
    spec = np.zeros(nat, dtype=int) if species is None else np.asarray(species)
    grid = max(10, int(round(0.5 / tol)))

    def key(f):
        f = f - np.floor(f)
        k = np.round(f * grid).astype(np.int64) % grid
        return (k[..., 0] * grid + k[..., 1]) * grid + k[..., 2]

    ref = key(frac)
    order = np.argsort(ref)
    ref_sorted = ref[order]
    rot = np.asarray(rotations, dtype=np.float64)
    tr = np.asarray(translations, dtype=np.float64)
    irt = np.empty((len(rot), nat), dtype=np.intc)
    for s in range(len(rot)):
        img = (pos @ rot[s].T + tr[s]) @ inv
        k = key(img)
        j = np.minimum(np.searchsorted(ref_sorted, k), nat - 1)
        b = order[j]
        d = img - frac[b]
        bad = (ref_sorted[j] != k) | (np.abs(d - np.round(d)).max(axis=1) > tol)
        for a in np.nonzero(bad)[0]:
            dd = img[a][None] - frac
            dd = np.abs(dd - np.round(dd)).max(axis=1)
            b[a] = int(np.argmin(dd))
            if dd[b[a]] > tol:
                raise ValueError("symmetry %d sends atom %d to no atom" % (s, a))
        if np.any(spec[b] != spec):
            a = int(np.nonzero(spec[b] != spec)[0][0])
            raise ValueError("symmetry %d sends atom %d to an atom of another species" % (s, a))
        irt[s] = b
    return irt


def symmetries_on_modes(rotations, irt, pols, block_start, device=None, stream=None):
    """Per block b: (n_sym, d_b, d_b) numpy arrays S_b[s] = E_b^T P_s E_b (FP64).

    rotations (n_sym, 3, 3) Cartesian; irt (n_sym, nat) permutations; pols (3 nat, n) the
    mass-scaled polarization vectors of the kept modes; block_start the n_blocks + 1 offsets
    of contiguous blocks (see degenerate_blocks)."""
    import torch

    rot = np.ascontiguousarray(rotations, dtype=np.float64)
    irt = np.ascontiguousarray(irt, dtype=np.intc)
    pols = np.ascontiguousarray(pols, dtype=np.float64)
    starts = np.ascontiguousarray(block_start, dtype=np.intc)
    n_sym, nat = irt.shape
    nc, n = pols.shape
    if nc != 3 * nat or rot.shape != (n_sym, 3, 3):
        raise ValueError("shapes: rotations (n_sym, 3, 3), irt (n_sym, nat), pols (3 nat, n)")
    if irt.min() < 0 or irt.max() >= nat or \
            np.any(np.sort(irt, axis=1) != np.arange(nat, dtype=np.intc)[None]):
        raise ValueError("every row of irt must be a permutation of the atoms")
    if starts[0] != 0 or starts[-1] > n or np.any(np.diff(starts) <= 0):
        raise ValueError("block_start must increase from 0 to at most n_modes")
    if isinstance(device, torch.device):
        dev = device
    else:
        dev = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))
    sizes = np.diff(starts)
    total = int(n_sym * np.sum(sizes.astype(np.int64) ** 2))
    R = torch.as_tensor(rot).to(dev)
    P = torch.as_tensor(irt).to(dev)
    E = torch.as_tensor(pols).to(dev)
    out = torch.empty(total, dtype=torch.float64, device=dev)
    st = stream if stream is not None else torch.cuda.current_stream(dev)
    L.check(L.lib().tdscha_symmetries_on_modes(
        n_sym, nat, n, R.data_ptr(), P.data_ptr(), E.data_ptr(), len(sizes),
        starts.ctypes.data_as(L._ip), out.data_ptr(), st.cuda_stream),
        "tdscha_symmetries_on_modes")
    st.synchronize()
    flat = out.cpu().numpy()
    blocks, off = [], 0
    for d in sizes:
        blocks.append(flat[off:off + n_sym * d * d].reshape(n_sym, d, d))
        off += n_sym * d * d
    return blocks


def block_arrays(block_start, n_modes):
    """(N_degeneracy, degenerate_space, block_id) of contiguous blocks (Py:741-776)."""
    starts = np.asarray(block_start)
    ndeg = np.diff(starts).astype(np.intc)
    degs = [np.arange(starts[b], starts[b + 1], dtype=np.intc) for b in range(len(ndeg))]
    bid = np.empty(n_modes, dtype=np.intc)
    for b, dg in enumerate(degs):
        bid[dg] = b
    return ndeg, degs, bid


def _main():
    """Time the block construction for the synthetic rock-salt L x L x L supercell:
    python -m tdscha_amd.symmetry [--L 6] [--cpu-syms K]. --cpu-syms times the same
    contraction on the host (numpy) for the first K operations."""
    import argparse
    import json
    import time

    import torch
    from . import synthetic as syn

    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=6)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--cpu-syms", type=int, default=0)
    a = ap.parse_args()
    cr = syn.make_crystal(a.L)
    w, E, blocks = syn.modes(cr, syn.force_constants(cr))
    rots, perms = syn.space_group(cr)
    starts = np.array([b0 for b0, _ in blocks] + [blocks[-1][1]], dtype=np.intc)
    dev = torch.device("cuda", 0)
    n_sym, nat = perms.shape
    n = E.shape[1]
    sizes = np.diff(starts)
    total = int(n_sym * np.sum(sizes.astype(np.int64) ** 2))
    R = torch.as_tensor(rots).to(dev)
    P = torch.as_tensor(perms.astype(np.intc)).to(dev)
    Ed = torch.as_tensor(np.ascontiguousarray(E)).to(dev)
    out = torch.empty(total, dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream(dev)

    def launch():
        L.check(L.lib().tdscha_symmetries_on_modes(
            n_sym, nat, n, R.data_ptr(), P.data_ptr(), Ed.data_ptr(), len(sizes),
            starts.ctypes.data_as(L._ip), out.data_ptr(), st.cuda_stream), "symmetries")

    launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(a.reps):
        launch()
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / a.reps
    flop = 2.0 * n_sym * float(np.sum(sizes.astype(np.int64) ** 2)) * 3 * nat
    res = {"kernel": "k_sym_block", "L": a.L, "n_sym": n_sym, "nat": nat, "n_modes": n,
           "n_blocks": len(sizes), "max_block": int(sizes.max()), "ms": ms,
           "tflops": flop / (ms * 1e-3) / 1e12, "out_bytes": 8 * total}
    if a.cpu_syms:
        k = min(a.cpu_syms, n_sym)
        Er = E.reshape(nat, 3, n)
        t0 = time.perf_counter()
        for s in range(k):
            for b0, b1 in zip(starts[:-1], starts[1:]):
                Eb = Er[:, :, b0:b1]
                RE = np.einsum("xy,ayj->axj", rots[s], Eb)
                np.einsum("axi,axj->ij", Eb[perms[s]], RE)
        res["cpu_ms_per_sym"] = (time.perf_counter() - t0) * 1e3 / k
        res["cpu_ms_all_syms_est"] = res["cpu_ms_per_sym"] * n_sym
    print(json.dumps(res))


if __name__ == "__main__":
    _main()


def translation_cosets(rotations, irt):
    """Split space-group operations into cosets of the pure translations (rotation = identity).

    Returns (representatives, translations): one operation index per distinct rotation and the
    indices of the pure translations, when every rotation appears once per translation and
    the operations with one rotation are exactly the representative composed with each
    translation (irt_g o irt_t), as for a space group; None otherwise (then no reduction
    applies)."""
    rot = np.asarray(rotations, dtype=np.float64)
    irt = np.asarray(irt, dtype=np.int64)
    eye = np.abs(rot - np.eye(3)[None]).max(axis=(1, 2)) < 1e-10
    tr = np.nonzero(eye)[0]
    if len(tr) < 2 or len(rot) % len(tr):
        return None
    key = np.round(rot.reshape(len(rot), 9), 8)
    _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inverse = inverse.ravel()
    reps = []
    for k, g in enumerate(first):
        members = np.nonzero(inverse == k)[0]
        if len(members) != len(tr):
            return None
        composed = irt[g][irt[tr]]                  # irt_g o irt_t for every translation t
        if not np.array_equal(np.unique(composed, axis=0), np.unique(irt[members], axis=0)):
            return None
        reps.append(g)
    return np.sort(np.array(reps)), tr
