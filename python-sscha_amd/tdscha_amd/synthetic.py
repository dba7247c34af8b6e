"""Seeded synthetic TD-SCHA ensembles of a rock-salt crystal (workload generator).

Crystal: fcc rock salt, conventional a = 6.3 A, species A (118.7 amu) at the origin and
B (127.6 amu) at (a/2, 0, 0); an L x L x L supercell of the 2-atom primitive cell
(2 L^3 atoms, 6 L^3 - 3 non-translational modes).

Harmonic force constants: central springs on the nearest-neighbour A-B bonds along <100>
(k1 = 1.5 eV/A^2). Springs on <100> bonds alone leave every rigid slide of an atomic line
along itself at zero frequency, so a second-shell central spring on the same-species
<110> bonds (k2 = 0.3 eV/A^2) is added to make every non-translational mode stable.

Anharmonicity: cubic and quartic bond potentials on the A->B bonds,
V = sum_bonds A_abc d_a d_b d_c / 6 + B_abcd d_a d_b d_c d_d / 24, d = u_B - u_A, with A and B
Gaussian random symmetric tensors (sigma 0.1 eV/A^3 and 0.05 eV/A^4) made invariant under
the bond's site group and carried to the six bond directions by the Oh operations, so the
potential has the full Fm-3m x translation symmetry.

Ensemble: displacements drawn from the SCHA Gaussian of the harmonic model at temperature T
(mode variance (1 + 2 n) / (2 w)), forces from the full model; X and Y follow the
reference's projection (Modules/DynamicalLanczos.py:380-442): Bohr * sqrt(m) and
Ry / Bohr / sqrt(m) on the polarisation vectors, SCHA (harmonic) force removed. The
symmetrised mean force vanishes on every site of this structure (each atom is an
inversion centre), so no mean is subtracted.

Symmetries: the 48 Oh operations about an A site times the L^3 supercell translations,
in the mode basis restricted to blocks of degenerate modes (frequencies grouped to 1e-7
relative and set equal inside a block), as Lanczos.prepare_symmetrization leaves them
(Modules/DynamicalLanczos.py:741-778).

Units as the reference: Ry, Bohr, masses in units of 2 m_e (1 amu = 911.444243 of them).
"""
import itertools
from dataclasses import dataclass

import numpy as np

RY_EV = 13.605698066          # Py:132
BOHR_A = 1.0 / 1.889725989    # Py:133 (Bohr per Angstrom inverse)
AMU_RY = 911.444243096        # atomic mass unit in Ry mass units (2 m_e)
RY_TO_K = 157887.32400374097
RY_TO_CM = 109737.31568160

EV_A2 = (1 / RY_EV) * BOHR_A ** 2   # eV/A^2 -> Ry/Bohr^2
EV_A3 = (1 / RY_EV) * BOHR_A ** 3
EV_A4 = (1 / RY_EV) * BOHR_A ** 4

PRIM = np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]])   # rows: primitive vectors in units a/2
PRIM_INV2 = np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1]])  # 2 * inverse


def oh_operations():
    """The 48 signed permutation matrices of the cubic point group Oh."""
    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3), dtype=int)
            for r, c in enumerate(perm):
                R[r, c] = signs[r]
            ops.append(R)
    # identity first
    ops.sort(key=lambda R: (not np.array_equal(R, np.eye(3, dtype=int)),))
    return np.array(ops)


@dataclass
class Crystal:
    L: int
    a: float
    pos: np.ndarray        # (N_at, 3) integer positions in units of a/2
    species: np.ndarray    # (N_at,) 0 = A, 1 = B
    cell: np.ndarray       # (N_at, 3) primitive-cell indices
    masses: np.ndarray     # (N_at,) Ry mass units

    @property
    def nat(self):
        return len(self.species)

    def index_of(self, p, spec):
        """Atom index of integer positions p (..., 3) of species spec (...)."""
        base = p - spec[..., None] * np.array([1, 0, 0])
        n = (base @ PRIM_INV2.T) // 2
        n %= self.L
        return ((n[..., 0] * self.L + n[..., 1]) * self.L + n[..., 2]) * 2 + spec


def make_crystal(L, a=6.3, m_a=118.7, m_b=127.6):
    cells = np.array(list(itertools.product(range(L), repeat=3)))
    pos, spec, cell = [], [], []
    for n in cells:
        base = n @ PRIM
        for s in (0, 1):
            pos.append(base + (np.array([1, 0, 0]) if s else 0))
            spec.append(s)
            cell.append(n)
    spec = np.array(spec)
    masses = np.where(spec == 0, m_a, m_b) * AMU_RY
    return Crystal(L, a, np.array(pos), spec, np.array(cell), masses)


def _bonds(cr: Crystal, shell):
    """Directed bonds (i, j, unit vector) of a shell: 1 = A->B <100>, 2 = same-species <110>."""
    out_i, out_j, out_e = [], [], []
    if shell == 1:
        dirs = [np.array(v) for v in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1),
                                      (0, 0, -1))]
        src = np.nonzero(cr.species == 0)[0]
        for d in dirs:
            j = cr.index_of(cr.pos[src] + d, np.ones(len(src), dtype=int))
            out_i.append(src)
            out_j.append(j)
            out_e.append(np.tile(d / np.linalg.norm(d), (len(src), 1)))
    else:
        dirs = []
        for v in itertools.product((-1, 0, 1), repeat=3):
            if sum(abs(x) for x in v) == 2:
                dirs.append(np.array(v))
        for s in (0, 1):
            src = np.nonzero(cr.species == s)[0]
            for d in dirs:
                if tuple(d) < tuple(-d):
                    continue   # each undirected bond once
                j = cr.index_of(cr.pos[src] + d, np.full(len(src), s))
                out_i.append(src)
                out_j.append(j)
                out_e.append(np.tile(d / np.linalg.norm(d), (len(src), 1)))
    return np.concatenate(out_i), np.concatenate(out_j), np.concatenate(out_e)


def force_constants(cr: Crystal, k1=1.5, k2=0.3):
    """Harmonic force constants (3N x 3N) in Ry/Bohr^2."""
    nat = cr.nat
    phi = np.zeros((nat, 3, nat, 3))
    for shell, k in ((1, k1), (2, k2)):
        i, j, e = _bonds(cr, shell)
        kk = k * EV_A2 * np.einsum("ba,bc->bac", e, e)
        np.add.at(phi, (i, slice(None), i, slice(None)), kk)
        np.add.at(phi, (j, slice(None), j, slice(None)), kk)
        np.add.at(phi, (i, slice(None), j, slice(None)), -kk)
        np.add.at(phi, (j, slice(None), i, slice(None)), -kk)
    return phi.reshape(3 * nat, 3 * nat)


def _sym_tensor(rng, rank, sigma):
    t = rng.standard_normal((3,) * rank) * sigma
    acc = np.zeros_like(t)
    perms = list(itertools.permutations(range(rank)))
    for p in perms:
        acc += np.transpose(t, p)
    return acc / len(perms)


def _rotate(t, R):
    rank = t.ndim
    for ax in range(rank):
        t = np.moveaxis(np.tensordot(R, t, axes=([1], [ax])), 0, ax)
    return t


def bond_tensors(seed=0, sigma3=0.1, sigma4=0.05):
    """Cubic / quartic tensors (Ry/Bohr^3, Ry/Bohr^4) for the six <100> bond directions."""
    rng = np.random.default_rng(seed)
    ops = oh_operations()
    ex = np.array([1, 0, 0])
    stab = [R for R in ops if np.array_equal(R @ ex, ex)]
    out = {}
    for rank, sigma, unit in ((3, sigma3, EV_A3), (4, sigma4, EV_A4)):
        t = _sym_tensor(rng, rank, sigma)
        t = sum(_rotate(t, R.astype(float)) for R in stab) / len(stab)
        for R in ops:
            d = tuple(R @ ex)
            if (rank, d) not in out:
                out[(rank, d)] = _rotate(t, R.astype(float)) * unit
    return out


def anharmonic_forces(cr: Crystal, u, tensors):
    """Forces (N, 3 nat) in Ry/Bohr from the bond cubic + quartic potentials, u (N, 3 nat) Bohr.

    u may be a numpy array or a torch tensor (then the result is a tensor on its device)."""
    import torch
    as_np = isinstance(u, np.ndarray)
    ut = torch.as_tensor(u)
    N = ut.shape[0]
    u3 = ut.reshape(N, cr.nat, 3)
    f = torch.zeros_like(u3)
    src_np = np.nonzero(cr.species == 0)[0]
    src = torch.as_tensor(src_np, device=ut.device)
    for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        j = torch.as_tensor(cr.index_of(cr.pos[src_np] + np.array(d),
                                        np.ones(len(src_np), dtype=int)), device=ut.device)
        dd = u3[:, j, :] - u3[:, src, :]
        A = torch.as_tensor(tensors[(3, d)], device=ut.device, dtype=ut.dtype)
        B = torch.as_tensor(tensors[(4, d)], device=ut.device, dtype=ut.dtype)
        q = torch.einsum("nkb,nkc->nkbc", dd, dd)
        g = 0.5 * torch.einsum("abc,nkbc->nka", A, q)
        g += torch.einsum("abcd,nkbc,nkd->nka", B, q, dd) / 6.0
        f.index_add_(1, j, -g)
        f.index_add_(1, src, g)
    f = f.reshape(N, 3 * cr.nat)
    return f.cpu().numpy() if as_np else f


@dataclass
class SyntheticEnsemble:
    """What the Lanczos needs: X, Y, w, rho, T, the symmetry blocks, pols and masses."""
    X: np.ndarray
    Y: np.ndarray
    w: np.ndarray
    rho: np.ndarray
    T: float
    symmetries: list
    N_degeneracy: np.ndarray
    degenerate_space: list
    block_id: np.ndarray
    pols: np.ndarray
    m: np.ndarray
    crystal: Crystal

    @property
    def n_syms(self):
        return self.symmetries[0].shape[0]

    def ir_perturbation(self, direction=0, zeff=2.0):
        """R-sector perturbation of an IR field along `direction` with Z* = +-zeff I
        (prepare_ir, Modules/DynamicalLanczos.py:1647-1689; masses_exp = -1)."""
        nat = self.crystal.nat
        v = np.zeros((nat, 3))
        v[:, direction] = np.where(self.crystal.species == 0, zeff, -zeff)
        return np.einsum("a,a,ab->b", 1 / np.sqrt(self.m), v.ravel(), self.pols)

    def raman_like_perturbation(self, component, seed=7):
        """R-sector perturbation of one polarizability component (0..5: xx, yy, zz, xy, xz,
        yz) from seeded synthetic Raman tensors d(alpha_ij)/dR_a, one per atom of the
        primitive cell and the same in every cell (a zone-centre perturbation, as the
        reference's prepare_raman builds it from the unit cell's tensor,
        Modules/DynamicalLanczos.py:1003-1124), in the form of Py:1693-1781. Rock salt has no
        first-order Raman activity, so these tensors are not symmetric: they only give
        twelve-chain batches realistic, distinct starting vectors."""
        pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
        i, j = pairs[component]
        t = np.random.default_rng(seed).standard_normal((2, 3, 3, 3))
        t = 0.5 * (t + t.transpose(0, 2, 1, 3))      # symmetric in the field indices
        v = t[self.crystal.species, i, j, :]
        return np.einsum("a,a,ab->b", 1 / np.sqrt(self.m), v.ravel(), self.pols)


def modes(cr: Crystal, phi, tol=1e-7):
    """Frequencies (Ry), mass-scaled polarisation vectors and degenerate blocks, translations
    removed. Inside a block the frequency is set to the block mean."""
    m3 = np.repeat(cr.masses, 3)
    D = phi / np.sqrt(np.outer(m3, m3))
    lam, E = np.linalg.eigh(D)
    # the three uniform translations are the smallest eigenvalues (zero by construction)
    trans = np.argsort(np.abs(lam))[:3]
    keep = np.setdiff1d(np.arange(len(lam)), trans)
    lam, E = lam[keep], E[:, keep]
    if lam.min() <= 0:
        raise ValueError("unstable harmonic model: smallest eigenvalue %g" % lam.min())
    w = np.sqrt(lam)
    blocks = []
    start = 0
    for k in range(1, len(w) + 1):
        if k == len(w) or (w[k] - w[k - 1]) > tol * w[k]:
            blocks.append((start, k))
            start = k
    for b0, b1 in blocks:
        w[b0:b1] = w[b0:b1].mean()
    return w, E, blocks


def space_group(cr: Crystal):
    """Cartesian rotations (n_sym, 3, 3) and atom permutations (n_sym, nat) of the
    Oh x translations operations of the supercell."""
    ops = oh_operations()
    trans = cr.cell[cr.species == 0] @ PRIM   # lattice translations in a/2 units, one per cell
    nat = cr.nat
    perms = np.empty((len(ops) * len(trans), nat), dtype=np.int64)
    rots = np.empty((len(ops) * len(trans), 3, 3))
    g = 0
    for R in ops:
        rp = cr.pos @ R.T
        for t in trans:
            perms[g] = cr.index_of(rp + t, cr.species)
            rots[g] = R
            g += 1
    return rots, perms


def symmetry_blocks(cr: Crystal, E, blocks, chunk=256, device=None):
    """S_g restricted to the degenerate blocks for g in Oh x translations (on a GPU device
    with the library's kernel, tdscha_symmetries_on_modes; on the CPU with torch)."""
    import torch
    dev = device if device is not None else torch.device("cpu")
    rots, perms = space_group(cr)
    n_sym, nat = perms.shape
    n = E.shape[1]
    if dev.type == "cuda":
        from . import symmetry as sym
        starts = np.array([b0 for b0, _ in blocks] + [blocks[-1][1]], dtype=np.intc)
        return sym.symmetries_on_modes(rots, perms, E, starts, device=dev), perms, rots
    Et = torch.as_tensor(E.reshape(nat, 3, n), device=dev)
    out = [np.empty((n_sym, b1 - b0, b1 - b0)) for b0, b1 in blocks]
    chunk = max(1, min(chunk, int(2e8 // (nat * 3 * n))))
    for c0 in range(0, n_sym, chunk):
        c1 = min(n_sym, c0 + chunk)
        G = c1 - c0
        P = torch.as_tensor(perms[c0:c1], device=dev)
        R = torch.as_tensor(rots[c0:c1], device=dev)
        RE = torch.matmul(R[:, None], Et[None]).reshape(G, 3 * nat, n)   # R_g E[j]
        Ep = Et[P].reshape(G, 3 * nat, n)                                # E[pi_g(j)]
        for k, (b0, b1) in enumerate(blocks):
            S = torch.bmm(Ep[:, :, b0:b1].transpose(1, 2), RE[:, :, b0:b1])   # E_b^T P_g E_b
            out[k][c0:c1] = S.cpu().numpy()
    return out, perms, rots


def make_ensemble(L, n_configs, T=300.0, weights="uniform", seed_fc=0, seed_cfg=1, seed_w=2,
                  device=None, chunk_configs=2000, sigma3=0.1, sigma4=0.05):
    """The synthetic ensemble of an L x L x L supercell (see the module docstring)."""
    cr = make_crystal(L)
    phi = force_constants(cr)
    w, E, blocks = modes(cr, phi)
    n = len(w)
    nocc = 1 / (np.exp(w * RY_TO_K / T) - 1) if T > 0 else np.zeros(n)
    sig = np.sqrt((1 + 2 * nocc) / (2 * w))
    import torch
    dev = device if device is not None else torch.device("cpu")
    rng = np.random.default_rng(seed_cfg)
    tens = bond_tensors(seed_fc, sigma3, sigma4)
    m3 = np.repeat(cr.masses, 3)
    Et = torch.as_tensor(E, device=dev)
    sm = torch.as_tensor(np.sqrt(m3), device=dev)
    X = np.empty((n_configs, n))
    Y = np.empty((n_configs, n))
    for c0 in range(0, n_configs, chunk_configs):
        c1 = min(n_configs, c0 + chunk_configs)
        x = rng.standard_normal((c1 - c0, n)) * sig
        xt = torch.as_tensor(x, device=dev)
        u = (xt @ Et.T) / sm                    # Bohr
        f = anharmonic_forces(cr, u, tens)      # f - f_SCHA, Ry/Bohr
        X[c0:c1] = x
        Y[c0:c1] = ((f / sm) @ Et).cpu().numpy()
    rng_w = np.random.default_rng(seed_w)
    if weights == "lognormal":
        rho = np.exp(0.2 * rng_w.standard_normal(n_configs))
    else:
        rho = np.ones(n_configs)
    syms, _, _ = symmetry_blocks(cr, E, blocks, device=device)
    ndeg = np.array([b1 - b0 for b0, b1 in blocks], dtype=np.intc)
    degs = [np.arange(b0, b1, dtype=np.intc) for b0, b1 in blocks]
    bid = np.empty(n, dtype=np.intc)
    for k, (b0, b1) in enumerate(blocks):
        bid[b0:b1] = k
    return SyntheticEnsemble(X, Y, w, rho, T, syms, ndeg, degs, bid, E / 1.0, m3, cr)
