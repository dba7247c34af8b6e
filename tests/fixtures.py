"""Loaders for the committed golden fixtures and seeded synthetic test ensembles."""
import os

import numpy as np

from oracle import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_npz(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def load_mode10():
    """The 2x2x2 SnTe ensemble of the reference's tests (45 modes, 10 configs, 384 syms)."""
    g = load_npz("mode10_ensemble.npz")
    nb = len(g["N_degeneracy"])
    sym = orc.SymBlocks([g["sym_%d" % b] for b in range(nb)], g["N_degeneracy"],
                        [g["deg_%d" % b] for b in range(nb)], g["block_id"])
    return orc.Ensemble(g["X"], g["Y"], g["w"], g["rho"], float(g["T"]), sym)


def random_orthogonal(rng, d):
    q, r = np.linalg.qr(rng.standard_normal((d, d)))
    return q * np.sign(np.diag(r))


def random_ensemble(seed, n, N, n_sym, max_block=6, T=250.0, weights="lognormal"):
    """Random block-diagonal symmetry set (blocks of equal frequency) and a random ensemble
    with the magnitudes of a real one (|X| ~ 1e2, |Y| ~ 1e-5, w ~ 1e-4..1e-3 Ry)."""
    rng = np.random.default_rng(seed)
    sizes = []
    left = n
    while left > 0:
        d = int(rng.integers(1, max_block + 1))
        d = min(d, left)
        sizes.append(d)
        left -= d
    w = np.empty(n)
    syms, degs = [], []
    bid = np.empty(n, dtype=np.intc)
    m0 = 0
    for b, d in enumerate(sizes):
        w[m0:m0 + d] = rng.uniform(2e-4, 2e-3)
        s = np.empty((n_sym, d, d))
        s[0] = np.eye(d)
        for k in range(1, n_sym):
            s[k] = random_orthogonal(rng, d)
        syms.append(s)
        degs.append(np.arange(m0, m0 + d, dtype=np.intc))
        bid[m0:m0 + d] = b
        m0 += d
    sym = orc.SymBlocks(syms, np.array(sizes), degs, bid)
    X = rng.standard_normal((N, n)) * 60.0
    Y = rng.standard_normal((N, n)) * 1e-5
    if weights == "lognormal":
        rho = np.exp(0.2 * rng.standard_normal(N))
    else:
        rho = np.ones(N)
    return orc.Ensemble(X, Y, w, rho, T, sym)


def random_perturbation(seed, n, symmetric=True):
    rng = np.random.default_rng(seed)
    R1 = rng.standard_normal(n)
    A = rng.standard_normal((n, n)) * 1e-4
    Y1 = 0.5 * (A + A.T) if symmetric else A
    return R1, Y1


def split_frequencies(ens, rel):
    """The same ensemble with the frequencies inside every block of more than one mode
    spread by up to `rel` relative (a block of modes that are not exactly degenerate)."""
    w = ens.w.copy()
    for d in ens.sym.degenerate_space:
        if len(d) > 1:
            w[d] *= 1 + rel * np.linspace(-1.0, 1.0, len(d))
    return orc.Ensemble(ens.X, ens.Y, w, ens.rho, ens.T, ens.sym)
