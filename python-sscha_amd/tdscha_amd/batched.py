"""Many Lanczos chains over several systems of uneven size, one graph launch per step.

The batched form of the TD-SCHA spectra: per system, IR (3), Raman-like (6) and random (3)
starting vectors advance in one DeviceLanczos (every pass over the symmetry images serves
all of them), and the systems advance together through a LanczosGroup (one hipGraph per
step with one branch per system). The pass size of each ensemble is picked by the
autotuner first.

    python -m tdscha_amd.batched --steps 5            (from python-sscha_amd/)
"""
import argparse
import json
import sys
import time

import numpy as np

SYSTEMS = [(2, 4000), (4, 10000), (5, 15000)]   # (supercell L, configurations)


def starting_vectors(ens, n_random=3, seed=3):
    n = len(ens.w)
    vecs = [ens.ir_perturbation(d) for d in range(3)]
    vecs += [ens.raman_like_perturbation(c) for c in range(6)]
    rng = np.random.default_rng(seed)
    vecs += [rng.standard_normal(n) for _ in range(n_random)]
    out = []
    for v in vecs:
        psi = np.zeros(n + n * (n + 1))
        psi[:n] = v / np.linalg.norm(v)
        out.append(psi)
    return out


def build(systems=SYSTEMS, precision="bf16x3", max_steps=20, device=None, autotune=True,
          log=None):
    import torch
    from . import synthetic as sy
    from .ensemble import DeviceEnsemble, SymmetryBlocks
    from .lanczos_device import DeviceLanczos, LanczosGroup
    dev = device if device is not None else torch.device("cuda", 0)
    members, ensembles, info = [], [], []
    for L, N in systems:
        t0 = time.time()
        ens = sy.make_ensemble(L, N, T=300.0, device=dev)
        sym = SymmetryBlocks(ens.symmetries, ens.N_degeneracy, ens.degenerate_space,
                             ens.block_id)
        d = DeviceEnsemble(ens.X, ens.Y, ens.w, ens.rho, ens.T, sym,
                           precisions=("fp64", precision))
        psis = starting_vectors(ens)
        tuned = d.autotune(precision, 2 * len(psis)) if autotune else (d.s_batch, float("nan"))
        lz = DeviceLanczos(d, len(psis), precision, max_steps=max_steps)
        lz.start(psis)
        members.append(lz)
        ensembles.append(d)
        info.append({"supercell": L, "n_modes": len(ens.w), "n_configs": N,
                     "n_syms": ens.n_syms, "chains": len(psis), "slots_per_pass": tuned[0],
                     "us_per_image": tuned[1]})
        if log:
            log("system %dx%dx%d: n=%d N=%d N_sym=%d, %d chains, pass slots %d (%.3g us/image), "
                "%.1f s" % (L, L, L, len(ens.w), N, ens.n_syms, len(psis), tuned[0], tuned[1],
                            time.time() - t0))
    return LanczosGroup(members), members, ensembles, info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--precision", default="bf16x3")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    import torch
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    group, members, ensembles, info = build(precision=args.precision,
                                            max_steps=args.steps + 1,
                                            autotune=not args.no_autotune, log=log)
    group.run(1)                       # capture and first replay
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        group.run(1)
        torch.cuda.synchronize()
        log("step %d done (%.1f s)" % (k, time.perf_counter() - t0))
    dt = (time.perf_counter() - t0) / max(args.steps, 1)
    n_apply = sum(2 * m.nc for m in members)
    flop = sum(2 * i["chains"] * 6.0 * i["n_configs"] * i["n_syms"] * i["n_modes"] ** 2
               for i in info)
    print(json.dumps({"workload": "uneven systems, 12 chains each, one graph per step",
                      "precision": args.precision, "systems": info, "s_per_step": dt,
                      "L_applications_per_s": n_apply / dt,
                      "algorithmic_TFLOPs": flop / dt / 1e12}), flush=True)


if __name__ == "__main__":
    main()
