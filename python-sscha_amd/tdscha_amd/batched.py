"""Many Lanczos chains over several systems of uneven size, one graph launch per step.

The batched form of the TD-SCHA spectra: per system, IR (3), Raman-like (6) and random (3)
starting vectors advance in one DeviceLanczos (every pass over the symmetry images serves
all of them), and the systems advance together through a LanczosGroup (one hipGraph per
step with one branch per system). The pass size of each ensemble is picked by the
autotuner first.

With zone_centre=True the nine IR and Raman chains of a system, which are translation
invariant, run on the zone-centre operator (gamma.GammaSectorOperator: one symmetry per
rotation, the translation average inside each captured step) and only the three random
chains on the full image set: two group members per system.

    python -m tdscha_amd.batched --steps 5 [--zone-centre]     (from python-sscha_amd/)
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
          log=None, zone_centre=False):
    import torch
    from . import synthetic as sy
    from .ensemble import DeviceEnsemble, SymmetryBlocks
    from .gamma import GammaSectorOperator
    from .lanczos_device import DeviceLanczos, LanczosGroup
    dev = device if device is not None else torch.device("cuda", 0)
    members, ensembles, info = [], [], []
    for L, N in systems:
        t0 = time.time()
        ens = sy.make_ensemble(L, N, T=300.0, device=dev)
        sym = SymmetryBlocks(ens.symmetries, ens.N_degeneracy, ens.degenerate_space,
                             ens.block_id)
        psis = starting_vectors(ens)
        parts = [(None, psis)]
        if zone_centre:   # IR + Raman (translation invariant) | random
            parts = [("zone_centre", psis[:9]), (None, psis[9:])]
        for kind, ps in parts:
            if kind == "zone_centre":
                rots, irt = sy.space_group(ens.crystal)
                d = GammaSectorOperator(ens.X, ens.Y, ens.w, ens.rho, ens.T, sym, rots, irt,
                                        ens.pols, precisions=("fp64", precision))
                n_syms = len(d.representatives)
            else:
                d = DeviceEnsemble(ens.X, ens.Y, ens.w, ens.rho, ens.T, sym,
                                   precisions=("fp64", precision))
                n_syms = ens.n_syms
            tuned = d.autotune(precision, 2 * len(ps)) if autotune else (
                d.ens.s_batch if kind else d.s_batch, float("nan"))
            lz = DeviceLanczos(d, len(ps), precision, max_steps=max_steps)
            lz.start(ps)
            members.append(lz)
            ensembles.append(d)
            info.append({"supercell": L, "n_modes": len(ens.w), "n_configs": N,
                         "n_syms": n_syms, "chains": len(ps), "slots_per_pass": tuned[0],
                         "us_per_image": tuned[1], "operator": kind or "full"})
            if log:
                log("system %dx%dx%d (%s): n=%d N=%d images per config %d, %d chains, pass "
                    "slots %d (%.3g us/image), %.1f s" % (
                        L, L, L, kind or "full", len(ens.w), N, n_syms, len(ps), tuned[0],
                        tuned[1], time.time() - t0))
    return LanczosGroup(members), members, ensembles, info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--precision", default="bf16x3")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--zone-centre", action="store_true",
                    help="IR and Raman chains on the zone-centre operator")
    args = ap.parse_args()
    import torch
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    group, members, ensembles, info = build(precision=args.precision,
                                            max_steps=args.steps + 1,
                                            autotune=not args.no_autotune, log=log,
                                            zone_centre=args.zone_centre)
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
    print(json.dumps({"workload": "uneven systems, 12 chains each, one graph per step%s"
                                  % (", IR and Raman chains on the zone-centre operator"
                                     if args.zone_centre else ""),
                      "precision": args.precision, "systems": info, "s_per_step": dt,
                      "L_applications_per_s": n_apply / dt,
                      "algorithmic_TFLOPs": flop / dt / 1e12}), flush=True)


if __name__ == "__main__":
    main()
