#!/usr/bin/env python
"""L-applications per second of the TD-SCHA ensemble L-application on one MI355X.

Workload (default): synthetic rock-salt 6x6x6 supercell, 432 atoms, 1293 modes, 20000
configurations, the full Fm-3m x translations image set (10368 symmetries), T = 300 K.
A step is one batched L application (tdscha_apply_full_L_batched) to the pair a Lanczos
step needs, L q and L^T p, on seeded random unit vectors: for each chain the harmonic part,
the pre-transform, the ensemble average over all 20000 x 10368 images and the
post-transform, with both chains sharing every pass over the symmetries. Every input is
resident in HBM; value counts L applications (2 per step) per second. Before the warm-up the
autotuner picks the contraction tile shape and the image slots per pass for this ensemble
(timed on two passes per candidate, outside the timed region).

Prints one JSON line. `roofline` is measured live with HIP events around every launch of
each kernel group over the timed steps, for the contraction with the larger share of the
time; `roofline_other_contraction` gives the other one. `frac` counts algorithmic flops
(FLOP_L = 6 N N_sym n^2, GEMM1 2/6 of it, GEMM2 4/6); `frac_executed` the flops the kernel
executes (GEMM2 forms B_s once and uses B_s + B_s^T, half its algorithmic count).
`fp64_plain` is the plain FP64 path on the same ensemble: one L application on one vector
over every image, timed once, with its contractions' event times. `cpu_baseline` times the
reference's own C trio (oracle/_ref, built from CModules/LanczosFunctions.c) on the host
cores as independent processes on configuration slices (the reference's MODE_FAST_MPI
split) over a bounded sample of images of the same ensemble, scaled linearly to the full
image count.

`gamma_sector` is a separate line, not the headline: the same two chains projected on the
translation-invariant sector (the zone-centre perturbations of IR and Raman spectra live
there) and applied by tdscha_amd.gamma.GammaSectorOperator, which sums one symmetry per
rotation (48) and averages the result over the 216 translations; it is checked once against
the full operator on the same vectors (max_rel_diff_vs_full_operator).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "python-sscha_amd"))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (supercell L, configs, weights)
    "rocksalt6": (6, 20000, "uniform"),
    "rocksalt5": (5, 15000, "uniform"),
    "rocksalt4": (4, 10000, "lognormal"),
    "rocksalt2": (2, 2000, "uniform"),
}
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3.0}
# contraction kernels per BF16x3 tile form (tdscha_ensemble_set_tuning), GEMM1 and GEMM2
BF16_KERNELS = {
    0: ("k_gemm1_bf16x3<256, 128>", "k_gemm2_bf16x3<192, 128>"),
    1: ("k_gemm1_bf16x3<256, 256>", "k_gemm2_bf16x3<256, 256>"),
    2: ("k_gemm1_bf16x3_dma", "k_gemm2_bf16x3_dma"),
    3: ("k_gemm1_bf16x3_rp<2>", "k_gemm2_bf16x3_rp<2, 3>"),
    4: ("k_gemm1_bf16x3_rp<1>", "k_gemm2_bf16x3_rp<1, 3>"),
    5: ("k_gemm1_bf16x3_rp<1>", "k_gemm2_bf16x3_rp<2, 3>"),
    6: ("k_gemm1_bf16x3_rp<1>", "k_gemm2_bf16x3_rp<1, 4>"),
    7: ("k_gemm1_bf16x3_rp<1>", "k_gemm2_bf16x3_r8<4>"),
    8: ("k_gemm1_bf16x3_rp<1>", "k_gemm2_bf16x3_r7"),
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_bf16x3_rocksalt6.md")


def measured_traffic(kernel, workload, slots):
    """HBM bytes per launch of `kernel` from the committed hardware-counter summary
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this benchmark, calibrated against a
    copy kernel), for the same workload and image slots per pass; None if not measured."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    for line in lines:
        f = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(f) >= 4 and f[0] == kernel and f[1] == workload and f[2] == str(slots):
            try:
                return float(f[3])
            except ValueError:
                return None
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _ref_slice_worker(X, Y, w, rho, R1, Y1, T, blocks, barrier, queue):
    """One MODE_FAST_MPI rank: the reference C trio on its configuration slice (a child
    process started with spawn; it never touches the GPU)."""
    os.environ["OMP_NUM_THREADS"] = "1"
    from oracle import oracle as orc
    sym = orc.SymBlocks(*blocks)
    orc.ref_perturb_average_sym(X[:1], Y[:1], w, rho[:1], R1, Y1, T, 1, sym.subset(1))
    barrier.wait()
    t0 = time.perf_counter()
    orc.ref_perturb_average_sym(X, Y, w, rho, R1, Y1, T, 1, sym)
    queue.put(time.perf_counter() - t0)


def cpu_baseline(ens, per_process_seconds=2.0, max_procs=16):
    """The reference C trio (oracle/_ref, compiled from CModules/LanczosFunctions.c) on the
    host cores, MODE_FAST_MPI style: P independent processes, each on its own configuration
    with the same first k symmetries (the reference's MPI split is over configurations,
    LanczosFunctions.c:2123-2136), all started together; the rate is P k images over the
    slowest process's time, scaled linearly to N N_sym images (the trio's cost is linear in
    the image count)."""
    import multiprocessing as mp
    import numpy as np
    from oracle import oracle as orc
    if not orc.ref_available():
        return None
    n = len(ens.w)
    rng = np.random.default_rng(3)
    R1 = rng.standard_normal(n)
    A = rng.standard_normal((n, n)) * 1e-4
    Y1 = A + A.T
    sym = orc.SymBlocks(ens.symmetries, ens.N_degeneracy, ens.degenerate_space, ens.block_id)
    os.environ["OMP_NUM_THREADS"] = "1"
    # one core: per-image time on 1 configuration
    k = 1
    while True:
        t0 = time.perf_counter()
        orc.ref_perturb_average_sym(ens.X, ens.Y, ens.w, ens.rho, R1, Y1, ens.T, 1, sym,
                                    n_configs=1, n_syms=k)
        t1 = time.perf_counter() - t0
        if t1 > 0.5 or k >= ens.n_syms:
            break
        k = min(ens.n_syms, max(2 * k, int(k * 0.6 / max(t1, 1e-3))))
    per_image_1 = t1 / k
    k = max(1, min(ens.n_syms, int(per_process_seconds / per_image_1)))
    try:
        P = len(os.sched_getaffinity(0))
    except AttributeError:
        P = os.cpu_count() or 1
    P = max(1, min(P, max_procs, ens.X.shape[0]))
    sub = sym.subset(k)
    blocks = (sub.symmetries, sub.N_degeneracy, sub.degenerate_space, sub.block_id)
    ctx = mp.get_context("spawn")
    barrier = ctx.Barrier(P)
    queue = ctx.Queue()
    procs = [ctx.Process(target=_ref_slice_worker,
                         args=(ens.X[r:r + 1].copy(), ens.Y[r:r + 1].copy(), ens.w,
                               ens.rho[r:r + 1].copy(), R1, Y1, ens.T, blocks, barrier, queue))
             for r in range(P)]
    for pr in procs:
        pr.start()
    times = [queue.get(timeout=600) for _ in procs]
    for pr in procs:
        pr.join()
    wall = max(times)
    rate = P * k / wall                      # images per second on P cores
    images = ens.X.shape[0] * ens.n_syms
    return {"value": rate / images, "unit": "L-applications/s", "cores": P,
            "kind": "reference",
            "sample": "reference C trio (get_f_average / get_d2v_dR2 from R / from Y, -O2, "
                      "serial per process) as %d independent processes (the MODE_FAST_MPI "
                      "split over configurations), each 1 config x the first %d symmetries "
                      "(slowest %.2f s, %.1f s of CPU work in all; one core alone: %.3g s per "
                      "image), scaled linearly to %d configs x %d symmetries" % (
                          P, k, wall, sum(times), per_image_1, ens.X.shape[0], ens.n_syms),
            "single_core_s_per_image": per_image_1}


def gamma_sector_line(args, ens, sym, precs, dens, psis, flags, shifts, full_value, reps=5):
    """The same chains projected on the translation-invariant sector (where IR and Raman
    perturbations at the zone centre live) and applied by the reduced operator (one symmetry
    per rotation, then the translation average); checked once against the full operator on
    the same vectors. A separate line: the headline value is the full operator on general
    vectors."""
    import torch
    from tdscha_amd import _lib as L
    from tdscha_amd import synthetic as sy
    from tdscha_amd.gamma import GammaSectorOperator
    rots, irt = sy.space_group(ens.crystal)
    t0 = time.time()
    gop = GammaSectorOperator(ens.X, ens.Y, ens.w, ens.rho, ens.T, sym, rots, irt, ens.pols,
                              precisions=precs, s_batch=args.s_batch)
    if args.precision == "bf16x3" and not args.no_autotune:
        gop.autotune(args.precision, args.chains)
    vs = []
    for p, f in zip(psis, flags):
        v = gop.proj.project(p, adjoint=bool(f & L.OP_TRANSPOSE))
        vs.append(v / v.norm())
    outs = [torch.empty_like(v) for v in vs]
    gop.apply_full_L_batched(vs, flags, shifts, precision=args.precision, outs=outs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        gop.apply_full_L_batched(vs, flags, shifts, precision=args.precision, outs=outs)
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    rate = args.chains * reps / el
    ref = dens.apply_full_L_batched(vs, flags, shifts, precision=args.precision)
    torch.cuda.synchronize()
    diff = max(float((o - r).abs().max() / r.abs().max()) for o, r in zip(outs, ref))
    log("zone-centre operator: %.4g L/s (%d representatives x %d translations), %.0fx the "
        "full operator, max rel diff %.2g, set-up %.1f s" % (
            rate, len(gop.representatives), len(gop.translations), rate / full_value, diff,
            time.time() - t0))
    res = {"value": rate, "unit": "L-applications/s",
           "sample": "the %d chains projected on the translation-invariant sector, %d batched "
                     "applications (%.3f s) over %d symmetries (one per rotation) and the "
                     "average over %d translations" % (args.chains, reps, el,
                                                       len(gop.representatives),
                                                       len(gop.translations)),
           "speedup_vs_full_operator": rate / full_value,
           "max_rel_diff_vs_full_operator": diff}
    gop.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rocksalt6", choices=sorted(WORKLOADS))
    ap.add_argument("--configs", type=int, default=0, help="override the number of configs")
    ap.add_argument("--precision", default="bf16x3", choices=sorted(PEAK_TFLOPS))
    ap.add_argument("--chains", type=int, default=2,
                    help="vectors per batched application (alternating L and L^T)")
    ap.add_argument("--s-batch", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile shape and pass size instead of timing them")
    ap.add_argument("--tiles", type=int, default=-1,
                    help="BF16x3 tile form 0-8 (with --pass-slots: no autotuning)")
    ap.add_argument("--pass-slots", type=int, default=0, help="image slots per pass")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="after the timed steps, copy a buffer of this many MiB on the device "
                         "(a kernel of known traffic for hardware-counter calibration)")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the plain FP64 comparison application")
    ap.add_argument("--sustained-seconds", type=float, default=8.0,
                    help="seconds of the pure MFMA stream that measures the sustained MFMA rate "
                         "under the board's power limit, per precision (0: skip)")
    ap.add_argument("--no-gamma", action="store_true",
                    help="skip the zone-centre (translation-invariant) operator line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this package has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from tdscha_amd import _lib as L
    from tdscha_amd import synthetic as sy
    from tdscha_amd.ensemble import DeviceEnsemble, SymmetryBlocks

    dev = torch.device("cuda", 0)
    Lc, N, weights = WORKLOADS[args.workload]
    if args.configs:
        N = args.configs
    t0 = time.time()
    ens = sy.make_ensemble(Lc, N, T=300.0, weights=weights, device=dev)
    n = len(ens.w)
    log("ensemble %s: n=%d N=%d N_sym=%d blocks=%d (max %d) in %.1f s" % (
        args.workload, n, N, ens.n_syms, len(ens.N_degeneracy), max(ens.N_degeneracy),
        time.time() - t0))
    sym = SymmetryBlocks(ens.symmetries, ens.N_degeneracy, ens.degenerate_space, ens.block_id)
    precs = ("fp64",) if args.precision == "fp64" else ("fp64", args.precision)
    dens = DeviceEnsemble(ens.X, ens.Y, ens.w, ens.rho, ens.T, sym, precisions=precs,
                          s_batch=args.s_batch)
    log("device ensemble: %.2f GB, s_batch=%d" % (dens.nbytes / 1e9, dens.s_batch))
    g = torch.Generator().manual_seed(3)
    psis = []
    for _ in range(args.chains):
        v = torch.randn(n + n * (n + 1), generator=g, dtype=torch.float64)
        psis.append((v / v.norm()).to(dev))
    outs = [torch.empty_like(p) for p in psis]
    flags = [L.OP_TRANSPOSE * (c % 2) for c in range(args.chains)]
    shifts = [0.0] * args.chains

    def step():
        dens.apply_full_L_batched(psis, flags, shifts, precision=args.precision, outs=outs)

    if args.tiles >= 0:
        dens.set_tuning(args.tiles, args.pass_slots)
        log("tuning set: tiles %d, %d slots per pass" % dens.tuning())
    elif not args.no_autotune:
        t0 = time.time()
        slots, per = dens.autotune(args.precision, args.chains)
        log("autotune: tiles %d, %d slots per pass (%.3g us per image) in %.1f s" % (
            dens.tuning()[0], slots, per, time.time() - t0))

    for k in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log("warmup %d done" % k)
    L.lib().tdscha_profile_reset()
    L.lib().tdscha_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        torch.cuda.synchronize()
        log("step %d done (%.1f s)" % (k, time.perf_counter() - t0))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    L.lib().tdscha_profile_enable(0)
    if not all(bool(torch.isfinite(o).all()) for o in outs):
        raise RuntimeError("non-finite output")

    if args.copy_calibration_mb:
        src = torch.ones(args.copy_calibration_mb * 131072, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: %d MiB read, %d MiB written" % (args.copy_calibration_mb,
                                                                args.copy_calibration_mb))
        del src, dst
    dens_tuning = dens.tuning()
    groups = {name: L.profile_query(gid) for gid, name in L.KGROUP_NAMES.items()}
    log("kernel groups (ms total, launches): %s" % groups)
    M = N * ens.n_syms
    flop_l = 6.0 * M * n * n
    # algorithmic flops per launch: GEMM1 carries the quadratic forms (2 N n^2 per symmetry),
    # GEMM2 the back-accumulation (4 N n^2 per symmetry)
    # (GEMM2 forms B_s once and uses B_s + B_s^T: it executes half its algorithmic flops)
    share = {"gemm1": 2.0 / 6.0, "gemm2": 4.0 / 6.0}
    executed = {"gemm1": 1.0, "gemm2": 0.5}
    peak = PEAK_TFLOPS[args.precision]
    slots_run = (dens.tuning()[1] // args.chains) * args.chains

    def kernel_roofline(grp):
        ms, cnt = groups[grp]
        per_launch_flop = flop_l * args.chains * share[grp] / (cnt / args.steps)
        achieved = per_launch_flop / (ms / cnt * 1e-3) / 1e12
        if args.precision == "bf16x3":
            kernel = BF16_KERNELS[dens.tuning()[0]][0 if grp == "gemm1" else 1]
        else:
            kernel = "k_%s<%s>" % (grp, "double" if args.precision == "fp64" else "float")
        return {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                "traffic": measured_traffic(kernel, args.workload, slots_run),
                "kernel": kernel, "slots_per_launch": slots_run,
                "ms_per_launch": round(ms / cnt, 3),
                "frac_executed": round(achieved * executed[grp] / peak, 4)}

    dom = max(("gemm1", "gemm2"), key=lambda k: groups[k][0])
    roofline = kernel_roofline(dom)
    roofline_other = kernel_roofline("gemm2" if dom == "gemm1" else "gemm1")
    value = args.chains * args.steps / elapsed
    gamma = None
    if not args.no_gamma:
        try:
            gamma = gamma_sector_line(args, ens, sym, precs, dens, psis, flags, shifts, value)
        except (ValueError, RuntimeError) as e:     # an auxiliary line: the headline stands
            log("zone-centre operator line failed: %s" % e)
            gamma = {"error": str(e)}
    fp64_ref = None
    if args.precision != "fp64" and not args.no_fp64:
        # the plain FP64 path: one L application on one vector over every image, timed once
        # (about 30 s at the default workload), with the contractions timed by events
        del dens
        torch.cuda.empty_cache()
        d64 = DeviceEnsemble(ens.X, ens.Y, ens.w, ens.rho, ens.T, sym, s_batch=args.s_batch)
        log("fp64 plain: device ensemble %.2f GB, s_batch %d" % (d64.nbytes / 1e9,
                                                                  d64.s_batch))
        L.lib().tdscha_profile_reset()
        L.lib().tdscha_profile_enable(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d64.apply_full_L(psis[0], precision="fp64", out=outs[0])
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        L.lib().tdscha_profile_enable(0)
        g64 = {name: L.profile_query(gid) for gid, name in L.KGROUP_NAMES.items()}
        v64 = 1.0 / t64
        rl64 = {}
        for grp in ("gemm1", "gemm2"):
            ms, cnt = g64[grp]
            if cnt:
                ach = flop_l * share[grp] / cnt / (ms / cnt * 1e-3) / 1e12
                rl64[grp] = {"kernel": "k_%s<double>" % grp, "ms_per_launch": round(ms / cnt, 3),
                             "launches": cnt, "achieved": round(ach, 3),
                             "frac": round(ach / PEAK_TFLOPS["fp64"], 4),
                             "frac_executed": round(ach * executed[grp] /
                                                    PEAK_TFLOPS["fp64"], 4)}
        fp64_ref = {"value": v64, "unit": "L-applications/s",
                    "sample": "plain FP64 path, one L application on one vector over all %d "
                              "configs x %d symmetries (%.2f s, timed once)" % (
                                  N, ens.n_syms, t64),
                    "speedup": value / v64, "kernel_groups_ms": g64, "roofline": rl64}
        log("fp64 plain: %.4g L/s (%.2f s) -> speedup %.2fx; groups %s" % (
            v64, t64, value / v64, g64))
        d64.close()
    if args.sustained_seconds > 0:
        # the sustained MFMA rate of this board under its power limit (random operands on
        # every CU), beside the published peak: issued MFMA work over it is the share of
        # what the power budget allows
        def stream(prec):
            tf, ms = ctypes.c_double(), ctypes.c_double()
            L.check(L.lib().tdscha_mfma_stream(L.PRECISIONS[prec], float(args.sustained_seconds),
                                               ctypes.byref(tf), ctypes.byref(ms)),
                    "tdscha_mfma_stream")
            log("sustained %s MFMA stream: %.1f TFLOP/s (%.2f ms per launch), ended at "
                "%.1f (epoch s)" % (prec, tf.value, ms.value, time.time()))
            return tf.value
        mfma_prec = "bf16x3" if args.precision == "bf16x3" else "fp64"
        issue = 3.0 if args.precision == "bf16x3" else 1.0   # MFMAs per product term
        if args.precision in ("bf16x3", "fp64"):
            sp = stream(mfma_prec)
            for rl, grp in ((roofline, dom), (roofline_other, "gemm2" if dom == "gemm1" else
                                              "gemm1")):
                rl["sustained_mfma_tflops"] = round(sp, 1)
                rl["frac_executed_sustained"] = round(rl["achieved"] * executed[grp] * issue / sp,
                                                      4)
        if fp64_ref is not None:
            sp64 = stream("fp64")
            for grp, r in fp64_ref["roofline"].items():
                r["sustained_mfma_tflops"] = round(sp64, 1)
                r["frac_executed_sustained"] = round(r["achieved"] * executed[grp] / sp64, 4)
    base = None
    if not args.no_cpu_baseline:
        base = cpu_baseline(ens)
    res = {
        "metric": "L-applications/sec at 432 atoms, 20k configs: BF16x3 batched vs fp64 plain",
        "value": value,
        "unit": "L-applications/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "bf16x3": "bf16x3"}[args.precision],
        "data": "synthetic",
        "config": {"workload": "rock-salt %dx%dx%d supercell, %d atoms, %d modes, %d configs, "
                               "%d symmetry images per config, %d-chain batched L application "
                               "(%s)" % (Lc, Lc, Lc, 2 * Lc ** 3, n, N, ens.n_syms, args.chains,
                                         args.precision),
                   "chains": args.chains,
                   "bf16_tiles": dens_tuning[0], "slots_per_pass": dens_tuning[1],
                   "n_modes": n, "n_configs": N, "n_syms": ens.n_syms,
                   "flop_per_L_application": flop_l,
                   "achieved_TFLOPs_whole_L": round(flop_l * value / 1e12, 3)},
        "roofline": roofline,
        "roofline_other_contraction": roofline_other,
        "cpu_baseline": base,
        "fp64_plain": fp64_ref,
        "gamma_sector": gamma,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
