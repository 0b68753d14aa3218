#!/usr/bin/env python3
"""Benchmark of the Wilson fermion-matrix hot path on one MI355X.

Workload: 32x32x32x16 lattice, kappa = 0.1355, synthetic gauge U = reunitarise(exp(i 0.5 H))
(seed 1), Gaussian source (seed 2), fermion boundary (1, 1, 1, -1).

A step is one application of the full-lattice Wilson D (the reference operator
_kernelDFermionWilsonSquareSU3) in FP64 to one right-hand side already resident in HBM.
`value` is the compulsory-byte Dslash bandwidth: the bytes no schedule can avoid moving between
HBM and the chip for one D (the site spinor in, 4 links per site, the spinor out: 120 reals x 8
bytes = 960 B per site) times the volume, divided by the measured time per step.  `roofline`
prices the same bytes against the kernel's HIP-event time and the 8 TB/s HBM peak, so its `frac`
cannot exceed what the hardware allows; the no-reuse count (384 reals per site: every neighbour
spinor and link counted at each use) is kept in `extra` only.  The same line carries FP32 /
12-real-link Dslash rates, the even-odd Schur operator, the staggered operator and the force
kernels (all on compulsory bytes), GMRES(12) solves at 32^3x16 (plain FP64 to 1e-10, plain FP32 to
1e-6, FP32/FP16-inner defect correction to 1e-10) with their iteration counts, the 32^3x32
kappa 0.137 mixed solves with their error against the CPU oracle's FP64 solution, the batched
12+4-system solve, the CPU oracle's Dslash on the host cores, and one staggered RHMC trajectory
at 24^3x8 (2+1 flavours, 12-pole rational functions on log-spaced shifts in [1e-4, 10], nested
force gradient 8 outer x 6 nested steps, beta 5.5) run once with FP32 multi-shift force solves
and once in plain FP64, with Delta H of both.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

DIMS = (32, 32, 32, 16)
KAPPA = 0.1355
HBM_PEAK_GBS = 8000.0
# Compulsory HBM bytes per unit (s = bytes per real): what one launch must read and write at least,
# with every reused neighbour spinor and link served on chip.
#   Wilson full D, per site: spinor in 24 + links 4 x 18 + spinor out 24 reals = 120 s
#   (12-real links: 24 + 4 x 12 + 24 = 96 s)
#   Wilson Schur 1 - kappa^2 H_eo H_oe, per even site: two half-lattice hops, each reading the
#   other parity's spinor and every link once (8 x 18 reals per output site): (24 + 144 + 24) +
#   (24 + 144 + 24 x-term + 24) = 408 s
#   staggered full D, per site: 6 + 4 x 18 + 6 = 84 s; Schur block per even site: two half hops
#   (6 + 144 + 6) + (6 + 144 + 6 + 6) = 318 s
#   Wilson force kernel, per site: X 24 + Y 24 + links 72 + force read and write 2 x 72 = 264 s
#   staggered force kernel, n poles, per site: links 72 + force 2 x 72 + n x (6 + 6) = (216 + 12 n) s
COMP_D = {8: 120 * 8, 4: 120 * 4}
COMP_D12 = {8: 96 * 8, 4: 96 * 4}
COMP_SCHUR = {8: 408 * 8, 4: 408 * 4}
COMP_KS = {8: 84 * 8, 4: 84 * 4}
COMP_KS_SCHUR = 318 * 8
COMP_WFORCE = 264 * 8
COMP_KSFORCE12 = (216 + 12 * 12) * 8
# no-reuse count (every neighbour spinor and link counted at each use), reported beside it
NOREUSE_D = {8: 384 * 8, 4: 384 * 4}
FLOPS_D = 1368


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-hmc", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="only run the FP64 D steps plus 5 in-place scalings of one FP64 Wilson field "
                         "(a read+write pass of known size used to calibrate rocprofv3 byte counters)")
    ap.add_argument("--only", choices=["rhmc_fp32", "rhmc_fp64", "rhmc_fom_fp32", "rhmc_fp16link", "gcrodr_nc"],
                    help="run one workload of the extras alone and log it (for a kernel-trace profile): an RHMC "
                         "trajectory variant or the near-critical GCRO-DR pair; no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "path is not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import cudalatticegauge_amd as clg

    V = int(np.prod(DIMS))
    ctx = clg.Context(0)
    log(f"device {ctx.info()}  lattice {DIMS}  V={V}")
    if args.only:
        if args.only.startswith("rhmc_"):
            res = rhmc_trajectory(ctx, clg, np, variants=(args.only[len("rhmc_"):],))
        else:
            res = gcrodr_near_critical(ctx, clg, np)
        log(json.dumps(res))
        ctx.close()
        return

    def make_inputs(prec, gkind):
        g = clg.Field(ctx, gkind, prec, DIMS)
        if prec == clg.F64 and gkind == clg.GAUGE:
            g.gauge_random(0.5, 1)
        else:
            g.copy_from(gauge64)
        src = clg.Field(ctx, clg.WILSON, prec, DIMS, bc=(1, 1, 1, -1))
        if prec == clg.F64:
            src.gaussian(2)
        else:
            src.copy_from(src64)
        out = clg.Field(ctx, clg.WILSON, prec, DIMS, bc=(1, 1, 1, -1))
        return g, src, out

    gauge64 = None
    src64 = None
    gauge64, src64, out64 = make_inputs(clg.F64, clg.GAUGE)
    ctx.sync()

    # ---------------- the timed steps: FP64 full-lattice D
    for _ in range(args.warmup):
        clg.wilson_apply(ctx, out64, src64, gauge64, KAPPA, clg.EFO_F_D)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        clg.wilson_apply(ctx, out64, src64, gauge64, KAPPA, clg.EFO_F_D)
    ctx.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    gbs = COMP_D[8] * V / (ms_per_step * 1e-3) / 1e9
    log(f"FP64 D: {ms_per_step * 1e3:.1f} us/step  {gbs:.0f} GB/s (compulsory bytes)")
    if args.profile:
        nbytes = out64.nbytes()
        for _ in range(5):
            clg.scale(ctx, out64, 1.0)
        ctx.sync()
        log(f"profile mode: 5 scalings of a {nbytes} byte field (read {nbytes} B, write {nbytes} B each)")
        ctx.close()
        return

    # kernel-only time with HIP events on the library stream (same launches)
    k64 = clg.time_wilson(ctx, out64, src64, gauge64, KAPPA, 0, args.steps)
    extra = {"fp64_D_kernel_us": round(k64 * 1e3, 2),
             "fp64_D_GBs": round(COMP_D[8] * V / (k64 * 1e-3) / 1e9, 1),
             "fp64_D_noreuse_GBs": round(NOREUSE_D[8] * V / (k64 * 1e-3) / 1e9, 1),
             "fp64_D_Gflops": round(FLOPS_D * V / (k64 * 1e-3) / 1e9, 1),
             "fp64_D_tuned_launch": clg.autotune_entries()}

    # ---------------- other precisions / layouts
    for name, prec, gk, nbytes in (("fp32_D", clg.F32, clg.GAUGE, COMP_D[4]),
                                   ("fp32_D_recon12", clg.F32, clg.GAUGE12, COMP_D12[4]),
                                   ("fp64_D_recon12", clg.F64, clg.GAUGE12, COMP_D12[8])):
        g, s, o = make_inputs(prec, gk)
        ms = clg.time_wilson(ctx, o, s, g, KAPPA, 0, args.steps)
        extra[name + "_kernel_us"] = round(ms * 1e3, 2)
        extra[name + "_GBs"] = round(nbytes * V / (ms * 1e-3) / 1e9, 1)
        log(f"{name}: {ms * 1e3:.1f} us  {extra[name + '_GBs']:.0f} GB/s")
        del g, s, o
    for name, prec, nb in (("fp64_schur", clg.F64, COMP_SCHUR[8]), ("fp32_schur", clg.F32, COMP_SCHUR[4])):
        g, s, o = make_inputs(prec, clg.GAUGE)
        e_in, e_out = s.parity(clg.EVEN), o.parity(clg.EVEN)
        ms = clg.time_wilson(ctx, e_out, e_in, g, KAPPA, 2, args.steps)
        extra[name + "_us"] = round(ms * 1e3, 2)
        extra[name + "_GBs"] = round(nb * (V // 2) / (ms * 1e-3) / 1e9, 1)
        log(f"{name}: {ms * 1e3:.1f} us  {extra[name + '_GBs']:.0f} GB/s")
        del e_in, e_out, g, s, o

    # ---------------- staggered full D (_kernelDFermionKS) at this lattice and at the RHMC lattice
    for tag, dims, prec in (("ks_fp64_D", DIMS, clg.F64), ("ks_fp32_D", DIMS, clg.F32),
                            ("ks_fp64_D_24x24x24x8", (24, 24, 24, 8), clg.F64)):
        g64 = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
        g64.gauge_random(0.5, 1)
        s64 = clg.Field(ctx, clg.KS, clg.F64, dims, bc=(1, 1, 1, -1))
        s64.gaussian(2)
        if prec == clg.F64:
            g, s = g64, s64
        else:
            g = clg.Field(ctx, clg.GAUGE, prec, dims)
            g.copy_from(g64)
            s = clg.Field(ctx, clg.KS, prec, dims, bc=(1, 1, 1, -1))
            s.copy_from(s64)
        o = clg.Field(ctx, clg.KS, prec, dims, bc=(1, 1, 1, -1))
        ms = clg.time_ks(ctx, o, s, g, 0.02, args.steps)
        vk = int(np.prod(dims))
        extra[tag + "_us"] = round(ms * 1e3, 2)
        extra[tag + "_GBs"] = round(COMP_KS[8 if prec == clg.F64 else 4] * vk / (ms * 1e-3) / 1e9, 1)
        log(f"{tag}: {ms * 1e3:.1f} us  {extra[tag + '_GBs']:.0f} GB/s")
        if prec == clg.F64:
            # the even-parity Schur block (2am)^2 - H_eo H_oe: two half-lattice hops
            se, oe = s.parity(clg.EVEN), o.parity(clg.EVEN)
            ms = clg.time_ks(ctx, oe, se, g, 0.02, args.steps, variant=1)
            stag = tag.replace("_D", "_schur")
            extra[stag + "_us"] = round(ms * 1e3, 2)
            extra[stag + "_GBs"] = round(COMP_KS_SCHUR * (vk // 2) / (ms * 1e-3) / 1e9, 1)
            log(f"{stag}: {ms * 1e3:.1f} us  {extra[stag + '_GBs']:.0f} GB/s")
            del se, oe
        del g, s, o, g64, s64

    # ---------------- HMC fermion force kernels (_kernelDWilsonForceSU3, _kernelDFermionKSForce)
    extra.update(force_kernels(ctx, clg, np, args.steps))

    # ---------------- solves to |r|/|b| < 1e-10 (EO-preconditioned GMRES(12))
    if not args.no_solve:
        x = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=(1, 1, 1, -1))
        results = {}
        for name, inner, inner_acc in (("fp64", clg.F64, 0.0), ("mixed_fp32", clg.F32, 1e-4),
                                       ("mixed_fp16link", clg.F16, 1e-3),
                                       ("mixed_fp16", clg.INNER_F16_VECTORS, 1e-2)):
            p = clg.solver_params(max_dim=12, restart=500, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL,
                                  even_odd=1, inner_prec=inner, inner_accuracy=inner_acc, max_outer=60)
            clg.solve(ctx, p, x, src64, gauge64, KAPPA, clg.EFO_F_InverseD)  # warm (tuning, scratch)
            best = None
            for _ in range(2):
                st, rep = clg.solve(ctx, p, x, src64, gauge64, KAPPA, clg.EFO_F_InverseD)
                if best is None or rep.seconds < best[1].seconds:
                    best = (st, rep)
            st, rep = best
            results[name] = rep
            extra[f"solve_{name}_s"] = round(rep.seconds, 4)
            extra[f"solve_{name}_cycles"] = rep.cycles
            extra[f"solve_{name}_iterations"] = rep.operator_applies
            extra[f"solve_{name}_outer"] = rep.outer
            extra[f"solve_{name}_true_res"] = float(f"{rep.true_residual:.3e}")
            log(f"solve {name}: {rep.seconds * 1e3:.1f} ms, cycles {rep.cycles}, operator applications "
                f"{rep.operator_applies}, outer {rep.outer}, res {rep.true_residual:.2e}, status {st}")
        # plain single-precision solve (FP32 fields, FP32 links, FP64 reductions and Givens) to 1e-6
        g32, s32, x32 = make_inputs(clg.F32, clg.GAUGE)
        p32 = clg.solver_params(max_dim=12, restart=500, accuracy=1e-6, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1)
        clg.solve(ctx, p32, x32, s32, g32, KAPPA, clg.EFO_F_InverseD)  # warm
        best = None
        for _ in range(2):
            st, rep = clg.solve(ctx, p32, x32, s32, g32, KAPPA, clg.EFO_F_InverseD)
            if best is None or rep.seconds < best[1].seconds:
                best = (st, rep)
        st, rep = best
        # the FP32 solution's residual in the FP64 operator
        x64c = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=(1, 1, 1, -1))
        x64c.copy_from(x32)
        r64 = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=(1, 1, 1, -1))
        clg.wilson_apply(ctx, r64, x64c, gauge64, KAPPA, clg.EFO_F_D)
        clg.axpy(ctx, r64, -1.0, src64)
        res64 = float(np.sqrt(clg.norm2(ctx, r64) / clg.norm2(ctx, src64)))
        extra.update({"solve_fp32_s": round(rep.seconds, 4), "solve_fp32_cycles": rep.cycles,
                      "solve_fp32_iterations": rep.operator_applies,
                      "solve_fp32_true_res": float(f"{rep.true_residual:.3e}"),
                      "solve_fp32_res_in_fp64_operator": float(f"{res64:.3e}")})
        log(f"solve fp32 (plain, to 1e-6): {rep.seconds * 1e3:.1f} ms, cycles {rep.cycles}, operator applications "
            f"{rep.operator_applies}, res {rep.true_residual:.2e} (FP64 operator: {res64:.2e}), status {st}")
        del g32, s32, x32, x64c, r64
        extra["solve_speedup_mixed_fp32_vs_fp64"] = round(results["fp64"].seconds / results["mixed_fp32"].seconds, 3)
        extra["solve_speedup_mixed_fp16link_vs_fp64"] = round(
            results["fp64"].seconds / results["mixed_fp16link"].seconds, 3)
        extra["solve_speedup_mixed_fp16_vs_fp64"] = round(results["fp64"].seconds / results["mixed_fp16"].seconds, 3)

        del x
        # harder system: 32^3x32, kappa 0.137, mixed precision against the FP64 solve
        extra.update(mixed_solves_32x32x32x32(ctx, clg, np, oracle_check=not args.no_cpu))
        # GCRO-DR with the recycled subspace carried across a sequence of right-hand sides
        extra.update(gcrodr_sequence(ctx, clg, np, gauge64))

    # ---------------- batched solves: 12 point sources + 4 uneven lattices in one batch
    if not args.no_batch:
        extra.update(batched_solves(ctx, clg, np))

    # ---------------- staggered RHMC trajectory (2+1 flavours)
    if not args.no_hmc:
        extra.update(rhmc_trajectory(ctx, clg, np))
        extra.update(wilson_hmc_trajectory(ctx, clg))

    # ---------------- CPU baseline: the oracle's Wilson D on the host cores (bounded sample)
    cpu = None
    if not args.no_cpu:
        import orc
        nth = orc.num_threads()
        t_one = orc.time_wilson_d(DIMS, 0, 1, nth)
        reps = max(1, min(2000, int(10.0 / max(t_one, 1e-3))))  # about 10 s of the host's threads
        t = orc.time_wilson_d(DIMS, 0, reps, nth)
        cpu_gbs = COMP_D[8] * V * reps / t / 1e9
        cpu = {"value": round(cpu_gbs, 2), "unit": "GB/s", "cores": nth, "kind": "port",
               "sample": f"{reps} FP64 Wilson D applications at 32x32x32x16 on {nth} OpenMP threads "
                         f"({t:.1f} s), oracle restatement of _kernelDFermionWilsonSquareSU3"}
        log(f"CPU oracle D: {cpu_gbs:.1f} GB/s on {nth} threads")

    # HBM bytes per launch from the committed rocprofv3 counter runs of this kernel (FETCH_SIZE and
    # WRITE_SIZE, each calibrated by the counters' reading for an in-place scale of known size): a
    # profile's figure, not a counter read in this run; the launch it was taken on is named beside it
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "wilson_dslash_fp64_32x32x32x16_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("hbm_bytes_per_launch")
        extra["traffic_source"] = f"profiles/{os.path.basename(tpath)}: {tj.get('kernel')}"
        extra["traffic_over_compulsory"] = round(traffic / (COMP_D[8] * V), 3) if traffic else None

    achieved = COMP_D[8] * V / (k64 * 1e-3) / 1e9
    line = {
        "metric": "Dslash GB/s (% HBM peak); 32\u00b3\u00d716 solve time to 1e-10: mixed/batched vs FP64",
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox-seeded gauge exp(i*0.5*H) reunitarised, Gaussian source)",
        "config": {"workload": "wilson_dslash_32x32x32x16_kappa0.1355", "lattice": list(DIMS), "kappa": KAPPA,
                   "gauge_eps": 0.5, "parallelism": "single GPU"},
        "pct_hbm_peak": round(100.0 * gbs / HBM_PEAK_GBS, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


HMC_DIMS = (24, 24, 24, 8)


def gcrodr_sequence(ctx, clg, np, gauge):
    """Three Gaussian sources at 32^3x16, kappa 0.1355, solved in turn to |r|/|b| < 1e-10 by EO GCRO-DR(12, 6)
    (FP64; first solve phase START, then IN: the deflation subspace of each solve seeds the next) and by
    EO GMRES(12) from scratch."""
    out = {}
    b = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=(1, 1, 1, -1))
    x = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=(1, 1, 1, -1))
    # the GCRO-DR stop rule compares |r| with Accuracy (AbsoluteAccuracy) or Accuracy |b|^2: ask for
    # |r| < 1e-10 |b| of these unit-variance sources
    b.gaussian(2)
    bn = float(np.sqrt(clg.norm2(ctx, b)))
    s = clg.Gcrodr(ctx, m_dim=12, k_dim=6, restart=300, accuracy=0.99e-10 * bn, absolute_accuracy=1,
                   deflate=clg.DEFLATE_SVD, even_odd=1)
    pg = clg.solver_params(max_dim=12, restart=500, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1)
    b.gaussian(9)
    s.solve(x, b, gauge, KAPPA, clg.EFO_F_D, phase=clg.PHASE_START)   # warm (tuning, scratch)
    clg.solve(ctx, pg, x, b, gauge, KAPPA, clg.EFO_F_InverseD)
    s.reset()
    t_g, t_m = [], []
    for i, seed in enumerate((2, 3, 4)):
        b.gaussian(seed)
        x.zero()
        st, rep = s.solve(x, b, gauge, KAPPA, clg.EFO_F_D, phase=clg.PHASE_START if i == 0 else clg.PHASE_IN)
        t_g.append(rep.seconds)
        out[f"gcrodr_solve{i}_s"] = round(rep.seconds, 4)
        out[f"gcrodr_solve{i}_cycles"] = rep.cycles
        out[f"gcrodr_solve{i}_true_res"] = float(f"{rep.true_residual:.3e}")
        st2, rep2 = clg.solve(ctx, pg, x, b, gauge, KAPPA, clg.EFO_F_InverseD)
        t_m.append(rep2.seconds)
        log(f"GCRO-DR solve {i}: {rep.seconds * 1e3:.1f} ms, {rep.cycles} cycles, res {rep.true_residual:.2e} "
            f"(GMRES(12): {rep2.seconds * 1e3:.1f} ms, {rep2.cycles} cycles)")
    out["gcrodr_gmres_s"] = [round(t, 4) for t in t_m]
    out["gcrodr_config"] = "32^3x16 EO GCRO-DR(12,6) SVD deflation, FP64, 3 sources in sequence vs EO GMRES(12)"
    s.close()
    out.update(gcrodr_near_critical(ctx, clg, np))
    return out


def gcrodr_near_critical(ctx, clg, np):
    """Where deflated restarting pays: a smooth (eps 0.1) gauge field at 32^3x16 with kappa 0.128 near its
    critical value, full-lattice D x = b to |r|/|b| < 1e-10 (FP64), GMRES(12) against GCRO-DR(30, 15)
    with SVD deflation: the first solve (phase START) and a second source recycling its subspace (IN)."""
    dims, kappa = (32, 32, 32, 16), 0.128
    out = {}
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.gauge_random(0.1, 1)
    b = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=(1, 1, 1, -1))
    x = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=(1, 1, 1, -1))
    b.gaussian(2)
    bn = float(np.sqrt(clg.norm2(ctx, b)))
    pg = clg.solver_params(max_dim=12, restart=5000, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=0)
    s = clg.Gcrodr(ctx, m_dim=30, k_dim=15, restart=3000, accuracy=0.99e-10 * bn, absolute_accuracy=1,
                   deflate=clg.DEFLATE_SVD, even_odd=0)
    s.solve(x, b, g, kappa, clg.EFO_F_D, phase=clg.PHASE_START)  # warm (tuning, scratch)
    s.reset()
    for i, seed in enumerate((2, 3)):
        b.gaussian(seed)
        x.zero()
        st, rep = s.solve(x, b, g, kappa, clg.EFO_F_D, phase=clg.PHASE_START if i == 0 else clg.PHASE_IN)
        x.zero()
        st2, rep2 = clg.solve(ctx, pg, x, b, g, kappa, clg.EFO_F_InverseD)
        tag = f"gcrodr_nc_solve{i}"
        out[f"{tag}_s"] = round(rep.seconds, 4)
        out[f"{tag}_applies"] = rep.operator_applies
        out[f"{tag}_true_res"] = float(f"{rep.true_residual:.3e}")
        out[f"{tag}_gmres_s"] = round(rep2.seconds, 4)
        out[f"{tag}_gmres_applies"] = rep2.operator_applies
        out[f"{tag}_speedup_vs_gmres"] = round(rep2.seconds / rep.seconds, 3)
        log(f"near-critical GCRO-DR(30,15) solve {i}: {rep.seconds * 1e3:.1f} ms, {rep.operator_applies} applies, "
            f"res {rep.true_residual:.2e}; GMRES(12): {rep2.seconds * 1e3:.1f} ms, {rep2.operator_applies} applies")
    out["gcrodr_nc_config"] = ("32^3x16 full-lattice Wilson D, kappa 0.128, eps 0.1 gauge (seed 1), Gaussian sources "
                               "(seeds 2, 3): GCRO-DR(30,15) SVD, START then IN, vs GMRES(12), FP64, |r|/|b| < 1e-10")
    s.close()
    del g, b, x
    return out


def force_kernels(ctx, clg, np, steps):
    """Kernel time of the fermion force kernels on resident inputs (host clock around `steps`
    back-to-back launches, device synchronised): the Wilson force at 32^3x16 and the staggered force
    with 12 poles at 24^3x8, priced on compulsory bytes (COMP_WFORCE, COMP_KSFORCE12)."""
    out = {}
    reps = max(10, steps // 4)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, DIMS)
    g.gauge_random(0.5, 1)
    F = clg.Field(ctx, clg.GAUGE, clg.F64, DIMS)
    F.zero()
    X = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=(1, 1, 1, -1))
    X.gaussian(5)
    Y = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=(1, 1, 1, -1))
    Y.gaussian(6)
    clg.wilson_force_kernel(ctx, F, g, X, Y, KAPPA)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        clg.wilson_force_kernel(ctx, F, g, X, Y, KAPPA)
    ctx.sync()
    us = (time.perf_counter() - t0) * 1e6 / reps
    V = int(np.prod(DIMS))
    out["wilson_force_kernel_us"] = round(us, 1)
    out["wilson_force_kernel_GBs"] = round(COMP_WFORCE * V / (us * 1e-6) / 1e9, 1)
    log(f"Wilson force kernel 32x32x32x16: {us:.1f} us  {out['wilson_force_kernel_GBs']:.0f} GB/s")
    del g, F, X, Y
    g = clg.Field(ctx, clg.GAUGE, clg.F64, HMC_DIMS)
    g.gauge_random(0.4, 1)
    F = clg.Field(ctx, clg.GAUGE, clg.F64, HMC_DIMS)
    F.zero()
    phis, ds = [], []
    for r in range(12):
        a = clg.Field(ctx, clg.KS, clg.F64, HMC_DIMS, bc=(1, 1, 1, -1))
        a.gaussian(10 + r)
        b = clg.Field(ctx, clg.KS, clg.F64, HMC_DIMS, bc=(1, 1, 1, -1))
        b.gaussian(40 + r)
        phis.append(a)
        ds.append(b)
    num = np.linspace(0.1, 1.0, 12)
    clg.ks_force_kernel(ctx, F, g, phis, ds, num)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        clg.ks_force_kernel(ctx, F, g, phis, ds, num)
    ctx.sync()
    us = (time.perf_counter() - t0) * 1e6 / reps
    V = int(np.prod(HMC_DIMS))
    out["ks_force_kernel_12poles_us"] = round(us, 1)
    out["ks_force_kernel_12poles_GBs"] = round(COMP_KSFORCE12 * V / (us * 1e-6) / 1e9, 1)
    log(f"staggered force kernel 24x24x24x8, 12 poles: {us:.1f} us  {out['ks_force_kernel_12poles_GBs']:.0f} GB/s")
    return out


def fit_rational(power, shifts, lo, hi):
    """[c, a_1..a_n, b_1..b_n] with fixed shifts b: least-squares fit of x^power on [lo, hi] in
    relative error (synthetic RHMC coefficients)."""
    import numpy as np
    x = np.logspace(np.log10(lo), np.log10(hi), 600)
    A = np.concatenate([np.ones((x.size, 1)), 1.0 / (x[:, None] + shifts[None, :])], axis=1)
    w = x ** (-power)
    coef, *_ = np.linalg.lstsq(A * w[:, None], np.ones_like(x), rcond=None)
    err = np.abs((A @ coef) * w - 1).max()
    return np.concatenate([coef, shifts]), float(err)


def rhmc_trajectory(ctx, clg, np, variants=("fp32", "fp64", "fom_fp32", "fp16link")):
    shifts = np.logspace(-4, 1, 12)
    flavours = []
    for mass, nf in ((0.01, 2), (0.05, 1)):
        md, err_md = fit_rational(-nf / 4.0, shifts, mass * mass, 20.0)
        mc, err_mc = fit_rational(nf / 8.0, shifts, mass * mass, 20.0)
        flavours.append((mass, md, mc, max(err_md, err_mc)))
    out = {}
    # the shifted BiCGStab (CMultiShiftBiCGStab, the multi-shift solver of the reference's staggered RHMC
    # configurations) for both force precisions and with FP16 links under FP32 vectors, and the multi-shift
    # FOM with FP32 forces for comparison
    for tag, prec, ms_kind in (("fp32", clg.F32, clg.MS_BICGSTAB), ("fp64", clg.F64, clg.MS_BICGSTAB),
                               ("fom_fp32", clg.F32, clg.MS_FOM), ("fp16link", clg.F16, clg.MS_BICGSTAB)):
        if tag not in variants:
            continue
        g = clg.Field(ctx, clg.GAUGE, clg.F64, HMC_DIMS)
        g.gauge_random(0.4, 1)
        acts = [dict(kind=clg.ACTION_GAUGE_PLAQUETTE, beta=5.5)]
        phis = []
        for mass, md, mc, _ in flavours:
            phi = clg.Field(ctx, clg.KS, clg.F64, HMC_DIMS, bc=(1, 1, 1, -1))
            phis.append(phi)
            sp = clg.solver_params(kind=clg.KS, max_dim=40, restart=200, absolute_accuracy=1,
                                   accuracy=1e-8 if prec == clg.F64 else 1e-5)
            acts.append(dict(kind=clg.ACTION_FERMION_KS, kappa_or_mass=mass, phi=phi, md=md, mc=mc, solver=sp,
                             ms_kind=ms_kind, ms_prec=prec, energy_accuracy=1e-9))
        hp = clg.hmc_params(integrator=clg.INT_NESTED_FORCE_GRADIENT, step_length=1.0, steps=8, nested_steps=6,
                            exp_precision=8, metropolis=1)
        rep = clg.hmc_trajectory(ctx, g, hp, acts, 1)
        log(f"RHMC {tag}: dH {rep.h_diff:.6f} (H {rep.h_before:.4f} -> {rep.h_after:.4f}), "
            f"{rep.seconds:.2f} s, {rep.force_evaluations} force passes, {rep.solver_iterations} solver cycles, "
            f"unconverged {rep.unconverged_solves}, accepted {rep.accepted}")
        out[f"rhmc_{tag}_dH"] = float(f"{rep.h_diff:.9g}")
        out[f"rhmc_{tag}_s"] = round(rep.seconds, 3)
        out[f"rhmc_{tag}_solver_cycles"] = rep.solver_iterations
        out[f"rhmc_{tag}_unconverged"] = rep.unconverged_solves
        del g, phis
    if "rhmc_fp32_dH" in out and "rhmc_fp64_dH" in out:
        out["rhmc_dH_abs_diff_fp32_vs_fp64"] = float(f"{abs(out['rhmc_fp32_dH'] - out['rhmc_fp64_dH']):.3g}")
    if "rhmc_fp16link_dH" in out and "rhmc_fp64_dH" in out:
        out["rhmc_dH_abs_diff_fp16link_vs_fp64"] = float(f"{abs(out['rhmc_fp16link_dH'] - out['rhmc_fp64_dH']):.3g}")
    out["rhmc_rational_fit_max_rel_err"] = float(f"{max(f[3] for f in flavours):.2g}")
    out["rhmc_config"] = ("24^3x8 KS 2+1 (Mass 0.01, 0.05), 12 poles, NFG 8x6, beta 5.5, eps 0.4; force solves: "
                          "shifted BiCGStab (fom_fp32: multi-shift FOM, MaxDim 40; fp16link: FP32 vectors, FP16 "
                          "12-real links)")
    return out


def mixed_solves_32x32x32x32(ctx, clg, np, oracle_check=True):
    """EO-GMRES(12) to |r|/|b| < 1e-10 at 32^3x32, kappa 0.137, eps 0.5 gauge, Gaussian source: plain FP64
    against defect correction with FP32 inner vectors and links, FP32 vectors with FP16 links, and FP16
    vectors and links; the mixed solutions are compared with the FP64 one (|x - x64| / |x64|).  With
    oracle_check, the oracle also solves the system on the host cores (full-lattice GMRES(12) to a true
    residual below 1e-12, about 20 s on 16 threads), and every solution is compared with that solution
    and its residual |b - D x| / |b| recomputed with the oracle's FP64 Wilson D."""
    dims, kappa = (32, 32, 32, 32), 0.137
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.gauge_random(0.5, 1)
    b = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=(1, 1, 1, -1))
    b.gaussian(2)
    x = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=(1, 1, 1, -1))
    out, sols, times = {}, {}, {}
    for name, inner, inner_acc in (("fp64", clg.F64, 0.0), ("mixed_fp32", clg.F32, 1e-4),
                                   ("mixed_fp16link", clg.F16, 1e-3), ("mixed_fp16", clg.INNER_F16_VECTORS, 1e-2)):
        p = clg.solver_params(max_dim=12, restart=2000, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1,
                              inner_prec=inner, inner_accuracy=inner_acc, max_outer=80)
        clg.solve(ctx, p, x, b, g, kappa, clg.EFO_F_InverseD)  # warm (tuning, scratch)
        best = None
        for _ in range(2):
            st, rep = clg.solve(ctx, p, x, b, g, kappa, clg.EFO_F_InverseD)
            if best is None or rep.seconds < best[1].seconds:
                best = (st, rep)
        st, rep = best
        sols[name] = x.download()
        times[name] = rep.seconds
        tag = f"c2_{name}"
        out[f"{tag}_s"] = round(rep.seconds, 4)
        out[f"{tag}_cycles"] = rep.cycles
        out[f"{tag}_outer"] = rep.outer
        out[f"{tag}_applies"] = rep.operator_applies
        out[f"{tag}_true_res"] = float(f"{rep.true_residual:.3e}")
        msg = f"32^3x32 kappa {kappa} solve {name}: {rep.seconds * 1e3:.1f} ms, cycles {rep.cycles}, outer {rep.outer}, "
        msg += f"res {rep.true_residual:.2e}, status {st}"
        if name != "fp64":
            err = float(np.linalg.norm(sols[name] - sols["fp64"]) / np.linalg.norm(sols["fp64"]))
            out[f"{tag}_err_vs_fp64"] = float(f"{err:.3e}")
            out[f"{tag}_speedup_vs_fp64"] = round(times["fp64"] / rep.seconds, 3)
            msg += f", |x-x64|/|x64| {err:.2e}, x{times['fp64'] / rep.seconds:.2f} vs FP64"
        log(msg)
    if oracle_check:
        import orc
        U, bh = g.download(), b.download()
        bn = np.linalg.norm(orc.as_complex(bh))
        # the oracle's own FP64 solution (full-lattice GMRES(12), true residual below 1e-12, on the host cores)
        t0 = time.perf_counter()
        xo, rc, cyc, res_o = orc.gmres(0, orc.D, dims, (1, 1, 1, -1), kappa, U, bh, max_dim=12, restart=400,
                                       accuracy=1e-12, stop_rule=1)
        out["c2_oracle_solve_s"] = round(time.perf_counter() - t0, 1)
        out["c2_oracle_true_res"] = float(f"{res_o:.3e}")
        log(f"32^3x32 oracle solve: {cyc} cycles, residual {res_o:.2e}, {out['c2_oracle_solve_s']} s on the host")
        for name, xs in sols.items():
            r = orc.as_complex(orc.wilson_op(dims, (1, 1, 1, -1), kappa, U, xs, orc.D)) - orc.as_complex(bh)
            res = float(np.linalg.norm(r) / bn)
            err = float(np.linalg.norm(xs - xo) / np.linalg.norm(xo))
            out[f"c2_{name}_res_oracle_D"] = float(f"{res:.3e}")
            out[f"c2_{name}_err_vs_oracle"] = float(f"{err:.3e}")
            log(f"32^3x32 solve {name}: |b - D x|/|b| with the oracle's D: {res:.2e}, |x - x_oracle|/|x_oracle| {err:.2e}")
    out["c2_config"] = "32^3x32 Wilson EO-GMRES(12), kappa 0.137, eps 0.5 gauge (seed 1), Gaussian source (seed 2)"
    del g, b, x
    return out


def batched_solves(ctx, clg, np):
    """12 point sources (every spin x colour at the origin) on 24^3x48 plus one Gaussian source each on
    8^4, 12^4, 16^3x32 and 20^3x40, kappa 0.1355, eps 0.5 gauge fields: one batched EO-GMRES(12)
    solve with FP32 (or FP16-storage) inner / FP64 outer to |r|/|b| < 1e-10, against the same 16
    systems solved one at a time with the same precision and in plain FP64."""
    lattices = [((24, 24, 24, 48), 12), ((8, 8, 8, 8), 1), ((12, 12, 12, 12), 1), ((16, 16, 16, 32), 1),
                ((20, 20, 20, 40), 1)]
    gauges, xs, bs, gl = {}, [], [], []
    site_rhs = 0
    for dims, nrhs in lattices:
        g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
        g.gauge_random(0.5, 1)
        gauges[dims] = g
        for r in range(nrhs):
            b = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=(1, 1, 1, -1))
            if nrhs > 1:
                b.point_source((0, 0, 0, 0), r // 3, r % 3)
            else:
                b.gaussian(2)
            bs.append(b)
            xs.append(clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=(1, 1, 1, -1)))
            gl.append(g)
            site_rhs += int(np.prod(dims))
    def timed(fn):
        fn()  # warm: tuning and scratch
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            r = fn()
            ctx.sync()
            dt = time.perf_counter() - t0
            if best is None or dt < best[0]:
                best = (dt, r)
        return best

    def separate(p):
        res = 0.0
        for i in range(len(xs)):
            _, rep1 = clg.solve(ctx, p, xs[i], bs[i], gl[i], KAPPA, clg.EFO_F_InverseD)
            res = max(res, rep1.true_residual)
        return res

    # plain FP64 solves, one at a time: the reference point
    p64 = clg.solver_params(max_dim=12, restart=2000, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1,
                            inner_prec=clg.F64, max_outer=1)
    t_64, res64 = timed(lambda: separate(p64))
    out = {"batch_systems": len(xs), "batch_site_rhs": site_rhs, "batch_separate_fp64_s": round(t_64, 4),
           "batch_separate_fp64_max_true_res": float(f"{res64:.3e}")}
    for tag, inner, acc in (("fp32", clg.F32, 1e-4), ("fp16", clg.INNER_F16_VECTORS, 1e-2)):
        p = clg.solver_params(max_dim=12, restart=500, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1,
                              inner_prec=inner, inner_accuracy=acc, max_outer=60)
        t_batch, (st, reps) = timed(lambda: clg.solve_batch(ctx, p, xs, bs, gl, KAPPA))
        t_sep, sep_res = timed(lambda: separate(p))
        max_res = max(r.true_residual for r in reps)
        log(f"batched solve {tag}: {len(xs)} systems, {site_rhs} site-RHS, {t_batch * 1e3:.1f} ms (separate mixed "
            f"{t_sep * 1e3:.1f} ms x{t_sep / t_batch:.2f}; separate FP64 {t_64 * 1e3:.1f} ms x{t_64 / t_batch:.2f}), "
            f"max |r|/|b| {max_res:.2e}, status {st}")
        out.update({f"batch_{tag}_solve_s": round(t_batch, 4), f"batch_{tag}_separate_s": round(t_sep, 4),
                    f"batch_{tag}_speedup_vs_separate": round(t_sep / t_batch, 3),
                    f"batch_{tag}_throughput_vs_fp64": round(t_64 / t_batch, 3),
                    f"batch_{tag}_max_true_res": float(f"{max_res:.3e}"),
                    f"batch_{tag}_separate_max_true_res": float(f"{sep_res:.3e}"),
                    f"batch_{tag}_site_rhs_per_s": round(site_rhs / t_batch, 1)})
    del xs, bs, gl, gauges
    return out


def wilson_hmc_trajectory(ctx, clg):
    """One Wilson Nf=2 HMC trajectory at 32^3x16 with the reference's published set-up (plaquette
    beta 2.5, kappa 0.1355, nested force gradient 8 outer x 6 nested steps, exponential order 8),
    on the synthetic eps 0.5 gauge field: even-odd GMRES(12) force solves to |r|/|b| < 1e-8, in
    plain FP64 and with FP32 or FP16-storage inner defect correction; energies to 1e-10."""
    out = {}
    for tag, inner, inner_acc in (("fp64", clg.F64, 0.0), ("mixed_fp32", clg.F32, 1e-4),
                                  ("mixed_fp16", clg.INNER_F16_VECTORS, 1e-2)):
        g = clg.Field(ctx, clg.GAUGE, clg.F64, DIMS)
        g.gauge_random(0.5, 1)
        phi = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=(1, 1, 1, -1))
        sp = clg.solver_params(max_dim=12, restart=500, accuracy=1e-8, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1,
                               inner_prec=inner, inner_accuracy=inner_acc, max_outer=60)
        acts = [dict(kind=clg.ACTION_GAUGE_PLAQUETTE, beta=2.5),
                dict(kind=clg.ACTION_FERMION_WILSON_NF2, kappa_or_mass=KAPPA, phi=phi, solver=sp,
                     energy_accuracy=1e-10)]
        hp = clg.hmc_params(integrator=clg.INT_NESTED_FORCE_GRADIENT, step_length=1.0, steps=8, nested_steps=6,
                            exp_precision=8, metropolis=1)
        rep = clg.hmc_trajectory(ctx, g, hp, acts, 1)
        log(f"Wilson HMC {tag}: dH {rep.h_diff:.6f} (H {rep.h_before:.3f} -> {rep.h_after:.3f}), {rep.seconds:.2f} s, "
            f"{rep.solver_iterations} solver cycles, unconverged {rep.unconverged_solves}, accepted {rep.accepted}")
        out[f"wilson_hmc_{tag}_s"] = round(rep.seconds, 3)
        out[f"wilson_hmc_{tag}_dH"] = float(f"{rep.h_diff:.9g}")
        out[f"wilson_hmc_{tag}_solver_cycles"] = rep.solver_iterations
        out[f"wilson_hmc_{tag}_unconverged"] = rep.unconverged_solves
        del g, phi
    out["wilson_hmc_config"] = "32^3x16 Wilson Nf=2 kappa 0.1355, beta 2.5, NFG 8x6, exp order 8, eps 0.5 gauge"
    return out


if __name__ == "__main__":
    main()
