#!/usr/bin/env python3
"""Benchmark: DOF/s to a 1e-10 relative residual on one MI355X.

Workload (default, --config 2 of the numbered workloads below): 2D adaptive Poisson, quadtree levels 4-9 refined
within 0.05 of 8 seeded random points, 2:1 balanced, 19,930 patches of 16x16
cells (5.1M DOF), BiQuadratic ghost filling, 5-point star operator, DFT patch
smoother (1 pre / 1 post / 1 coarse sweep) in a 10-level V-cycle preconditioning
an FP64 BiCGStab to 1e-10.

A "step" is one complete solve from x = 0 to ||r||/||b|| <= 1e-10 with the
V-cycle in FP32 (coarsest level FP64) under FP64 BiCGStab refinement; `value`
is finest-level DOF x steps / wall time of the timed steps, inputs resident in
HBM.  The plain FP64 GPU solve is timed the same way and reported beside it.
The BiCGStab early exit uses x += alpha M p (the reference adds alpha p, which
leaves the true residual above the tolerance whenever that exit fires).

Prints one JSON line.  --gpus N > 1 exits: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def true_residual(R, filler, dom, b, x, r):
    """||b - A x|| / ||b|| in FP64 on the device."""
    filler.residual(b, x, r)
    return dom.two_norm(r) / dom.two_norm(b)


def time_launches(R, fn, n, stream):
    e0, e1 = R.Event(), R.Event()
    fn()  # warm
    stream.sync()
    e0.record(stream)
    for _ in range(n):
        fn()
    e1.record(stream)
    return R.Event.elapsed_ms(e0, e1) / n


def kernel_rooflines(R, P, stream, reps=50):
    """Average launch time (HIP events on the launch stream) of the finest-level kernels
    of the mixed-precision V-cycle, launched standalone through the C ABI with the same
    dtype combination as inside the solve (the BiCGStab hands the V-cycle FP32 copies of p
    and s and keeps M p, M s in FP32), and their algorithmic bytes per launch (every input
    read once, every output written once; ghost values computed from neighbours are cache
    traffic, not algorithmic)."""
    import numpy as np

    dim, ns = P.dim, P.ns
    lv0, lv1 = P.levels[0], P.levels[1]
    d0, d1 = R.Domain(dim, ns, lv0), R.Domain(dim, ns, lv1)
    n0, n1 = len(lv0), len(lv1)
    N = int(np.prod(ns))
    rng = np.random.default_rng(1)
    out = {}
    g = R.GhostFiller(d0, P.filler, R.FACES)
    link = R.LevelLink(d0, d1)
    f64 = R.DeviceArray.from_host(rng.standard_normal(d0.shape(1)))
    f32 = R.DeviceArray.from_host(rng.standard_normal(d0.shape(1)), np.float32)
    u32 = R.DeviceArray.from_host(rng.standard_normal(d0.shape(1)), np.float32)
    o32 = R.DeviceArray(d0.shape(1), np.float32)
    m32 = R.DeviceArray(d0.shape(1), np.float32)
    uc32 = R.DeviceArray.from_host(rng.standard_normal(d1.shape(1)), np.float32)
    fc32 = R.DeviceArray(d1.shape(1), np.float32)

    def rec(name, fn, byt, flops=None):
        t = time_launches(R, fn, reps, stream)
        out[name] = {"ms": t, "bytes": byt, "GBs": byt / t / 1e6}
        if flops:
            out[name].update({"flops": flops, "TFs": flops / t / 1e9})

    if dim == 2 and P.smoother == 0:
        plan = R.DftPlan(g, op_kind=P.op_kind)
        fl = n0 * (8 * ns[0] ** 3 + N)  # four n^3 products of 2 n^3 flops, n^2 divides
        # pre-smoothing from u = 0: f read, u written (FP32)
        rec("dft_apply_f32", lambda: plan.apply_mixed(R.F32, f32, o32, stream=stream.h), n0 * N * (4 + 4), fl)
        # Cycle::restrict: u, f read; coarse f written (FP32)
        rec("residual_restrict_f32", lambda: link.residual_restrict(g, f32, u32, fc32, stream=stream.h),
            n0 * N * (4 + 4) + n1 * N * 4)
        # post-smoothing with the interpolation: f, u, coarse u read; M f written (FP32)
        rec("dft_smooth_interp_f32",
            lambda: plan.smooth_interp(link, f32, u32, uc32, m32, stream=stream.h), n0 * N * (4 + 4 + 4) + n1 * N * 4,
            fl)
    # copy of a vector of known size (calibrates the HBM byte counters)
    c64 = R.DeviceArray(d0.shape(1), np.float64)
    rec("vec_copy_f64", lambda: R.check(R.lib().te_vec_copy(d0.h, R.F64, c64.ptr, f64.ptr, 1, stream.h)),
        2 * n0 * N * 8)
    return out


def cpu_baseline(P, gpu_iters, sample_iters=5):
    """The oracle (FP64 C restatement of the reference algorithm, one core) on the same
    problem: a bounded sample of `sample_iters` outer BiCGStab iterations, extrapolated to
    the GPU FP64 solve's iteration count."""
    from oracle import oracle as O

    od = [O.Domain(P.dim, P.ns, lv) for lv in P.levels]
    kw = {}
    if P.op_kind == 1:
        coeffs = []
        for lv, d in zip(P.levels, od):
            c = P.coeffs[len(coeffs)].copy()
            O.fill_ghost(d, c, kind=P.filler)
            coeffs.append(c)
        kw = {"op_kind": 1, "coeffs": coeffs}
    L = O.Levels(od, filler_kind=P.filler, smoother_kind=P.smoother, **kw)
    t0 = time.perf_counter()
    _, its, _ = L.bicgstab(P.rhs.copy(), tol=1e-30, precond=2, pre=P.pre, post=P.post, coarse=P.coarse,
                           max_it=sample_iters)
    dt = time.perf_counter() - t0
    per_it = dt / max(its, 1)
    est = per_it * max(gpu_iters, 1)
    return {
        "value": P.dof / est,
        "unit": "DOF/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle FP64 BiCGStab+V-cycle on the full {P.dof}-DOF problem, {its} outer iterations "
                   f"in {dt:.2f} s ({per_it:.3f} s/iteration), extrapolated to the {gpu_iters} iterations "
                   f"the FP64 GPU solve needs"),
    }


def run_many_problems(args):
    """Config 5: independent problems solved concurrently (one stream per problem); value =
    aggregate finest-level DOF x rounds / wall time, every problem to 1e-10, mixed precision,
    with the FP64 aggregate beside it."""
    import numpy as np

    from thunderegg_amd import configs
    from thunderegg_amd import runtime as R
    from thunderegg_amd import solver

    R.lib()
    t0 = time.perf_counter()
    probs = [configs.small_adaptive2d_problem(k) for k in range(args.problems)]
    t_mesh = time.perf_counter() - t0
    dof = sum(P.dof for P in probs)
    log(f"many problems: {len(probs)} problems, {dof} DOF, patch sizes "
        f"{sorted({P.ns[0] for P in probs})}; meshes {t_mesh:.1f}s")

    def run(prec):
        gs = [solver.GMG(P.dim, P.ns, P.levels, filler_kind=P.filler, smoother=P.smoother, precision=prec,
                         use_graph=True) for P in probs]
        bs = [R.DeviceArray.from_host(P.rhs) for P in probs]
        xs = [R.DeviceArray(g.shape, np.float64) for g in gs]

        def round_():
            for x in xs:
                x.zero()
            R.sync()
            return solver.solve_many(gs, bs, xs, tol=1e-10, max_it=500, precond=2, threads=args.threads)

        for _ in range(args.warmup):
            round_()
        R.sync()
        t = time.perf_counter()
        its = None
        for _ in range(args.steps):
            its = round_()
        R.sync()
        return time.perf_counter() - t, its

    T32, its32 = run(R.F32)
    T64, its64 = run(R.F64)
    out = {
        "metric": "DOF/s to 1e-10 rel residual, FP32-precond+FP64 refine vs FP64 GPU; HBM GB/s",
        "value": dof * args.steps / T32,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": T32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 BiCGStab",
        "data": "synthetic (seeded meshes and right-hand sides)",
        "config": {"workload": f"{len(probs)} independent 2D adaptive problems , solved concurrently",
                   "problems": len(probs), "dof": dof, "threads": args.threads,
                   "patch_sizes": sorted({P.ns[0] for P in probs})},
        "fp64_dof_per_s": dof * args.steps / T64,
        "mixed_over_fp64": T64 / T32,
        "iterations_mixed": its32,
        "iterations_fp64": its64,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2,
                    help="workload: 1 uniform 2D (64 x 32^2, BiCGStab smoother), 2 adaptive 2D (19,930 x 16^2, "
                         "DFT smoother), 3 uniform 3D (32,768 x 16^3), 4 VarPoisson 2D, 5 independent small problems")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison solve")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the standalone kernel timings (for counter runs under rocprofv3)")
    ap.add_argument("--problems", type=int, default=64, help="many problems: number of problems (512 in full)")
    ap.add_argument("--threads", type=int, default=8, help="many problems: concurrent solves")
    ap.add_argument("--timer", metavar="FILE",
                    help="after the timed solves, one more mixed-precision solve with a Timer attached: its "
                         "timing tree (Iteration, per-level Total Patch Smooth) saved as JSON to FILE and "
                         "the report printed on stderr")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)
    if args.config == 5:
        run_many_problems(args)
        return

    import numpy as np

    from thunderegg_amd import configs
    from thunderegg_amd import runtime as R
    from thunderegg_amd import solver

    R.lib()  # fails loudly when the HIP library is missing
    t0 = time.perf_counter()
    P = configs.CONFIGS[args.config]()
    t_mesh = time.perf_counter() - t0
    log(f"config {args.config}: {P.name}; levels {[len(l) for l in P.levels]}; DOF {P.dof}; mesh {t_mesh:.1f}s")
    if args.kernels_only:
        kr = kernel_rooflines(R, P, R.Stream(), reps=args.steps)
        print(json.dumps({"kernels": kr}), flush=True)
        return

    def make(precision):
        return solver.GMG(P.dim, P.ns, P.levels, filler_kind=P.filler, smoother=P.smoother, op_kind=P.op_kind,
                          coeffs=P.coeffs, pre_sweeps=P.pre, post_sweeps=P.post, coarse_sweeps=P.coarse,
                          inner_max_it=P.inner_max_it, inner_tol=P.inner_tol, precision=precision, use_graph=True)

    t0 = time.perf_counter()
    g_mixed = make(R.F32)
    g_fp64 = None if args.no_fp64 else make(R.F64)
    t_setup = time.perf_counter() - t0
    dom = g_mixed.domains[0]
    b = R.DeviceArray.from_host(P.rhs)
    xm = R.DeviceArray(dom.shape(1), np.float64)
    x64 = R.DeviceArray(dom.shape(1), np.float64)
    r = R.DeviceArray(dom.shape(1), np.float64)
    filler = R.GhostFiller(dom, P.filler, R.FACES)

    def run(g, x, k):
        its = []
        R.sync()
        ts = []
        t_start = time.perf_counter()
        for _ in range(k):
            t1 = time.perf_counter()
            x.zero()
            it, _ = g.solve_device(b, x, tol=1e-10, max_it=500, precond=2)
            R.sync()
            ts.append(time.perf_counter() - t1)
            its.append(it)
        R.sync()
        return time.perf_counter() - t_start, its, ts

    # warmup (includes hipGraph capture of both BiCGStab phases)
    run(g_mixed, xm, args.warmup)
    if g_fp64:
        run(g_fp64, x64, args.warmup)
    log("warmup done")
    T_mixed, its_mixed, ts_mixed = run(g_mixed, xm, args.steps)
    log(f"mixed: {args.steps} solves in {T_mixed:.3f}s, iterations {its_mixed}")
    res = {"mixed_true_residual": true_residual(R, filler, dom, b, xm, r)}
    if g_fp64:
        T_64, its_64, ts_64 = run(g_fp64, x64, args.steps)
        log(f"fp64: {args.steps} solves in {T_64:.3f}s, iterations {its_64}")
        res["fp64_true_residual"] = true_residual(R, filler, dom, b, x64, r)
        a, c = xm.get(), x64.get()
        from thunderegg_amd import mesh

        ia, ic = mesh.interior(a, P.dim), mesh.interior(c, P.dim)
        res["mixed_vs_fp64_rel_l2"] = float(np.sqrt(((ia - ic) ** 2).sum() / (ic ** 2).sum()))
    if args.timer:
        # outside the timed region: the timings come from events on the solver's stream
        timer = R.Timer()
        for k, d in enumerate(g_mixed.domains):
            timer.add_domain(k, d)
        g_mixed.set_timer(timer, iterations=True, domain_ids=list(range(len(g_mixed.domains))))
        xm.zero()
        g_mixed.solve_device(b, xm, tol=1e-10, max_it=500, precond=2)
        g_mixed.set_timer(None)
        timer.save_to_file(args.timer)
        log(timer.report())
        tj = timer.to_json()
        for it in tj["timings"]:
            log(f"{it['name']}: {it['num_calls']} calls, {it['sum'] * 1e3:.3f} ms")
            for t in it.get("timings", []):
                k = t["domain_id"]
                log(f"  level {k} ({len(P.levels[k])} patches) {t['name']}: {t['num_calls']} calls, "
                    f"{t['sum'] * 1e3:.3f} ms, {t['sum'] / t['num_calls'] * 1e6:.1f} us each")
    ms = T_mixed / args.steps * 1e3
    value = P.dof * args.steps / T_mixed
    out = {
        "metric": "DOF/s to 1e-10 rel residual, FP32-precond+FP64 refine vs FP64 GPU; HBM GB/s",
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 BiCGStab",
        "data": "synthetic (seeded mesh and right-hand side)",
        "config": {
            "workload": P.name,
            "patches_finest": len(P.levels[0]),
            "levels": [len(l) for l in P.levels],
            "dof": P.dof,
            "cells_per_patch": list(P.ns),
            "tolerance": 1e-10,
            "smoothing": [P.pre, P.post, P.coarse],
        },
        "iterations_mixed": its_mixed,
        "solve_ms_median_mixed": statistics.median(ts_mixed) * 1e3,
        "setup_seconds": {"mesh": t_mesh, "plans": t_setup},
    }
    out.update(res)
    if g_fp64:
        out["fp64_dof_per_s"] = P.dof * args.steps / T_64
        out["iterations_fp64"] = its_64
        out["solve_ms_median_fp64"] = statistics.median(ts_64) * 1e3
        out["mixed_over_fp64"] = out["value"] / out["fp64_dof_per_s"]
    if not args.no_roofline:
        st = R.Stream()
        kr = kernel_rooflines(R, P, st)
        out["kernels"] = kr
        # the dominant kernel: the longest of the finest-level V-cycle launches
        cands = [k for k in kr if k != "vec_copy_f64"]
        dom_k = max(cands, key=lambda k: kr[k]["ms"])
        k = kr[dom_k]
        traffic = None
        tf = os.path.join(ROOT, "profiles", f"traffic_config{args.config}.json")
        if os.path.exists(tf):
            with open(tf) as fh:
                traffic = json.load(fh).get(dom_k, {}).get("bytes_per_launch")
        out["roofline"] = {"bound": "hbm", "achieved": k["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": k["GBs"] / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom_k,
                           "bytes_per_launch": k["bytes"], "ms_per_launch": k["ms"]}
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(P, out.get("iterations_fp64", its_mixed)[0])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
