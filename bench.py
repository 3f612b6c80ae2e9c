#!/usr/bin/env python3
"""Benchmark: DOF/s to a 1e-10 relative residual on one MI355X.

Workload (default): 3D Poisson on a uniform octree at level 5, 32,768 patches of 16^3
cells (134,217,728 DOF), TriLinear ghost filling, 7-point star operator, DFT patch
smoother with 2 pre / 2 post sweeps and 1 coarsest sweep in a 5-level V-cycle (32,768 ->
8 patches) preconditioning an FP64 BiCGStab to 1e-10; right-hand side
-12 pi^2 prod sin(2 pi x_i) plus N(0, 1e-3) noise.  The largest single-GPU solve of the
benchmark workloads.  `--config N` selects another one (1 uniform 2D with the BiCGStab
patch smoother, 2 adaptive 2D, 4 VarPoisson 2D, 5 independent small problems).

A "step" is one complete solve from x = 0 to ||r||/||b|| <= 1e-10 with the V-cycle in
FP32 (coarsest level FP64) under FP64 BiCGStab refinement; `value` is finest-level DOF x
steps / wall time of the timed steps, inputs resident in HBM.  The plain FP64 GPU solve is
timed the same way and reported beside it (`fp64_dof_per_s`, `mixed_over_fp64`), with the
true FP64 residuals of both.  The BiCGStab early exit uses x += alpha M p (the reference
adds alpha p, which leaves the true residual above the tolerance whenever that exit fires).
The adaptive 2D workload (19,930 patches of 16^2) is measured the same way and reported
under `secondary`.

Prints one JSON line.  --gpus N > 1 exits: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFS = {"f32": 157.3, "f64": 78.6}
METRIC = "DOF/s to 1e-10 rel residual, FP32-precond+FP64 refine vs FP64 GPU; HBM GB/s"


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


def kernel_rooflines_2d(R, P, stream, reps=50):
    """Average launch time (HIP events on the launch stream) of the finest-level kernels of
    the 2D mixed-precision V-cycle, launched standalone through the C ABI with the dtype
    combination of the solve, and their algorithmic bytes per launch (every input read
    once, every output written once)."""
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

    if P.smoother == 0:
        plan = R.DftPlan(g, op_kind=P.op_kind)
        fl = n0 * (8 * ns[0] ** 3 + N)  # four n^3 products of 2 n^3 flops, n^2 divides
        rec("dft_apply_f32", lambda: plan.apply_mixed(R.F32, f32, o32, stream=stream.h), n0 * N * (4 + 4), fl)
        rec("residual_restrict_f32", lambda: link.residual_restrict(g, f32, u32, fc32, stream=stream.h),
            n0 * N * (4 + 4) + n1 * N * 4)
        rec("dft_smooth_interp_f32",
            lambda: plan.smooth_interp(link, f32, u32, uc32, m32, stream=stream.h), n0 * N * (4 + 4 + 4) + n1 * N * 4,
            fl)
    c64 = R.DeviceArray(d0.shape(1), np.float64)
    rec("vec_copy_f64", lambda: R.check(R.lib().te_vec_copy(d0.h, R.F64, c64.ptr, f64.ptr, 1, stream.h)),
        2 * n0 * N * 8)
    # the standalone ghost fill (the path's other configurations use it; the fused cycle
    # computes ghosts inside its launches)
    u64 = R.DeviceArray.from_host(rng.standard_normal(d0.shape(1)))
    rec("ghost_fill_f64", lambda: g.fill_ghost(u64, stream=stream.h), ghost_fill_bytes(lv0, ns, P.filler, 8))
    out["copy_f64"] = flat_copy(R, n0 * N, reps, stream)
    return out


def ghost_fill_bytes(level, ns, kind, s):
    """Algorithmic bytes of one 2D side-ghost fill: per side with a neighbour, its n ghost
    cells written once and every cell their formulas read, once: the neighbour's facing row
    (same level), the patch's own boundary row (and the one behind it, BiQuadratic) with the
    n / 2 coarse cells (coarser neighbour), or the own row and the two finer neighbours'
    facing rows (and the rows behind them, BiQuadratic)."""
    import numpy as np

    types = np.asarray(level["nbr_type"])[:, 4:8]
    total = 0
    for side in range(4):
        n = ns[1] if side < 2 else ns[0]
        t = types[:, side]
        deep = 2 if kind == 1 else 1
        total += int((t == 0).sum()) * 2 * n
        total += int((t == 1).sum()) * (n + deep * n + n // 2)
        total += int((t == 2).sum()) * (n + n + deep * 2 * n)
    return total * s


def kernel_rooflines_3d(g, P, its, reps=20):
    """The finest-level kernels of the compact 3D cycle and the outer solver, each launched
    standalone on the mixed-precision solver's own buffers and timed with HIP events on its
    stream (te_gmg_time_kernel), with algorithmic bytes and flops per launch.  Per 16^3
    patch: N = 4096 cells, B = 1352 cells of the boundary layer (the cells a neighbour's
    ghost formula or the patch's own interface terms read), Bc = 296 coarse cells under the
    boundary layer; FP32 levels (s = 4), FP64 outer vectors.  Bytes count every value the
    result depends on once:
      sweep_zero      f, out                        s (2N)
      sweep           f, u's boundary layer, out    s (2N + B)
      sweep_interp    + the coarse boundary cells   s (2N + B + Bc)
      resid_restrict  f, u, coarse f                s (2N + N/8)
      operator        M s (FP32), s (FP64)          4N + 8N (+ 8N: A M s stored, TE_NO_XRFLY=1)
      operator_p      M p (FP32), rhat (FP32)       4N + 4N (TE_NO_RHAT32=1: rhat FP64, + 4N;
                                                    TE_NO_XRFLY=1: that, and A M p stored, + 8N)
      update_xr       r written; M s (FP32); s (FP64), rhat (FP32): 8 + 4 + 8 + 4 per cell, x
                      deferred to x_combine (TE_NO_XDEFER=1: x read + written and M p read,
                      8 2 + 8 + 4 2 + 8 + 4; rhat FP64: + 4; TE_NO_XRFLY=1: that, and r read,
                      A M p and A M s instead of s: + 8 + 8)
      x_combine       x read + written, M p and M s of K iterations (FP32): 8 2 + 4 2 K, once
                      per K iterations (K = 8 slots)
      update_s        r, M p (FP32) -> s, s32        8 + 4 + 8 + 4 (M p -> A M p, TE_NO_XRFLY=1: + 4)
      update_p        p32, M p (FP32), r -> p32      4 + 4 + 8 + 4 (TE_NO_P32ONLY=1: p FP64 read and
                      written as well, 8 + 4 + 8 + 8 + 4; TE_NO_XRFLY=1: + 4)
    Flops of the patch solve: six 16 x 16 x 16 transform products per axis pass, 12 n^4."""
    np_ = len(P.levels[0])
    N, B, Bc = 4096, 16 ** 3 - 14 ** 3, 8 ** 3 - 6 ** 3
    s = 4
    fl = np_ * 12 * 16 ** 4
    pre, post = P.pre, P.post
    xrfly = os.environ.get("TE_NO_XRFLY") != "1"  # the updates form A M p and A M s themselves
    rh = 8 if (not xrfly or os.environ.get("TE_NO_RHAT32") == "1") else 4  # bytes per rhat value
    xdefer = rh == 4 and os.environ.get("TE_NO_XDEFER") != "1"  # x updated by x_combine
    ks = max(1, min(8, int(os.environ.get("TE_XDEFER_SLOTS", "8"))))
    p32only = rh == 4 and os.environ.get("TE_NO_P32ONLY") != "1"  # p only as its FP32 copy
    specs = {
        # name: (bytes, flops, launches per solve)
        "sweep_zero": (np_ * s * 2 * N, fl, 2 * its),
        "sweep": (np_ * s * (2 * N + B), fl, 2 * its * max(pre - 1 + post - 1, 0)),
        "sweep_interp": (np_ * s * (2 * N + B + Bc), fl, 2 * its),
        "resid_restrict": (np_ * s * (2 * N + N // 8), None, 2 * its),
        "operator": (np_ * N * (4 + 8 + (0 if xrfly else 8)), None, its),
        "operator_p": (np_ * N * (4 + rh + (0 if xrfly else 8)), None, its),
        "update_xr": (np_ * N * ((8 + 4 + 8 + rh) if xdefer else (8 * 2 + 8 + 4 * 2 + 8 + rh + (0 if xrfly else 16))),
                      None, its),
        "update_s": (np_ * N * (8 + 4 + 8 + 4 + (0 if xrfly else 4)), None, its),
        "update_p": (np_ * N * ((4 + 4 + 8 + 4) if p32only else (8 + 4 + 8 + 8 + 4 + (0 if xrfly else 4))), None, its),
    }
    if xdefer:
        specs["x_combine"] = (np_ * N * (8 * 2 + 4 * 2 * ks), None, -(-its // ks))
    out = {}
    for name, (byt, flops, per_solve) in specs.items():
        ms = g.time_kernel(name, reps)
        d = {"ms": ms, "bytes": byt, "GBs": byt / ms / 1e6, "launches_per_solve": per_solve,
             "ms_per_solve": ms * per_solve}
        if flops:
            d.update({"flops": flops, "TFs": flops / ms / 1e9})
        out[name] = d
    # a plain copy of the interior of one FP64 vector of the finest level: the reference
    # point that scales the HBM counters (and the bandwidth a plain stream reaches)
    import numpy as np

    from thunderegg_amd import runtime as R

    d0 = g.domains[0]
    stream = R.Stream()
    a = R.DeviceArray(d0.shape(1), np.float64)
    c = R.DeviceArray(d0.shape(1), np.float64)
    a.zero()
    byt = 2 * np_ * N * 8
    ms = time_launches(R, lambda: R.check(R.lib().te_vec_copy(d0.h, R.F64, c.ptr, a.ptr, 1, stream.h)), reps, stream)
    out["vec_copy_f64"] = {"ms": ms, "bytes": byt, "GBs": byt / ms / 1e6, "launches_per_solve": 0,
                           "ms_per_solve": 0.0}
    out["copy_f64"] = flat_copy(R, np_ * N, reps, stream)
    out["copy_f64"].update({"launches_per_solve": 0, "ms_per_solve": 0.0})
    return out


def flat_copy(R, count, reps, stream):
    """A plain device-to-device copy of `count` contiguous doubles (hipMemcpyAsync on the
    stream): the copy of known size whose HBM counters calibrate the other kernels'."""
    import numpy as np

    a = R.DeviceArray((count,), np.float64)
    c = R.DeviceArray((count,), np.float64)
    a.zero(stream.h)
    byt = 2 * count * 8
    ms = time_launches(R, lambda: R.check(R.lib().te_memcpy_d2d(c.ptr, a.ptr, count * 8, stream.h)), reps, stream)
    stream.sync()
    return {"ms": ms, "bytes": byt, "GBs": byt / ms / 1e6}


def tune_report(R, shapes):
    """The autotuner's table as the library loaded it, and the launch variant each of the
    given (dim, n) patch shapes runs, per precision (variant 0 = the hand default)."""
    out = {"table": os.path.relpath(R.tune_table_path(), ROOT) if R.tune_table_path() else None, "variants": {}}
    for dim, n in sorted(set(shapes)):
        fam = R.tune_family(dim, n)
        out["variants"][f"{dim}d_n{n}"] = {
            "family": ["dft_generic", "dft_mfma_2d", "c3_sweep"][fam],
            "f32": R.tune_get(fam, dim, n, R.F32),
            "f64": R.tune_get(fam, dim, n, R.F64),
        }
    return out


def cpu_info():
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(P, gpu_iters, budget_s=30.0):
    """The oracle (the FP64 C restatement of the reference algorithm; its per-patch loops run
    under OpenMP, the ghost fills and reductions serially) on the same problem, on this
    host's cores: outer BiCGStab iterations until about `budget_s` of CPU work (at least
    one), extrapolated to the GPU FP64 solve's iteration count."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
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
    _, its, _ = L.bicgstab(P.rhs.copy(), tol=1e-30, precond=2, pre=P.pre, post=P.post, coarse=P.coarse, max_it=1)
    dt = time.perf_counter() - t0
    k = max(its, 1)
    if dt * 2 < budget_s:
        more = max(1, min(int(budget_s / max(dt, 1e-3)) - 1, 5))
        t0 = time.perf_counter()
        _, its2, _ = L.bicgstab(P.rhs.copy(), tol=1e-30, precond=2, pre=P.pre, post=P.post, coarse=P.coarse,
                                max_it=more)
        dt2 = time.perf_counter() - t0
        if its2 > 0:
            dt, k = dt2, its2
    per_it = dt / k
    est = per_it * max(gpu_iters, 1)
    model, ncpu = cpu_info()
    return {
        "value": P.dof / est,
        "unit": "DOF/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": model,
        "host_logical_cpus": ncpu,
        "sample": (f"oracle FP64 BiCGStab + V-cycle on the full {P.dof}-DOF problem with {threads} OpenMP threads "
                   f"on {model}: {k} outer iteration(s) in {dt:.2f} s ({per_it:.2f} s/iteration, setup excluded), "
                   f"extrapolated to the {gpu_iters} iterations of the FP64 GPU solve"),
    }


def run_many_problems(args, nprob=None, emit=True):
    """The independent small problems (512 in full): patch counts log-uniform in [16, 4096],
    patch sizes 8, 16 and 32.  Batched (te_batch: one launch per cycle and BiCGStab step for
    all problems of a patch size, one graph per outer iteration) in mixed precision and FP64,
    beside the same problems solved one te_gmg per problem on concurrent streams (the
    per-problem GPU solve).  value = aggregate finest-level DOF x rounds / wall time of the
    batched mixed-precision rounds, every problem to 1e-10."""
    import numpy as np

    from thunderegg_amd import configs
    from thunderegg_amd import runtime as R
    from thunderegg_amd import solver

    R.lib()
    nprob = nprob or args.problems
    t0 = time.perf_counter()
    probs = [configs.small_adaptive2d_problem(k) for k in range(nprob)]
    t_mesh = time.perf_counter() - t0
    dof = sum(P.dof for P in probs)
    log(f"many problems: {len(probs)} problems, {dof} DOF, patch sizes "
        f"{sorted({P.ns[0] for P in probs})}; meshes {t_mesh:.1f}s")
    bs = [R.DeviceArray.from_host(P.rhs) for P in probs]

    def timed(round_):
        for _ in range(args.warmup):
            round_()
        R.sync()
        t = time.perf_counter()
        its = None
        for _ in range(args.steps):
            its = round_()
        R.sync()
        return time.perf_counter() - t, its

    def batched(prec):
        t = time.perf_counter()
        B = solver.Batch([(P.ns, P.levels) for P in probs], filler_kind=probs[0].filler, precision=prec)
        setup = time.perf_counter() - t
        xs = [R.DeviceArray(P.rhs.shape, np.float64) for P in probs]

        def round_():
            for x in xs:
                x.zero()
            R.sync()
            its, _ = B.solve_device(bs, xs, tol=1e-10, max_it=500, precond=2)
            return its

        T, its = timed(round_)
        # every problem's true residual after the last round
        worst = 0.0
        for P, b, x in zip(probs, bs, xs):
            dom = R.Domain(2, P.ns, P.levels[0])
            filler = R.GhostFiller(dom, P.filler, R.FACES)
            r = R.DeviceArray(dom.shape(1), np.float64)
            filler.residual(b, x, r)
            worst = max(worst, dom.two_norm(r) / dom.two_norm(b))
        return T, [int(i) for i in its], setup, worst

    def concurrent(prec):
        gs = [solver.GMG(P.dim, P.ns, P.levels, filler_kind=P.filler, smoother=P.smoother, precision=prec,
                         use_graph=True) for P in probs]
        xs = [R.DeviceArray(g.shape, np.float64) for g in gs]

        def round_():
            for x in xs:
                x.zero()
            R.sync()
            return solver.solve_many(gs, bs, xs, tol=1e-10, max_it=500, precond=2, threads=args.threads)

        T, its = timed(round_)
        del gs, xs
        return T, [int(i) for i in its]

    Tb32, itb32, setup32, worst32 = batched(R.F32)
    log(f"many problems: batched mixed {args.steps} rounds in {Tb32:.3f}s")
    Tb64, itb64, _, worst64 = batched(R.F64)
    log(f"many problems: batched fp64 {args.steps} rounds in {Tb64:.3f}s")
    Tc64 = Tc32 = itc64 = None
    if not args.batched_only:
        Tc64, itc64 = concurrent(R.F64)
        log(f"many problems: per-problem fp64 (concurrent streams) {args.steps} rounds in {Tc64:.3f}s")
        Tc32, _ = concurrent(R.F32)
        log(f"many problems: per-problem mixed (concurrent streams) {args.steps} rounds in {Tc32:.3f}s")
    out = {
        "metric": METRIC,
        "value": dof * args.steps / Tb32,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": Tb32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle / f64 BiCGStab",
        "data": "synthetic (seeded meshes and right-hand sides)",
        "config": {"workload": f"{len(probs)} independent 2D adaptive problems, batched",
                   "problems": len(probs), "dof": dof,
                   "patch_sizes": sorted({P.ns[0] for P in probs}),
                   "patches_min_max": [min(len(P.levels[0]) for P in probs), max(len(P.levels[0]) for P in probs)]},
        "fp64_dof_per_s": dof * args.steps / Tb64,
        "mixed_over_fp64": Tb64 / Tb32,
        "per_problem_fp64_dof_per_s": dof * args.steps / Tc64 if Tc64 else None,
        "per_problem_mixed_dof_per_s": dof * args.steps / Tc32 if Tc32 else None,
        "batched_over_per_problem_fp64": Tc64 / Tb64 if Tc64 else None,
        "batched_over_per_problem_mixed": Tc32 / Tb32 if Tc32 else None,
        "concurrent_threads": args.threads,
        "batch_setup_seconds": setup32,
        "fp64_iterations_equal_per_problem": itb64 == itc64 if itc64 else None,
        "tune": tune_report(R, [(2, P.ns[0]) for P in probs]),
        "worst_true_residual_mixed": worst32,
        "worst_true_residual_fp64": worst64,
        "iterations_mixed": itb32,
        "iterations_fp64": itb64,
        "iterations_per_problem_fp64": itc64,
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def measure(P, args, R, solver, with_fp64=True):
    """Timed mixed-precision (and FP64) solves of problem P; returns the numbers and the solvers."""
    import numpy as np

    from thunderegg_amd import mesh

    def make(precision):
        return solver.GMG(P.dim, P.ns, P.levels, filler_kind=P.filler, smoother=P.smoother, op_kind=P.op_kind,
                          coeffs=P.coeffs, pre_sweeps=P.pre, post_sweeps=P.post, coarse_sweeps=P.coarse,
                          inner_max_it=P.inner_max_it, inner_tol=P.inner_tol, precision=precision, use_graph=True)

    t0 = time.perf_counter()
    g_mixed = make(R.F32)
    g_fp64 = make(R.F64) if with_fp64 else None
    t_setup = time.perf_counter() - t0
    dom = g_mixed.domains[0]
    b = R.DeviceArray.from_host(P.rhs)
    xm = R.DeviceArray(dom.shape(1), np.float64)
    x64 = R.DeviceArray(dom.shape(1), np.float64)
    r = R.DeviceArray(dom.shape(1), np.float64)
    filler = R.GhostFiller(dom, P.filler, R.FACES)

    def run(g, x, k):
        its, ts = [], []
        R.sync()
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

    # warmup (includes the hipGraph capture of the iteration)
    run(g_mixed, xm, args.warmup)
    if g_fp64:
        run(g_fp64, x64, args.warmup)
    T_mixed, its_mixed, ts_mixed = run(g_mixed, xm, args.steps)
    log(f"{P.name}: mixed {args.steps} solves in {T_mixed:.3f}s, iterations {its_mixed}")
    res = {
        "value": P.dof * args.steps / T_mixed,
        "ms_per_step": T_mixed / args.steps * 1e3,
        "iterations_mixed": its_mixed,
        "solve_ms_median_mixed": statistics.median(ts_mixed) * 1e3,
        "setup_seconds": {"plans": t_setup},
        "mixed_true_residual": true_residual(R, filler, dom, b, xm, r),
    }
    if g_fp64:
        T_64, its_64, ts_64 = run(g_fp64, x64, args.steps)
        log(f"{P.name}: fp64 {args.steps} solves in {T_64:.3f}s, iterations {its_64}")
        res["fp64_true_residual"] = true_residual(R, filler, dom, b, x64, r)
        ia, ic = mesh.interior(xm.get(), P.dim), mesh.interior(x64.get(), P.dim)
        res["mixed_vs_fp64_rel_l2"] = float(np.sqrt(((ia - ic) ** 2).sum() / (ic ** 2).sum()))
        res["fp64_dof_per_s"] = P.dof * args.steps / T_64
        res["iterations_fp64"] = its_64
        res["solve_ms_median_fp64"] = statistics.median(ts_64) * 1e3
        res["mixed_over_fp64"] = res["value"] / res["fp64_dof_per_s"]
        if P.exact is not None:
            e = mesh.interior(x64.get(), P.dim) - mesh.interior(P.exact, P.dim)
            res["fp64_max_error_vs_exact"] = float(np.abs(e).max())
    return res, g_mixed, g_fp64, (b, xm, r, filler, dom)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3,
                    help="workload: 1 uniform 2D (64 x 32^2, BiCGStab smoother), 2 adaptive 2D (19,930 x 16^2), "
                         "3 uniform 3D (32,768 x 16^3, default), 4 VarPoisson 2D, 5 independent small problems")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison solve")
    ap.add_argument("--no-secondary", action="store_true", help="skip the adaptive 2D secondary measurement")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the standalone kernel timings (for counter runs under rocprofv3)")
    ap.add_argument("--problems", type=int, default=512, help="many problems: number of problems")
    ap.add_argument("--threads", type=int, default=8, help="many problems: concurrent solves")
    ap.add_argument("--batched-only", action="store_true",
                    help="many problems: the batched solves only (no per-problem comparison)")
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

    from thunderegg_amd import configs
    from thunderegg_amd import runtime as R
    from thunderegg_amd import solver

    R.lib()  # fails loudly when the HIP library is missing
    t0 = time.perf_counter()
    P = configs.CONFIGS[args.config]()
    t_mesh = time.perf_counter() - t0
    log(f"config {args.config}: {P.name}; levels {[len(l) for l in P.levels]}; DOF {P.dof}; mesh {t_mesh:.1f}s")
    if args.kernels_only:
        if P.dim == 3:
            g = solver.GMG(P.dim, P.ns, P.levels, filler_kind=P.filler, smoother=P.smoother, pre_sweeps=P.pre,
                           post_sweeps=P.post, coarse_sweeps=P.coarse, precision=R.F32)
            print(json.dumps({"kernels": kernel_rooflines_3d(g, P, 9, reps=args.steps)}), flush=True)
        else:
            print(json.dumps({"kernels": kernel_rooflines_2d(R, P, R.Stream(), reps=args.steps)}), flush=True)
        return

    res, g_mixed, g_fp64, (b, xm, r, filler, dom) = measure(P, args, R, solver, with_fp64=not args.no_fp64)
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
    out = {
        "metric": METRIC,
        "value": res.pop("value"),
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # the compact 3D path keeps the shadow residual rhat (a fixed vector) and the direction p
        # (the V-cycle's input) as FP32 copies; x, r and s are FP64
        "dtype": "f32 V-cycle / f64 BiCGStab" + (" (f32 shadow residual and direction)" if args.config == 3 else ""),
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
    }
    res["setup_seconds"]["mesh"] = t_mesh
    out.update(res)
    out["tune"] = tune_report(R, [(P.dim, P.ns[0])] + ([(2, 16)] if args.config == 3 and not args.no_secondary else []))
    its_m = out["iterations_mixed"][0]
    if not args.no_roofline:
        if P.dim == 3:
            kr = kernel_rooflines_3d(g_mixed, P, its_m)
            dom_k = max(kr, key=lambda k: kr[k]["ms_per_solve"])
        else:
            kr = kernel_rooflines_2d(R, P, R.Stream())
            cands = [k for k in kr if k not in ("vec_copy_f64", "copy_f64", "ghost_fill_f64")]  # off the timed path
            dom_k = max(cands, key=lambda k: kr[k]["ms"])
        out["kernels"] = kr
        k = kr[dom_k]
        traffic = None
        tf = os.path.join(ROOT, "profiles", f"traffic_{'uniform3d' if P.dim == 3 else 'adaptive2d'}.json")
        if os.path.exists(tf):
            with open(tf) as fh:
                traffic = json.load(fh).get(dom_k, {}).get("bytes_per_launch")
        out["roofline"] = {"bound": "hbm", "achieved": k["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": k["GBs"] / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom_k,
                           "bytes_per_launch": k["bytes"], "ms_per_launch": k["ms"]}
        if "TFs" in k:
            out["roofline"]["mfma_tflops"] = k["TFs"]
            out["roofline"]["mfma_frac"] = k["TFs"] / MFMA_PEAK_TFS["f32"]
    del g_mixed, g_fp64
    if args.config == 3 and not args.no_secondary:
        P2 = configs.adaptive2d_dft()
        sargs = argparse.Namespace(steps=max(args.steps, 10), warmup=max(args.warmup, 2))
        r2, g2m, g2d, _ = measure(P2, sargs, R, solver)
        out["secondary"] = {"workload": P2.name, "dof": P2.dof, "value": r2["value"],
                            "fp64_dof_per_s": r2["fp64_dof_per_s"], "mixed_over_fp64": r2["mixed_over_fp64"],
                            "iterations_mixed": r2["iterations_mixed"][0], "iterations_fp64": r2["iterations_fp64"][0],
                            "solve_ms_median_mixed": r2["solve_ms_median_mixed"],
                            "solve_ms_median_fp64": r2["solve_ms_median_fp64"],
                            "mixed_true_residual": r2["mixed_true_residual"],
                            "fp64_true_residual": r2["fp64_true_residual"]}
        del g2m, g2d
        try:
            r5 = run_many_problems(argparse.Namespace(steps=args.steps, warmup=args.warmup, problems=512,
                                                      threads=args.threads, batched_only=False), emit=False)
            keep = ("value", "fp64_dof_per_s", "mixed_over_fp64", "per_problem_fp64_dof_per_s",
                    "per_problem_mixed_dof_per_s", "batched_over_per_problem_fp64", "batched_over_per_problem_mixed",
                    "fp64_iterations_equal_per_problem", "ms_per_step", "worst_true_residual_mixed",
                    "worst_true_residual_fp64", "tune")
            out["batched_small_problems"] = {"workload": r5["config"]["workload"], "dof": r5["config"]["dof"],
                                             **{k: r5[k] for k in keep},
                                             "iterations_mixed_min_max_sum": [min(r5["iterations_mixed"]),
                                                                              max(r5["iterations_mixed"]),
                                                                              sum(r5["iterations_mixed"])],
                                             "iterations_fp64_min_max_sum": [min(r5["iterations_fp64"]),
                                                                             max(r5["iterations_fp64"]),
                                                                             sum(r5["iterations_fp64"])]}
        except Exception as e:  # the headline line must not depend on this measurement
            log(f"many problems: failed: {e!r}")
            out["batched_small_problems"] = {"error": repr(e)}
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(P, out.get("iterations_fp64", out["iterations_mixed"])[0])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
