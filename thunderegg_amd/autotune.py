"""Autotuner for the DFT patch-solve launches.

For every patch shape the solvers run with the DFT smoother -- 2D n = 8, 16, 32 and 3D
n = 8, in FP32 and FP64 -- it times each launch variant of the kernel that shape uses
(tune.h: patches per workgroup x threads per patch for the VALU kernel, the workgroup count
of the 2D matrix-core kernel) on a uniform level of about 4-16 thousand patches, with HIP
events (te_tune_time_dft), checks that the variant's result has the same bits as the hand
default's, and writes the table the library loads on first use (tune_table.txt beside
libthunderegg_amd.so).  A variant replaces the default only when it is faster by more than
`--margin` (3 %).  The variants are timed in interleaved rounds (each round times every
variant once, starting at a different one) and each keeps its median over the rounds, so
that clock drift during the search favours none.

Usage: python -m thunderegg_amd.autotune [--out PATH] [--reps N] [--rounds R] [--margin F]
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

from . import runtime as R
from .solver import uniform_levels

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "tune_table.txt")

# (dim, n, uniform level): about 4-16 thousand patches each
SHAPES = [(2, 8, 7), (2, 16, 7), (2, 32, 6), (3, 8, 4)]
# the 3D compact cycle (16^3 patches): its finest-level sweeps on a uniform level
C3_LEVEL = 4


def _level(dim, n, level):
    lv = uniform_levels(dim, level, (n,) * dim)[0]
    dom = R.Domain(dim, (n,) * dim, lv)
    filler = R.GhostFiller(dom, R.BILINEAR if dim == 2 else R.TRILINEAR, R.FACES)
    return dom, filler, R.DftPlan(filler)


def interleaved(nvar, time_one, rounds):
    """{variant: median time} over `rounds` rounds, each timing every variant once, the
    starting variant rotating from round to round."""
    samples = {v: [] for v in range(nvar)}
    for r in range(rounds):
        for k in range(nvar):
            v = (k + r) % nvar
            samples[v].append(time_one(v))
    return {v: float(np.median(t)) for v, t in samples.items()}


def tune_shape(dim, n, level, dtype, reps=20, margin=0.03, log=print, rounds=5):
    """Times every variant on one shape; returns (family, best variant, {variant: ms})."""
    dom, filler, plan = _level(dim, n, level)
    fam = R.tune_family(dim, n)
    npdt = np.float64 if dtype == R.F64 else np.float32
    rng = np.random.default_rng(1000 * dim + n)
    shape = dom.shape(1)
    f = R.DeviceArray.from_host(rng.standard_normal(shape).astype(npdt))
    u = R.DeviceArray.from_host(rng.standard_normal(shape).astype(npdt))
    filler.fill_ghost(u)
    outs = {}
    nvar = R.tune_num_variants(fam)
    out = R.DeviceArray(shape, npdt)
    times = interleaved(nvar, lambda v: plan.time_variant(f, u, out, v, reps), rounds)
    for v in range(nvar):
        out.zero()
        plan.time_variant(f, u, out, v, 1)
        outs[v] = out.get()
    ref = outs[0]
    for v, o in outs.items():
        if not np.array_equal(o, ref):
            raise RuntimeError(f"variant {v} of family {fam} ({dim}D n={n}) changed the result")
    best = min(times, key=times.get)
    if times[best] > times[0] * (1.0 - margin):
        best = 0
    log(f"[autotune] {dim}D n={n} {'f64' if dtype == R.F64 else 'f32'} family {fam} ({dom.np} patches): "
        + ", ".join(f"v{v} {t * 1e3:.1f}us" for v, t in times.items()) + f" -> v{best}")
    return fam, best, times


def tune_c3(dtype, reps=20, margin=0.03, log=print, rounds=5):
    """The 3D fused sweeps' variants (slices per load batch, registers sized for more waves
    per SIMD) timed on the compact cycle of a uniform level of 16^3 patches
    (te_tune_time_c3): the finest level's sweeps of one 2 + 2 V-cycle, i.e. the sweep from
    zero, two interface sweeps and the interpolating one; each variant's cycle result must
    have the default's bits."""
    from . import solver

    levels = uniform_levels(3, C3_LEVEL, (16, 16, 16))
    g = solver.GMG(3, (16, 16, 16), levels, filler_kind=R.TRILINEAR, smoother=0, pre_sweeps=2, post_sweeps=2,
                   precision=dtype)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(g.shape)
    fam = R.TUNE_C3_SWEEP
    def time_one(v):
        t = 0.0
        for name, per_cycle in (("sweep_zero", 1), ("sweep", 2), ("sweep_interp", 1)):
            ms = C.c_double()
            R.check(R.lib().te_tune_time_c3(g.h, g.KERNELS[name], v, reps, C.byref(ms)))
            t += per_cycle * ms.value
        return t

    nvar = R.tune_num_variants(fam)
    times = interleaved(nvar, time_one, rounds)
    outs = {}
    for v in range(nvar):
        R.tune_set(fam, 3, 16, dtype, v)
        outs[v] = g.vcycle(f)
    R.tune_clear()
    for v, o in outs.items():
        if not np.array_equal(o, outs[0]):
            raise RuntimeError(f"3D sweep variant {v} changed the cycle's result")
    best = min(times, key=times.get)
    if times[best] > times[0] * (1.0 - margin):
        best = 0
    log(f"[autotune] 3D n=16 {'f64' if dtype == R.F64 else 'f32'} family {fam} ({len(levels[0])} patches, "
        "finest-level sweeps of a 2 + 2 cycle): " + ", ".join(f"v{v} {t * 1e3:.1f}us" for v, t in times.items()) + f" -> v{best}")
    return fam, best, times


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--margin", type=float, default=0.03)
    ap.add_argument("--rounds", type=int, default=5)
    args = ap.parse_args(argv)
    R.lib()
    R.tune_clear()  # time against the hand defaults, not a table loaded earlier
    lines = ["# DFT patch-solve launch variants: family dim n dtype variant (thunderegg_amd/autotune.py)",
             f"# written {time.strftime('%Y-%m-%d')} by: python -m thunderegg_amd.autotune --reps {args.reps} "
             f"--rounds {args.rounds} --margin {args.margin}",
             "# timings per variant in microseconds (v0 = hand default)"]
    for dim, n, level in SHAPES:
        for dtype in (R.F32, R.F64):
            fam, best, times = tune_shape(dim, n, level, dtype, args.reps, args.margin,
                                          log=lambda *a: print(*a, file=sys.stderr, flush=True), rounds=args.rounds)
            lines.append("# " + " ".join(f"v{v}={t * 1e3:.1f}" for v, t in times.items())
                         + f"  gain {times[0] / times[best]:.3f}x")
            lines.append(f"{fam} {dim} {n} {dtype} {best}")
    for dtype in (R.F32, R.F64):
        fam, best, times = tune_c3(dtype, args.reps, args.margin, log=lambda *a: print(*a, file=sys.stderr, flush=True),
                                   rounds=args.rounds)
        lines.append("# " + " ".join(f"v{v}={t * 1e3:.1f}" for v, t in times.items())
                     + f"  gain {times[0] / times[best]:.3f}x")
        lines.append(f"{fam} 3 16 {dtype} {best}")
    with open(args.out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(args.out)


if __name__ == "__main__":
    main()
