"""Synthetic problem generators for the five benchmark workloads.

Each returns a Problem: dimension, cells per patch side, the patch tables of
every multigrid level (finest first), the right-hand side (padded FP64 vector of
the finest level), solver settings, and the exact solution when one is known.
Seeds: numpy default_rng(20261200 + workload number).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import mesh
from . import runtime as R
from .solver import Forest

SEED0 = 20261200


@dataclass
class Problem:
    name: str
    dim: int
    ns: tuple
    levels: list
    rhs: np.ndarray
    filler: int
    smoother: int
    pre: int = 1
    post: int = 1
    coarse: int = 1
    op_kind: int = 0
    coeffs: list | None = None
    inner_max_it: int = 1000  # Iterative::PatchSolver's inner BiCGStab (smoother 1)
    inner_tol: float = 1e-12
    exact: np.ndarray | None = None
    info: dict = field(default_factory=dict)

    @property
    def dof(self):
        return len(self.levels[0]) * int(np.prod(self.ns))


def _centers(levels, dim, ns):
    return mesh.cell_centers(levels[0], dim, ns)


def uniform2d_bicgstab_smoother():
    """2D Poisson, uniform level 3 (64 patches of 32x32), BiLinear, BiCGStab patch
    smoother, f = -8 pi^2 sin(2 pi x) sin(2 pi y), u = sin(2 pi x) sin(2 pi y)."""
    ns = (32, 32)
    f = Forest(2, 3)
    levels = f.build_domains(ns)
    X = _centers(levels, 2, ns)
    s = np.sin(2 * np.pi * X[..., 0]) * np.sin(2 * np.pi * X[..., 1])
    rhs = (-8 * np.pi**2 * s)[:, None].copy()
    return Problem("2D uniform 64x32^2 BiLinear/BiCGStab-smoother", 2, ns, levels, rhs, R.BILINEAR, 1,
                   exact=s[:, None].copy())


def _gaussian_points(rng, k, dim):
    pts = rng.uniform(0.0, 1.0, size=(k, dim))
    amp = rng.uniform(-1.0, 1.0, size=k)
    wid = rng.uniform(0.005, 0.02, size=k)
    return pts, amp, wid


def _refine_near_points(forest: Forest, dim, pts, radius, min_level, max_level):
    for lev in range(min_level, max_level):
        lo, size = forest.leaf_boxes()
        lv, _ = forest.leaves()
        hi = lo + size[:, None]
        flag = np.zeros(len(lv), dtype=bool)
        for p in pts:
            q = np.clip(p[None, :], lo, hi)
            d = np.sqrt(((q - p[None, :]) ** 2).sum(axis=1))
            flag |= d <= radius
        flag &= lv == lev
        if not flag.any():
            continue
        forest.refine(flag.astype(np.uint8))
    forest.balance()


def _gaussian_rhs(levels, dim, ns, pts, amp, wid):
    X = _centers(levels, dim, ns)
    axes = [np.ascontiguousarray(X[..., a]) for a in range(dim)]
    f = np.zeros(X.shape[:-1])
    r2 = np.empty_like(f)
    t = np.empty_like(f)
    for p, a, w in zip(pts, amp, wid):
        np.subtract(axes[0], p[0], out=r2)
        np.multiply(r2, r2, out=r2)
        for ax in range(1, dim):
            np.subtract(axes[ax], p[ax], out=t)
            np.multiply(t, t, out=t)
            r2 += t
        np.divide(r2, -(2 * w * w), out=r2)
        np.exp(r2, out=r2)
        f += a * r2
    return f[:, None].copy()


def adaptive2d_dft(ns=(16, 16), max_level=9, seed=SEED0 + 2):
    """2D adaptive Poisson: levels 4-9 refined within 0.05 of 8 random points, 2:1
    balanced, 16x16 patches (~20k patches), BiQuadratic filling, DFT smoother, Gaussian-sum RHS."""
    rng = np.random.default_rng(seed)
    pts, amp, wid = _gaussian_points(rng, 8, 2)
    forest = Forest(2, 4)
    _refine_near_points(forest, 2, pts, 0.05, 4, max_level)
    levels = forest.build_domains(ns)
    rhs = _gaussian_rhs(levels, 2, ns, pts, amp, wid)
    return Problem("2D adaptive levels 4-9, 16x16 patches, BiQuadratic/DFT", 2, ns, levels, rhs, R.BIQUADRATIC, 0,
                   info={"points": pts.tolist(), "amplitudes": amp.tolist(), "widths": wid.tolist(), "seed": seed})


def uniform3d_dft(level=5, ns=(16, 16, 16), seed=SEED0 + 3):
    """3D Poisson, uniform octree level 5 (32^3 patches of 16^3), TriLinear, DFT smoother,
    2 pre / 2 post sweeps, f = -12 pi^2 prod sin(2 pi x_i) + N(0, 1e-3) noise, V-cycle over
    5 levels (32768 -> 8 patches)."""
    rng = np.random.default_rng(seed)
    forest = Forest(3, level)
    levels = forest.build_domains(ns)[:5]
    X = _centers(levels, 3, ns)
    s = np.sin(2 * np.pi * X[..., 0]) * np.sin(2 * np.pi * X[..., 1]) * np.sin(2 * np.pi * X[..., 2])
    f = -12 * np.pi**2 * s + rng.normal(0.0, 1e-3, size=s.shape)
    return Problem("3D uniform 32^3 x 16^3, TriLinear/DFT", 3, ns, levels, f[:, None].copy(), R.TRILINEAR, 0, pre=2,
                   post=2, coarse=1, info={"seed": seed, "noise_sigma": 1e-3})


def varpoisson2d_bicgstab_smoother(ns=(32, 32), seed=SEED0 + 4, min_level=5, max_level=9, contrast=1e3):
    """2D variable-coefficient Poisson: beta = exp(g), g a 64-mode random Fourier field
    (|k| <= 16) scaled to max/min beta ~ contrast; levels 5-9 refined where |grad g| is above
    its 90th percentile; 32x32 patches; BiLinear filling; random RHS, zero mean.

    Smoother: Iterative::PatchSolver (an inner BiCGStab on each patch's VarPoisson operator,
    at most 50 iterations to 1e-8).  The DFT patch solver inverts the constant-coefficient
    Laplacian whatever beta is, and at a 1e3 contrast the BiCGStab it preconditions does not
    converge."""
    rng = np.random.default_rng(seed)
    kx = rng.integers(-16, 17, size=64)
    ky = rng.integers(-16, 17, size=64)
    ph = rng.uniform(0, 2 * np.pi, size=64)
    am = rng.normal(size=64) / np.maximum(1.0, np.hypot(kx, ky))

    def g_raw(x, y):
        out = np.zeros(np.broadcast(x, y).shape)
        for a, k1, k2, p in zip(am, kx, ky, ph):
            out += a * np.cos(2 * np.pi * (k1 * x + k2 * y) + p)
        return out

    gx = np.linspace(0, 1, 257)
    G = g_raw(gx[:, None], gx[None, :])
    scale = np.log(contrast) / (G.max() - G.min())

    def gfun(x, y):
        return g_raw(x, y) * scale

    forest = Forest(2, min_level)
    thr = None
    for lev in range(min_level, max_level):
        lo, size = forest.leaf_boxes()
        lv, _ = forest.leaves()
        c = lo + size[:, None] / 2
        h = 1e-4
        grad = np.hypot(gfun(c[:, 0] + h, c[:, 1]) - gfun(c[:, 0] - h, c[:, 1]),
                        gfun(c[:, 0], c[:, 1] + h) - gfun(c[:, 0], c[:, 1] - h)) / (2 * h)
        if thr is None:
            thr = np.percentile(grad, 90)
        flag = (grad > thr) & (lv == lev)
        if flag.any():
            forest.refine(flag.astype(np.uint8))
    forest.balance()
    levels = forest.build_domains(ns)

    def g_cells(X):
        # g at the padded cell centres of every patch: the cell centres of a patch are a
        # tensor grid, so each mode cos(2 pi (k1 x + k2 y) + p) = Re(e^{2 pi i k1 x} e^{i (2 pi k2 y + p)})
        # and the 64-mode sum is one small complex matrix product per patch
        out = np.empty(X.shape[:-1])
        for a in range(0, len(X), 2048):
            xs = X[a:a + 2048, 0, :, 0]  # (P, nx + 2)
            ys = X[a:a + 2048, :, 0, 1]  # (P, ny + 2)
            ex = np.exp(2j * np.pi * kx[None, :, None] * xs[:, None, :])  # (P, M, nx + 2)
            ey = am[None, :, None] * np.exp(1j * (2 * np.pi * ky[None, :, None] * ys[:, None, :] + ph[None, :, None]))
            out[a:a + 2048] = np.real(np.matmul(ey.transpose(0, 2, 1), ex)) * scale
        return out

    coeffs = []
    for lv in levels:
        X = mesh.cell_centers(lv, 2, ns)
        coeffs.append(np.exp(g_cells(X))[:, None].copy())
    f = rng.uniform(-1, 1, size=mesh.padded_shape(len(levels[0]), 2, ns))
    inner = mesh.interior(f, 2)
    inner -= inner.mean()
    return Problem("2D VarPoisson levels 5-9, 32x32 patches, BiLinear/BiCGStab-smoother", 2, ns, levels, f,
                   R.BILINEAR, 1, op_kind=1, coeffs=coeffs, inner_max_it=50, inner_tol=1e-8,
                   info={"seed": seed, "contrast": contrast})


def small_adaptive2d_problem(k, seed=SEED0 + 5, max_level=11):
    """Problem k of the 512 independent small problems: seed + 1000 k; a quadtree refined
    level by level around 8 random points (nearest leaves first) until it holds about a
    patch count drawn log-uniform in [16, 4096]; patch size drawn from {8, 16, 32};
    BiQuadratic filling, DFT smoother, the Gaussian-sum right-hand side of adaptive2d_dft."""
    rng = np.random.default_rng(seed + 1000 * k)
    target = int(round(np.exp(rng.uniform(np.log(16), np.log(4096)))))
    n = int(rng.choice([8, 16, 32]))
    pts, amp, wid = _gaussian_points(rng, 8, 2)
    forest = Forest(2, 2)
    for lev in range(2, max_level):
        lv, _ = forest.leaves()
        if len(lv) >= target:
            break
        lo, size = forest.leaf_boxes()
        c = lo + size[:, None] / 2
        dist = np.min(np.sqrt(((c[:, None, :] - pts[None, :, :]) ** 2).sum(axis=-1)), axis=1)
        cand = np.flatnonzero(lv == lev)
        cand = cand[np.argsort(dist[cand], kind="stable")]
        take = cand[: max(1, (target - len(lv) + 2) // 3)]  # a refinement adds 3 leaves
        flag = np.zeros(len(lv), dtype=np.uint8)
        flag[take] = 1
        forest.refine(flag)
    forest.balance()
    ns = (n, n)
    levels = forest.build_domains(ns)
    rhs = _gaussian_rhs(levels, 2, ns, pts, amp, wid)
    return Problem(f"2D adaptive, {len(levels[0])} patches of {n}x{n}, BiQuadratic/DFT", 2, ns, levels, rhs,
                   R.BIQUADRATIC, 0, info={"seed": seed + 1000 * k, "target_patches": target})


def small_adaptive2d_problems(num=512, seed=SEED0 + 5):
    """`num` independent small problems (small_adaptive2d_problem)."""
    return [small_adaptive2d_problem(k, seed) for k in range(num)]


CONFIGS = {1: uniform2d_bicgstab_smoother, 2: adaptive2d_dft, 3: uniform3d_dft, 4: varpoisson2d_bicgstab_smoother}
