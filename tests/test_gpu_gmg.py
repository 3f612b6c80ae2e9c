"""GPU parity of the multigrid path against the oracle: the per-patch BiCGStab
smoother, the V-cycle (GMG::Cycle::apply) and the V-cycle-preconditioned
BiCGStab solve, FP64 and mixed precision, through the C ABI."""
import numpy as np
import pytest

from helpers import load, random_field, relerr
from oracle import oracle as O
from thunderegg_amd import mesh
from thunderegg_amd import runtime as R
from thunderegg_amd import solver

pytestmark = pytest.mark.gpu


def _print_case(*a):
    print("case:", *a, flush=True)


def _rhs_sin(levels, dim, ns):
    X = mesh.cell_centers(levels[0], dim, ns)
    k = 2 * np.pi
    if dim == 2:
        f = -2 * k * k * np.sin(k * X[..., 0]) * np.sin(k * X[..., 1])
    else:
        f = -3 * k * k * np.sin(k * X[..., 0]) * np.sin(k * X[..., 1]) * np.sin(k * X[..., 2])
    return f[:, None].copy()


def _fixture_levels(name, ns):
    dim, levels = load(name, ns)
    return dim, levels


CASES = [
    # (label, dim, ns, levels-builder, filler, smoother)
    ("uniform2d_dft", 2, (16, 16), lambda ns: solver.uniform_levels(2, 2, ns), R.BILINEAR, 0),
    ("uniform2d_bicg", 2, (8, 8), lambda ns: solver.uniform_levels(2, 2, ns), R.BILINEAR, 1),
    ("cross2d_biquad_dft", 2, (8, 8), lambda ns: _fixture_levels("2d_uniform_8x8_refined_cross_mpi1.json", ns)[1],
     R.BIQUADRATIC, 0),
    ("cross2d_biquad_dft16", 2, (16, 16), lambda ns: _fixture_levels("2d_uniform_8x8_refined_cross_mpi1.json", ns)[1],
     R.BIQUADRATIC, 0),
    ("cross2d_bilin_bicg", 2, (8, 8), lambda ns: _fixture_levels("2d_uniform_8x8_refined_cross_mpi1.json", ns)[1],
     R.BILINEAR, 1),
    ("mid3d_dft", 3, (4, 4, 4), lambda ns: _fixture_levels("3d_mid_refine_4x4x4_mpi1.json", ns)[1], R.TRILINEAR, 0),
]


def _oracle_levels(dim, ns, levels, filler, smoother, coeffs=None):
    od = [O.Domain(dim, ns, lv) for lv in levels]
    return O.Levels(od, filler_kind=filler, smoother_kind=smoother, op_kind=1 if coeffs is not None else 0,
                    coeffs=coeffs)


@pytest.mark.parametrize("label,dim,ns,mk,filler,smoother", CASES)
def test_vcycle_fp64(label, dim, ns, mk, filler, smoother):
    _print_case("vcycle", label)
    levels = mk(ns)
    f = random_field(mesh.padded_shape(len(levels[0]), dim, ns), 7)
    ol = _oracle_levels(dim, ns, levels, filler, smoother)
    uref = ol.vcycle(f.copy())
    for graph in (False, True):
        g = solver.GMG(dim, ns, levels, filler_kind=filler, smoother=smoother, precision=R.F64, use_graph=graph)
        u = g.vcycle(f)
        assert relerr(mesh.interior(u, dim), mesh.interior(uref, dim)) <= 1e-12, graph
        u2 = g.vcycle(f)  # replay
        assert np.array_equal(mesh.interior(u2, dim), mesh.interior(u, dim))


@pytest.mark.parametrize("label,dim,ns,mk,filler,smoother", CASES)
def test_bicgstab_fp64_and_mixed(label, dim, ns, mk, filler, smoother):
    _print_case("bicgstab", label)
    levels = mk(ns)
    f = _rhs_sin(levels, dim, ns)
    ol = _oracle_levels(dim, ns, levels, filler, smoother)
    xref, its_ref, hist_ref = ol.bicgstab(f.copy(), tol=1e-10)
    g = solver.GMG(dim, ns, levels, filler_kind=filler, smoother=smoother, precision=R.F64)
    x, its, hist = g.solve(f, tol=1e-10)
    assert its == its_ref
    assert np.allclose(hist, hist_ref, rtol=1e-6, atol=0)
    assert relerr(mesh.interior(x, dim), mesh.interior(xref, dim)) <= 1e-10
    # corrected early exit (x += alpha M p), FP64: parity with the oracle's same variant
    xref2, its_ref2, hist_ref2 = ol.bicgstab(f.copy(), tol=1e-10, precond=2)
    x2, its2, hist2 = g.solve(f, tol=1e-10, precond=2)
    assert its2 == its_ref2
    assert relerr(mesh.interior(x2, dim), mesh.interior(xref2, dim)) <= 1e-10
    od = ol.domains[0]

    def true_resid(xs):
        ax = O.poisson_apply(od, xs.copy(), kind=filler)
        r = mesh.interior(f - ax, dim)
        return np.sqrt((r * r).sum()) / np.sqrt((mesh.interior(f, dim) ** 2).sum())

    assert true_resid(x2) <= 1.5e-10
    # mixed precision: FP32 V-cycle under FP64 BiCGStab reaches the same tolerance
    gm = solver.GMG(dim, ns, levels, filler_kind=filler, smoother=smoother, precision=R.F32)
    xm, itm, histm = gm.solve(f, tol=1e-10, precond=2)
    assert itm <= its_ref2 + 2
    assert true_resid(xm) <= 1.5e-10
    assert relerr(mesh.interior(xm, dim), mesh.interior(x2, dim)) <= 1e-8


def test_bicgstab_unpreconditioned():
    _print_case("bicgstab-noprecond")
    ns = (8, 8)
    levels = solver.uniform_levels(2, 1, ns)
    # a sine right-hand side is an eigenvector of this operator (one step); use a random one
    f = random_field(mesh.padded_shape(len(levels[0]), 2, ns), 17)
    ol = _oracle_levels(2, ns, levels, R.BILINEAR, 0)
    xref, its_ref, hist_ref = ol.bicgstab(f.copy(), tol=1e-9, precond=False)
    assert its_ref > 10
    g = solver.GMG(2, ns, levels, filler_kind=R.BILINEAR, smoother=0, precision=R.F64)
    x, its, hist = g.solve(f, tol=1e-9, precond=False)
    assert abs(its - its_ref) <= 1
    assert relerr(mesh.interior(x, 2), mesh.interior(xref, 2)) <= 1e-6


def test_bicgstab_early_exit():
    """The s-norm test (BiCGStab.h:162-167) decided on the device: with a right-hand side
    that is an eigenvector of A, unpreconditioned, s = r - alpha A r vanishes on the first
    iteration, x += alpha p and the solve stops with 0 iterations counted.  A later solve
    on the same (cached) iteration graph starts from the running state again."""
    _print_case("bicgstab-early-exit")
    ns = (8, 8)
    levels = solver.uniform_levels(2, 1, ns)
    f = _rhs_sin(levels, 2, ns)
    ol = _oracle_levels(2, ns, levels, R.BILINEAR, 0)
    xref, its_ref, _ = ol.bicgstab(f.copy(), tol=1e-9, precond=False)
    assert its_ref == 0
    for graph in (False, True):
        g = solver.GMG(2, ns, levels, filler_kind=R.BILINEAR, smoother=0, precision=R.F64, use_graph=graph)
        x, its, _ = g.solve(f, tol=1e-9, precond=False)
        assert its == 0, graph
        assert relerr(mesh.interior(x, 2), mesh.interior(xref, 2)) <= 1e-12, graph
        f2 = random_field(f.shape, 17)
        xr2, its_r2, _ = ol.bicgstab(f2.copy(), tol=1e-9, precond=False)
        x2, its2, _ = g.solve(f2, tol=1e-9, precond=False)
        assert abs(its2 - its_r2) <= 1, graph
        assert relerr(mesh.interior(x2, 2), mesh.interior(xr2, 2)) <= 1e-6, graph


def test_bicgstab_zero_rhs_and_converged_guess():
    """test/Iterative/BiCGStab_MPI1.cpp:132-154 (zero RHS -> 0 iterations, x untouched)
    and :202-240 (converged initial guess -> 0 iterations)."""
    _print_case("bicgstab-zero")
    ns = (8, 8)
    levels = solver.uniform_levels(2, 1, ns)
    g = solver.GMG(2, ns, levels, filler_kind=R.BILINEAR, smoother=0, precision=R.F64)
    z = np.zeros(mesh.padded_shape(len(levels[0]), 2, ns))
    x0 = random_field(z.shape, 3)
    x, its, _ = g.solve(z, x0=x0)
    assert its == 0 and np.array_equal(mesh.interior(x, 2), mesh.interior(x0, 2))
    f = _rhs_sin(levels, 2, ns)
    xs, _, _ = g.solve(f, tol=1e-12)
    x2, its2, _ = g.solve(f, x0=xs, tol=1e-6)
    assert its2 == 0


@pytest.mark.parametrize("name,ns,op_kind", [("2d_uniform_2x2_refined_nw_mpi1.json", (8, 8), 0),
                                             ("2d_uniform_8x8_refined_cross_mpi1.json", (16, 16), 0),
                                             ("2d_uniform_8x8_refined_cross_mpi1.json", (8, 8), 1),
                                             ("3d_refined_bnw_2x2x2_mpi1.json", (4, 4, 4), 0)])
def test_patch_bicgstab_smoother(name, ns, op_kind):
    _print_case("patch-bicgstab", name, ns, op_kind)
    dim, levels = load(name, ns)
    lv = levels[0]
    kind = R.TRILINEAR if dim == 3 else R.BILINEAR
    od = O.Domain(dim, ns, lv)
    dd = R.Domain(dim, ns, lv)
    g = R.GhostFiller(dd, kind, R.FACES)
    f0 = random_field(od.shape(1), 41)
    u0 = random_field(od.shape(1), 42)
    coeffs = None
    if op_kind == 1:
        X = mesh.cell_centers(lv, dim, ns)
        coeffs = (1.0 + 0.5 * np.sin(3 * X[..., 0]) * np.cos(2 * X[..., 1]))[:, None].copy()
        O.fill_ghost(od, coeffs, kind=kind)
    uref = u0.copy()
    it_ref = O.bicgstab_smooth(od, f0.copy(), uref, kind=kind, op_kind=op_kind, coeffs=coeffs, max_it=1000,
                               tol=1e-12)
    du = R.DeviceArray.from_host(u0)
    iters = R.DeviceArray((len(lv),), np.int32)
    dc = R.DeviceArray.from_host(coeffs) if coeffs is not None else None
    g.patch_bicgstab_smooth(R.DeviceArray.from_host(f0), du, op_kind=op_kind, coeffs=dc, max_it=1000, tol=1e-12,
                            iters=iters)
    got = du.get()
    # inner solves run to 1e-12: the iteration at which each crosses the tolerance depends
    # on rounding (reduction order), so counts agree only to a few iterations
    assert np.abs(iters.get() - it_ref).max() <= max(1, int(0.15 * it_ref.max()))
    assert relerr(mesh.interior(got, dim), mesh.interior(uref, dim)) <= 1e-9
    # FP32 smoother runs and reduces the patch residuals (tolerance 1e-6)
    du32 = R.DeviceArray.from_host(u0.astype(np.float32))
    g.patch_bicgstab_smooth(R.DeviceArray.from_host(f0.astype(np.float32)), du32, op_kind=op_kind,
                            coeffs=R.DeviceArray.from_host(coeffs.astype(np.float32)) if coeffs is not None else None,
                            max_it=1000, tol=1e-6, continue_on_breakdown=True)
    assert relerr(mesh.interior(du32.get(), dim), mesh.interior(uref, dim)) <= 1e-4


def test_varpoisson_gmg():
    _print_case("varpoisson-gmg")
    ns = (16, 16)
    dim, levels = load("2d_uniform_8x8_refined_cross_mpi1.json", ns)
    coeffs = []
    ods = [O.Domain(dim, ns, lv) for lv in levels]
    for lv, od in zip(levels, ods):
        X = mesh.cell_centers(lv, dim, ns)
        c = np.exp(np.sin(3 * X[..., 0]) * np.cos(4 * X[..., 1]))[:, None].copy()
        O.fill_ghost(od, c, kind=R.BILINEAR)
        coeffs.append(c)
    f = _rhs_sin(levels, dim, ns)
    ol = O.Levels(ods, filler_kind=R.BILINEAR, smoother_kind=0, op_kind=1, coeffs=coeffs)
    g = solver.GMG(dim, ns, levels, filler_kind=R.BILINEAR, smoother=0, op_kind=1, coeffs=coeffs, precision=R.F64)
    # one V-cycle: parity at the FP64 bar
    r0 = random_field(f.shape, 9)
    assert relerr(mesh.interior(g.vcycle(r0), dim), mesh.interior(ol.vcycle(r0.copy()), dim)) <= 1e-12
    # the DFT smoother inverts the constant-coefficient Laplacian, so the solve takes
    # ~10^2 iterations and the exact count depends on rounding: compare convergence
    xref, its_ref, _ = ol.bicgstab(f.copy(), tol=1e-10, precond=2)
    x, its, _ = g.solve(f, tol=1e-10, precond=2)
    assert abs(its - its_ref) <= max(2, int(0.2 * its_ref))
    assert relerr(mesh.interior(x, dim), mesh.interior(xref, dim)) <= 1e-8
    gm = solver.GMG(dim, ns, levels, filler_kind=R.BILINEAR, smoother=0, op_kind=1, coeffs=coeffs, precision=R.F32)
    xm, itm, histm = gm.solve(f, tol=1e-10, precond=2)
    ax = O.varpoisson_apply(ods[0], coeffs[0], xm.copy(), kind=R.BILINEAR)
    r = mesh.interior(f - ax, dim)
    assert np.sqrt((r * r).sum()) / np.sqrt((mesh.interior(f, dim) ** 2).sum()) <= 1.5e-10


@pytest.mark.parametrize("split", ["bicg_fine", "dft_fine"])
@pytest.mark.parametrize("op_kind", [0, 1])
def test_heterogeneous_smoothers(split, op_kind):
    """A smoother per level (CycleBuilder.h:97-160): the BiCGStab patch solver on the two
    finest levels and the DFT patch solver below (config 4's cycle), and the reverse, against
    the oracle's cycle with the same per-level smoothers (FP64; 1e-10, the bar of cycles with
    inner BiCGStab solves), and the preconditioned solve (FP64 and mixed) to the
    true-residual bar."""
    _print_case("heterogeneous", split, op_kind)
    ns = (16, 16)
    dim, levels = load("2d_uniform_8x8_refined_cross_mpi1.json", ns)
    nl = len(levels)
    kinds = [1 if (k < 2) == (split == "bicg_fine") else 0 for k in range(nl)]
    assert 0 in kinds and 1 in kinds
    ods = [O.Domain(dim, ns, lv) for lv in levels]
    coeffs = None
    if op_kind == 1:
        coeffs = []
        for lv, od in zip(levels, ods):
            X = mesh.cell_centers(lv, dim, ns)
            c = np.exp(np.sin(3 * X[..., 0]) * np.cos(4 * X[..., 1]))[:, None].copy()
            O.fill_ghost(od, c, kind=R.BILINEAR)
            coeffs.append(c)
    ol = O.Levels(ods, filler_kind=R.BILINEAR, smoother_kind=kinds, op_kind=op_kind, coeffs=coeffs,
                  inner_max_it=[200] * nl, inner_tol=[1e-12] * nl)
    kw = dict(filler_kind=R.BILINEAR, smoother=kinds, op_kind=op_kind, coeffs=coeffs, inner_max_it=[200] * nl,
              inner_tol=[1e-12] * nl)
    g = solver.GMG(dim, ns, levels, precision=R.F64, **kw)
    r0 = random_field(mesh.padded_shape(len(levels[0]), dim, ns), 31)
    # the inner patch solves stop at 1e-12 on a rounding-dependent iteration (as in
    # test_patch_bicgstab_smoother), which moves the cycle's result by ~1e-11
    assert relerr(mesh.interior(g.vcycle(r0), dim), mesh.interior(ol.vcycle(r0.copy()), dim)) <= 1e-10
    f = _rhs_sin(levels, dim, ns)
    for prec in (R.F64, R.F32):
        gp = solver.GMG(dim, ns, levels, precision=prec, **kw)
        x, its, _ = gp.solve(f, tol=1e-10, precond=2)
        if op_kind == 0:
            ax = O.poisson_apply(ods[0], x.copy(), kind=R.BILINEAR)
        else:
            ax = O.varpoisson_apply(ods[0], coeffs[0], x.copy(), kind=R.BILINEAR)
        r = mesh.interior(f - ax, dim)
        assert np.sqrt((r * r).sum()) / np.sqrt((mesh.interior(f, dim) ** 2).sum()) <= 1.5e-10, (prec, its)


@pytest.mark.parametrize("filler", [R.BILINEAR, R.BIQUADRATIC])
@pytest.mark.parametrize("prec", [R.F64, R.F32])
@pytest.mark.parametrize("op_kind", [0, 1])
@pytest.mark.parametrize("ns", [(16, 16), (8, 8), (10, 6)])
def test_fused_vcycle_matches_unfused(monkeypatch, filler, prec, op_kind, ns):
    """The fused 2D launches (ghosts computed inside the smoother / residual-restriction /
    outer-operator kernels; the matrix-core patch solve for 16 x 16, the generic one with
    on-the-fly ghosts for other even sizes) give the same V-cycle and solve as one kernel
    per step."""
    _print_case("fused-vs-unfused", filler, prec, op_kind, ns)
    dim, levels = load("2d_uniform_8x8_refined_cross_mpi1.json", ns)
    coeffs = None
    if op_kind == 1:
        coeffs = []
        for lv in levels:
            X = mesh.cell_centers(lv, dim, ns)
            coeffs.append(np.exp(np.sin(3 * X[..., 0]) * np.cos(4 * X[..., 1]))[:, None].copy())
    kw = dict(filler_kind=filler, smoother=0, op_kind=op_kind, coeffs=coeffs, precision=prec)
    gf = solver.GMG(dim, ns, levels, **kw)
    monkeypatch.setenv("TE_NO_RRPRE", "1")  # fused, but restriction and coarse sweep separate
    gs = solver.GMG(dim, ns, levels, **kw)
    monkeypatch.setenv("TE_NO_FUSE", "1")
    gu = solver.GMG(dim, ns, levels, **kw)
    r0 = random_field(mesh.padded_shape(len(levels[0]), dim, ns), 23)
    tol = 1e-12 if prec == R.F64 else 2e-5
    vu = mesh.interior(gu.vcycle(r0), dim)
    assert relerr(mesh.interior(gf.vcycle(r0), dim), vu) <= tol
    assert relerr(mesh.interior(gs.vcycle(r0), dim), vu) <= tol
    f = _rhs_sin(levels, dim, ns)
    xf, itf, _ = gf.solve(f, tol=1e-10, precond=2)
    xu, itu, _ = gu.solve(f, tol=1e-10, precond=2)
    # VarPoisson takes ~10^2 iterations (the DFT smoother inverts the constant-coefficient
    # Laplacian), so rounding-level differences shift the count by a few
    slack = 0 if (prec == R.F64 and op_kind == 0) else (2 if op_kind == 0 else max(2, int(0.2 * itu)))
    assert abs(itf - itu) <= slack
    assert relerr(mesh.interior(xf, dim), mesh.interior(xu, dim)) <= 1e-8


@pytest.mark.parametrize("prec", [R.F64, R.F32])
@pytest.mark.parametrize("name,base_level", [("2d_uniform_8x8_refined_cross_mpi1.json", None), (None, 4)])
def test_tail_matches_per_level_launches(monkeypatch, prec, name, base_level):
    """The coarsest levels in one launch (levels of at most 16 patches of 16x16, kept in LDS)
    against the per-level launches (TE_NO_TAIL=1): the same arithmetic in the same order,
    so V-cycles and solves agree bit for bit."""
    _print_case("tail", prec, name, base_level)
    ns = (16, 16)
    if name:
        dim, levels = load(name, ns)
    else:  # a uniform level-4 quadtree: 256, 64, 16, 4, 1 patches
        dim, levels = 2, solver.uniform_levels(2, base_level, ns)
    nps = [len(lv) for lv in levels]
    assert any(n <= 16 for n in nps[1:-1]), nps  # a tail of two or more levels exists
    kw = dict(filler_kind=R.BIQUADRATIC, smoother=0, precision=prec)
    gt = solver.GMG(dim, ns, levels, **kw)
    monkeypatch.setenv("TE_NO_TAIL", "1")
    gn = solver.GMG(dim, ns, levels, **kw)
    f = random_field(mesh.padded_shape(nps[0], dim, ns), 51)
    vt, vn = gt.vcycle(f), gn.vcycle(f)
    assert np.array_equal(mesh.interior(vt, dim), mesh.interior(vn, dim))
    assert np.array_equal(gt.vcycle(f), vt)  # graph replay
    b = _rhs_sin(levels, dim, ns)
    xt, itt, ht = gt.solve(b, tol=1e-10, precond=2)
    xn, itn, hn = gn.solve(b, tol=1e-10, precond=2)
    assert itt == itn and np.array_equal(ht, hn)
    assert np.array_equal(mesh.interior(xt, dim), mesh.interior(xn, dim))
    # pre / post / coarse sweep counts other than one
    kw2 = dict(kw, pre_sweeps=2, post_sweeps=0, coarse_sweeps=3)
    monkeypatch.delenv("TE_NO_TAIL")
    g2t = solver.GMG(dim, ns, levels, **kw2)
    monkeypatch.setenv("TE_NO_TAIL", "1")
    g2n = solver.GMG(dim, ns, levels, **kw2)
    assert np.array_equal(mesh.interior(g2t.vcycle(f), dim), mesh.interior(g2n.vcycle(f), dim))


@pytest.mark.parametrize("name", ["2d_uniform_8x8_refined_cross_mpi1.json", None])
def test_fused_update_sweep_matches_separate(monkeypatch, name):
    """Mixed precision, 16x16 patches: the BiCGStab s and p updates that also run the finest
    level's zero-guess sweep give the same solve, bit for bit, as the separate update and
    sweep launches (TE_NO_BSWEEP=1); only the grouping of the s.s partial sums differs, and
    that feeds the early-exit test alone."""
    _print_case("bsweep", name)
    ns = (16, 16)
    if name:
        dim, levels = _fixture_levels(name, ns)
    else:  # a uniform level-4 quadtree
        dim, levels = 2, solver.uniform_levels(2, 4, ns)
    kw = dict(filler_kind=R.BIQUADRATIC, smoother=0, precision=R.F32)
    gf = solver.GMG(dim, ns, levels, **kw)
    monkeypatch.setenv("TE_NO_BSWEEP", "1")
    gs = solver.GMG(dim, ns, levels, **kw)
    for b in (_rhs_sin(levels, dim, ns), random_field(mesh.padded_shape(len(levels[0]), dim, ns), 61)):
        xf, itf, hf = gf.solve(b, tol=1e-10, precond=2)
        xs, its, hs = gs.solve(b, tol=1e-10, precond=2)
        assert itf == its and np.array_equal(hf, hs)
        assert np.array_equal(mesh.interior(xf, dim), mesh.interior(xs, dim))
        # and a second solve from the cached graph
        xf2, itf2, _ = gf.solve(b, tol=1e-10, precond=2)
        assert itf2 == itf and np.array_equal(xf2, xf)


def test_uniform3d_mfma_cycle_and_solve():
    """A uniform octree of 16^3 patches (64, 8, 1), the matrix-core 3D patch solve on every
    level: the V-cycle matches the oracle's and the solve (2 + 2 sweeps) reaches the
    tolerance in the true residual."""
    _print_case("uniform3d")
    ns = (16, 16, 16)
    levels = solver.uniform_levels(3, 2, ns)
    assert [len(lv) for lv in levels] == [64, 8, 1]
    f = random_field(mesh.padded_shape(64, 3, ns), 81)
    ol = _oracle_levels(3, ns, levels, R.TRILINEAR, 0)
    vref = ol.vcycle(f.copy())
    for prec, tol in ((R.F64, 1e-12), (R.F32, 2e-5)):
        g = solver.GMG(3, ns, levels, filler_kind=R.TRILINEAR, smoother=0, precision=prec, pre_sweeps=2,
                       post_sweeps=2)
        g1 = solver.GMG(3, ns, levels, filler_kind=R.TRILINEAR, smoother=0, precision=prec)
        v1 = g1.vcycle(f)
        assert relerr(mesh.interior(v1, 3), mesh.interior(vref, 3)) <= tol, prec
        b = _rhs_sin(levels, 3, ns)
        x, its, _ = g.solve(b, tol=1e-10, precond=2)
        dom = g.domains[0]
        filler = R.GhostFiller(dom, R.TRILINEAR, R.FACES)
        db, dx = R.DeviceArray.from_host(b), R.DeviceArray.from_host(x)
        r = R.DeviceArray(dom.shape(1), np.float64)
        filler.residual(db, dx, r)
        assert dom.two_norm(r) / dom.two_norm(db) <= 1.5e-10, prec


@pytest.mark.parametrize("cycle_type", ["W", "F"])
@pytest.mark.parametrize("label,dim,ns,mk,filler,smoother", [CASES[0], CASES[1]] + [c for c in CASES if "3d" in c[0]][:1])
def test_w_and_f_cycles(cycle_type, label, dim, ns, mk, filler, smoother):
    """GMG::WCycle / GMG::FMGCycle (GMG/WCycle.h:50-84, GMG/FMGCycle.h:43-104) against the
    oracle: FP64 to 1e-12, the FP32 cycle to 2e-5; a BiCGStab solve preconditioned by them
    takes the oracle's iteration count."""
    _print_case("cycle", cycle_type, label)
    levels = mk(ns)
    f = random_field(mesh.padded_shape(len(levels[0]), dim, ns), 31)
    ol = _oracle_levels(dim, ns, levels, filler, smoother)
    uref = ol.cycle(f.copy(), cycle_type)
    for prec, tol in ((R.F64, 1e-12), (R.F32, 2e-5)):
        g = solver.GMG(dim, ns, levels, filler_kind=filler, smoother=smoother, precision=prec, cycle_type=cycle_type)
        u = g.vcycle(f)
        assert relerr(mesh.interior(u, dim), mesh.interior(uref, dim)) <= tol, prec
        u2 = g.vcycle(f)  # graph replay
        assert np.array_equal(u2, u)
    b = _rhs_sin(levels, dim, ns)
    g = solver.GMG(dim, ns, levels, filler_kind=filler, smoother=smoother, precision=R.F64, cycle_type=cycle_type)
    x, its, hist = g.solve(b, tol=1e-10, precond=2)
    od = [O.Domain(dim, ns, lv) for lv in levels]
    L = O.Levels(od, filler_kind=filler, smoother_kind=smoother)
    # the oracle's BiCGStab takes its V-cycle: the W / F preconditioned solve needs no more
    # iterations (to rounding) and reaches the same solution
    xv, itv, _ = L.bicgstab(b.copy(), tol=1e-10, precond=2)
    assert its <= itv + 1
    assert relerr(mesh.interior(x, dim), mesh.interior(xv, dim)) <= 1e-8


@pytest.mark.parametrize("cycle_type", ["V", "W"])
def test_cg_matches_oracle(cycle_type):
    """Iterative::CG::solve (Iterative/CG.h:126-194) through te_gmg_cg against the oracle's
    CG on a uniform mesh (a symmetric operator): plain and right-preconditioned, the same
    iteration count and solution; a zero right-hand side leaves x at zero."""
    _print_case("cg", cycle_type)
    ns = (8, 8)
    dim, levels = load("2d_uniform_4x4_mpi1.json", ns)
    b = _rhs_sin(levels, dim, ns) + 0.1 * random_field(mesh.padded_shape(len(levels[0]), dim, ns), 32)
    ol = _oracle_levels(dim, ns, levels, R.BILINEAR, 0)
    g = solver.GMG(dim, ns, levels, filler_kind=R.BILINEAR, smoother=0, precision=R.F64, cycle_type=cycle_type)
    for precond in (False, True):
        xr, itr, hr = ol.cg(b.copy(), precond=precond, cycle_type=cycle_type, tol=1e-10)
        x, its, hist = g.cg(b, tol=1e-10, precond=precond)
        assert abs(its - itr) <= 1, precond
        assert relerr(mesh.interior(x, dim), mesh.interior(xr, dim)) <= 1e-8, precond
        n = min(len(hist), len(hr))
        assert np.allclose(hist[:n], hr[:n], rtol=1e-6, atol=1e-13), precond
    x0 = random_field(b.shape, 33)
    x, its, _ = g.cg(np.zeros_like(b), x0=x0)
    assert its == 0 and not mesh.interior(x, dim).any()
    # mixed-precision cycle as the preconditioner: the recursive residual converges as fast;
    # the solution is the cycle applied to the CG iterate at the end (x = M y, CG.h:186-189),
    # so it carries the FP32 cycle's rounding
    gm = solver.GMG(dim, ns, levels, filler_kind=R.BILINEAR, smoother=0, precision=R.F32, cycle_type=cycle_type)
    xm, itm, _ = gm.cg(b, tol=1e-10)
    xr, itr, _ = ol.cg(b.copy(), precond=True, cycle_type=cycle_type, tol=1e-10)
    assert itm <= itr + 2
    assert relerr(mesh.interior(xm, dim), mesh.interior(xr, dim)) <= 2e-5
