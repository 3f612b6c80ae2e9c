"""GPU parity of the compact 3D path (uniform octree levels of 16^3 patches: the V-cycle and
the outer BiCGStab on interior-only work vectors, ghost values computed in the launches)
against the oracle and against the one-launch-per-step cycle, through the C ABI.

FP64: max-norm relative difference <= 1e-12 against the oracle's cycle; FP32 cycle (FP64
coarsest level): <= 2e-5 (the patch solves' FP32 matrix-core transforms)."""
import numpy as np
import pytest

from helpers import load, random_field, relerr
from oracle import oracle as O
from thunderegg_amd import mesh
from thunderegg_amd import runtime as R
from thunderegg_amd import solver

pytestmark = pytest.mark.gpu

NS = (16, 16, 16)


def _print_case(*a):
    print("case:", *a, flush=True)


def _levels(base):
    return solver.uniform_levels(3, base, NS)


def _rhs_sin(levels):
    X = mesh.cell_centers(levels[0], 3, NS)
    k = 2 * np.pi
    f = -3 * k * k * np.sin(k * X[..., 0]) * np.sin(k * X[..., 1]) * np.sin(k * X[..., 2])
    return f[:, None].copy()


@pytest.mark.parametrize("prec,tol", [(R.F64, 1e-12), (R.F32, 2e-5)])
@pytest.mark.parametrize("sweeps", [(1, 1, 1), (2, 2, 1), (0, 1, 2), (1, 0, 1), (2, 1, 0)])
def test_compact3d_vcycle_matches_oracle(prec, tol, sweeps):
    pre, post, coarse = sweeps
    _print_case("compact3d vcycle", prec, sweeps)
    levels = _levels(2)
    assert [len(lv) for lv in levels] == [64, 8, 1]
    f = random_field(mesh.padded_shape(64, 3, NS), 91)
    od = [O.Domain(3, NS, lv) for lv in levels]
    ol = O.Levels(od, filler_kind=R.TRILINEAR, smoother_kind=0)
    vref = ol.vcycle(f.copy(), pre=pre, post=post, coarse=coarse)
    for graph in (False, True):
        g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=prec, pre_sweeps=pre,
                       post_sweeps=post, coarse_sweeps=coarse, use_graph=graph)
        v = g.vcycle(f)
        assert relerr(mesh.interior(v, 3), mesh.interior(vref, 3)) <= tol, graph
        v2 = g.vcycle(f)  # replay (a cached graph with use_graph)
        assert np.array_equal(mesh.interior(v2, 3), mesh.interior(v, 3))


@pytest.mark.parametrize("prec,tol", [(R.F64, 1e-12), (R.F32, 2e-5)])
def test_compact3d_neumann_vcycle_matches_oracle(prec, tol):
    """Neumann physical sides: the operator's ghosts are +u there and the patch solves use
    the DCT / DST-IV transforms on those sides (and drop the zero mode of the one patch of the
    coarsest level, which has no neighbours)."""
    _print_case("compact3d neumann vcycle", prec)
    levels = _levels(2)
    f = random_field(mesh.padded_shape(64, 3, NS), 93)
    od = [O.Domain(3, NS, lv) for lv in levels]
    for mask in (0b111111, 0b000101):
        ol = O.Levels(od, filler_kind=R.TRILINEAR, smoother_kind=0, neumann=True, neumann_mask=mask)
        vref = ol.vcycle(f.copy())
        g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=prec, neumann=True,
                       neumann_mask=mask)
        v = g.vcycle(f)
        assert relerr(mesh.interior(v, 3), mesh.interior(vref, 3)) <= tol, mask


def test_compact3d_bicgstab_matches_oracle():
    """FP64: the oracle's iteration count and residual history, both early-exit variants;
    mixed precision: the tolerance in the true residual and the FP64 solution."""
    _print_case("compact3d bicgstab")
    levels = _levels(2)
    f = _rhs_sin(levels)
    od = [O.Domain(3, NS, lv) for lv in levels]
    ol = O.Levels(od, filler_kind=R.TRILINEAR, smoother_kind=0)
    g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=R.F64, pre_sweeps=2, post_sweeps=2)
    for precond in (True, 2):
        xref, its_ref, hist_ref = ol.bicgstab(f.copy(), tol=1e-10, precond=precond, pre=2, post=2)
        x, its, hist = g.solve(f, tol=1e-10, precond=precond)
        assert its == its_ref, precond
        assert np.allclose(hist, hist_ref, rtol=1e-6, atol=0)
        assert relerr(mesh.interior(x, 3), mesh.interior(xref, 3)) <= 1e-10
    x2, _, _ = g.solve(f, tol=1e-10, precond=2)
    x0, its0, _ = g.solve(f, tol=1e-10, precond=False, max_it=400)
    xr0, itsr0, _ = ol.bicgstab(f.copy(), tol=1e-10, precond=False, max_it=400)
    # unpreconditioned: a long, rounding-sensitive iteration (sums in a different order)
    assert abs(its0 - itsr0) <= 2 and relerr(mesh.interior(x0, 3), mesh.interior(xr0, 3)) <= 1e-6

    def true_resid(xs):
        ax = O.poisson_apply(od[0], xs.copy(), kind=R.TRILINEAR)
        r = mesh.interior(f - ax, 3)
        return np.sqrt((r * r).sum()) / np.sqrt((mesh.interior(f, 3) ** 2).sum())

    assert true_resid(x2) <= 1.5e-10
    gm = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=R.F32, pre_sweeps=2,
                    post_sweeps=2)
    xm, itm, _ = gm.solve(f, tol=1e-10, precond=2)
    assert true_resid(xm) <= 1.5e-10
    assert relerr(mesh.interior(xm, 3), mesh.interior(x2, 3)) <= 1e-8


@pytest.mark.parametrize("prec", [R.F64, R.F32])
def test_compact3d_matches_unfused(monkeypatch, prec):
    """The compact path and the one-launch-per-step cycle (TE_NO_FUSE=1) on a 4-level
    hierarchy (512 patches): the same cycle to rounding, and FP64 solves with the same
    iteration count."""
    _print_case("compact3d vs unfused", prec)
    levels = _levels(3)
    assert [len(lv) for lv in levels] == [512, 64, 8, 1]
    f = random_field(mesh.padded_shape(512, 3, NS), 95)
    b = _rhs_sin(levels)
    kw = dict(filler_kind=R.TRILINEAR, smoother=0, precision=prec, pre_sweeps=2, post_sweeps=2)
    gc = solver.GMG(3, NS, levels, **kw)
    vc = gc.vcycle(f)
    xc, itc, _ = gc.solve(b, tol=1e-10, precond=2)
    monkeypatch.setenv("TE_NO_FUSE", "1")
    gu = solver.GMG(3, NS, levels, **kw)
    vu = gu.vcycle(f)
    xu, itu, _ = gu.solve(b, tol=1e-10, precond=2)
    tol = 1e-12 if prec == R.F64 else 2e-5
    assert relerr(mesh.interior(vc, 3), mesh.interior(vu, 3)) <= tol
    if prec == R.F64:
        assert itc == itu
        assert relerr(mesh.interior(xc, 3), mesh.interior(xu, 3)) <= 1e-9
    else:
        assert abs(itc - itu) <= 1
        assert relerr(mesh.interior(xc, 3), mesh.interior(xu, 3)) <= 1e-8


def test_compact3d_fixture_mesh_and_cg():
    """A reference fixture mesh (3d_uniform_2x2x2: 8 patches, two levels) through the compact
    path, and the CG driver on top of the compact cycle."""
    _print_case("compact3d fixture + cg")
    dim, levels = load("3d_uniform_2x2x2_mpi1.json", NS)
    f = random_field(mesh.padded_shape(len(levels[0]), 3, NS), 97)
    od = [O.Domain(3, NS, lv) for lv in levels]
    ol = O.Levels(od, filler_kind=R.TRILINEAR, smoother_kind=0)
    vref = ol.vcycle(f.copy())
    g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=R.F64)
    assert relerr(mesh.interior(g.vcycle(f), 3), mesh.interior(vref, 3)) <= 1e-12
    b = _rhs_sin(levels)
    xr, itr, _ = ol.cg(b.copy(), tol=1e-10)
    x, it, _ = g.cg(b, tol=1e-10)
    assert it == itr
    assert relerr(mesh.interior(x, 3), mesh.interior(xr, 3)) <= 1e-8


@pytest.mark.parametrize("prec", [R.F64, R.F32])
def test_compact3d_face_columns_bit_identical(monkeypatch, prec):
    """The face columns the compact path keeps behind its solution buffers (the x = 0 and
    x = 15 cells of every row, so a neighbour's x face is a contiguous read) hold exact
    copies: cycles (pre-smoothing from zero, without pre-smoothing, without post-smoothing)
    and solves give the same bits as reading the x faces from the cells (TE_NO_XCOL=1)."""
    _print_case("compact3d face columns", prec)
    levels = _levels(3)
    f = random_field(mesh.padded_shape(512, 3, NS), 97)
    b = _rhs_sin(levels)
    for pre, post in ((2, 2), (0, 1), (1, 0)):
        out = []
        for nx in ("0", "1"):
            monkeypatch.setenv("TE_NO_XCOL", nx)
            g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=prec, pre_sweeps=pre,
                           post_sweeps=post)
            v = g.vcycle(f)
            x, its, hist = g.solve(b, tol=1e-10, precond=2)
            out.append((v, x, its, hist))
        (v0, x0, i0, h0), (v1, x1, i1, h1) = out
        assert np.array_equal(mesh.interior(v0, 3), mesh.interior(v1, 3)), (pre, post)
        assert i0 == i1 and np.array_equal(np.asarray(h0), np.asarray(h1)), (pre, post)
        assert np.array_equal(mesh.interior(x0, 3), mesh.interior(x1, 3)), (pre, post)


@pytest.mark.parametrize("prec", [R.F64, R.F32])
def test_compact3d_xr_update_computes_operator(monkeypatch, prec):
    """The s, x / r and p updates that form A M p and A M s themselves from M p and M s (the
    operator launches then store nothing; r = s - omega A M s) against the updates that read
    the stored products (TE_NO_XRFLY=1, r - alpha A M p - omega A M s as the reference):
    the products have the same bits either way; r is rounded in another order and the
    partial sums are summed per patch instead of per grid stride, so FP64 solves take the
    same iterations to rounding-level differences, and both reach the tolerance in the true
    residual.  Also with the unpreconditioned iteration (M s = s, FP64, no face columns) and
    the early exit."""
    _print_case("compact3d xr fly", prec)
    levels = _levels(3)
    b = _rhs_sin(levels)
    od = O.Domain(3, NS, levels[0])
    out = []
    for nx in ("0", "1"):
        monkeypatch.setenv("TE_NO_XRFLY", nx)
        g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=prec, pre_sweeps=2,
                       post_sweeps=2)
        runs = [g.solve(b, tol=1e-10, precond=pc) for pc in (2, True)]
        runs.append(g.solve(b, tol=1e-6, precond=False, max_it=300))
        out.append(runs)
    # preconditioned (8 iterations): equal counts in FP64; unpreconditioned (a long,
    # rounding-sensitive iteration to 1e-6): within a few iterations
    for k, ((x0, i0, h0), (x1, i1, h1)) in enumerate(zip(out[0], out[1])):
        assert abs(i0 - i1) <= (3 if k == 2 else (0 if prec == R.F64 else 1)), k
        assert relerr(mesh.interior(x0, 3), mesh.interior(x1, 3)) <= (1e-4 if k == 2 else 1e-8), k
    for x, _, _ in out[0][:1]:
        ax = O.poisson_apply(od, x.copy(), kind=R.TRILINEAR)
        r = mesh.interior(b - ax, 3)
        assert np.sqrt((r * r).sum() / (mesh.interior(b, 3) ** 2).sum()) <= 1.5e-10


def test_compact3d_fp32_shadow_residual(monkeypatch):
    """Mixed precision keeps BiCGStab's shadow residual rhat as an FP32 copy of r0 (rho_0 and
    every rhat.ap, rhat.r formed from the same widened values): the same iteration with the
    fixed vector fl32(r0) in place of r0.  Against the FP64 rhat (TE_NO_RHAT32=1): equal
    iteration counts, solutions to 1e-8, both at the tolerance in the true residual; with a
    zero and a random initial guess, both early-exit variants."""
    _print_case("compact3d fp32 rhat")
    levels = _levels(3)
    b = _rhs_sin(levels)
    od = O.Domain(3, NS, levels[0])
    x0 = random_field(mesh.padded_shape(len(levels[0]), 3, NS), 7)
    out = []
    for nr in ("0", "1"):
        monkeypatch.setenv("TE_NO_RHAT32", nr)
        g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=R.F32, pre_sweeps=2,
                       post_sweeps=2)
        out.append([g.solve(b, tol=1e-10, precond=pc) for pc in (2, True)] +
                   [g.solve(b, x0=x0, tol=1e-10, precond=2)])
    for k, ((xa, ia, ha), (xb, ib, hb)) in enumerate(zip(out[0], out[1])):
        assert ia == ib, k
        assert np.allclose(ha[:ia + 1], hb[:ib + 1], rtol=1e-3, atol=0), k
        assert relerr(mesh.interior(xa, 3), mesh.interior(xb, 3)) <= 1e-8, k
    for x, _, _ in (out[0][0], out[0][2], out[1][0]):
        ax = O.poisson_apply(od, x.copy(), kind=R.TRILINEAR)
        r = mesh.interior(b - ax, 3)
        assert np.sqrt((r * r).sum() / (mesh.interior(b, 3) ** 2).sum()) <= 1.5e-10


def test_compact3d_fp32_direction(monkeypatch):
    """Mixed precision keeps the direction p only as its FP32 copy (the cycle's input, fl32(p),
    which the FP64 p was rounded to anyway): the p recurrence then rounds to FP32 each
    iteration, x, r and s stay FP64.  Against an FP64 p (TE_NO_P32ONLY=1): iteration counts
    within one, solutions to 1e-8, both at the tolerance in the true residual; zero and
    random initial guesses, both early-exit variants."""
    _print_case("compact3d fp32 direction")
    levels = _levels(3)
    b = _rhs_sin(levels)
    od = O.Domain(3, NS, levels[0])
    x0 = random_field(mesh.padded_shape(len(levels[0]), 3, NS), 9)
    out = []
    for nv in ("0", "1"):
        monkeypatch.setenv("TE_NO_P32ONLY", nv)
        g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=R.F32, pre_sweeps=2,
                       post_sweeps=2)
        out.append([g.solve(b, tol=1e-10, precond=2), g.solve(b, tol=1e-10, precond=True),
                    g.solve(b, x0=x0, tol=1e-10, precond=2)])
    for j, ((xa, ia, _), (xb, ib, _)) in enumerate(zip(out[0], out[1])):
        assert abs(ia - ib) <= 1, j
        assert relerr(mesh.interior(xa, 3), mesh.interior(xb, 3)) <= 1e-8, j
    for x, _, _ in (out[0][0], out[0][2], out[1][0], out[1][2]):
        ax = O.poisson_apply(od, x.copy(), kind=R.TRILINEAR)
        r = mesh.interior(b - ax, 3)
        assert np.sqrt((r * r).sum() / (mesh.interior(b, 3) ** 2).sum()) <= 1.5e-10


def test_compact3d_deferred_x_update(monkeypatch):
    """Mixed precision defers the x update: the x / r update keeps alpha and omega, M p and M s
    rotate through K slots, and x += alpha M p + omega M s is added slot by slot every K
    iterations and at the end of the solve.  Against the x update in the x / r update
    (TE_NO_XDEFER=1): the same iterations and residual history (the r recurrence is
    untouched) and x to 1e-13, with K = 8, 3 and 1, both early-exit variants, a random
    initial guess, and an iteration limit that ends the solve inside a group of slots."""
    _print_case("compact3d deferred x")
    levels = _levels(3)
    b = _rhs_sin(levels)
    od = O.Domain(3, NS, levels[0])
    x0 = random_field(mesh.padded_shape(len(levels[0]), 3, NS), 5)

    def runs():
        g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=R.F32, pre_sweeps=2,
                       post_sweeps=2)
        return [g.solve(b, tol=1e-10, precond=2), g.solve(b, tol=1e-10, precond=True),
                g.solve(b, x0=x0, tol=1e-10, precond=2), g.solve(b, tol=1e-10, precond=2, max_it=5)]

    monkeypatch.setenv("TE_NO_XDEFER", "1")
    ref = runs()
    monkeypatch.delenv("TE_NO_XDEFER")
    for k in ("8", "3", "1"):
        monkeypatch.setenv("TE_XDEFER_SLOTS", k)
        for j, ((xa, ia, ha), (xb, ib, hb)) in enumerate(zip(runs(), ref)):
            assert ia == ib, (k, j)
            assert np.allclose(ha, hb, rtol=1e-13, atol=0), (k, j)
            assert relerr(mesh.interior(xa, 3), mesh.interior(xb, 3)) <= 1e-13, (k, j)
            if j in (0, 2):  # precond True's early exit adds alpha p, as the reference's
                ax = O.poisson_apply(od, xa.copy(), kind=R.TRILINEAR)
                r = mesh.interior(b - ax, 3)
                assert np.sqrt((r * r).sum() / (mesh.interior(b, 3) ** 2).sum()) <= 1.5e-10, (k, j)
    assert ref[3][1] == 5


def test_compact3d_mixed_solve_without_graphs(monkeypatch):
    """The mixed-precision iteration (FP32 rhat and p, deferred x updates) enqueued launch by
    launch (use_graph=False: each slot's launches directly, the x combines at the last slot
    and after the loop) gives the bits of the captured graphs (one per slot), with 3 and 8
    slots and an iteration limit inside a group of slots."""
    _print_case("compact3d mixed without graphs")
    levels = _levels(3)
    b = _rhs_sin(levels)
    for k in ("3", "8"):
        monkeypatch.setenv("TE_XDEFER_SLOTS", k)
        out = []
        for graph in (True, False):
            g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=R.F32, pre_sweeps=2,
                           post_sweeps=2, use_graph=graph)
            out.append([g.solve(b, tol=1e-10, precond=2), g.solve(b, tol=1e-10, precond=2, max_it=4)])
        for (xa, ia, ha), (xb, ib, hb) in zip(out[0], out[1]):
            assert ia == ib and np.array_equal(ha, hb), k
            assert np.array_equal(mesh.interior(xa, 3), mesh.interior(xb, 3)), k


def test_compact3d_bicgstab_initial_guess_matches_oracle(monkeypatch):
    """A nonzero initial guess: the setup forms r = b - A x from the caller's padded b and x
    in one launch (face ghosts from the neighbours' cells of x or the physical condition);
    FP64 against the oracle's BiCGStab (same iterations, history to 1e-6, solution to 1e-10),
    and against the setup through a ghost fill and the padded residual (TE_NO_XRFLY=1)."""
    _print_case("compact3d bicgstab initial guess")
    levels = _levels(2)
    b = _rhs_sin(levels)
    x0 = random_field(mesh.padded_shape(64, 3, NS), 99)
    od = [O.Domain(3, NS, lv) for lv in levels]
    ol = O.Levels(od, filler_kind=R.TRILINEAR, smoother_kind=0)
    xref, its_ref, hist_ref = ol.bicgstab(b.copy(), x=x0.copy(), tol=1e-10, precond=2, pre=2, post=2)
    sols = []
    for nx in ("0", "1"):
        monkeypatch.setenv("TE_NO_XRFLY", nx)
        g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=R.F64, pre_sweeps=2,
                       post_sweeps=2)
        x, its, hist = g.solve(b, x0=x0, tol=1e-10, precond=2)
        assert its == its_ref, nx
        assert np.allclose(hist[:its + 1], hist_ref[:its + 1], rtol=1e-6, atol=0), nx
        assert relerr(mesh.interior(x, 3), mesh.interior(xref, 3)) <= 1e-10, nx
        sols.append(mesh.interior(x, 3))
    assert relerr(sols[0], sols[1]) <= 1e-11


@pytest.mark.parametrize("prec", [R.F64, R.F32])
def test_compact3d_boundary_interpolation_bit_identical(monkeypatch, prec):
    """On levels of 4096 or more patches the first post-smoothing sweep runs as a launch that
    writes the boundary layer of u + P uc (and its face columns) and a plain sweep on it: the
    values the interpolating sweep gathers itself (TE_NO_FACEINTERP=1), so cycles and solves
    give the same bits.  Five levels, 4096 finest patches."""
    _print_case("compact3d boundary interpolation", prec)
    levels = _levels(4)
    assert len(levels[0]) == 4096
    f = random_field(mesh.padded_shape(4096, 3, NS), 101)
    b = _rhs_sin(levels)
    out = []
    for nf in ("0", "1"):
        monkeypatch.setenv("TE_NO_FACEINTERP", nf)
        g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=prec, pre_sweeps=2,
                       post_sweeps=2)
        v = g.vcycle(f)
        x, its, hist = g.solve(b, tol=1e-10, precond=2)
        out.append((mesh.interior(v, 3), mesh.interior(x, 3), its, np.asarray(hist)))
    (v0, x0, i0, h0), (v1, x1, i1, h1) = out
    assert np.array_equal(v0, v1)
    assert i0 == i1 and np.array_equal(h0, h1)
    assert np.array_equal(x0, x1)


def test_compact3d_neumann_bicgstab_matches_oracle():
    """Neumann physical sides on every face (the operator's ghosts +u, the patch solves' DCT
    transforms, the coarsest patch's zero mode dropped) with a zero-mean right-hand side:
    the compact setup, operator and update launches against the oracle's BiCGStab in FP64
    (same iterations, history to 1e-6, solution to 1e-10)."""
    _print_case("compact3d neumann bicgstab")
    levels = _levels(2)
    X = mesh.cell_centers(levels[0], 3, NS)
    k = 2 * np.pi
    b = (np.cos(k * X[..., 0]) * np.cos(k * X[..., 1]) * np.cos(2 * k * X[..., 2]))[:, None].copy()
    od = [O.Domain(3, NS, lv) for lv in levels]
    ol = O.Levels(od, filler_kind=R.TRILINEAR, smoother_kind=0, neumann=True, neumann_mask=0b111111)
    xref, its_ref, hist_ref = ol.bicgstab(b.copy(), tol=1e-10, precond=2, pre=2, post=2)
    g = solver.GMG(3, NS, levels, filler_kind=R.TRILINEAR, smoother=0, precision=R.F64, pre_sweeps=2, post_sweeps=2,
                   neumann=True, neumann_mask=0b111111)
    x, its, hist = g.solve(b, tol=1e-10, precond=2)
    assert its == its_ref
    assert np.allclose(hist[:its + 1], hist_ref[:its + 1], rtol=1e-6, atol=0)
    assert relerr(mesh.interior(x, 3), mesh.interior(xref, 3)) <= 1e-10
