"""GPU parity of the single-level kernels against the oracle, through the C ABI.

FP64: max-norm relative difference <= 1e-12 (the hot path's FP64 parity bar).
FP32: the oracle runs on the float-rounded inputs in FP64; max-norm relative
difference <= 2e-6 (a few float ulps accumulated over the stencil / transforms).
"""
import numpy as np
import pytest

from helpers import MESH_2D, MESH_3D, face_ghost_mask, load, random_field, relerr
from oracle import oracle as O
from thunderegg_amd import mesh
from thunderegg_amd import runtime as R

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6


def _cases_2d():
    for name in MESH_2D:
        for ns in [(8, 8), (16, 16), (10, 6)]:
            yield name, ns


def _dev(dim, ns, lv):
    return R.Domain(dim, ns, lv)


def _print_case(*a):
    print("case:", *a, flush=True)


@pytest.mark.parametrize("name,ns", list(_cases_2d()))
@pytest.mark.parametrize("kind", [R.BILINEAR, R.BIQUADRATIC])
@pytest.mark.parametrize("fill_type", [R.FACES, R.CORNERS])
def test_ghost_fill_2d(name, ns, kind, fill_type):
    _print_case("ghost2d", name, ns, kind, fill_type)
    dim, levels = load(name, ns)
    lv = levels[0]
    od = O.Domain(dim, ns, lv)
    dd = _dev(dim, ns, lv)
    g = R.GhostFiller(dd, kind, fill_type)
    for dtype, tol in [(R.F64, TOL64), (R.F32, TOL32)]:
        for ncomp in (1, 2):
            u0 = random_field(od.shape(ncomp), 11 + ncomp, R.np_dtype(dtype))
            ref = u0.astype(np.float64)
            O.fill_ghost(od, ref, kind=kind, fill_type=fill_type)
            du = R.DeviceArray.from_host(u0)
            g.fill_ghost(du)
            got = du.get()
            assert relerr(got, ref) <= tol, (dtype, ncomp)
            # interior cells are untouched
            assert np.array_equal(mesh.interior(got, dim), mesh.interior(u0, dim))


@pytest.mark.parametrize("name", MESH_3D)
@pytest.mark.parametrize("ns", [(4, 4, 4), (6, 4, 8)])
@pytest.mark.parametrize("fill_type", [R.FACES, R.EDGES, R.CORNERS])
def test_ghost_fill_3d(name, ns, fill_type):
    _print_case("ghost3d", name, ns, fill_type)
    dim, levels = load(name, ns)
    lv = levels[0]
    od = O.Domain(dim, ns, lv)
    dd = _dev(dim, ns, lv)
    g = R.GhostFiller(dd, R.TRILINEAR, fill_type)
    for dtype, tol in [(R.F64, TOL64), (R.F32, TOL32)]:
        u0 = random_field(od.shape(1), 5, R.np_dtype(dtype))
        ref = u0.astype(np.float64)
        O.fill_ghost(od, ref, kind=2, fill_type=fill_type)
        du = R.DeviceArray.from_host(u0)
        g.fill_ghost(du)
        got = du.get()
        assert relerr(got, ref) <= tol
        assert face_ghost_mask(lv, dim, ns, fill_type).sum() > 0


@pytest.mark.parametrize("name,ns", [(n, s) for n, s in _cases_2d() if s != (10, 6)] +
                         [(n, (4, 4, 4)) for n in MESH_3D])
@pytest.mark.parametrize("neumann", [False, True])
def test_star_apply(name, ns, neumann):
    _print_case("star", name, ns, neumann)
    dim, levels = load(name, ns)
    lv = levels[0]
    kind = R.TRILINEAR if dim == 3 else R.BILINEAR
    od = O.Domain(dim, ns, lv)
    dd = _dev(dim, ns, lv)
    g = R.GhostFiller(dd, kind, R.FACES)
    for dtype, tol in [(R.F64, TOL64), (R.F32, 5e-6)]:
        u0 = random_field(od.shape(1), 3, R.np_dtype(dtype))
        ref = O.poisson_apply(od, u0.astype(np.float64), kind=kind, neumann=neumann)
        du = R.DeviceArray.from_host(u0)
        df = dd.vector(1, dtype)
        g.star_apply(du, df, neumann=neumann)
        got = df.get()
        assert relerr(got, ref) <= tol, dtype
        # residual r = f - A u
        f0 = random_field(od.shape(1), 4, R.np_dtype(dtype))
        dfin = R.DeviceArray.from_host(f0)
        dr = dd.vector(1, dtype)
        g.residual(dfin, du, dr, op_kind=0, neumann=neumann)
        rref = np.zeros_like(ref)
        mesh.interior(rref, dim)[...] = mesh.interior(f0.astype(np.float64) - ref, dim)
        assert relerr(dr.get(), rref) <= tol


@pytest.mark.parametrize("name,ns", [("2d_uniform_2x2_refined_nw_mpi1.json", (8, 8)),
                                     ("2d_uniform_8x8_refined_cross_mpi1.json", (16, 16)),
                                     ("3d_refined_bnw_2x2x2_mpi1.json", (4, 4, 4))])
def test_varpoisson_apply(name, ns):
    _print_case("varpoisson", name, ns)
    dim, levels = load(name, ns)
    lv = levels[0]
    kind = R.TRILINEAR if dim == 3 else R.BILINEAR
    od = O.Domain(dim, ns, lv)
    dd = _dev(dim, ns, lv)
    g = R.GhostFiller(dd, kind, R.FACES)
    X = mesh.cell_centers(lv, dim, ns)
    beta = np.exp(np.sin(3 * X[..., 0]) * np.cos(2 * X[..., 1]))[:, None]
    O.fill_ghost(od, beta, kind=kind)
    for dtype, tol in [(R.F64, TOL64), (R.F32, 5e-6)]:
        u0 = random_field(od.shape(1), 8, R.np_dtype(dtype))
        b = beta.astype(R.np_dtype(dtype))
        ref = O.varpoisson_apply(od, b.astype(np.float64), u0.astype(np.float64), kind=kind)
        du = R.DeviceArray.from_host(u0)
        dc = R.DeviceArray.from_host(b)
        df = dd.vector(1, dtype)
        g.varstar_apply(dc, du, df)
        assert relerr(df.get(), ref) <= tol


@pytest.mark.parametrize("name,ns", [("2d_uniform_2x2_refined_nw_mpi1.json", (8, 8)),
                                     ("2d_uniform_8x8_refined_cross_mpi1.json", (16, 16)),
                                     ("2d_uniform_8x8_refined_cross_mpi1.json", (6, 10)),
                                     ("3d_mid_refine_4x4x4_mpi1.json", (4, 4, 4)),
                                     ("3d_refined_bnw_2x2x2_mpi1.json", (6, 4, 4))])
def test_restrict_interpolate(name, ns):
    _print_case("transfer", name, ns)
    dim, levels = load(name, ns)
    of, oc = O.Domain(dim, ns, levels[0]), O.Domain(dim, ns, levels[1])
    of.link_parents(oc)
    df, dc = _dev(dim, ns, levels[0]), _dev(dim, ns, levels[1])
    link = R.LevelLink(df, dc)
    assert np.array_equal(link.parents(), [of.parent_local(p) for p in range(of.np)])
    for dtype, tol in [(R.F64, TOL64), (R.F32, TOL32)]:
        for ncomp in (1, 2):
            fv = random_field(of.shape(ncomp), 21, R.np_dtype(dtype))
            cref = O.restrict(of, oc, fv.astype(np.float64))
            dfv = R.DeviceArray.from_host(fv)
            dcv = R.DeviceArray.from_host(random_field(oc.shape(ncomp), 1, R.np_dtype(dtype)))
            link.restrict(dfv, dcv)
            assert relerr(dcv.get(), cref) <= tol
            cv = random_field(oc.shape(ncomp), 22, R.np_dtype(dtype))
            fref = fv.astype(np.float64).copy()
            O.interpolate(oc, of, cv.astype(np.float64), fref)
            dcv.upload(cv)
            link.interpolate(dcv, dfv)
            assert relerr(dfv.get(), fref) <= tol


@pytest.mark.parametrize("name,ns", [("2d_uniform_2x2_mpi1.json", (8, 8)),
                                     ("2d_uniform_2x2_refined_nw_mpi1.json", (16, 16)),
                                     ("2d_uniform_8x8_refined_cross_mpi1.json", (32, 32)),
                                     ("2d_uniform_8x8_refined_cross_mpi1.json", (10, 6)),
                                     ("3d_uniform_2x2x2_mpi1.json", (8, 8, 8)),
                                     ("3d_refined_bnw_2x2x2_mpi1.json", (16, 16, 16))])
@pytest.mark.parametrize("neumann_mask", [0, 0b0101])
def test_dft_smooth(name, ns, neumann_mask):
    _print_case("dft", name, ns, neumann_mask)
    dim, levels = load(name, ns)
    lv = levels[0]
    kind = R.TRILINEAR if dim == 3 else R.BIQUADRATIC
    od = O.Domain(dim, ns, lv)
    dd = _dev(dim, ns, lv)
    g = R.GhostFiller(dd, kind, R.FACES)
    plan = R.DftPlan(g, neumann_mask=neumann_mask)
    for dtype, tol in [(R.F64, TOL64), (R.F32, 2e-5)]:
        f0 = random_field(od.shape(1), 31, R.np_dtype(dtype))
        u0 = random_field(od.shape(1), 32, R.np_dtype(dtype))
        uref = u0.astype(np.float64)
        O.dft_smooth(od, f0.astype(np.float64), uref, kind=kind, neumann_mask=neumann_mask)
        du = R.DeviceArray.from_host(u0)
        plan.smooth(R.DeviceArray.from_host(f0), du)
        got = du.get()
        assert relerr(mesh.interior(got, dim), mesh.interior(uref, dim)) <= tol, dtype


@pytest.mark.parametrize("name", ["3d_refined_bnw_2x2x2_mpi1.json", "3d_mid_refine_4x4x4_mpi1.json"])
@pytest.mark.parametrize("op_kind", [0, 1])
@pytest.mark.parametrize("neumann_mask", [0, 0b100101])
def test_dft3_mfma_matches_generic(name, op_kind, neumann_mask, monkeypatch):
    """The 3D matrix-core patch solve (16^3 cells) against the generic kernel and the oracle:
    Dirichlet and mixed Neumann transform types, Poisson and VarPoisson interface terms, the
    smoothing sweep and the zero-u apply."""
    _print_case("dft3-mfma", name, op_kind, neumann_mask)
    ns = (16, 16, 16)
    dim, levels = load(name, ns)
    lv = levels[0]
    od = O.Domain(dim, ns, lv)
    dd = _dev(dim, ns, lv)
    g = R.GhostFiller(dd, R.TRILINEAR, R.FACES)
    coeffs = None
    if op_kind == 1:
        X = mesh.cell_centers(lv, dim, ns)
        coeffs = np.exp(np.sin(3 * X[..., 0]) * np.cos(2 * X[..., 1]) + 0.5 * X[..., 2])[:, None].copy()
        O.fill_ghost(od, coeffs, kind=R.TRILINEAR)
    dc = R.DeviceArray.from_host(coeffs) if coeffs is not None else None
    plan_mfma = R.DftPlan(g, neumann_mask=neumann_mask, op_kind=op_kind, coeffs=dc)
    monkeypatch.setenv("TE_DFT_VALU", "1")
    plan_valu = R.DftPlan(g, neumann_mask=neumann_mask, op_kind=op_kind, coeffs=dc)
    for dtype, tol in [(R.F64, 1e-12), (R.F32, 2e-5)]:
        f0 = random_field(od.shape(1), 71, R.np_dtype(dtype))
        u0 = random_field(od.shape(1), 72, R.np_dtype(dtype))
        uref = u0.astype(np.float64)
        O.dft_smooth(od, f0.astype(np.float64), uref, kind=R.TRILINEAR, neumann_mask=neumann_mask, op_kind=op_kind,
                     coeffs=coeffs)
        outs = []
        for plan in (plan_mfma, plan_valu):
            du = R.DeviceArray.from_host(u0)
            plan.smooth(R.DeviceArray.from_host(f0), du)
            outs.append(mesh.interior(du.get(), dim))
        assert relerr(outs[0], mesh.interior(uref, dim)) <= tol, dtype
        assert relerr(outs[0], outs[1]) <= tol, dtype
        du = R.DeviceArray.from_host(u0)
        plan_mfma.apply(R.DeviceArray.from_host(f0), du)
        uz = np.zeros_like(uref)
        O.dft_smooth(od, f0.astype(np.float64), uz, kind=R.TRILINEAR, neumann_mask=neumann_mask, op_kind=op_kind,
                     coeffs=coeffs)
        assert relerr(mesh.interior(du.get(), dim), mesh.interior(uz, dim)) <= tol, dtype


@pytest.mark.parametrize("ns", [(16, 16), (32, 32)])
@pytest.mark.parametrize("op_kind", [0, 1])
def test_dft_mfma_matches_generic(ns, op_kind, monkeypatch):
    """The matrix-core patch solve (2D, n = 16/32, Dirichlet) against the generic kernel
    and against the oracle."""
    _print_case("dft-mfma", ns, op_kind)
    dim, levels = load("2d_uniform_8x8_refined_cross_mpi1.json", ns)
    lv = levels[0]
    od = O.Domain(dim, ns, lv)
    dd = _dev(dim, ns, lv)
    g = R.GhostFiller(dd, R.BIQUADRATIC, R.FACES)
    coeffs = None
    if op_kind == 1:
        X = mesh.cell_centers(lv, dim, ns)
        coeffs = np.exp(np.sin(3 * X[..., 0]) * np.cos(2 * X[..., 1]))[:, None].copy()
        O.fill_ghost(od, coeffs, kind=R.BIQUADRATIC)
    dc = R.DeviceArray.from_host(coeffs) if coeffs is not None else None
    plan_mfma = R.DftPlan(g, op_kind=op_kind, coeffs=dc)
    monkeypatch.setenv("TE_DFT_VALU", "1")
    plan_valu = R.DftPlan(g, op_kind=op_kind, coeffs=dc)
    for dtype, tol in [(R.F64, 1e-12), (R.F32, 2e-5)]:
        f0 = random_field(od.shape(1), 51, R.np_dtype(dtype))
        u0 = random_field(od.shape(1), 52, R.np_dtype(dtype))
        uref = u0.astype(np.float64)
        O.dft_smooth(od, f0.astype(np.float64), uref, kind=R.BIQUADRATIC, op_kind=op_kind, coeffs=coeffs)
        outs = []
        for plan in (plan_mfma, plan_valu):
            du = R.DeviceArray.from_host(u0)
            plan.smooth(R.DeviceArray.from_host(f0), du)
            outs.append(mesh.interior(du.get(), dim))
        assert relerr(outs[0], mesh.interior(uref, dim)) <= tol, dtype
        assert relerr(outs[0], outs[1]) <= tol, dtype
        # apply (u := 0 first): the zero-u variant of the kernel
        du = R.DeviceArray.from_host(u0)
        plan_mfma.apply(R.DeviceArray.from_host(f0), du)
        uz = np.zeros_like(uref)
        O.dft_smooth(od, f0.astype(np.float64), uz, kind=R.BIQUADRATIC, op_kind=op_kind, coeffs=coeffs)
        assert relerr(mesh.interior(du.get(), dim), mesh.interior(uz, dim)) <= tol


@pytest.mark.parametrize("name,ns", [("2d_uniform_8x8_refined_cross_mpi1.json", (16, 16)),
                                     ("2d_uniform_2x2_refined_nw_mpi1.json", (32, 32)),
                                     ("2d_uniform_8x8_refined_cross_mpi1.json", (8, 8)),
                                     ("2d_uniform_2x2_refined_nw_mpi1.json", (10, 6))])
@pytest.mark.parametrize("kind", [R.BILINEAR, R.BIQUADRATIC])
@pytest.mark.parametrize("op_kind", [0, 1])
def test_fused_vcycle_launches(name, ns, kind, op_kind):
    """te_dft_apply_mixed, te_dft_smooth_interp and te_residual_restrict against the oracle's
    unfused sequences (PatchSolver::apply; interpolate then smooth; residual then restrict),
    FP64 and FP32 arithmetic, including the mixed-precision input / output types; 16 x 16 and
    32 x 32 on the matrix cores, 8 x 8 and 10 x 6 in the generic patch-solve kernel."""
    _print_case("fused", name, ns, kind, op_kind)
    dim, levels = load(name, ns)
    fl, cl = levels[0], levels[1]
    fo, co = O.Domain(dim, ns, fl), O.Domain(dim, ns, cl)
    fo.link_parents(co)
    fd, cd = _dev(dim, ns, fl), _dev(dim, ns, cl)
    link = R.LevelLink(fd, cd)
    g = R.GhostFiller(fd, kind, R.FACES)
    coeffs = None
    if op_kind == 1:
        X = mesh.cell_centers(fl, dim, ns)
        coeffs = np.exp(np.sin(3 * X[..., 0]) * np.cos(2 * X[..., 1]))[:, None].copy()
        O.fill_ghost(fo, coeffs, kind=kind)
    f0 = random_field(fo.shape(1), 61)
    u0 = random_field(fo.shape(1), 62)
    uc0 = random_field(co.shape(1), 63)
    # oracle references, FP64
    ua = np.zeros_like(f0)
    O.dft_smooth(fo, f0.copy(), ua, kind=kind, op_kind=op_kind, coeffs=coeffs)
    ui = u0.copy()
    O.interpolate(co, fo, uc0, ui)
    O.dft_smooth(fo, f0.copy(), ui, kind=kind, op_kind=op_kind, coeffs=coeffs)
    if op_kind == 0:
        r = f0 - O.poisson_apply(fo, u0.copy(), kind=kind)
    else:
        r = f0 - O.varpoisson_apply(fo, coeffs, u0.copy(), kind=kind)
    fcr = O.restrict(fo, co, r)
    for dt, tol in [(R.F64, TOL64), (R.F32, 2e-5)]:
        npd = R.np_dtype(dt)
        dc = R.DeviceArray.from_host(coeffs, npd) if coeffs is not None else None
        dc64 = R.DeviceArray.from_host(coeffs) if coeffs is not None else None
        plan = R.DftPlan(g, op_kind=op_kind, coeffs=dc64)
        for tf, to in [(dt, dt), (R.F64, dt), (R.F64, R.F64)]:
            f = R.DeviceArray.from_host(f0, R.np_dtype(tf))
            out = R.DeviceArray(fo.shape(1), R.np_dtype(to))
            plan.apply_mixed(dt, f, out)
            assert relerr(mesh.interior(out.get(), dim), mesh.interior(ua, dim)) <= tol, (dt, tf, to)
            u = R.DeviceArray.from_host(u0, npd)
            for tc in ([R.F64, R.F32] if dt == R.F32 else [R.F64]):
                uc = R.DeviceArray.from_host(uc0, R.np_dtype(tc))
                out = R.DeviceArray(fo.shape(1), R.np_dtype(to))
                plan.smooth_interp(link, f, u, uc, out)
                assert relerr(mesh.interior(out.get(), dim), mesh.interior(ui, dim)) <= tol, (dt, tf, to, tc)
                fc = R.DeviceArray(co.shape(1), R.np_dtype(tc))
                fc.upload(np.zeros(co.shape(1)))
                link.residual_restrict(g, f, u, fc, op_kind=op_kind, coeffs=dc)
                assert relerr(mesh.interior(fc.get(), dim), mesh.interior(fcr, dim)) <= tol, (dt, tf, tc)
            # u and uc are inputs only
            assert np.array_equal(u.get(), u0.astype(npd))


@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_vector_algebra(dtype):
    _print_case("vector", dtype)
    dim, levels = load("2d_uniform_8x8_refined_cross_mpi1.json", (16, 16))
    lv = levels[0]
    od = O.Domain(dim, (16, 16), lv)
    dd = _dev(dim, (16, 16), lv)
    nt = R.np_dtype(dtype)
    tol = TOL64 if dtype == R.F64 else 1e-6
    a = random_field(od.shape(2), 1, nt)
    b = random_field(od.shape(2), 2, nt)
    da, db = R.DeviceArray.from_host(a), R.DeviceArray.from_host(b)
    ia, ib = mesh.interior(a, 2).astype(np.float64), mesh.interior(b, 2).astype(np.float64)
    assert abs(dd.dot(da, db) - (ia * ib).sum()) <= tol * np.abs(ia * ib).sum()
    assert abs(dd.two_norm(da) - np.sqrt((ia * ia).sum())) <= tol * np.sqrt((ia * ia).sum())
    assert dd.inf_norm(da) == pytest.approx(max(ia.max(), 0.0), rel=1e-15)
    neg = R.DeviceArray.from_host(-np.abs(a))
    assert dd.inf_norm(neg) == 0.0  # Vector::infNorm is max(0, max v)
    L = R.lib()
    x = da
    R.check(L.te_vec_add_scaled2(dd.h, dtype, x.ptr, 0.5, db.ptr, -2.0, db.ptr, 2, None))
    exp = a.astype(np.float64).copy()
    mesh.interior(exp, 2)[...] += ib * 0.5 + ib * -2.0
    assert relerr(x.get(), exp) <= tol
    R.check(L.te_vec_scale_then_add(dd.h, dtype, x.ptr, 3.0, db.ptr, 2, None))
    mesh.interior(exp, 2)[...] = mesh.interior(exp, 2) * 3.0 + ib
    assert relerr(x.get(), exp) <= tol
    R.check(L.te_vec_set_with_ghost(dd.h, dtype, x.ptr, 1.5, 2, None))
    assert np.all(x.get() == 1.5)
    R.check(L.te_vec_set(dd.h, dtype, x.ptr, 0.0, 2, None))
    got = x.get()
    assert np.all(mesh.interior(got, 2) == 0.0) and got.max() == 1.5


@pytest.mark.parametrize("name,dim,ns", [("2d_uniform_2x2_refined_nw_mpi1.json", 2, (10, 6)),
                                         ("2d_uniform_8x8_refined_cross_mpi1.json", 2, (16, 16)),
                                         ("3d_refined_bnw_2x2x2_mpi1.json", 3, (4, 6, 8))])
def test_boundary_data_rhs(name, dim, ns):
    """te_star_add_dirichlet_bc (Poisson and VarPoisson forms) and te_star_add_neumann_bc
    against the oracle on the same sampled boundary data (FP64 <= 1e-14, FP32 storage
    <= 1e-6), and the reference's linear-exactness identity on the device operator: for a
    linear g, A g (physical ghosts -g_b, or +g_b with Neumann) equals the boundary terms."""
    from thunderegg_amd import domain_tools as DT

    _print_case("boundary data", name, ns)
    _, levels = load(name, ns)
    lv = levels[0]
    od = O.Domain(dim, ns, lv)
    dd = _dev(dim, ns, lv)
    pts = DT.boundary_points(lv, dim, ns)
    assert len(pts) == dd.boundary_length()
    rng = np.random.default_rng(3)
    g = rng.standard_normal(len(pts))
    cval = rng.uniform(0.5, 2.0, len(pts))
    dg, dc = R.DeviceArray.from_host(g), R.DeviceArray.from_host(cval)
    for dtype, tol in [(R.F64, 1e-14), (R.F32, 1e-6)]:
        f0 = random_field(od.shape(1), 21, R.np_dtype(dtype))
        for form in ("dirichlet", "varpoisson", "neumann"):
            ref = f0.astype(np.float64)
            df = R.DeviceArray.from_host(f0)
            if form == "dirichlet":
                O.add_bc(od, ref, g)
                dd.add_dirichlet_bc(df, dg)
            elif form == "varpoisson":
                O.add_bc(od, ref, g, c=cval)
                dd.add_dirichlet_bc(df, dg, dc)
            else:
                O.add_bc(od, ref, g, neumann=True)
                dd.add_neumann_bc(df, dg)
            assert relerr(df.get(), ref) <= tol, (dtype, form)
    # linear exactness through the device operator
    kind = R.TRILINEAR if dim == 3 else R.BILINEAR
    coef = [5.0, 13.0, 7.0][:dim]
    Xg = mesh.cell_centers(lv, dim, ns, with_ghost=True)
    gv = (1 + sum(coef[a] * Xg[..., a] for a in range(dim)))[:, None].copy()
    filler = R.GhostFiller(dd, kind, R.FACES)
    for neumann in (False, True):
        du = R.DeviceArray.from_host(gv)
        df = R.DeviceArray(dd.shape(1), np.float64)
        filler.star_apply(du, df, neumann=neumann)
        fe = R.DeviceArray.from_host(np.zeros_like(gv))
        if neumann:
            DT.add_neumann_bc(dd, fe, [(lambda c, k=k: np.full_like(c[0], coef[k])) for k in range(dim)])
        else:
            DT.add_dirichlet_bc(dd, fe, lambda c: 1 + sum(coef[a] * c[a] for a in range(dim)))
        a, b = mesh.interior(df.get(), dim), mesh.interior(fe.get(), dim)
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max(), neumann


@pytest.mark.parametrize("name,ns", [("2d_uniform_4x4_mpi1.json", (10, 2)), ("2d_uniform_2x2_refined_nw_mpi1.json", (4, 4)),
                                     ("2d_uniform_8x8_refined_cross_mpi1.json", (8, 8)),
                                     ("3d_refined_bnw_2x2x2_mpi1.json", (4, 4, 4)),
                                     ("3d_mid_refine_4x4x4_mpi1.json", (2, 4, 6))])
@pytest.mark.parametrize("ncomp", [1, 2])
def test_restrict_extrapolate(name, ns, ncomp):
    """LinearRestrictor with extrapolate_boundary_ghosts (te_restrict_extrapolate) against the
    oracle, every cell of the coarse vector including its ghosts, FP64 bit-exact (the same
    terms summed in the same order) and FP32 to rounding; without the flag the ghosts stay 0."""
    print("case: restrict_extrapolate", name, ns, ncomp, flush=True)
    dim, levels = load(name, ns)
    of, oc = O.Domain(dim, ns, levels[0]), O.Domain(dim, ns, levels[1])
    of.link_parents(oc)
    df, dc = R.Domain(dim, ns, levels[0]), R.Domain(dim, ns, levels[1])
    link = R.LevelLink(df, dc)
    fv = random_field(of.shape(ncomp), 17)
    ref = O.restrict(of, oc, fv, extrapolate=True)
    assert np.any(ref[..., 0, 1:-1]) if dim == 2 else np.any(ref[..., 0, 1:-1, 1:-1])
    dfv = R.DeviceArray.from_host(fv)
    dcv = R.DeviceArray(oc.shape(ncomp), np.float64)
    link.restrict_extrapolate(dfv, dcv)
    assert np.array_equal(dcv.get(), ref)
    d32 = R.DeviceArray.from_host(fv.astype(np.float32))
    c32 = R.DeviceArray(oc.shape(ncomp), np.float32)
    link.restrict_extrapolate(d32, c32)
    assert np.max(np.abs(c32.get() - ref)) <= 1e-5 * np.max(np.abs(ref))
    link.restrict(dfv, dcv)
    assert np.array_equal(dcv.get(), O.restrict(of, oc, fv))
