"""CPU tests that pin the oracle (oracle/te_oracle.c) before it is trusted as the
checker of the HIP path.  Each check is one the reference's own tests make, or one
the mathematics fixes, chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails it:

  ghost fillers   polynomial exactness on the reference's fixture meshes
                  (test/BiLinearGhostFiller_MPI1.cpp, test/BiQuadraticGhostFiller_MPI1.cpp,
                  test/TriLinearGhostFiller_MPI1.cpp), also with ghosts already set;
  operators       exactness of ghost fill + star stencil on polynomials at every cell
                  away from the physical boundary, second-order convergence
                  (test/Poisson/StarPatchOperator_MPI1.cpp), VarPoisson reducing to
                  Poisson for constant beta;
  DFT solver      DST matrices against scipy, exact inverse of the single-patch
                  operator with internal boundary conditions, second-order convergence
                  (test/Poisson/DFTPatchSolver_MPI1.cpp);
  transfers       LinearRestrictor exact on linear functions
                  (test/GMG/LinearRestrictor_MPI1.cpp), DirectInterpolator known answer
                  (test/GMG/DirectInterpolator_MPI1.cpp);
  BiCGStab        against a dense direct solve of the assembled operator (brute force
                  on a 64-DOF mesh), zero right-hand side, converged initial guess
                  (test/Iterative/BiCGStab_MPI1.cpp);
  Neumann DFT     the solve with any set of Neumann sides (2D, 3D) is the exact inverse of
                  an independently assembled dense one-patch operator (zero-mean solution
                  when every side is Neumann);
  patch BiCGStab  one sweep of the inner-BiCGStab patch smoother at a tight tolerance
                  equals one DFT sweep on refined meshes (Dirichlet, Neumann, and
                  VarPoisson with beta = 1);
  V-cycle         one patch and one level: the cycle is the exact inverse; two levels:
                  the cycle equals its definition (VCycle.h:46-70) composed from the
                  separately pinned smoother, operator, restrictor and interpolator;
  reductions      dot / twoNorm / infNorm: ghosts excluded, symmetry, bilinearity,
                  max(0, max v);
  Integrate       the midpoint rule's closed forms on the fixture meshes (exact for
                  products of affine functions, the known h^2/12 shortfall for x^2).
"""
import numpy as np
import pytest
import scipy.fft

from helpers import MESH_2D, MESH_3D, load, random_field, relerr
from oracle import oracle as O
from thunderegg_amd import mesh

BIL, BIQ, TRI = 0, 1, 2
FACES, EDGES, CORNERS = 0, 1, 2

MESH_2D_REF = ["2d_uniform_2x2_mpi1.json", "2d_uniform_2x2_refined_nw_mpi1.json",
               "2d_uniform_8x8_refined_cross_mpi1.json"]


def _fun2(kind):
    if kind == BIL:
        return lambda x, y: 1 + ((x * 0.3) + y)
    return lambda x, y: 1 + ((x * x + 0.3 * x + 0.3) + (0.5 * y * y))


def _check_filled_2d(lv, dim, ns, u, expected, fill_type):
    err = 0.0
    for p in range(len(lv)):
        t = lv["nbr_type"][p]
        for s in range(4):
            if t[mesh.side_slot(2, s)] < 0:
                continue
            sl = [(slice(1, -1), 0), (slice(1, -1), -1), (0, slice(1, -1)), (-1, slice(1, -1))][s]
            err = max(err, np.abs(u[p, 0][sl] - expected[p, 0][sl]).max())
        if fill_type == CORNERS:
            for c in range(4):
                if t[c] < 0:
                    continue
                iy, ix = (-1 if c & 2 else 0), (-1 if c & 1 else 0)
                err = max(err, abs(u[p, 0, iy, ix] - expected[p, 0, iy, ix]))
    return err


@pytest.mark.parametrize("name", MESH_2D_REF)
@pytest.mark.parametrize("kind,nss", [(BIL, [(2, 2), (10, 10), (10, 2)]), (BIQ, [(10, 10), (13, 13), (10, 13)])])
@pytest.mark.parametrize("fill_type", [FACES, CORNERS])
def test_ghost_filler_2d_exact(name, kind, nss, fill_type):
    fun = _fun2(kind)
    for ns in nss:
        dim, levels = load(name, ns)
        lv = levels[0]
        d = O.Domain(dim, ns, lv)
        X = mesh.cell_centers(lv, dim, ns)
        expected = fun(X[..., 0], X[..., 1])[:, None]
        for ghosts_preset in (False, True):
            u = expected.copy()
            if not ghosts_preset:  # ghosts start at zero
                u[..., 0, :] = 0
                u[..., -1, :] = 0
                u[..., :, 0] = 0
                u[..., :, -1] = 0
            else:  # "ghost already set": garbage in the ghosts must not leak into the result
                g = random_field(u.shape, 5)
                for sl in [(Ellipsis, 0, slice(None)), (Ellipsis, -1, slice(None)), (Ellipsis, slice(None), 0),
                           (Ellipsis, slice(None), -1)]:
                    u[sl] = g[sl]
            O.fill_ghost(d, u, kind=kind, fill_type=fill_type)
            assert np.array_equal(mesh.interior(u, dim), mesh.interior(expected, dim))
            err = _check_filled_2d(lv, dim, ns, u, expected, fill_type)
            assert err <= 1e-12 * np.abs(expected).max(), (ns, ghosts_preset, err)


def test_biquadratic_is_not_exact_for_cubics():
    """The exactness test has teeth: a cubic breaks BiQuadratic at coarse-fine faces."""
    ns = (10, 10)
    dim, levels = load("2d_uniform_2x2_refined_nw_mpi1.json", ns)
    lv = levels[0]
    d = O.Domain(dim, ns, lv)
    X = mesh.cell_centers(lv, dim, ns)
    expected = (X[..., 0] ** 3 + X[..., 1] ** 3)[:, None]
    u = expected.copy()
    O.fill_ghost(d, u, kind=BIQ, fill_type=FACES)
    assert _check_filled_2d(lv, dim, ns, u, expected, FACES) > 1e-6


@pytest.mark.parametrize("name", MESH_3D)
@pytest.mark.parametrize("fill_type", [FACES, EDGES, CORNERS])
def test_trilinear_exact_3d(name, fill_type):
    ns = (4, 6, 8)
    dim, levels = load(name, ns)
    lv = levels[0]
    d = O.Domain(dim, ns, lv)
    X = mesh.cell_centers(lv, dim, ns)
    expected = (1 + 0.3 * X[..., 0] + X[..., 1] - 0.7 * X[..., 2])[:, None]
    u = expected.copy()
    mask = np.zeros(u.shape, dtype=bool)
    mask[..., 1:-1, 1:-1, 1:-1] = True
    u[~mask] = 0.0
    O.fill_ghost(d, u, kind=TRI, fill_type=fill_type)
    from helpers import face_ghost_mask
    gm = face_ghost_mask(lv, dim, ns, fill_type)
    assert gm.any()
    assert np.abs(u[gm] - expected[gm]).max() <= 1e-12 * np.abs(expected).max()


def _away_from_boundary(lv, dim, ns):
    """Mask of interior cells whose stencil does not touch a physical boundary."""
    shp = mesh.padded_shape(len(lv), dim, ns)
    m = np.zeros(shp, dtype=bool)
    m[(Ellipsis,) + (slice(1, -1),) * dim] = True
    for p in range(len(lv)):
        for s in range(2 * dim):
            if lv["nbr_type"][p][mesh.side_slot(dim, s)] >= 0:
                continue
            a, up = s >> 1, s & 1
            idx = [slice(None)] * dim
            idx[dim - 1 - a] = -2 if up else 1
            m[(p, 0) + tuple(idx)] = False
    return m


@pytest.mark.parametrize("name", MESH_2D)
def test_poisson_apply_exact_on_quadratics(name):
    """Ghost fill (BiQuadratic) + 5-point stencil reproduce the Laplacian of a quadratic at
    every cell not next to the physical boundary, including coarse-fine interfaces."""
    ns = (10, 10)
    dim, levels = load(name, ns)
    lv = levels[0]
    d = O.Domain(dim, ns, lv)
    X = mesh.cell_centers(lv, dim, ns)
    u = (1 + X[..., 0] * X[..., 0] + 0.3 * X[..., 0] + 0.5 * X[..., 1] * X[..., 1] - 0.2 * X[..., 0] * X[..., 1])[:, None]
    f = O.poisson_apply(d, u.copy(), kind=BIQ)
    m = _away_from_boundary(lv, dim, ns)
    assert np.abs(f[m] - 3.0).max() <= 1e-8  # 2 + 1
    # a linear function is harmonic: BiLinear suffices
    ul = (1 + 5 * X[..., 0] + 13 * X[..., 1])[:, None]
    fl = O.poisson_apply(d, ul.copy(), kind=BIL)
    assert np.abs(fl[m]).max() <= 1e-8


def _sin_problem(lv, dim, ns):
    X = mesh.cell_centers(lv, dim, ns)
    x, y = X[..., 0], X[..., 1]
    g = np.sin(np.pi * x) * np.sin(2 * np.pi * y)
    return g[:, None], (-5 * np.pi ** 2 * g)[:, None]


@pytest.mark.parametrize("name", ["2d_uniform_2x2_mpi1.json", "2d_uniform_4x4_mpi1.json"])
def test_poisson_second_order(name):
    """Uniform meshes (test/Poisson/StarPatchOperator_MPI1.cpp sizes 32 -> 64).  On refined
    meshes the linear interface interpolation leaves an O(1) truncation error on a layer
    of O(1/h) cells, so the global order there is 1/2 (BiLinear) or about 3/2 (BiQuadratic);
    those meshes are pinned by the polynomial exactness tests instead."""
    errs = []
    for i in (1, 2):
        ns = (32 * i, 32 * i)
        dim, levels = load(name, ns)
        lv = levels[0]
        d = O.Domain(dim, ns, lv)
        g, fexact = _sin_problem(lv, dim, ns)
        f = O.poisson_apply(d, g.copy(), kind=BIL)
        e = mesh.interior(f - fexact, dim)
        errs.append(np.sqrt((e * e).sum() / (mesh.interior(fexact, dim) ** 2).sum()))
    assert np.log(errs[0] / errs[1]) / np.log(2) > 1.8


def test_varpoisson_constant_beta_is_poisson():
    """beta = 1: (c+ + c)(u+ - u) - (c- + c)(u - u-) over 2h^2 is the 5-point Laplacian."""
    ns = (8, 8)
    dim, levels = load("2d_uniform_2x2_refined_nw_mpi1.json", ns)
    lv = levels[0]
    d = O.Domain(dim, ns, lv)
    u = random_field(d.shape(), 4)
    c = np.ones_like(u)
    a = O.poisson_apply(d, u.copy(), kind=BIL)
    b = O.varpoisson_apply(d, c, u.copy(), kind=BIL)
    assert relerr(mesh.interior(b, dim), mesh.interior(a, dim)) <= 1e-13
    # beta = 3 scales the operator
    b3 = O.varpoisson_apply(d, 3 * c, u.copy(), kind=BIL)
    assert relerr(mesh.interior(b3, dim), 3 * mesh.interior(a, dim)) <= 1e-13


@pytest.mark.parametrize("n", [4, 10, 13, 16])
def test_dft_matrices_match_scipy(n):
    """DST-II (kind 3 in the plan's table) and DST-III (kind 4) against scipy's
    unnormalised transforms: the reference's matrices are scipy's / 2."""
    e = np.eye(n)
    s2 = O.dft_matrix(3, n)
    s3 = O.dft_matrix(4, n)
    ref2 = scipy.fft.dst(e, type=2, axis=0) / 2
    ref3 = scipy.fft.dst(e, type=3, axis=0) / 2
    assert np.allclose(s2, ref2, atol=1e-13)
    assert np.allclose(s3, ref3, atol=1e-13)
    # S3 S2 = (n / 2) I: the transform pair inverts up to the scale prod(2/n)
    assert np.allclose(s3 @ s2, (n / 2) * np.eye(n), atol=1e-11)


def test_dft_solve_is_exact_inverse_single_patch():
    """One patch, homogeneous Dirichlet: the DFT smooth solves A u = f exactly."""
    ns = (12, 12)
    dim, levels = load("2d_uniform_2x2_mpi1.json", ns)
    lv = levels[-1]  # coarsest level: one patch
    assert len(lv) == 1
    d = O.Domain(dim, ns, lv)
    f = random_field(d.shape(), 8)
    u = np.zeros_like(f)
    O.dft_smooth(d, f, u, kind=BIL)
    back = O.poisson_apply(d, u.copy(), kind=BIL)
    assert relerr(mesh.interior(back, dim), mesh.interior(f, dim)) <= 1e-12


@pytest.mark.parametrize("name", ["2d_uniform_2x2_mpi1.json", "2d_uniform_8x8_refined_cross_mpi1.json"])
@pytest.mark.parametrize("nx", [10, 13])
def test_dft_second_order(name, nx):
    """test/Poisson/DFTPatchSolver_MPI1.cpp: one block-Jacobi sweep from the exact solution
    with interface ghosts filled from it converges at second order (homogeneous
    Dirichlet variant of the reference problem)."""
    errs = []
    for i in (1, 2):
        ns = (nx * i, nx * i)
        dim, levels = load(name, ns)
        lv = levels[0]
        d = O.Domain(dim, ns, lv)
        g, f = _sin_problem(lv, dim, ns)
        u = g.copy()
        O.dft_smooth(d, f, u, kind=BIL)
        e = mesh.interior(u - g, dim)
        errs.append(np.sqrt((e * e).sum() / (mesh.interior(g, dim) ** 2).sum()))
    assert np.log(errs[0] / errs[1]) / np.log(2) > 1.8


@pytest.mark.parametrize("name", ["2d_uniform_4x4_mpi1.json", "2d_uniform_2x2_refined_nw_mpi1.json"])
@pytest.mark.parametrize("ns", [(2, 2), (10, 10), (2, 10)])
def test_restrict_exact_on_linear(name, ns):
    dim, levels = load(name, ns)
    fine, coarse = O.Domain(dim, ns, levels[0]), O.Domain(dim, ns, levels[1])
    fine.link_parents(coarse)
    fun = _fun2(BIL)
    Xf = mesh.cell_centers(levels[0], dim, ns)
    Xc = mesh.cell_centers(levels[1], dim, ns)
    fv = fun(Xf[..., 0], Xf[..., 1])[:, None]
    cexp = fun(Xc[..., 0], Xc[..., 1])[:, None]
    cv = O.restrict(fine, coarse, fv)
    assert relerr(mesh.interior(cv, dim), mesh.interior(cexp, dim)) <= 1e-13


def _physical_side_ghosts(level, dim, ns, v):
    """(patch, side, ghost values) of every side without a neighbour, 2D"""
    out = []
    for p in range(len(level)):
        for s in range(4):
            if level["nbr_type"][p][4 + s] >= 0:
                continue
            a, up = s >> 1, s & 1
            if a == 0:
                out.append((p, s, v[p, :, 1:-1, -1 if up else 0]))
            else:
                out.append((p, s, v[p, :, -1 if up else 0, 1:-1]))
    return out


@pytest.mark.parametrize("name", ["2d_uniform_4x4_mpi1.json", "2d_uniform_2x2_refined_nw_mpi1.json"])
@pytest.mark.parametrize("nx,ny", [(2, 2), (10, 10), (2, 10), (10, 2)])
@pytest.mark.parametrize("ncomp", [1, 2])
def test_restrict_extrapolation_known_answer(name, nx, ny, ncomp):
    """test/GMG/LinearRestrictor_MPI1.cpp "Linear Test LinearRestrictor" (and its two-component
    case): 1 + 0.3 x + y (and 9 + 0.9 x + 4 y) set with ghosts on the fine level restricts with
    extrapolate_boundary_ghosts to the same functions on the coarse interior and on the
    ghosts of every coarse side without a neighbour; without extrapolation those ghosts stay 0."""
    ns = (nx, ny)
    dim, levels = load(name, ns)
    fine, coarse = O.Domain(dim, ns, levels[0]), O.Domain(dim, ns, levels[1])
    fine.link_parents(coarse)
    funs = [lambda x, y: 1 + (x * 0.3 + y), lambda x, y: 9 + (x * 0.9 + y * 4)][:ncomp]
    Xf = mesh.cell_centers(levels[0], dim, ns)
    Xc = mesh.cell_centers(levels[1], dim, ns)
    fv = np.stack([f(Xf[..., 0], Xf[..., 1]) for f in funs], axis=1)
    cexp = np.stack([f(Xc[..., 0], Xc[..., 1]) for f in funs], axis=1)
    cv = O.restrict(fine, coarse, fv, extrapolate=True)
    assert relerr(mesh.interior(cv, dim), mesh.interior(cexp, dim)) <= 1e-13
    ghosts = _physical_side_ghosts(levels[1], dim, ns, cv)
    assert ghosts
    for (p, s, got), (_, _, exp) in zip(ghosts, _physical_side_ghosts(levels[1], dim, ns, cexp)):
        assert np.allclose(got, exp, rtol=1e-13, atol=1e-13), (p, s)
    plain = O.restrict(fine, coarse, fv)
    assert all(not np.any(g) for _, _, g in _physical_side_ghosts(levels[1], dim, ns, plain))
    assert np.array_equal(mesh.interior(plain, dim), mesh.interior(cv, dim))


def test_restrict_extrapolation_corners_and_copies():
    """The other two pieces of extrapolateBoundaries / copyToParent (LinearRestrictor.h:79-91,
    138-155) on the nw-refined mesh, where the coarse level holds both coarsened patches and
    copies: a coarsened parent's corner ghost with no neighbour across it gets 3 g - i of the
    fine corner (for a linear function: twice its value at the coarse corner ghost); a copy
    keeps every fine value, ghosts included; 3D edge ghosts get nothing."""
    ns = (4, 4)
    dim, levels = load("2d_uniform_2x2_refined_nw_mpi1.json", ns)
    fl, cl = levels[0], levels[1]
    fine, coarse = O.Domain(dim, ns, fl), O.Domain(dim, ns, cl)
    fine.link_parents(coarse)
    fun = lambda x, y: 1 + (x * 0.3 + y)
    Xf = mesh.cell_centers(fl, dim, ns)
    Xc = mesh.cell_centers(cl, dim, ns)
    fv = fun(Xf[..., 0], Xf[..., 1])[:, None]
    cv = O.restrict(fine, coarse, fv, extrapolate=True)
    ids = {int(i): k for k, i in enumerate(cl["id"])}
    seen_corner = seen_copy = False
    for p in range(len(fl)):
        cp = ids[int(fl["parent_id"][p])]
        o = int(fl["orth_on_parent"][p])
        if o < 0:
            assert np.array_equal(cv[cp], fv[p])
            seen_copy = True
            continue
        if fl["nbr_type"][p][o] >= 0:
            continue
        iy, ix = (-1 if o & 2 else 0), (-1 if o & 1 else 0)
        assert np.isclose(cv[cp, 0, iy, ix], 2 * fun(Xc[cp, iy, ix, 0], Xc[cp, iy, ix, 1]), rtol=1e-13)
        seen_corner = True
    assert seen_corner and seen_copy
    dim3, lv3 = load("3d_refined_bnw_2x2x2_mpi1.json", (2, 2, 2))
    f3, c3 = O.Domain(3, (2, 2, 2), lv3[0]), O.Domain(3, (2, 2, 2), lv3[1])
    f3.link_parents(c3)
    v3 = O.restrict(f3, c3, random_field(f3.shape(1), 3), extrapolate=True)
    # edge ghosts (two axes outside the patch, the third inside) of the coarsened parents
    ids3 = {int(i): k for k, i in enumerate(lv3[1]["id"])}
    par = sorted({ids3[int(lv3[0]["parent_id"][p])] for p in range(len(lv3[0])) if lv3[0]["orth_on_parent"][p] >= 0})
    assert par
    for cp in par:
        w = v3[cp, 0]
        for a, b in ((0, 0), (0, -1), (-1, 0), (-1, -1)):
            assert not np.any(w[a, b, 1:-1]) and not np.any(w[a, 1:-1, b]) and not np.any(w[1:-1, a, b])


@pytest.mark.parametrize("name", ["2d_uniform_4x4_mpi1.json", "2d_uniform_2x2_refined_nw_mpi1.json"])
@pytest.mark.parametrize("nx,ny", [(2, 2), (10, 10), (2, 10)])
def test_interpolate_known_answer(name, nx, ny):
    """test/GMG/DirectInterpolator_MPI1.cpp: coarse cell value 1 + id nx ny + i + j nx lands
    on its fine cells (piecewise-constant prolongation), added to what is there."""
    ns = (nx, ny)
    dim, levels = load(name, ns)
    fl, cl = levels[0], levels[1]
    fine, coarse = O.Domain(dim, ns, fl), O.Domain(dim, ns, cl)
    fine.link_parents(coarse)
    cv = coarse.zeros()
    for p in range(len(cl)):
        for j in range(ny):
            for i in range(nx):
                cv[p, 0, j + 1, i + 1] = 1 + cl["id"][p] * nx * ny + i + j * nx
    base = random_field(fine.shape(), 2)
    fv = base.copy()
    O.interpolate(coarse, fine, cv, fv)
    exp = base.copy()
    for p in range(len(fl)):
        o = fl["orth_on_parent"][p]
        pid = fl["parent_id"][p]
        sx = 0 if o < 0 or not (o & 1) else nx
        sy = 0 if o < 0 or not (o & 2) else ny
        if o < 0:
            sx = sy = 0
        for j in range(ny):
            for i in range(nx):
                if o < 0:
                    ci, cj = i, j
                else:
                    ci, cj = (i + sx) // 2, (j + sy) // 2
                exp[p, 0, j + 1, i + 1] += 1 + pid * nx * ny + ci + cj * nx
    assert relerr(mesh.interior(fv, dim), mesh.interior(exp, dim)) <= 1e-15


def _dense_operator(d, dim, ns, kind=BIL):
    """Assemble A column by column from the oracle's apply (interior DOF only)."""
    shp = d.shape()
    inner = np.zeros(shp, dtype=bool)
    inner[(Ellipsis,) + (slice(1, -1),) * dim] = True
    pos = np.flatnonzero(inner.ravel())
    A = np.zeros((len(pos), len(pos)))
    for k, q in enumerate(pos):
        e = np.zeros(shp)
        e.ravel()[q] = 1.0
        A[:, k] = O.poisson_apply(d, e, kind=kind).ravel()[pos]
    return A, pos


def test_bicgstab_against_dense_solve():
    ns = (4, 4)
    dim, levels = load("2d_uniform_2x2_refined_nw_mpi1.json", ns)
    ods = [O.Domain(dim, ns, lv) for lv in levels]
    A, pos = _dense_operator(ods[0], dim, ns)
    assert A.shape[0] == len(levels[0]) * 16
    b = random_field(ods[0].shape(), 12)
    xd = np.linalg.solve(A, b.ravel()[pos])
    L = O.Levels(ods, filler_kind=BIL, smoother_kind=0)
    for precond in (False, True, 2):
        x, its, hist = L.bicgstab(b.copy(), precond=precond, tol=1e-13)
        assert its > 0 and hist[-1] <= 1e-13
        assert relerr(x.ravel()[pos], xd) <= 1e-9, precond


def test_bicgstab_zero_rhs_and_converged_guess():
    ns = (8, 8)
    dim, levels = load("2d_uniform_2x2_mpi1.json", ns)
    ods = [O.Domain(dim, ns, lv) for lv in levels]
    L = O.Levels(ods, filler_kind=BIL, smoother_kind=0)
    z = np.zeros(ods[0].shape())
    x0 = random_field(z.shape, 3)
    x, its, _ = L.bicgstab(z, x=x0.copy())
    # zero iterations and x untouched (its ghost cells are workspace of the operator)
    assert its == 0 and np.array_equal(mesh.interior(x, dim), mesh.interior(x0, dim))
    g, f = _sin_problem(levels[0], dim, ns)
    # (the corrected early exit, so that xs is as good as its recursive residual says)
    xs, _, _ = L.bicgstab(f.copy(), tol=1e-12, precond=2)
    _, its2, _ = L.bicgstab(f.copy(), x=xs.copy(), tol=1e-6)
    assert its2 == 0


def test_vcycle_one_patch_is_exact_inverse():
    ns = (8, 8)
    dim, levels = load("2d_uniform_2x2_mpi1.json", ns)
    d = O.Domain(dim, ns, levels[-1])
    L = O.Levels([d], filler_kind=BIL, smoother_kind=0)
    f = random_field(d.shape(), 6)
    u = L.vcycle(f.copy())
    back = O.poisson_apply(d, u.copy(), kind=BIL)
    assert relerr(mesh.interior(back, dim), mesh.interior(f, dim)) <= 1e-12


@pytest.mark.parametrize("name,kind", [("2d_uniform_2x2_refined_nw_mpi1.json", BIL),
                                       ("2d_uniform_8x8_refined_cross_mpi1.json", BIQ)])
def test_vcycle_two_levels_matches_definition(name, kind):
    """VCycle::visit (GMG/VCycle.h:46-70) with one pre, one post and one coarse sweep:
    u1 = S(f, 0); r = f - A u1; fc = R r; uc = S_c(fc, 0); u2 = u1 + P uc; u = S(f, u2)."""
    ns = (8, 8)
    dim, levels = load(name, ns)
    levels = levels[-2:]
    fd, cd = O.Domain(dim, ns, levels[0]), O.Domain(dim, ns, levels[1])
    L = O.Levels([fd, cd], filler_kind=kind, smoother_kind=0)
    f = random_field(fd.shape(), 9)
    got = L.vcycle(f.copy())
    u1 = np.zeros_like(f)
    O.dft_smooth(fd, f, u1, kind=kind)
    r = f - O.poisson_apply(fd, u1.copy(), kind=kind)
    fc = O.restrict(fd, cd, r)
    uc = np.zeros_like(fc)
    O.dft_smooth(cd, fc, uc, kind=kind)
    u2 = u1.copy()
    O.interpolate(cd, fd, uc, u2)
    O.dft_smooth(fd, f, u2, kind=kind)
    assert relerr(mesh.interior(got, dim), mesh.interior(u2, dim)) <= 1e-13


def _cycle_by_definition(ds, kind, f, typ):
    """VCycle / WCycle / FMGCycle::visit (GMG/VCycle.h:46-70, GMG/WCycle.h:50-84,
    GMG/FMGCycle.h:43-104) composed from the separately pinned smoother, operator,
    restrictor and interpolator, one sweep each (pre, mid, post, coarse)."""

    def visit(li, f, u, vonly):
        d = ds[li]
        O.dft_smooth(d, f, u, kind=kind)
        if li == len(ds) - 1:
            return
        c = ds[li + 1]
        fc = O.restrict(d, c, f - O.poisson_apply(d, u.copy(), kind=kind))
        uc = np.zeros_like(fc)
        visit(li + 1, fc, uc, False)
        O.interpolate(c, d, uc, u)
        if typ != "V" and not vonly:
            O.dft_smooth(d, f, u, kind=kind)  # mid sweep
            fc = O.restrict(d, c, f - O.poisson_apply(d, u.copy(), kind=kind))
            visit(li + 1, fc, uc, typ == "F")  # uc keeps the first visit's result
            O.interpolate(c, d, uc, u)
        O.dft_smooth(d, f, u, kind=kind)  # post sweep

    u = np.zeros_like(f)
    visit(0, f, u, False)
    return u


@pytest.mark.parametrize("name,kind", [("2d_uniform_2x2_refined_nw_mpi1.json", BIL),
                                       ("2d_uniform_8x8_refined_cross_mpi1.json", BIQ)])
def test_w_and_f_cycles_match_definition(name, kind):
    ns = (8, 8)
    dim, levels = load(name, ns)
    levels = levels[-3:]
    ds = [O.Domain(dim, ns, lv) for lv in levels]
    L = O.Levels(ds, filler_kind=kind, smoother_kind=0)
    f = random_field(ds[0].shape(), 19)
    out = {}
    for typ in ("V", "W", "F"):
        got = L.cycle(f.copy(), typ)
        want = _cycle_by_definition(ds, kind, f, typ)
        assert relerr(mesh.interior(got, dim), mesh.interior(want, dim)) <= 1e-13, typ
        out[typ] = mesh.interior(got, dim)
    assert np.array_equal(out["V"], mesh.interior(L.vcycle(f.copy()), dim))
    # three levels: the W- and F-cycles differ below the second level's second visit
    assert relerr(out["W"], out["F"]) > 1e-8
    # two levels: the F-cycle's V-visit of the coarsest level is its W-visit
    L2 = O.Levels(ds[1:], filler_kind=kind, smoother_kind=0)
    f2 = random_field(ds[1].shape(), 20)
    assert np.array_equal(L2.cycle(f2.copy(), "W"), L2.cycle(f2.copy(), "F"))
    # every cycle is linear in f for fixed sweeps
    g = random_field(f.shape, 21)
    for typ in ("W", "F"):
        lhs = mesh.interior(L.cycle(2 * f + 3 * g, typ), dim)
        rhs = 2 * out[typ] + 3 * mesh.interior(L.cycle(g.copy(), typ), dim)
        assert relerr(lhs, rhs) <= 1e-12, typ


def test_cg_against_dense_solve():
    """Iterative::CG::solve (Iterative/CG.h:126-194) reaches the dense direct solve, plain and
    right-preconditioned by a V- or W-cycle; on 64 unknowns plain CG needs at most about 64
    iterations.  A zero right-hand side returns at once with x set to zero (the reference
    zeroes x before it checks ||b||)."""
    ns = (4, 4)
    dim, levels = load("2d_uniform_2x2_mpi1.json", ns)
    ods = [O.Domain(dim, ns, lv) for lv in levels]
    A, pos = _dense_operator(ods[0], dim, ns)
    assert np.allclose(A, A.T)  # the uniform-mesh operator is symmetric (CG's premise)
    b = random_field(ods[0].shape(), 22)
    xd = np.linalg.solve(A, b.ravel()[pos])
    L = O.Levels(ods, filler_kind=BIL, smoother_kind=0)
    x, its, hist = L.cg(b.copy(), precond=False, tol=1e-13)
    assert 0 < its <= 80 and hist[-1] <= 1e-13
    assert relerr(x.ravel()[pos], xd) <= 1e-9
    for typ in ("V", "W"):
        x, its_p, hist = L.cg(b.copy(), precond=True, cycle_type=typ, tol=1e-13)
        assert 0 < its_p < its and hist[-1] <= 1e-13, typ
        assert relerr(x.ravel()[pos], xd) <= 1e-9, typ
    # initial guess: CG solves for the correction and adds the guess back
    x0 = random_field(b.shape, 23)
    x, _, _ = L.cg(b.copy(), x=x0.copy(), precond=True, tol=1e-13)
    assert relerr(x.ravel()[pos], xd) <= 1e-9
    z = np.zeros_like(b)
    x, its, _ = L.cg(z, x=x0.copy())
    assert its == 0 and not mesh.interior(x, dim).any()


@pytest.mark.parametrize("name", MESH_2D + MESH_3D)
def test_enumerated_calls_match_tables(name):
    """test/MPIGhostFiller_MPI1.cpp: the local face calls are one per (patch, side) with a
    normal or coarse neighbour and one per fine neighbour (2^(D-1) per side), ordered by
    patch, then face value, then orthant."""
    dim = 2 if name.startswith("2d") else 3
    ns = (4,) * dim
    _, levels = load(name, ns)
    lv = levels[0]
    d = O.Domain(dim, ns, lv)
    calls = d.calls(FACES)
    sides = calls[calls[:, 0] == dim - 1]
    expect = []
    for p in range(len(lv)):
        for s in range(2 * dim):
            t = int(lv["nbr_type"][p][mesh.side_slot(dim, s)])
            if t < 0:
                continue
            for o in range(1 << (dim - 1) if t == 2 else 1):
                expect.append((p, s, t))
    got = [(int(c[1]), int(c[2]), int(c[3])) for c in sides]
    assert got == expect
    # neighbour local indexes are those of the tables
    for c in sides:
        p, s, t, o, nl = int(c[1]), int(c[2]), int(c[3]), int(c[4]), int(c[5])
        k = o if t == 2 else 0
        assert d.nbr_local(p, mesh.side_slot(dim, s), k) == nl


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_operator_is_symmetric_negative_definite(dim):
    ns = (4,) * dim
    name = "2d_uniform_2x2_mpi1.json" if dim == 2 else "3d_uniform_2x2x2_mpi1.json"
    _, levels = load(name, ns)
    d = O.Domain(dim, ns, levels[0])
    shp = d.shape()
    inner = np.zeros(shp, dtype=bool)
    inner[(Ellipsis,) + (slice(1, -1),) * dim] = True
    pos = np.flatnonzero(inner.ravel())
    A = np.zeros((len(pos), len(pos)))
    for k, q in enumerate(pos):
        e = np.zeros(shp)
        e.ravel()[q] = 1.0
        A[:, k] = O.poisson_apply(d, e, kind=BIL if dim == 2 else TRI).ravel()[pos]
    assert np.allclose(A, A.T, atol=1e-9 * np.abs(A).max())
    assert np.linalg.eigvalsh(A).max() < 0


@pytest.mark.parametrize("n", [4, 9, 16])
def test_dct_and_dst4_matrices_match_scipy(n):
    """The Neumann-side transforms of the DFT plan (kinds 0-2, 5: DCT-II, DCT-III, DCT-IV,
    DST-IV) against scipy's unnormalised transforms (= 2 x the plan's matrices)."""
    e = np.eye(n)
    for kind, typ, fn in [(0, 2, scipy.fft.dct), (1, 3, scipy.fft.dct), (2, 4, scipy.fft.dct), (5, 4, scipy.fft.dst)]:
        assert np.allclose(O.dft_matrix(kind, n), fn(e, type=typ, axis=0) / 2, atol=1e-13), kind


def test_dft_solve_is_exact_inverse_single_patch_3d():
    ns = (4, 6, 8)
    dim, levels = load("3d_uniform_2x2x2_mpi1.json", ns)
    lv = levels[-1]
    assert len(lv) == 1
    d = O.Domain(dim, ns, lv)
    f = random_field(d.shape(), 18)
    u = np.zeros_like(f)
    O.dft_smooth(d, f, u, kind=TRI)
    back = O.poisson_apply(d, u.copy(), kind=TRI)
    assert relerr(mesh.interior(back, dim), mesh.interior(f, dim)) <= 1e-12


def test_varpoisson_second_order():
    """test/VarPoisson/StarPatchOperator_MPI1.cpp: beta = 1 + x y, u = sin(pi x) sin(2 pi y),
    f = div(beta grad u), uniform mesh, sizes 32 -> 64."""
    errs = []
    for n in (32, 64):
        ns = (n, n)
        dim, levels = load("2d_uniform_2x2_mpi1.json", ns)
        lv = levels[0]
        d = O.Domain(dim, ns, lv)
        X = mesh.cell_centers(lv, dim, ns)
        x, y = X[..., 0], X[..., 1]
        beta = (1 + x * y)[:, None]
        O.fill_ghost(d, beta, kind=BIL)
        u = (np.sin(np.pi * x) * np.sin(2 * np.pi * y))[:, None]
        ux = np.pi * np.cos(np.pi * x) * np.sin(2 * np.pi * y)
        uy = 2 * np.pi * np.sin(np.pi * x) * np.cos(2 * np.pi * y)
        lap = -5 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(2 * np.pi * y)
        fe = ((1 + x * y) * lap + y * ux + x * uy)[:, None]
        f = O.varpoisson_apply(d, beta, u.copy(), kind=BIL)
        e = mesh.interior(f - fe, dim)
        errs.append(np.sqrt((e * e).sum() / (mesh.interior(fe, dim) ** 2).sum()))
    assert np.log(errs[0] / errs[1]) / np.log(2) > 1.8


def _box(level, dim, ns):
    lo = level["starts"][:, :dim].min(axis=0)
    hi = (level["starts"][:, :dim] + level["spacings"][:, :dim] * np.array(ns[:dim])).max(axis=0)
    return lo, hi


@pytest.mark.parametrize("name", MESH_2D + MESH_3D)
def test_integrate_midpoint_rule(name):
    """DomainTools::Integrate (DomainTools.h:513-536) is the midpoint rule over each cell,
    with each patch's own cell volume.  Every level of a fixture mesh tiles its bounding
    box, so the integral of 1 is the box volume; products of affine functions of each
    coordinate are integrated exactly; x0^2 is short by h0^2 / 12 times each patch's
    volume (the midpoint rule's known error, which differs between refinement levels);
    ghost values and the component split do not enter."""
    dim = 2 if name.startswith("2d") else 3
    ns = (4, 6) if dim == 2 else (4, 4, 2)
    dim, levels = load(name, ns)
    coef = (1.0, 2.0, -3.0)
    for lv in levels:
        dom = O.Domain(dim, ns, lv)
        lo, hi = _box(lv, dim, ns)
        X = mesh.cell_centers(lv, dim, ns)
        vol = np.prod(hi - lo)
        one = np.ones(mesh.padded_shape(len(lv), dim, ns))
        assert abs(O.integrate(dom, one) - vol) <= 1e-13 * vol
        f = np.ones(X.shape[:-1])
        exact = 1.0
        for a in range(dim):
            f = f * (1.0 + coef[a] * X[..., a])
            exact *= (hi[a] - lo[a]) + coef[a] * (hi[a] ** 2 - lo[a] ** 2) / 2.0
        fv = f[:, None].copy()
        garbage = random_field(fv.shape, 7) * 1e6
        fg = garbage.copy()
        mesh.interior(fg, dim)[...] = mesh.interior(fv, dim)
        assert abs(O.integrate(dom, fg) - exact) <= 1e-12 * abs(exact)
        # x0^2: exact minus h0^2/12 per unit volume of each patch
        q = (X[..., 0] ** 2)[:, None].copy()
        pvol = np.prod(lv["spacings"][:, :dim] * np.array(ns[:dim]), axis=1)
        exact_q = (hi[0] ** 3 - lo[0] ** 3) / 3.0 * np.prod((hi - lo)[1:])
        midpoint_q = exact_q - np.sum(lv["spacings"][:, 0] ** 2 / 12.0 * pvol)
        assert abs(O.integrate(dom, q) - midpoint_q) <= 1e-12 * abs(exact_q)
        if len(set(lv["refine_level"])) > 1:
            assert abs(midpoint_q - exact_q) > 1e-6 * abs(exact_q)
        # two components: the sum of the two integrals
        two = np.concatenate([fg, q], axis=1)
        assert abs(O.integrate(dom, two) - (exact + midpoint_q)) <= 1e-12 * (abs(exact) + abs(midpoint_q))


def _lap1d(n, h, lo_neumann, hi_neumann):
    """Second difference on n cell centres; a Dirichlet end reflects with ghost = -u,
    a Neumann end with ghost = u (the star operator's physical-boundary ghosts)."""
    t = np.diag(-2.0 * np.ones(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    t[0, 0] += 1.0 if lo_neumann else -1.0
    t[-1, -1] += 1.0 if hi_neumann else -1.0
    return t / (h * h)


def _dense_laplacian(ns, hs, mask):
    """Dense n x n operator on one patch, x fastest; side s is Neumann when bit s of mask
    is set (sides W, E, S, N, B, T: axis s // 2, lower when s is even)."""
    dim = len(ns)
    A = np.zeros((int(np.prod(ns)),) * 2)
    for a in range(dim):
        t = _lap1d(ns[a], hs[a], (mask >> (2 * a)) & 1, (mask >> (2 * a + 1)) & 1)
        term = np.ones((1, 1))
        for b in reversed(range(dim)):  # slowest axis first in the Kronecker product
            term = np.kron(term, t if b == a else np.eye(ns[b]))
        A += term
    return A


@pytest.mark.parametrize("dim,mask", [(2, 0b0001), (2, 0b0010), (2, 0b0101), (2, 0b0011), (2, 0b1100),
                                      (2, 0b1111), (3, 0b000011), (3, 0b100101), (3, 0b111111)])
def test_dft_neumann_solve_is_exact_inverse(dim, mask):
    """The DFT patch solve with Neumann sides (DCT-II/III, DCT-IV / DST-IV transforms and
    their eigenvalues, Poisson/DFTPatchSolver.h:272-361, 424-440) is the exact inverse of
    the one-patch operator assembled here independently (Dirichlet ghost = -u, Neumann
    ghost = u), anisotropic spacing.  All sides Neumann: the operator is singular; for a
    zero-mean f the solve gives the zero-mean solution (the constant mode set to 0)."""
    if dim == 2:
        ns, name = (12, 8), "2d_uniform_2x2_mpi1.json"
    else:
        ns, name = (4, 6, 8), "3d_uniform_2x2x2_mpi1.json"  # TriLinear: even n
    _, levels = load(name, ns)
    lv = levels[-1]
    assert len(lv) == 1
    d = O.Domain(dim, ns, lv)
    hs = [lv["spacings"][0][a] for a in range(dim)]
    A = _dense_laplacian(ns, hs, mask)
    f = random_field(d.shape(), 30 + mask)
    fi = mesh.interior(f, dim)[0, 0].reshape(-1)
    singular = mask == (1 << (2 * dim)) - 1
    if singular:
        fi -= fi.mean()
        mesh.interior(f, dim)[0, 0] = fi.reshape(mesh.interior(f, dim)[0, 0].shape)
    u = np.zeros_like(f)
    O.dft_smooth(d, f, u, kind=BIL if dim == 2 else TRI, neumann_mask=mask)
    ui = mesh.interior(u, dim)[0, 0].reshape(-1)
    if singular:
        expect = np.linalg.lstsq(A, fi, rcond=None)[0]
        expect -= expect.mean()
        assert abs(ui.mean()) <= 1e-12 * np.abs(ui).max()
    else:
        expect = np.linalg.solve(A, fi)
    assert relerr(ui, expect) <= 1e-11
    assert relerr(A @ ui, fi) <= 1e-11


@pytest.mark.parametrize("name", ["2d_uniform_2x2_refined_nw_mpi1.json", "2d_uniform_8x8_refined_cross_mpi1.json"])
@pytest.mark.parametrize("neumann", [False, True])
def test_patch_bicgstab_smoother_matches_dft_smoother(name, neumann):
    """Iterative::PatchSolver (inner BiCGStab per patch, PatchSolver.h:153-168) and the DFT
    patch solver solve the same local problems (interface terms from the neighbours'
    ghosts, internal boundary conditions): with a tight inner tolerance one Poisson sweep
    of each agrees on refined meshes, Dirichlet or Neumann physical sides; the DFT sweep
    is pinned above (exact inverse, second-order convergence)."""
    ns = (8, 6)
    dim, levels = load(name, ns)
    lv = levels[0]
    d = O.Domain(dim, ns, lv)
    f = random_field(d.shape(), 51)
    u0 = random_field(d.shape(), 52)
    ud = u0.copy()
    O.dft_smooth(d, f, ud, kind=BIQ, neumann_mask=0b1111 if neumann else 0)
    ub = u0.copy()
    its = O.bicgstab_smooth(d, f, ub, kind=BIQ, neumann=neumann, max_it=2000, tol=1e-14)
    assert (its > 0).all()
    assert relerr(mesh.interior(ub, dim), mesh.interior(ud, dim)) <= 1e-9
    if not neumann:
        # VarPoisson with beta = 1 (ghosts included) is the same local problem
        uv = u0.copy()
        ones = np.ones(d.shape())
        O.bicgstab_smooth(d, f, uv, kind=BIQ, op_kind=1, coeffs=ones, max_it=2000, tol=1e-14)
        assert relerr(mesh.interior(uv, dim), mesh.interior(ud, dim)) <= 1e-9


def test_vector_reductions():
    """Vector::dot / twoNorm / infNorm (Vector.h:391-611) over interior cells only: ghost
    values do not enter; twoNorm^2 = dot(a, a); dot is symmetric and bilinear; infNorm is
    max(0, max v), so an all-negative vector has infNorm 0 (Vector.h:568-579)."""
    ns = (6, 4)
    dim, levels = load("2d_uniform_2x2_refined_nw_mpi1.json", ns)
    d = O.Domain(dim, ns, levels[0])
    a = random_field(d.shape(2), 61)
    b = random_field(d.shape(2), 62)
    ga, gb = a.copy(), b.copy()
    ga[...] = 1e9
    gb[...] = -1e9
    mesh.interior(ga, dim)[...] = mesh.interior(a, dim)
    mesh.interior(gb, dim)[...] = mesh.interior(b, dim)
    assert O.dot(d, ga, gb) == pytest.approx(O.dot(d, a, b), rel=1e-14)
    assert O.dot(d, a, b) == pytest.approx(O.dot(d, b, a), rel=1e-14)
    assert O.dot(d, a, 2 * a + b) == pytest.approx(2 * O.dot(d, a, a) + O.dot(d, a, b), rel=1e-12)
    assert O.two_norm(d, ga) ** 2 == pytest.approx(O.dot(d, a, a), rel=1e-14)
    ones = np.ones(d.shape(2))
    assert O.dot(d, ones, ones) == 2 * len(levels[0]) * ns[0] * ns[1]
    neg = -np.abs(a) - 1.0
    neg_g = neg.copy()
    neg_g[...] = 5.0
    mesh.interior(neg_g, dim)[...] = mesh.interior(neg, dim)
    assert O.inf_norm(d, neg_g) == 0.0
    mesh.interior(neg_g, dim)[1, 1, 2, 3] = 0.25
    assert O.inf_norm(d, neg_g) == 0.25


# ---- boundary data on the right-hand side (addDrichletBCToRHS / addNeumannBCToRHS) --------
MESHES_BC = ["2d_uniform_2x2_mpi1.json", "2d_uniform_4x4_mpi1.json", "2d_uniform_2x2_refined_nw_mpi1.json",
             "2d_uniform_8x8_refined_cross_mpi1.json"]


def _bc_samples(lv, dim, ns, func):
    from thunderegg_amd.domain_tools import boundary_points

    pts = boundary_points(lv, dim, ns)
    return np.ascontiguousarray(func(*[pts[:, a] for a in range(dim)]), dtype=np.float64), pts


def _side_axis(lv, dim, ns):
    """the axis of the side of every boundary-data entry"""
    per = np.concatenate([np.full(int(np.prod([ns[b] for b in range(dim) if b != (s >> 1)])), s >> 1)
                          for s in range(2 * dim)])
    return np.tile(per, len(lv))


@pytest.mark.parametrize("name", MESHES_BC)
@pytest.mark.parametrize("ns", [(2, 2), (2, 10), (10, 2), (10, 10)])
@pytest.mark.parametrize("neumann", [False, True])
def test_bc_rhs_linear_exactness_2d(name, ns, neumann):
    """test/Poisson/StarPatchOperator_MPI1.cpp, 'apply on linear lhs constant coeff' (and its
    Neumann twin): for g = 1 + 5x + 13y the operator with the physical ghosts -g_b (+g_b)
    gives exactly the boundary terms, 0 + addDrichletBCToRHS(g) (addNeumannBCToRHS(g,
    grad g)), at every cell -- interfaces, refined meshes and corners included."""
    dim, levels = load(name, ns)
    lv = levels[0]
    d = O.Domain(dim, ns, lv)
    Xg = mesh.cell_centers(lv, dim, ns, with_ghost=True)
    g = (1 + 5 * Xg[..., 0] + 13 * Xg[..., 1])[:, None].copy()
    f = O.poisson_apply(d, g.copy(), kind=BIL, neumann=neumann)
    fexp = np.zeros_like(g)
    if neumann:
        axis = _side_axis(lv, dim, ns)
        O.add_bc(d, fexp, np.where(axis == 0, 5.0, 13.0), neumann=True)
    else:
        gs, _ = _bc_samples(lv, dim, ns, lambda x, y: 1 + 5 * x + 13 * y)
        O.add_bc(d, fexp, gs)
    a, b = mesh.interior(f, dim), mesh.interior(fexp, dim)
    assert np.abs(b).max() > 0
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()


@pytest.mark.parametrize("name", MESH_3D)
@pytest.mark.parametrize("neumann", [False, True])
def test_bc_rhs_linear_exactness_3d(name, neumann):
    """The same identity in 3D (TriLinear ghosts, 7-point operator), g = 1 + 5x + 13y + 7z."""
    ns = (4, 6, 8)
    dim, levels = load(name, ns)
    lv = levels[0]
    d = O.Domain(dim, ns, lv)
    Xg = mesh.cell_centers(lv, dim, ns, with_ghost=True)
    g = (1 + 5 * Xg[..., 0] + 13 * Xg[..., 1] + 7 * Xg[..., 2])[:, None].copy()
    f = O.poisson_apply(d, g.copy(), kind=TRI, neumann=neumann)
    fexp = np.zeros_like(g)
    if neumann:
        axis = _side_axis(lv, dim, ns)
        O.add_bc(d, fexp, np.choose(axis, [5.0, 13.0, 7.0]), neumann=True)
    else:
        gs, _ = _bc_samples(lv, dim, ns, lambda x, y, z: 1 + 5 * x + 13 * y + 7 * z)
        O.add_bc(d, fexp, gs)
    a, b = mesh.interior(f, dim), mesh.interior(fexp, dim)
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()


def test_bc_rhs_varpoisson_constant_coefficient():
    """The VarPoisson form -2 g c / h^2 with c = 1 is the Poisson form, and it scales with c."""
    ns = (8, 8)
    dim, levels = load("2d_uniform_2x2_refined_nw_mpi1.json", ns)
    lv = levels[0]
    d = O.Domain(dim, ns, lv)
    gs, _ = _bc_samples(lv, dim, ns, lambda x, y: np.sin(3 * x) + y * y)
    f0 = random_field(mesh.padded_shape(len(lv), dim, ns), 5)
    a, b, c = f0.copy(), f0.copy(), f0.copy()
    O.add_bc(d, a, gs)
    O.add_bc(d, b, gs, c=np.ones_like(gs))
    O.add_bc(d, c, gs, c=np.full_like(gs, 2.5))
    assert np.allclose(a, b, rtol=1e-15, atol=0)
    assert np.allclose(c - f0, 2.5 * (a - f0), rtol=1e-13, atol=1e-12)
    # interior cells change (on physical sides), ghost cells do not
    changed = np.abs(a - f0) > 0
    inner = np.zeros(a.shape, dtype=bool)
    mesh.interior(inner, dim)[...] = True
    assert changed.any() and not changed[~inner].any()


@pytest.mark.parametrize("name", MESHES_BC)
@pytest.mark.parametrize("neumann", [False, True])
def test_bc_rhs_second_order(name, neumann):
    """test/Poisson/StarPatchOperator_MPI1.cpp, 'gets 2nd order convergence' (with Neumann
    BC, BiQuadratic ghosts): A g against f + the boundary terms for g = sin(pi y) cos(2 pi x),
    f = -5 pi^2 g, at n = 32 and 64 -- order > 1.8."""
    errs = []
    for n in (32, 64):
        ns = (n, n)
        dim, levels = load(name, ns)
        lv = levels[0]
        d = O.Domain(dim, ns, lv)
        X = mesh.cell_centers(lv, dim, ns)
        x, y = X[..., 0], X[..., 1]
        g = (np.sin(np.pi * y) * np.cos(2 * np.pi * x))[:, None]
        fexp = (-5 * np.pi ** 2 * np.sin(np.pi * y) * np.cos(2 * np.pi * x))[:, None].copy()
        if neumann:
            axis = _side_axis(lv, dim, ns)
            _, pts = _bc_samples(lv, dim, ns, lambda x, y: x)
            px, py = pts[:, 0], pts[:, 1]
            dg = np.where(axis == 0, -2 * np.pi * np.sin(np.pi * py) * np.sin(2 * np.pi * px),
                          np.pi * np.cos(np.pi * py) * np.cos(2 * np.pi * px))
            O.add_bc(d, fexp, dg, neumann=True)
            f = O.poisson_apply(d, g.copy(), kind=BIQ, neumann=True)
        else:
            gs, _ = _bc_samples(lv, dim, ns, lambda x, y: np.sin(np.pi * y) * np.cos(2 * np.pi * x))
            O.add_bc(d, fexp, gs)
            f = O.poisson_apply(d, g.copy(), kind=BIL)
        e = mesh.interior(f - fexp, dim)
        errs.append(np.sqrt((e * e).sum() / (mesh.interior(fexp, dim) ** 2).sum()))
    assert np.log(errs[0] / errs[1]) / np.log(2) > 1.8
