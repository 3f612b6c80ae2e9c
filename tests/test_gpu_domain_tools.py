"""DomainTools (reference DomainTools.h:36-537) on the device.

Integrate (te_vec_integrate) against the oracle's restatement on the fixture meshes
(random interiors, garbage ghosts, one and two components, FP64 and FP32 storage) and
against the midpoint rule's closed forms on a large refined forest; SetValues /
SetValuesWithGhost restating test/DomainTools_MPI1.cpp (one patch, the reference's
ny / start / spacing sweeps, one and two components, the error when more functions than
components are given); the C++ header's device functions (tests/cpp/domain_tools_test.cpp);
a manufactured-solution solve built from them (second-order convergence).
"""
import itertools
import subprocess

import numpy as np
import pytest

from helpers import MESH_2D, MESH_3D, load, random_field
from oracle import oracle as O
from test_host import _build_api_test
from thunderegg_amd import domain_tools as DT
from thunderegg_amd import mesh
from thunderegg_amd import runtime as R
from thunderegg_amd import solver

pytestmark = pytest.mark.gpu


def _print_case(*a):
    print("case:", *a, flush=True)


@pytest.mark.parametrize("name", MESH_2D + MESH_3D)
@pytest.mark.parametrize("ncomp", [1, 2])
def test_integrate_matches_oracle(name, ncomp):
    _print_case("integrate", name, ncomp)
    dim = 2 if name.startswith("2d") else 3
    ns = (16, 12) if dim == 2 else (8, 6, 4)
    dim, levels = load(name, ns)
    for k, lv in enumerate(levels):
        dom, odom = R.Domain(dim, ns, lv), O.Domain(dim, ns, lv)
        u = random_field(mesh.padded_shape(len(lv), dim, ns, ncomp), 100 + k)
        g = u.copy()
        g[...] = 1e6  # ghosts must not enter
        mesh.interior(g, dim)[...] = mesh.interior(u, dim)
        scale = O.integrate(odom, np.abs(g))
        ref = O.integrate(odom, g)
        got = dom.integrate(R.DeviceArray.from_host(g))
        assert abs(got - ref) <= 1e-14 * scale, (k, got, ref)
        # FP32 storage: the device accumulates in FP64, as the oracle on the rounded values
        g32 = g.astype(np.float32)
        ref32 = O.integrate(odom, g32.astype(np.float64))
        got32 = DT.integrate(dom, R.DeviceArray.from_host(g32))
        assert abs(got32 - ref32) <= 1e-14 * scale, (k, got32, ref32)


def test_integrate_closed_form_large_forest():
    """(1 + x)(3 - 2y) and x^2 over the unit square on a 2:1 refined forest of 16 x 16
    patches (about 1.4 M cells): the midpoint rule is exact for the first and short by
    h^2 / 12 per unit volume of each patch for the second."""
    _print_case("integrate-forest")
    ns = (16, 16)
    forest = solver.Forest(2, 6)
    lo, size = forest.leaf_boxes()
    forest.refine(((lo[:, 0] - 0.3) ** 2 + (lo[:, 1] - 0.6) ** 2 < 0.05).astype(np.uint8))
    forest.balance()
    lv = forest.build_domains(ns)[0]
    assert len(set(lv["refine_level"])) == 2
    dom = R.Domain(2, ns, lv)
    u = dom.vector()
    DT.set_values(dom, u, lambda c: (1 + c[0]) * (3 - 2 * c[1]))
    assert abs(DT.integrate(dom, u) - 3.0) <= 1e-12
    DT.set_values(dom, u, lambda c: c[0] ** 2)
    pvol = np.prod(lv["spacings"][:, :2] * 16, axis=1)
    expect = 1.0 / 3.0 - np.sum(lv["spacings"][:, 0] ** 2 / 12.0 * pvol)
    assert abs(DT.integrate(dom, u) - expect) <= 1e-12


def _one_patch(nx, ny, sx, sy, hx, hy):
    p = mesh.empty_patches(1)
    p["id"] = 0
    p["rank"] = 0
    p["refine_level"] = 0
    p["starts"][0, :2] = (sx, sy)
    p["spacings"][0] = (hx, hy, 1.0)
    return R.Domain(2, (nx, ny), p)


@pytest.mark.parametrize("ny", [1, 2, 10, 13])
def test_set_values_2d(ny):
    """DomainTools::setValues / setValuesWithGhost 2D f=x+y, g=x*y (DomainTools_MPI1.cpp)."""
    _print_case("set-values", ny)
    nx = 3
    f = lambda c: c[0] + c[1]  # noqa: E731
    g = lambda c: c[0] * c[1]  # noqa: E731
    for sy, hy in itertools.product((0.0, -1.0, 1.0, 0.23, -0.23, 3.0, -3.0), (0.01, 1.0, 3.14)):
        dom = _one_patch(nx, ny, 0.0, sy, 0.1, hy)
        x = (0.5 + np.arange(-1, nx + 1)) * 0.1
        y = sy + (0.5 + np.arange(-1, ny + 1)) * hy
        X, Y = np.meshgrid(x, y)
        inside = np.zeros_like(X, dtype=bool)
        inside[1:-1, 1:-1] = True
        v = dom.vector(2)
        DT.set_values(dom, v, f, g)
        h = v.get()[0]
        assert np.allclose(h[0], np.where(inside, X + Y, 0.0), rtol=1e-13, atol=1e-13)
        assert np.allclose(h[1], np.where(inside, X * Y, 0.0), rtol=1e-13, atol=1e-13)
        vg = dom.vector(2)
        DT.DomainTools.SetValuesWithGhost(dom, vg, f, g)
        h = vg.get()[0]
        assert np.allclose(h[0], X + Y, rtol=1e-13, atol=1e-13)
        assert np.allclose(h[1], X * Y, rtol=1e-13, atol=1e-13)
        # one component, two functions: component 0 set, then RuntimeError
        one = dom.vector(1)
        with pytest.raises(RuntimeError, match="Invalid component"):
            DT.set_values(dom, one, f, g)
        assert np.allclose(one.get()[0, 0], np.where(inside, X + Y, 0.0), rtol=1e-13, atol=1e-13)
        with pytest.raises(RuntimeError, match="Invalid component"):
            DT.set_values_with_ghost(dom, one, f, g)
        # the per-cell form gives the same values
        w = dom.vector(1)
        DT.set_values(dom, w, lambda c: c[0] * c[1] if c[0] > 0.1 else -1.0, vectorized=False)
        e = np.where(inside, np.where(X > 0.1, X * Y, -1.0), 0.0)
        assert np.allclose(w.get()[0, 0], e, rtol=1e-13, atol=1e-13)


def test_set_values_3d_on_mesh():
    """SetValues on a refined 3D fixture mesh: every interior cell holds f at its centre
    (mesh.cell_centers), ghosts untouched; component_index selects the component."""
    _print_case("set-values-3d")
    ns = (4, 6, 2)
    dim, levels = load("3d_mid_refine_4x4x4_mpi1.json", ns)
    dom = R.Domain(dim, ns, levels[0])
    v = dom.vector(2)
    v.upload(np.full(dom.shape(2), 5.0))
    DT.set_values(dom, v, lambda c: c[0] - 2 * c[1] + 3 * c[2], component_index=1)
    h = v.get()
    X = mesh.cell_centers(levels[0], dim, ns, with_ghost=False)
    assert np.array_equal(h[:, 0], np.full_like(h[:, 0], 5.0))
    got = mesh.interior(h[:, 1:2], dim)[:, 0]
    assert np.allclose(got, X[..., 0] - 2 * X[..., 1] + 3 * X[..., 2], rtol=1e-13, atol=1e-13)
    ghost = h[:, 1].copy()
    mesh.interior(ghost[:, None], dim)[...] = 5.0
    assert np.array_equal(ghost, np.full_like(ghost, 5.0))


def test_cpp_domain_tools_on_device(tmp_path):
    exe = _build_api_test(tmp_path, "domain_tools_test")
    _print_case("cpp domain_tools_test device")
    r = subprocess.run([exe, "device"], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "domain_tools_test: ok" in r.stdout


@pytest.mark.parametrize("prec", [R.F64, R.F32])
def test_manufactured_solution_second_order(prec):
    """The way the reference's Poisson examples use DomainTools: f = -2 pi^2 sin(pi x)
    sin(pi y) set at the cell centres, the GMG-preconditioned BiCGStab solve, the error
    against u = sin(pi x) sin(pi y) (zero Dirichlet boundary) set the same way.  On uniform
    meshes of 16 x 16 patches the error falls by about 4 per halving of h (second order),
    and the integral of the solution tends to 4 / pi^2."""
    _print_case("manufactured", prec)
    ns = (16, 16)
    exact = lambda c: np.sin(np.pi * c[0]) * np.sin(np.pi * c[1])  # noqa: E731
    errs, ints = [], []
    for base in (2, 3, 4):
        levels = solver.uniform_levels(2, base, ns)
        g = solver.GMG(2, ns, levels, filler_kind=R.BIQUADRATIC, smoother=0, precision=prec)
        dom = g.domains[0]
        f = dom.vector()
        DT.set_values(dom, f, lambda c: -2 * np.pi ** 2 * exact(c))
        x = dom.vector()
        its, _ = g.solve_device(f, x, tol=1e-11, precond=2)
        assert 0 < its <= 12
        ue = dom.vector()
        DT.set_values(dom, ue, exact)
        errs.append(np.abs(mesh.interior(x.get(), 2) - mesh.interior(ue.get(), 2)).max())
        ints.append(DT.integrate(dom, x))
    print("errors", errs, "integrals", ints)
    for a, b in zip(errs, errs[1:]):
        assert 3.5 < a / b < 4.5, errs
    for a, b in zip(ints, ints[1:]):
        assert abs(b - 4 / np.pi ** 2) < abs(a - 4 / np.pi ** 2)
    assert abs(ints[-1] - 4 / np.pi ** 2) < 1e-4
