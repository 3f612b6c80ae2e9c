"""``DomainTools`` (reference ``DomainTools.h:36-537``): coordinates of a patch's cells and
interfaces, vectors set from functions of the coordinates, and the integral of a vector
over a domain.

The coordinate helpers take a ``PatchInfo`` (the fields of ``PatchInfo<D>`` that locate
cells).  ``set_values`` / ``set_values_with_ghost`` evaluate host functions at the cell
centres of every patch and upload the result (setup work, as in the reference); by
default a function is called once per component with arrays: ``coord[a]`` holds axis
``a``'s coordinate of every cell, so ``lambda c: c[0] + c[1]`` or
``lambda c: np.sin(c[0])`` work unchanged; ``vectorized=False`` calls it once per cell
with a tuple.  ``integrate`` is a device reduction (``te_vec_integrate``).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import mesh
from . import runtime as R

# Side<D> values (Face.h:41-584): axis = value // 2, lower side = even value
WEST, EAST, SOUTH, NORTH, BOTTOM, TOP = range(6)


@dataclass
class PatchInfo:
    """The cell-locating fields of ``PatchInfo<D>`` (``PatchInfo.h:50-140``); the default
    is one cell of unit size at the origin, as the reference's default constructor."""

    dim: int = 2
    ns: tuple = None
    starts: tuple = None
    spacings: tuple = None
    num_ghost_cells: int = 1
    id: int = 0
    local_index: int = 0
    refine_level: int = -1

    def __post_init__(self):
        d = self.dim
        self.ns = tuple(int(v) for v in (self.ns if self.ns is not None else (1,) * d))
        self.starts = tuple(float(v) for v in (self.starts if self.starts is not None else (0.0,) * d))
        self.spacings = tuple(float(v) for v in (self.spacings if self.spacings is not None else (1.0,) * d))
        if not (len(self.ns) == len(self.starts) == len(self.spacings) == d):
            raise ValueError("ns, starts and spacings need one entry per axis")


def patch_info(domain: R.Domain, local_index: int) -> PatchInfo:
    """``domain.getPatchInfoVector()[local_index]``'s cell-locating fields."""
    rec = domain.patches[local_index]
    d = domain.dim
    return PatchInfo(dim=d, ns=domain.ns, starts=tuple(rec["starts"][:d]), spacings=tuple(rec["spacings"][:d]),
                     num_ghost_cells=1, id=int(rec["id"]), local_index=local_index,
                     refine_level=int(rec["refine_level"]))


def _on_axis(p: PatchInfo, a: int, i: int) -> float:
    # DomainTools.h:139-147: -1 is the lower interface, ns[a] the upper, else a cell centre
    if i == -1:
        return p.starts[a]
    if i == p.ns[a]:
        return p.starts[a] + p.spacings[a] * p.ns[a]
    return p.starts[a] + p.spacings[a] / 2.0 + p.spacings[a] * i


def get_real_coord(pinfo: PatchInfo, coord) -> tuple:
    """``GetRealCoord`` (``DomainTools.h:134-149``)."""
    return tuple(_on_axis(pinfo, a, int(coord[a])) for a in range(pinfo.dim))


def get_real_coord_ghost(pinfo: PatchInfo, coord) -> tuple:
    """``GetRealCoordGhost`` (``DomainTools.h:158-168``): cell centres, ghosts included."""
    return tuple(pinfo.starts[a] + pinfo.spacings[a] / 2.0 + pinfo.spacings[a] * int(coord[a])
                 for a in range(pinfo.dim))


def get_real_coord_bound(pinfo: PatchInfo, coord, side: int) -> tuple:
    """``GetRealCoordBound`` (``DomainTools.h:178-210``): a point of ``side``; ``coord``
    indexes the other axes in increasing order, as ``get_real_coord``."""
    axis = side >> 1
    out, k = [], 0
    for a in range(pinfo.dim):
        if a == axis:
            lower = (side & 1) == 0
            out.append(pinfo.starts[a] if lower else pinfo.starts[a] + pinfo.spacings[a] * pinfo.ns[a])
        else:
            out.append(_on_axis(pinfo, a, int(coord[k])))
            k += 1
    return tuple(out)


def _evaluate(func, centres: np.ndarray, vectorized: bool) -> np.ndarray:
    """func at every point of centres (shape (..., dim)), as an array of shape centres[..., 0]."""
    shape = centres.shape[:-1]
    if vectorized:
        coord = np.moveaxis(centres, -1, 0)
        return np.broadcast_to(np.asarray(func(coord), dtype=np.float64), shape)
    flat = centres.reshape(-1, centres.shape[-1])
    return np.array([float(func(tuple(pt))) for pt in flat], dtype=np.float64).reshape(shape)


def _set(domain: R.Domain, vec: R.DeviceArray, funcs, first: int, ghost: bool, vectorized: bool):
    ncomp = vec.shape[1]
    if tuple(vec.shape) != tuple(domain.shape(ncomp)):
        raise ValueError(f"vector shape {vec.shape} is not this domain's {domain.shape(ncomp)}")
    h = vec.get()
    centres = mesh.cell_centers(domain.patches, domain.dim, domain.ns, with_ghost=ghost)
    err = None
    for k, func in enumerate(funcs):
        c = first + k
        if c < 0 or c >= ncomp:
            err = RuntimeError("Invalid component to set")
            break
        vals = _evaluate(func, centres, vectorized).astype(h.dtype)
        if ghost:
            h[:, c] = vals
        else:
            mesh.interior(h[:, c:c + 1], domain.dim)[:, 0] = vals
    vec.upload(h)
    if err is not None:
        raise err


def set_values(domain: R.Domain, vec: R.DeviceArray, *funcs, component_index: int = 0, vectorized: bool = True):
    """``SetValues`` (``DomainTools.h:220-346``): one function per component from
    ``component_index``; ghost values are left as they were.  More functions than
    components raise ``RuntimeError`` once the components run out (the earlier ones are
    set, as in the reference)."""
    _set(domain, vec, funcs, component_index, False, vectorized)


def set_values_with_ghost(domain: R.Domain, vec: R.DeviceArray, *funcs, component_index: int = 0,
                          vectorized: bool = True):
    """``SetValuesWithGhost`` (``DomainTools.h:357-506``): as ``set_values``, ghost cells
    included (their centres, ``get_real_coord_ghost``)."""
    _set(domain, vec, funcs, component_index, True, vectorized)


def boundary_points(patches, dim: int, ns) -> np.ndarray:
    """The face points of every side of every patch (``GetRealCoordBound``), in the layout of
    the boundary data that ``te_star_add_dirichlet_bc`` / ``te_star_add_neumann_bc`` read:
    per patch, per side (west, east, south, north, bottom, top), the side's face cells with
    the first of the other axes fastest.  Shape (te_domain_boundary_length, dim)."""
    starts = np.array([r["starts"][:dim] for r in patches], dtype=np.float64).reshape(-1, dim)
    h = np.array([r["spacings"][:dim] for r in patches], dtype=np.float64).reshape(-1, dim)
    sides = []
    for s in range(2 * dim):
        a = s >> 1
        free = [b for b in range(dim) if b != a]
        idx = np.indices([ns[b] for b in free][::-1]).reshape(len(free), -1)[::-1]
        pts = np.empty((len(starts), idx.shape[1], dim))
        pts[:, :, a] = starts[:, a:a + 1] + (h[:, a:a + 1] * ns[a] if s & 1 else 0.0)
        for k, b in enumerate(free):
            pts[:, :, b] = starts[:, b:b + 1] + h[:, b:b + 1] / 2.0 + h[:, b:b + 1] * idx[k][None, :]
        sides.append(pts)
    return np.concatenate(sides, axis=1).reshape(-1, dim)


def boundary_values(domain: R.Domain, func, vectorized: bool = True) -> np.ndarray:
    """func at the boundary face points of ``domain`` (``boundary_points``)."""
    pts = boundary_points(domain.patches, domain.dim, domain.ns)
    return np.ascontiguousarray(_evaluate(func, pts, vectorized), dtype=np.float64)


def add_dirichlet_bc(domain: R.Domain, f: R.DeviceArray, gfunc, hfunc=None, vectorized: bool = True):
    """``Poisson::StarPatchOperator::addDrichletBCToRHS(f, gfunc)`` (and the VarPoisson form
    with the coefficient function hfunc): the functions are sampled here at the face points
    and the right-hand side is updated on the device."""
    g = R.DeviceArray.from_host(boundary_values(domain, gfunc, vectorized))
    c = R.DeviceArray.from_host(boundary_values(domain, hfunc, vectorized)) if hfunc is not None else None
    domain.add_dirichlet_bc(f, g, c)
    R.sync()


def add_neumann_bc(domain: R.Domain, f: R.DeviceArray, gfunc_grad, vectorized: bool = True):
    """``Poisson::StarPatchOperator::addNeumannBCToRHS(f, gfunc, gfunc_grad)``: gfunc_grad is
    one function per axis (the reference's gradient array); each side uses its axis' one."""
    pts = boundary_points(domain.patches, domain.dim, domain.ns)
    stride = len(pts) // max(len(domain.patches), 1)
    side = np.concatenate([np.full(int(np.prod([domain.ns[b] for b in range(domain.dim) if b != (s >> 1)])), s)
                           for s in range(2 * domain.dim)])
    side = np.tile(side, len(domain.patches))
    assert len(side) == len(pts) and stride * len(domain.patches) == len(pts)
    dg = np.empty(len(pts))
    for a in range(domain.dim):
        m = (side >> 1) == a
        dg[m] = _evaluate(gfunc_grad[a], pts[m], vectorized)
    d = R.DeviceArray.from_host(dg)
    domain.add_neumann_bc(f, d)
    R.sync()


def integrate(domain: R.Domain, u: R.DeviceArray) -> float:
    """``Integrate`` (``DomainTools.h:513-536``): interior values times the patch's cell
    volume, summed over patches and components on the device."""
    return domain.integrate(u)


class DomainTools:
    """The reference's static-method spelling of the functions above."""

    GetRealCoord = staticmethod(get_real_coord)
    GetRealCoordGhost = staticmethod(get_real_coord_ghost)
    GetRealCoordBound = staticmethod(get_real_coord_bound)
    SetValues = staticmethod(set_values)
    SetValuesWithGhost = staticmethod(set_values_with_ghost)
    Integrate = staticmethod(integrate)
    BoundaryPoints = staticmethod(boundary_points)
