"""DomainTools (reference DomainTools.h:36-537) on the host: the coordinate helpers of
thunderegg_amd.domain_tools and of include/ThunderEgg/DomainTools.h, restating
test/DomainTools_MPI1.cpp (GetRealCoord, GetRealCoordGhost, GetRealCoordBound in 2D and
3D).  Expected values are written as start + (0.5 + i) * length / n, a different
arithmetic path from the helpers' start + h / 2 + h * i.  The device parts (SetValues,
Integrate) are in test_gpu_domain_tools.py."""
import itertools
import subprocess

import numpy as np
import pytest

from test_host import _build_api_test
from thunderegg_amd import domain_tools as DT

STARTS = (0.0, -1.0, 1.0, 0.23, -0.23, 3.0, -3.0)
LENGTHS = (1.0, 0.23, 3.0)


def _expect(i, n, start, length, interfaces=True):
    if interfaces and i == -1:
        return start
    if interfaces and i == n:
        return start + length
    return start + (0.5 + i) * length / n


def _close(a, b):
    return np.allclose(np.asarray(a) + 100, np.asarray(b) + 100, rtol=1e-13, atol=0)


@pytest.mark.parametrize("ny", [1, 2, 3])
def test_get_real_coord_2d(ny):
    nx = 2
    for sy, ly in itertools.product(STARTS, LENGTHS):
        p = DT.PatchInfo(dim=2, ns=(nx, ny), spacings=(1.0 / nx, ly / ny), starts=(0.0, sy))
        for j in range(-1, ny + 1):
            for i in range(-1, nx + 1):
                assert _close(DT.get_real_coord(p, (i, j)), (_expect(i, nx, 0.0, 1.0), _expect(j, ny, sy, ly)))
                assert _close(DT.get_real_coord_ghost(p, (i, j)),
                              (_expect(i, nx, 0.0, 1.0, False), _expect(j, ny, sy, ly, False)))


@pytest.mark.parametrize("nz", [1, 2, 3])
def test_get_real_coord_3d(nz):
    for sz, lz in itertools.product(STARTS, LENGTHS):
        p = DT.PatchInfo(dim=3, ns=(2, 2, nz), spacings=(0.5, 0.5, lz / nz), starts=(0.0, 0.0, sz))
        for k, j, i in itertools.product(range(-1, nz + 1), range(-1, 3), range(-1, 3)):
            e = (_expect(i, 2, 0.0, 1.0), _expect(j, 2, 0.0, 1.0), _expect(k, nz, sz, lz))
            assert _close(DT.get_real_coord(p, (i, j, k)), e)
            e = (_expect(i, 2, 0.0, 1.0, False), _expect(j, 2, 0.0, 1.0, False), _expect(k, nz, sz, lz, False))
            assert _close(DT.DomainTools.GetRealCoordGhost(p, (i, j, k)), e)


@pytest.mark.parametrize("ny", [1, 2, 3])
def test_get_real_coord_bound_2d(ny):
    nx = 2
    for sy, ly in itertools.product(STARTS, LENGTHS):
        p = DT.PatchInfo(dim=2, ns=(nx, ny), spacings=(1.0 / nx, ly / ny), starts=(0.0, sy))
        for i in range(-1, nx + 1):
            x = _expect(i, nx, 0.0, 1.0)
            assert _close(DT.get_real_coord_bound(p, (i,), DT.SOUTH), (x, sy))
            assert _close(DT.get_real_coord_bound(p, (i,), DT.NORTH), (x, sy + ly))
        for j in range(-1, ny + 1):
            y = _expect(j, ny, sy, ly)
            assert _close(DT.get_real_coord_bound(p, (j,), DT.WEST), (0.0, y))
            assert _close(DT.get_real_coord_bound(p, (j,), DT.EAST), (1.0, y))


def test_get_real_coord_bound_3d():
    p = DT.PatchInfo(dim=3, ns=(2, 3, 4), spacings=(0.5, 2.0 / 3, 0.25), starts=(0.0, -1.0, 0.5))
    for a, b in itertools.product(range(-1, 4), range(-1, 5)):
        # west / east: coord indexes y then z; south / north: x then z; bottom / top: x then y
        if a <= 3 and b <= 4:
            y, z = _expect(a, 3, -1.0, 2.0), _expect(b, 4, 0.5, 1.0)
            assert _close(DT.get_real_coord_bound(p, (a, b), DT.EAST), (1.0, y, z))
        if a <= 2:
            x, z = _expect(a, 2, 0.0, 1.0), _expect(b, 4, 0.5, 1.0)
            assert _close(DT.get_real_coord_bound(p, (a, b), DT.SOUTH), (x, -1.0, z))
        if a <= 2 and b <= 3:
            x, y = _expect(a, 2, 0.0, 1.0), _expect(b, 3, -1.0, 2.0)
            assert _close(DT.get_real_coord_bound(p, (a, b), DT.TOP), (x, y, 1.5))
            assert _close(DT.get_real_coord_bound(p, (a, b), DT.BOTTOM), (x, y, 0.5))


def test_patch_info_defaults_and_checks():
    p = DT.PatchInfo(dim=3)
    assert p.ns == (1, 1, 1) and p.starts == (0.0, 0.0, 0.0) and p.spacings == (1.0, 1.0, 1.0)
    with pytest.raises(ValueError):
        DT.PatchInfo(dim=2, ns=(2, 2, 2))


def test_cpp_coordinates(tmp_path):
    """include/ThunderEgg/DomainTools.h's coordinate helpers (host code only)."""
    exe = _build_api_test(tmp_path, "domain_tools_test")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "domain_tools_test: ok" in r.stdout
