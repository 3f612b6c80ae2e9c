"""te_vec_dot_noalloc (the allocation-free variant of te_vec_dot) against te_vec_dot,
through the C ABI: two calls in a row, the second on a smaller domain, FP64 and FP32
storage, one and two components.  Both are FP64 sums of the same products in different
orders, so they agree to rounding: relative difference <= 1e-12 of the sum of |a b|."""
import ctypes as C

import numpy as np
import pytest

from helpers import load, random_field
from thunderegg_amd import mesh
from thunderegg_amd import runtime as R

pytestmark = pytest.mark.gpu


def _noalloc_dot(dom, a, b, out):
    R.check(R.lib().te_vec_dot_noalloc(dom.h, R.F64 if a.dtype == np.float64 else R.F32, a.ptr, b.ptr, a.shape[1],
                                       out.ptr, None))
    R.sync()
    return float(out.get()[0])


def _case(dom, dim, ncomp, nt, seed):
    a = random_field(dom.shape(ncomp), seed, nt)
    b = random_field(dom.shape(ncomp), seed + 1, nt)
    return a, b


@pytest.mark.parametrize("dtype", [R.F64, R.F32])
@pytest.mark.parametrize("ncomp", [1, 2])
def test_noalloc_dot_matches_dot(dtype, ncomp):
    print("case: noalloc dot", dtype, ncomp, flush=True)
    nt = R.np_dtype(dtype)
    out = R.DeviceArray((1,), np.float64)
    # first call: a refined 2D mesh of 16x16 patches; second: a smaller 3D mesh
    dim1, lv1 = load("2d_uniform_8x8_refined_cross_mpi1.json", (16, 16))
    dim2, lv2 = load("3d_uniform_2x2x2_mpi1.json", (4, 6, 8))
    d1 = R.Domain(dim1, (16, 16), lv1[0])
    d2 = R.Domain(dim2, (4, 6, 8), lv2[0])
    assert d2.np < d1.np
    for dom, dim, seed in [(d1, dim1, 3), (d2, dim2, 7)]:
        a, b = _case(dom, dim, ncomp, nt, seed)
        da, db = R.DeviceArray.from_host(a), R.DeviceArray.from_host(b)
        ref = dom.dot(da, db)
        got = _noalloc_dot(dom, da, db, out)
        ia, ib = mesh.interior(a, dim).astype(np.float64), mesh.interior(b, dim).astype(np.float64)
        scale = np.abs(ia * ib).sum()
        assert abs(got - ref) <= 1e-12 * scale, (dom.np, got, ref)
        assert abs(got - (ia * ib).sum()) <= 1e-12 * scale


def test_noalloc_dot_rejects_too_many_patches():
    print("case: noalloc dot size limit", flush=True)
    n = 65536 + 1
    lv = np.zeros(n, dtype=mesh.PATCH_DTYPE)
    lv["id"] = np.arange(n)
    lv["nbr_type"] = -1
    lv["nbr_orth"] = -1
    lv["orth_on_parent"] = -1
    lv["spacings"] = 1.0
    dom = R.Domain(2, (1, 1), lv)
    a = R.DeviceArray(dom.shape(1), np.float64)
    out = R.DeviceArray((1,), np.float64)
    rc = R.lib().te_vec_dot_noalloc(dom.h, R.F64, a.ptr, a.ptr, 1, out.ptr, None)
    assert rc == -1  # TE_ERR_ARG
    assert b"65536" in R.lib().te_last_error()


def test_noalloc_dot_empty_domain():
    print("case: noalloc dot empty", flush=True)
    dom = R.Domain(2, (4, 4), np.zeros(0, dtype=mesh.PATCH_DTYPE))
    a = R.DeviceArray((1,), np.float64)
    out = R.DeviceArray.from_host(np.array([7.0]))
    R.check(R.lib().te_vec_dot_noalloc(dom.h, R.F64, a.ptr, a.ptr, 1, out.ptr, None))
    R.sync()
    assert out.get()[0] == 0.0
