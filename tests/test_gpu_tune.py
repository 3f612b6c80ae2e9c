"""Every launch variant of the autotuner's families gives the hand default's bits: the DFT
patch solve on fixture and synthetic meshes (2D n = 8, 16, 32; 3D n = 8), FP32 and FP64, with
each variant forced through the table, for the unfused solve and for the fused sweeps of the
matrix-core kernel; and the table shipped beside the library is the one loaded."""
import os

import numpy as np
import pytest

from thunderegg_amd import runtime as R
from thunderegg_amd.solver import uniform_levels

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _print_case(*a):
    print("case:", *a, flush=True)


@pytest.fixture
def clean_table():
    R.tune_clear()
    yield
    R.tune_clear()


@pytest.mark.parametrize("dim,n,level", [(2, 8, 5), (2, 16, 5), (2, 32, 4), (3, 8, 3), (2, 12, 4)])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_variants_give_identical_results(dim, n, level, dtype, clean_table):
    _print_case("tune", dim, n, dtype)
    lv = uniform_levels(dim, level, (n,) * dim)[0]
    dom = R.Domain(dim, (n,) * dim, lv)
    filler = R.GhostFiller(dom, R.BILINEAR if dim == 2 else R.TRILINEAR, R.FACES)
    plan = R.DftPlan(filler)
    fam = R.tune_family(dim, n)
    npdt = np.float64 if dtype == R.F64 else np.float32
    rng = np.random.default_rng(7)
    f = R.DeviceArray.from_host(rng.standard_normal(dom.shape(1)).astype(npdt))
    u0 = rng.standard_normal(dom.shape(1)).astype(npdt)
    outs = []
    for v in range(R.tune_num_variants(fam)):
        R.tune_set(fam, dim, n, dtype, v)
        u = R.DeviceArray.from_host(u0)
        plan.smooth(f, u)  # ghost fill + patch solve
        a = R.DeviceArray(dom.shape(1), npdt)
        a.zero()
        plan.apply(f, a)  # from u = 0
        outs.append((u.get(), a.get()))
        ms = plan.time_variant(f, u, a, v, reps=2)
        assert ms > 0
    for k, (s, a) in enumerate(outs[1:], 1):
        assert np.array_equal(s, outs[0][0]), f"variant {k}"
        assert np.array_equal(a, outs[0][1]), f"variant {k}"


@pytest.mark.parametrize("n", [16, 32])
def test_mfma_variants_in_the_fused_cycle(n, clean_table):
    """the fused V-cycle (every sweep mode of the matrix-core kernel) gives the same bits
    under every workgroup-count variant"""
    from thunderegg_amd import solver

    _print_case("tune fused cycle", n)
    levels = uniform_levels(2, 6 if n == 16 else 5, (n, n))
    rng = np.random.default_rng(3)
    f = rng.standard_normal((len(levels[0]), 1, n + 2, n + 2))
    res = []
    for v in range(R.tune_num_variants(R.TUNE_DFT_MFMA_2D)):
        for dt in (R.F32, R.F64):
            R.tune_set(R.TUNE_DFT_MFMA_2D, 2, n, dt, v)
        g = solver.GMG(2, (n, n), levels, filler_kind=R.BIQUADRATIC, smoother=0, precision=R.F32)
        res.append(g.vcycle(f))
    for k, r in enumerate(res[1:], 1):
        assert np.array_equal(r, res[0]), f"variant {k}"


@pytest.mark.parametrize("prec", [R.F32, R.F64])
def test_c3_sweep_variants_identical(prec, clean_table):
    """the 3D compact cycle's sweeps (modes 1 and 2) under every load-batch variant give the
    hand default's bits; 4096 finest patches (smaller levels issue all their loads in one batch
    whatever the variant)"""
    from thunderegg_amd import solver

    _print_case("tune c3", prec)
    levels = uniform_levels(3, 4, (16, 16, 16))
    assert len(levels[0]) == 4096
    rng = np.random.default_rng(11)
    f = rng.standard_normal((len(levels[0]), 1, 18, 18, 18))
    res = []
    for v in range(R.tune_num_variants(R.TUNE_C3_SWEEP)):
        R.tune_set(R.TUNE_C3_SWEEP, 3, 16, prec, v)
        g = solver.GMG(3, (16, 16, 16), levels, filler_kind=R.TRILINEAR, smoother=0, pre_sweeps=2, post_sweeps=2,
                       precision=prec)
        res.append(g.vcycle(f))
    for k, r in enumerate(res[1:], 1):
        assert np.array_equal(r, res[0]), f"variant {k}"
