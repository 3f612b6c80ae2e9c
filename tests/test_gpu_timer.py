"""A Timer attached to the solver (te_gmg_set_timer): the timings are recorded by events on
the solver's stream and leave the arithmetic unchanged.

Checked: the same iteration count, residual history and bit-identical solution with the
timer attached; one "Iteration" timing per outer iteration (BiCGStab.h:145-200,
CG.h:158-186); under it, per level, a "Total Patch Smooth" domain timing for every
smoothing sweep (PatchSolver.h:153-168), with the call counts the V-cycle's sweeps give;
device time nested inside the host's wall time; stream timings of user code.
"""
import numpy as np
import pytest

from helpers import load, random_field
from thunderegg_amd import mesh
from thunderegg_amd import runtime as R
from thunderegg_amd import solver

pytestmark = pytest.mark.gpu


def _print_case(*a):
    print("case:", *a, flush=True)


CASES = [
    # label, dim, ns, mesh file, filler, smoother, precision (fused 2D path for DFT in 2D)
    ("cross2d_biquad_dft16_mixed", 2, (16, 16), "2d_uniform_8x8_refined_cross_mpi1.json", R.BIQUADRATIC, 0, R.F32),
    ("cross2d_biquad_dft16_fp64", 2, (16, 16), "2d_uniform_8x8_refined_cross_mpi1.json", R.BIQUADRATIC, 0, R.F64),
    ("cross2d_bilin_bicg", 2, (8, 8), "2d_uniform_8x8_refined_cross_mpi1.json", R.BILINEAR, 1, R.F64),
    ("mid3d_dft", 3, (4, 4, 4), "3d_mid_refine_4x4x4_mpi1.json", R.TRILINEAR, 0, R.F64),
]


def _timings(j, name):
    return [t for t in j.get("timings", []) if t["name"] == name]


@pytest.mark.parametrize("label,dim,ns,meshfile,filler,smoother,prec", CASES)
def test_bicgstab_with_timer(label, dim, ns, meshfile, filler, smoother, prec):
    _print_case("timer-bicgstab", label)
    dim, levels = load(meshfile, ns)
    nl = len(levels)
    g = solver.GMG(dim, ns, levels, filler_kind=filler, smoother=smoother, precision=prec)
    f = random_field(mesh.padded_shape(len(levels[0]), dim, ns), 41)
    b = R.DeviceArray.from_host(f)
    x = R.DeviceArray(g.shape, np.float64)
    x.zero()
    its0, hist0 = g.solve_device(b, x, tol=1e-10, precond=2, history_cap=502)
    x0 = x.get()

    t = R.Timer()
    for k, d in enumerate(g.domains):
        t.add_domain(k, d)
    with pytest.raises(R.TEError, match="not added"):
        g.set_timer(t, domain_ids=[k + 100 for k in range(nl)])
    g.set_timer(t, iterations=True, domain_ids=list(range(nl)))
    x.zero()
    its1, hist1 = g.solve_device(b, x, tol=1e-10, precond=2, history_cap=502)
    wall = g.last_solve_seconds()
    assert its1 == its0
    assert np.array_equal(hist1, hist0)
    assert np.array_equal(x.get(), x0)

    j = t.to_json()
    assert [d[0] for d in j["domains"]] == list(range(nl))
    assert [tt["name"] for tt in j["timings"]] == ["Iteration"]
    it = j["timings"][0]
    passes = it["num_calls"]
    assert passes in (its0, its0 + 1)  # + 1 when the s-norm test ends the solve
    assert 0 < it["min"] <= it["max"] and it["sum"] <= wall * 1.05
    smooth = _timings(it, "Total Patch Smooth")
    assert sorted(s["domain_id"] for s in smooth) == list(range(nl))
    total = 0.0
    for s in smooth:
        k = s["domain_id"]
        per_cycle = 1 if k == nl - 1 else 2  # coarse_sweeps = 1; pre + post = 2
        assert s["num_calls"] == 2 * passes * per_cycle, (k, s["num_calls"], passes)
        assert s["min"] > 0
        total += s["sum"]
    assert total <= it["sum"] * 1.05
    rep = t.report()
    assert "Iteration -> Total Patch Smooth" in rep

    # detached: a further solve records nothing
    g.set_timer(None)
    x.zero()
    g.solve_device(b, x, tol=1e-10, precond=2)
    assert t.to_json()["timings"][0]["num_calls"] == passes


def test_cg_and_vcycle_with_timer():
    _print_case("timer-cg")
    ns = (8, 8)
    dim, levels = load("2d_uniform_4x4_mpi1.json", ns)
    g = solver.GMG(dim, ns, levels, filler_kind=R.BILINEAR, smoother=0, precision=R.F64)
    f = random_field(mesh.padded_shape(len(levels[0]), dim, ns), 42)
    x0, its0, h0 = g.cg(f, tol=1e-10)
    t = R.Timer()
    t.add_domain(0, g.domains[0])
    # only the finest level timed
    ids = [0] + [R.TIMER_NONE] * (len(levels) - 1)
    g.set_timer(t, iterations=True, domain_ids=ids)
    x1, its1, h1 = g.cg(f, tol=1e-10)
    assert its1 == its0 and np.array_equal(x1, x0) and np.array_equal(h1, h0)
    j = t.to_json()
    names = [tt["name"] for tt in j["timings"]]
    # CG's final x = M x (CG.h:186-189) runs one more cycle after the last iteration
    assert names == ["Iteration", "Total Patch Smooth"]
    it = j["timings"][0]
    assert it["num_calls"] == its0
    (s,) = _timings(it, "Total Patch Smooth")
    assert s["domain_id"] == 0 and s["num_calls"] == 2 * its0
    assert j["timings"][1]["num_calls"] == 2
    # a cycle alone (Cycle::apply) records only level timings
    g.set_timer(t, iterations=False, domain_ids=ids)
    g.vcycle(f)
    assert t.to_json()["timings"][1]["num_calls"] == 4


def test_user_stream_timings():
    _print_case("timer-stream")
    t = R.Timer()
    s = R.Stream()
    a = R.DeviceArray((1 << 26,), np.float64)  # 512 MiB
    t.start("outer")  # host clock around stream timings
    for _ in range(3):
        t.start("fill", stream=s)
        R.check(R.lib().te_memset(a.ptr, 0, a.nbytes, s.h))
        t.stop("fill", stream=s)
    s.sync()
    t.stop("outer")
    with pytest.raises(R.TEError, match="stopped the same way"):
        t.start("mixed", stream=s)
        t.stop("mixed")
    t.stop("mixed", stream=s)
    j = t.to_json()
    outer = j["timings"][0]
    (fill,) = _timings(outer, "fill")
    assert fill["num_calls"] == 3
    # 512 MiB at no more than 10 TB/s takes at least 50 us
    assert fill["min"] > 5e-5 and fill["sum"] <= outer["sum"] * 1.05
