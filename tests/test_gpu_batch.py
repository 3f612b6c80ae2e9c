"""Batches of independent problems (te_batch) against the same problems solved one at a
time.  A batch runs every problem's V-cycle and BiCGStab with one launch per step for all
problems of a patch size; each problem's inner products are folded from the partial sums of
the workgroups its own solve would use.  So a batch must reproduce each problem's own solve
bit for bit when that solve uses the cycle's one-launch-per-level forms (TE_NO_TAIL,
TE_NO_RRPRE, TE_NO_BSWEEP: the batch cannot use the fused tail, the restriction fused with
the coarse sweep, or the sweep fused into the BiCGStab updates, which assume one problem per
level), and agree with the default single solve to rounding.

The full set of 512 small adaptive problems (patch sizes 8, 16 and 32, 16 to 4096 patches)
runs as one batch: every problem reaches 1e-10 in its true residual, FP64 iteration counts
equal those of one-at-a-time solves on a sample, and the smallest problems match the
oracle's BiCGStab iteration for iteration.
"""
import numpy as np
import pytest

from oracle import oracle as O
from thunderegg_amd import configs, mesh
from thunderegg_amd import runtime as R
from thunderegg_amd import solver

pytestmark = pytest.mark.gpu

UNFUSED_ENV = ("TE_NO_TAIL", "TE_NO_RRPRE", "TE_NO_BSWEEP")


def _print_case(*a):
    print("case:", *a, flush=True)


def _single(P, precision):
    return solver.GMG(P.dim, P.ns, P.levels, filler_kind=P.filler, smoother=P.smoother, pre_sweeps=P.pre,
                      post_sweeps=P.post, coarse_sweeps=P.coarse, precision=precision)


def _batch(probs, precision):
    return solver.Batch([(P.ns, P.levels) for P in probs], filler_kind=probs[0].filler, pre_sweeps=probs[0].pre,
                        post_sweeps=probs[0].post, coarse_sweeps=probs[0].coarse, precision=precision)


def _true_residual(P, b, x):
    dom = R.Domain(2, P.ns, P.levels[0])
    filler = R.GhostFiller(dom, P.filler, R.FACES)
    r = R.DeviceArray(dom.shape(1), np.float64)
    filler.residual(b, x, r)
    return dom.two_norm(r) / dom.two_norm(b)


def _some_problems():
    """16 problems: all three patch sizes, 4 to 7 levels deep, 31 to 3775 patches"""
    probs = [configs.small_adaptive2d_problem(k) for k in list(range(12)) + [12, 13, 17, 31]]
    assert {P.ns[0] for P in probs} == {8, 16, 32}
    assert len({len(P.levels) for P in probs}) >= 3
    return probs


@pytest.mark.parametrize("prec", [R.F64, R.F32])
def test_batch_vcycle_equals_single(prec, monkeypatch):
    _print_case("batch vcycle", "f64" if prec == R.F64 else "mixed")
    probs = _some_problems()
    B = _batch(probs, prec)
    assert B.num_groups() == 3
    fs = [R.DeviceArray.from_host(P.rhs) for P in probs]
    us = [R.DeviceArray(P.rhs.shape, np.float64) for P in probs]
    for u in us:
        u.zero()
    B.vcycle_device(fs, us)
    got = [mesh.interior(u.get(), 2) for u in us]
    for P, g_ in zip(probs, got):
        ref = mesh.interior(_single(P, prec).vcycle(P.rhs), 2)
        assert np.max(np.abs(g_ - ref)) <= (1e-12 if prec == R.F64 else 1e-5) * np.max(np.abs(ref))
    for k in UNFUSED_ENV:
        monkeypatch.setenv(k, "1")
    for P, g_ in zip(probs, got):
        ref = mesh.interior(_single(P, prec).vcycle(P.rhs), 2)
        assert np.array_equal(g_, ref), P.name


@pytest.mark.parametrize("prec", [R.F64, R.F32])
def test_batch_bicgstab_equals_single(prec, monkeypatch):
    _print_case("batch bicgstab", "f64" if prec == R.F64 else "mixed")
    probs = _some_problems()
    for k in UNFUSED_ENV:
        monkeypatch.setenv(k, "1")
    B = _batch(probs, prec)
    bs = [R.DeviceArray.from_host(P.rhs) for P in probs]
    xs = [R.DeviceArray(P.rhs.shape, np.float64) for P in probs]
    for x in xs:
        x.zero()
    its, res = B.solve_device(bs, xs, tol=1e-10, max_it=500, precond=2)
    print("batch iterations", list(its), flush=True)
    for P, b, x, it, rs in zip(probs, bs, xs, its, res):
        g = _single(P, prec)
        x1 = R.DeviceArray(P.rhs.shape, np.float64)
        x1.zero()
        it1, hist = g.solve_device(b, x1, tol=1e-10, max_it=500, precond=2, history_cap=502)
        assert it == it1, P.name
        assert rs == hist[-1]
        assert np.array_equal(mesh.interior(x.get(), 2), mesh.interior(x1.get(), 2)), P.name
        assert _true_residual(P, b, x) <= 1.5e-10


def test_batch_repeat_and_stream_concurrent_baseline():
    """a second solve of the same batch gives the same bits; the stream-concurrent solves
    (one te_gmg per problem) reach the same iteration counts"""
    _print_case("batch repeat")
    probs = _some_problems()
    B = _batch(probs, R.F32)
    bs = [R.DeviceArray.from_host(P.rhs) for P in probs]
    outs = []
    for rep in range(2):
        xs = [R.DeviceArray(P.rhs.shape, np.float64) for P in probs]
        for x in xs:
            x.zero()
        its, _ = B.solve_device(bs, xs, tol=1e-10, max_it=500, precond=2)
        outs.append((its, [x.get() for x in xs]))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert all(np.array_equal(a, b) for a, b in zip(outs[0][1], outs[1][1]))
    gs = [_single(P, R.F32) for P in probs]
    xs = []
    for g in gs:
        x = R.DeviceArray(g.shape, np.float64)
        x.zero()
        xs.append(x)
    its_c = solver.solve_many(gs, bs, xs, tol=1e-10, max_it=500, precond=2, threads=4)
    assert all(abs(a - b) <= 1 for a, b in zip(its_c, outs[0][0]))


def test_batch_max_it_and_zero_rhs():
    """max_it ends every problem's loop (BiCGStab.h:148), and a zero right-hand side takes
    no iteration (r0 = 0) and leaves x as it was"""
    _print_case("batch limits")
    probs = _some_problems()[:4]
    B = _batch(probs, R.F64)
    bs = [R.DeviceArray.from_host(P.rhs) for P in probs]
    bs[1] = R.DeviceArray.from_host(np.zeros_like(probs[1].rhs))
    xs = [R.DeviceArray(P.rhs.shape, np.float64) for P in probs]
    for x in xs:
        x.zero()
    its, res = B.solve_device(bs, xs, tol=1e-10, max_it=3, precond=2)
    assert its[1] == 0 and res[1] == 0.0 and not np.any(xs[1].get())
    assert all(its[k] == 3 for k in (0, 2, 3)) and all(res[k] > 1e-10 for k in (0, 2, 3))


def test_small_adaptive2d_all_512_problems_batched():
    """All 512 problems in one batch, FP64 and mixed: every true residual <= 1.5e-10; FP64
    iteration counts equal one-at-a-time solves on every 16th problem; the 4 smallest
    problems' FP64 solves match the oracle's BiCGStab iteration for iteration."""
    _print_case("small_adaptive2d 512 batched")
    probs = configs.small_adaptive2d_problems(512)
    sizes = np.array([len(P.levels[0]) for P in probs])
    print("patches", sizes.min(), sizes.max(), "dof", sum(P.dof for P in probs), flush=True)
    bs = [R.DeviceArray.from_host(P.rhs) for P in probs]
    its_by = {}
    for prec in (R.F64, R.F32):
        B = _batch(probs, prec)
        xs = [R.DeviceArray(P.rhs.shape, np.float64) for P in probs]
        for x in xs:
            x.zero()
        its, res = B.solve_device(bs, xs, tol=1e-10, max_it=500, precond=2)
        its_by[prec] = its
        print("512 batched", "f64" if prec == R.F64 else "mixed", "iterations min/max/sum", its.min(), its.max(),
              its.sum(), flush=True)
        # (res is the last relative residual of the loop; a problem that ends through the
        # s-norm test keeps the previous iteration's, as the reference's loop does)
        worst = max(_true_residual(P, b, x) for P, b, x in zip(probs, bs, xs))
        print("worst true residual", worst, flush=True)
        assert worst <= 1.5e-10
        del B, xs
    assert np.all(np.abs(its_by[R.F32].astype(int) - its_by[R.F64]) <= 2)
    for k in range(0, 512, 16):
        P = probs[k]
        x = R.DeviceArray(P.rhs.shape, np.float64)
        x.zero()
        it1, _ = _single(P, R.F64).solve_device(bs[k], x, tol=1e-10, max_it=500, precond=2)
        assert it1 == its_by[R.F64][k], k
    for k in np.argsort(sizes, kind="stable")[:4]:
        P = probs[k]
        od = [O.Domain(2, P.ns, lv) for lv in P.levels]
        L = O.Levels(od, filler_kind=P.filler, smoother_kind=P.smoother)
        _, its_ref, hist_ref = L.bicgstab(P.rhs.copy(), precond=2, tol=1e-10)
        x = R.DeviceArray(P.rhs.shape, np.float64)
        x.zero()
        it1, hist = _single(P, R.F64).solve_device(bs[k], x, tol=1e-10, max_it=500, precond=2, history_cap=502)
        assert it1 == its_ref == its_by[R.F64][k]
        assert np.allclose(hist, hist_ref, rtol=1e-6, atol=0)
