"""Host-side driver objects over the C ABI: the forest (mesh generator), the
multigrid hierarchy and the outer BiCGStab.

These mirror the reference's setup sequence (GMG/CycleBuilder.h:97-192 with
Iterative/BiCGStab.h:111-203): a domain generator produces the finest and
successively coarser domains; every level gets
a ghost filler, operator and smoother; CycleBuilder assembles the V-cycle;
BiCGStab solves with the cycle as right preconditioner.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import mesh
from . import runtime as R


class Forest:
    """Quadtree (2D) / octree (3D) forest on the unit square / cube (te_forest),
    the replacement of P4estDomainGenerator / P8estDomainGenerator."""

    def __init__(self, dim: int, base_level: int):
        self.dim = dim
        h = C.c_void_p()
        R.check(R.lib().te_forest_create(dim, base_level, C.byref(h)))
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            R.lib().te_forest_destroy(self.h)
            self.h = None

    def leaves(self):
        n = R.lib().te_forest_num_leaves(self.h)
        lv = np.zeros(n, dtype=np.int32)
        co = np.zeros((n, self.dim), dtype=np.int32)
        R.check(R.lib().te_forest_get_leaves(self.h, lv.ctypes.data, co.ctypes.data))
        return lv, co

    def leaf_boxes(self):
        """(lower corner, size) of every leaf in unit coordinates."""
        lv, co = self.leaves()
        size = 1.0 / (1 << lv.astype(np.int64))
        return co * size[:, None], size

    def refine(self, flags):
        flags = np.ascontiguousarray(flags, dtype=np.uint8)
        R.check(R.lib().te_forest_refine(self.h, flags.ctypes.data))

    def balance(self):
        R.check(R.lib().te_forest_balance(self.h))

    def build_domains(self, ns):
        """Patch tables of every level, finest first (getFinestDomain, getCoarserDomain...)."""
        nsa = (C.c_int * 3)(*(list(ns) + [1] * (3 - len(ns))))
        nl = C.c_int()
        R.check(R.lib().te_forest_build_domains(self.h, nsa, C.byref(nl)))
        out = []
        for k in range(nl.value):
            n = C.c_int()
            R.check(R.lib().te_forest_domain_size(self.h, k, C.byref(n)))
            arr = np.zeros(n.value, dtype=mesh.PATCH_DTYPE)
            R.check(R.lib().te_forest_domain_patches(self.h, k, arr.ctypes.data))
            out.append(arr)
        return out


def uniform_levels(dim: int, level: int, ns):
    f = Forest(dim, level)
    return f.build_domains(ns)


class _Opts(C.Structure):
    _fields_ = [
        ("filler_kind", C.c_int),
        ("fill_type", C.c_int),
        ("op_kind", C.c_int),
        ("neumann", C.c_int),
        ("smoother", C.c_int),
        ("neumann_mask", C.c_uint),
        ("inner_max_it", C.c_int),
        ("inner_tol", C.c_double),
        ("continue_on_breakdown", C.c_int),
        ("pre_sweeps", C.c_int),
        ("post_sweeps", C.c_int),
        ("coarse_sweeps", C.c_int),
        ("precision", C.c_int),
        ("use_graph", C.c_int),
        ("cycle_type", C.c_int),
        ("mid_sweeps", C.c_int),
        ("level_smoother", C.POINTER(C.c_int)),
        ("level_inner_max_it", C.POINTER(C.c_int)),
        ("level_inner_tol", C.POINTER(C.c_double)),
    ]


CYCLE_TYPES = {"V": 0, "W": 1, "F": 2}


class GMG:
    """Multigrid hierarchy + outer BiCGStab / CG on the device (te_gmg).

    levels: patch tables finest first.  smoother 0 = DFTPatchSolver,
    1 = Iterative::PatchSolver(BiCGStab), or a list with one kind per level (finest first;
    inner_max_it / inner_tol may be lists too).  precision F64 or F32 (FP32 cycle,
    FP64 coarsest solve, FP64 outer iteration).  cycle_type "V", "W" or "F"
    (GMG::VCycle, WCycle, FMGCycle; mid_sweeps for W and F)."""

    def __init__(self, dim, ns, levels, filler_kind=R.BILINEAR, fill_type=R.FACES, op_kind=0, neumann=False,
                 smoother=0, neumann_mask=0, inner_max_it=1000, inner_tol=1e-12, continue_on_breakdown=False,
                 pre_sweeps=1, post_sweeps=1, coarse_sweeps=1, precision=R.F64, use_graph=True, coeffs=None,
                 cycle_type="V", mid_sweeps=1):
        self.dim, self.ns = dim, tuple(ns)
        self.domains = [R.Domain(dim, ns, lv) for lv in levels]
        self.coeffs = None
        cptrs = None
        if coeffs is not None:
            self.coeffs = [R.DeviceArray.from_host(np.ascontiguousarray(c, dtype=np.float64)) for c in coeffs]
            cptrs = (C.c_void_p * len(levels))(*[c.ptr for c in self.coeffs])
        nl = len(levels)

        def per_level(v, ctype):
            if np.isscalar(v):
                return v, None
            if len(v) != nl:
                raise ValueError("per-level settings need one entry per level")
            return v[0], (ctype * nl)(*v)

        smoother, self._lsm = per_level(smoother, C.c_int)
        inner_max_it, self._lmi = per_level(inner_max_it, C.c_int)
        inner_tol, self._ltol = per_level(inner_tol, C.c_double)
        o = _Opts(filler_kind, fill_type, op_kind, int(neumann), smoother, neumann_mask, inner_max_it, inner_tol,
                  int(continue_on_breakdown), pre_sweeps, post_sweeps, coarse_sweeps, precision, int(use_graph),
                  CYCLE_TYPES[cycle_type], mid_sweeps,
                  C.cast(self._lsm, C.POINTER(C.c_int)) if self._lsm is not None else None,
                  C.cast(self._lmi, C.POINTER(C.c_int)) if self._lmi is not None else None,
                  C.cast(self._ltol, C.POINTER(C.c_double)) if self._ltol is not None else None)
        hs = (C.c_void_p * len(self.domains))(*[d.h for d in self.domains])
        h = C.c_void_p()
        R.check(R.lib().te_gmg_create(hs, len(self.domains), C.byref(o), cptrs, C.byref(h)))
        self.h = h.value
        self.shape = self.domains[0].shape(1)

    def __del__(self):
        if getattr(self, "h", None):
            R.lib().te_gmg_destroy(self.h)
            self.h = None

    def vcycle_device(self, f: R.DeviceArray, u: R.DeviceArray):
        R.check(R.lib().te_gmg_vcycle(self.h, f.ptr, u.ptr, None))

    def vcycle(self, f: np.ndarray) -> np.ndarray:
        df = R.DeviceArray.from_host(np.ascontiguousarray(f, dtype=np.float64))
        du = R.DeviceArray(self.shape, np.float64)
        du.zero()
        self.vcycle_device(df, du)
        return du.get()

    def solve_device(self, b: R.DeviceArray, x: R.DeviceArray, tol=1e-10, max_it=1000, precond=True,
                     history_cap=0):
        """precond: False (none), True (the reference's solve), 2 (V-cycle with the
        early exit corrected to x += alpha M p)."""
        its = C.c_int(0)
        hist = np.full(max(history_cap, 1), np.nan)
        R.check(R.lib().te_gmg_bicgstab(self.h, b.ptr, x.ptr, int(precond), max_it, tol, C.byref(its),
                                        hist.ctypes.data_as(C.POINTER(C.c_double)), history_cap, None))
        return its.value, hist[: its.value + 1] if history_cap else None

    def solve(self, b: np.ndarray, x0: np.ndarray | None = None, tol=1e-10, max_it=1000, precond=True):
        db = R.DeviceArray.from_host(np.ascontiguousarray(b, dtype=np.float64))
        dx = R.DeviceArray.from_host(np.zeros(self.shape) if x0 is None else np.ascontiguousarray(x0, np.float64))
        its, hist = self.solve_device(db, dx, tol=tol, max_it=max_it, precond=precond, history_cap=max_it + 2)
        return dx.get(), its, hist

    def cg_device(self, b: R.DeviceArray, x: R.DeviceArray, tol=1e-10, max_it=1000, precond=True, history_cap=0):
        """Iterative::CG::solve, right-preconditioned by the cycle (precond True) or plain."""
        its = C.c_int(0)
        hist = np.full(max(history_cap, 1), np.nan)
        R.check(R.lib().te_gmg_cg(self.h, b.ptr, x.ptr, int(bool(precond)), max_it, tol, C.byref(its),
                                  hist.ctypes.data_as(C.POINTER(C.c_double)), history_cap, None))
        return its.value, hist[: its.value + 1] if history_cap else None

    def cg(self, b: np.ndarray, x0: np.ndarray | None = None, tol=1e-10, max_it=1000, precond=True):
        db = R.DeviceArray.from_host(np.ascontiguousarray(b, dtype=np.float64))
        dx = R.DeviceArray.from_host(np.zeros(self.shape) if x0 is None else np.ascontiguousarray(x0, np.float64))
        its, hist = self.cg_device(db, dx, tol=tol, max_it=max_it, precond=precond, history_cap=max_it + 2)
        return dx.get(), its, hist

    def last_solve_seconds(self):
        return R.lib().te_gmg_last_solve_seconds(self.h)

    # te_kernel_id: the finest level's kernels of the compact 3D cycle and the solver
    KERNELS = {"sweep_zero": 0, "sweep": 1, "sweep_interp": 2, "resid_restrict": 3, "operator": 4, "update_xr": 5,
               "operator_p": 6, "update_s": 7, "update_p": 8, "x_combine": 9}

    def time_kernel(self, name, reps=20):
        """Average milliseconds of `reps` standalone launches of one kernel on the solver's
        own buffers and stream (te_gmg_time_kernel; overwrites the work buffers)."""
        ms = C.c_double()
        R.check(R.lib().te_gmg_time_kernel(self.h, self.KERNELS[name], reps, C.byref(ms)))
        return ms.value

    def set_timer(self, timer: "R.Timer | None", iterations=True, domain_ids=None):
        """Attach a Timer (None detaches): "Iteration" timings of the outer solves
        (iterations=True) and "Total Patch Smooth" domain timings of every sweep of the
        levels whose entry in domain_ids (one per level, TIMER_NONE = untimed; None: no
        level timings) names a domain added to the timer.  Keeps a reference to the timer."""
        ids = None
        if timer is not None and domain_ids is not None:
            if len(domain_ids) != len(self.domains):
                raise ValueError("domain_ids needs one entry per level")
            ids = (C.c_int * len(domain_ids))(*[int(v) for v in domain_ids])
        R.check(R.lib().te_gmg_set_timer(self.h, timer.h if timer is not None else None, int(bool(iterations)), ids))
        self._timer = timer


class Batch:
    """Many independent 2D problems solved together (te_batch): one launch per cycle step and
    per BiCGStab step for all problems of a patch size, one graph per outer iteration.

    problems: a list of (ns, levels) with levels the patch tables finest first; the options
    are GMG's (V-cycle, Poisson operator, DFT smoother).  Each problem keeps its own vectors
    (shape of its finest level); its iterations and result are those of solving it alone."""

    def __init__(self, problems, filler_kind=R.BIQUADRATIC, fill_type=R.FACES, neumann=False, neumann_mask=0,
                 pre_sweeps=1, post_sweeps=1, coarse_sweeps=1, precision=R.F64, use_graph=True):
        self.domains = []
        nlev = []
        for ns, levels in problems:
            nlev.append(len(levels))
            self.domains += [R.Domain(2, ns, lv) for lv in levels]
        self.n = len(problems)
        firsts = np.concatenate([[0], np.cumsum(nlev)[:-1]]).astype(int)
        self.shapes = [self.domains[f].shape(1) for f in firsts]
        o = _Opts(filler_kind, fill_type, 0, int(neumann), 0, neumann_mask, 1000, 1e-12, 0, pre_sweeps, post_sweeps,
                  coarse_sweeps, precision, int(use_graph), CYCLE_TYPES["V"], 1, None, None, None)
        hs = (C.c_void_p * len(self.domains))(*[d.h for d in self.domains])
        nl = (C.c_int * self.n)(*nlev)
        h = C.c_void_p()
        R.check(R.lib().te_batch_create(self.n, hs, nl, C.byref(o), C.byref(h)))
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            R.lib().te_batch_destroy(self.h)
            self.h = None

    def num_groups(self):
        return R.lib().te_batch_num_groups(self.h)

    def _ptrs(self, arrays):
        if len(arrays) != self.n:
            raise ValueError("one vector per problem")
        for a, shp in zip(arrays, self.shapes):
            if a.dtype != np.float64 or tuple(a.shape) != tuple(shp):
                raise ValueError(f"each problem's vector must be FP64 of its finest level's shape {shp}")
        return (C.c_void_p * self.n)(*[a.ptr for a in arrays])

    def vcycle_device(self, fs, us):
        R.check(R.lib().te_batch_vcycle(self.h, self._ptrs(fs), self._ptrs(us), None))

    def solve_device(self, bs, xs, tol=1e-10, max_it=1000, precond=2):
        """BiCGStab on every problem (precond as GMG.solve_device); returns the iteration
        counts and the last relative residual norms."""
        its = np.zeros(self.n, dtype=np.int32)
        res = np.zeros(self.n)
        R.check(R.lib().te_batch_bicgstab(self.h, self._ptrs(bs), self._ptrs(xs), int(precond), max_it, tol,
                                          its.ctypes.data_as(C.POINTER(C.c_int)),
                                          res.ctypes.data_as(C.POINTER(C.c_double)), None))
        return its, res


def solve_many(solvers, bs, xs, tol=1e-10, max_it=1000, precond=2, threads=8):
    """Independent solves run concurrently: each GMG owns a stream, and the C calls release
    the interpreter lock, so up to `threads` solves keep the GPU busy together (many small
    independent problems).  Returns the iteration counts."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=threads) as ex:
        futs = [ex.submit(g.solve_device, b, x, tol, max_it, precond) for g, b, x in zip(solvers, bs, xs)]
        return [f.result()[0] for f in futs]
