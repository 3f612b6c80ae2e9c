"""ctypes binding of libthunderegg_amd.so (the C ABI in include/thunderegg_amd.h).

The library must be built in-tree (``__graft_entry__.build()``); importing this
module on a machine without it raises instead of falling back to anything.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import mesh

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libthunderegg_amd.so")

F64, F32 = 0, 1
FACES, EDGES, CORNERS = 0, 1, 2
BILINEAR, BIQUADRATIC, TRILINEAR = 0, 1, 2
TE_ERR_BREAKDOWN = -3

_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)

# name: (restype, [argtypes]) -- every function declared in include/thunderegg_amd.h
SIGNATURES = {
    "te_last_error": (C.c_char_p, []),
    "te_device_count": (C.c_int, [_ip]),
    "te_device_alloc": (C.c_int, [C.c_size_t, C.POINTER(_vp)]),
    "te_device_free": (C.c_int, [_vp]),
    "te_memcpy_h2d": (C.c_int, [_vp, _vp, C.c_size_t]),
    "te_memcpy_d2h": (C.c_int, [_vp, _vp, C.c_size_t]),
    "te_memcpy_d2d": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "te_memset": (C.c_int, [_vp, C.c_int, C.c_size_t, _vp]),
    "te_stream_synchronize": (C.c_int, [_vp]),
    "te_stream_create": (C.c_int, [C.POINTER(_vp)]),
    "te_stream_destroy": (C.c_int, [_vp]),
    "te_event_create": (C.c_int, [C.POINTER(_vp)]),
    "te_event_destroy": (C.c_int, [_vp]),
    "te_event_record": (C.c_int, [_vp, _vp]),
    "te_event_elapsed_ms": (C.c_int, [_vp, _vp, C.POINTER(C.c_float)]),
    "te_domain_create": (C.c_int, [C.c_int, _ip, C.c_int, C.c_int, _vp, C.POINTER(_vp)]),
    "te_domain_destroy": (C.c_int, [_vp]),
    "te_domain_num_patches": (C.c_int, [_vp]),
    "te_domain_vector_length": (C.c_int64, [_vp, C.c_int]),
    "te_domain_get_nbr_local": (C.c_int, [_vp, _vp]),
    "te_level_link_create": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "te_level_link_destroy": (C.c_int, [_vp]),
    "te_level_link_get_parents": (C.c_int, [_vp, _vp]),
    "te_ghost_filler_create": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(_vp)]),
    "te_ghost_filler_destroy": (C.c_int, [_vp]),
    "te_ghost_fill": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _vp]),
    "te_star_apply": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int, _vp]),
    "te_star_apply_patches": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int, _vp]),
    "te_varstar_apply": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_int, _vp]),
    "te_domain_boundary_length": (C.c_int64, [_vp]),
    "te_star_add_dirichlet_bc": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
    "te_star_add_neumann_bc": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    "te_residual": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, C.c_int, _vp]),
    "te_dft_plan_create": (C.c_int, [_vp, C.c_uint, C.c_int, _vp, C.POINTER(_vp)]),
    "te_dft_plan_destroy": (C.c_int, [_vp]),
    "te_dft_smooth": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    "te_dft_apply": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    "te_dft_patch_solve": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    "te_dft_apply_mixed": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "te_dft_smooth_interp": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_int, _vp, _vp, _vp]),
    "te_residual_restrict": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp, _vp, C.c_int, _vp,
                                       _vp]),
    "te_patch_bicgstab_smooth": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, C.c_double, C.c_int,
                                           _vp, _vp, _vp, _vp]),
    "te_restrict": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    "te_restrict_extrapolate": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    "te_interpolate": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    "te_vec_set": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_int, _vp]),
    "te_vec_set_with_ghost": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_int, _vp]),
    "te_vec_scale": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_int, _vp]),
    "te_vec_shift": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_int, _vp]),
    "te_vec_copy": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    "te_vec_add": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    "te_vec_add_scaled": (C.c_int, [_vp, C.c_int, _vp, C.c_double, _vp, C.c_int, _vp]),
    "te_vec_add_scaled2": (C.c_int, [_vp, C.c_int, _vp, C.c_double, _vp, C.c_double, _vp, C.c_int, _vp]),
    "te_vec_scale_then_add": (C.c_int, [_vp, C.c_int, _vp, C.c_double, _vp, C.c_int, _vp]),
    "te_vec_scale_then_add_scaled": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_double, _vp, C.c_int, _vp]),
    "te_vec_scale_then_add_scaled2": (C.c_int, [_vp, C.c_int, _vp, C.c_double, C.c_double, _vp, C.c_double, _vp,
                                                C.c_int, _vp]),
    "te_vec_dot": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _dp, _vp]),
    "te_vec_dot_noalloc": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _vp, _vp]),
    "te_vec_two_norm": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _dp, _vp]),
    "te_vec_inf_norm": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _dp, _vp]),
    "te_vec_integrate": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _dp, _vp]),
    "te_vec_convert": (C.c_int, [_vp, C.c_int, _vp, C.c_int, _vp, C.c_int, _vp]),
    "te_gmg_create": (C.c_int, [C.POINTER(_vp), C.c_int, _vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "te_gmg_destroy": (C.c_int, [_vp]),
    "te_gmg_vcycle": (C.c_int, [_vp, _vp, _vp, _vp]),
    "te_gmg_bicgstab": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_double, _ip, _dp, C.c_int, _vp]),
    "te_gmg_cg": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_double, _ip, _dp, C.c_int, _vp]),
    "te_gmg_last_solve_seconds": (C.c_double, [_vp]),
    "te_gmg_time_kernel": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "te_tune_num_variants": (C.c_int, [C.c_int]),
    "te_tune_load": (C.c_int, [C.c_char_p]),
    "te_tune_clear": (C.c_int, []),
    "te_tune_set": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "te_tune_get": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "te_tune_table_path": (C.c_int, [C.c_char_p, C.c_size_t]),
    "te_tune_time_dft": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, _dp]),
    "te_tune_time_c3": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    "te_batch_create": (C.c_int, [C.c_int, C.POINTER(_vp), _ip, _vp, C.POINTER(_vp)]),
    "te_batch_destroy": (C.c_int, [_vp]),
    "te_batch_num_groups": (C.c_int, [_vp]),
    "te_batch_vcycle": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), _vp]),
    "te_batch_bicgstab": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.c_int, C.c_int, C.c_double, _ip, _dp,
                                    _vp]),
    "te_gmg_set_timer": (C.c_int, [_vp, _vp, C.c_int, _ip]),
    "te_timer_create": (C.c_int, [C.POINTER(_vp)]),
    "te_timer_destroy": (C.c_int, [_vp]),
    "te_timer_add_domain": (C.c_int, [_vp, C.c_int, C.c_int, _ip, C.c_int, _vp]),
    "te_timer_start": (C.c_int, [_vp, C.c_char_p, C.c_int, C.c_int]),
    "te_timer_stop": (C.c_int, [_vp, C.c_char_p, C.c_int, C.c_int]),
    "te_timer_start_on": (C.c_int, [_vp, C.c_char_p, C.c_int, C.c_int, _vp]),
    "te_timer_stop_on": (C.c_int, [_vp, C.c_char_p, C.c_int, C.c_int, _vp]),
    "te_timer_add_int_info": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "te_timer_add_double_info": (C.c_int, [_vp, C.c_char_p, C.c_double]),
    "te_timer_to_json": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "te_timer_print": (C.c_int, [_vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "te_timer_save": (C.c_int, [_vp, C.c_char_p]),
    "te_forest_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(_vp)]),
    "te_forest_destroy": (C.c_int, [_vp]),
    "te_forest_num_leaves": (C.c_int, [_vp]),
    "te_forest_get_leaves": (C.c_int, [_vp, _vp, _vp]),
    "te_forest_refine": (C.c_int, [_vp, _vp]),
    "te_forest_balance": (C.c_int, [_vp]),
    "te_forest_build_domains": (C.c_int, [_vp, _ip, _ip]),
    "te_forest_domain_size": (C.c_int, [_vp, C.c_int, _ip]),
    "te_forest_domain_patches": (C.c_int, [_vp, C.c_int, _vp]),
}

_LIB = None


class TEError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"{msg} (status {rc})")
        self.rc = rc


class BreakdownError(TEError):
    """Iterative::BreakdownError (reference Iterative/BreakdownError.h)."""


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        msg = lib().te_last_error().decode(errors="replace")
        if rc == TE_ERR_BREAKDOWN:
            raise BreakdownError(rc, msg)
        raise TEError(rc, msg)


def sync(stream=None):
    check(lib().te_stream_synchronize(stream))


def np_dtype(dtype):
    return np.float64 if dtype == F64 else np.float32


class DeviceArray:
    """Device buffer holding a host-shaped array (allocated with te_device_alloc)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().te_device_alloc(max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().te_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes))

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().te_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self, stream=None):
        check(lib().te_memset(self.ptr, 0, self.nbytes, stream))

    def free(self):
        if getattr(self, "ptr", None):
            lib().te_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Domain:
    """``Domain<D>`` on the device (te_domain)."""

    def __init__(self, dim, ns, patches: np.ndarray, num_ghost: int = 1):
        self.dim = dim
        self.ns = tuple(int(v) for v in ns)
        self.patches = np.ascontiguousarray(patches, dtype=mesh.PATCH_DTYPE)
        self.np = len(self.patches)
        nsa = (C.c_int * 3)(*(list(self.ns) + [1] * (3 - dim)))
        h = C.c_void_p()
        check(lib().te_domain_create(dim, nsa, num_ghost, self.np, self.patches.ctypes.data, C.byref(h)))
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().te_domain_destroy(self.h)
            self.h = None

    def shape(self, ncomp=1):
        return mesh.padded_shape(self.np, self.dim, self.ns, ncomp)

    def vector(self, ncomp=1, dtype=F64) -> DeviceArray:
        d = DeviceArray(self.shape(ncomp), np_dtype(dtype))
        d.zero()
        return d

    def nbr_local(self):
        out = np.zeros((self.np, 26, 4), dtype=np.int32)
        check(lib().te_domain_get_nbr_local(self.h, out.ctypes.data))
        return out

    # Vector<D> algebra -------------------------------------------------------
    def dot(self, a: DeviceArray, b: DeviceArray, stream=None):
        out = C.c_double()
        check(lib().te_vec_dot(self.h, _dt(a), a.ptr, b.ptr, a.shape[1], C.byref(out), stream))
        return out.value

    def two_norm(self, a: DeviceArray, stream=None):
        out = C.c_double()
        check(lib().te_vec_two_norm(self.h, _dt(a), a.ptr, a.shape[1], C.byref(out), stream))
        return out.value

    def inf_norm(self, a: DeviceArray, stream=None):
        out = C.c_double()
        check(lib().te_vec_inf_norm(self.h, _dt(a), a.ptr, a.shape[1], C.byref(out), stream))
        return out.value

    def boundary_length(self):
        """Length of the boundary data (te_domain_boundary_length)."""
        return int(lib().te_domain_boundary_length(self.h))

    def add_dirichlet_bc(self, f: DeviceArray, g: DeviceArray, c: DeviceArray | None = None, stream=None):
        """StarPatchOperator::addDrichletBCToRHS with sampled boundary data g (and the
        VarPoisson coefficient c)."""
        check(lib().te_star_add_dirichlet_bc(self.h, _dt(f), f.ptr, g.ptr, c.ptr if c is not None else None, stream))

    def add_neumann_bc(self, f: DeviceArray, dg: DeviceArray, stream=None):
        """StarPatchOperator::addNeumannBCToRHS with the sampled normal-axis derivative dg."""
        check(lib().te_star_add_neumann_bc(self.h, _dt(f), f.ptr, dg.ptr, stream))

    def integrate(self, u: DeviceArray, stream=None):
        """DomainTools::Integrate (DomainTools.h:513-536): interior values times cell
        volume, summed over patches and components."""
        out = C.c_double()
        check(lib().te_vec_integrate(self.h, _dt(u), u.ptr, u.shape[1], C.byref(out), stream))
        return out.value


def _dt(a: DeviceArray):
    return F64 if a.dtype == np.float64 else F32


class LevelLink:
    """``GMG::InterLevelComm`` for one rank (te_level_link)."""

    def __init__(self, fine: Domain, coarse: Domain):
        self.fine, self.coarse = fine, coarse
        h = C.c_void_p()
        check(lib().te_level_link_create(fine.h, coarse.h, C.byref(h)))
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().te_level_link_destroy(self.h)
            self.h = None

    def parents(self):
        out = np.zeros(self.fine.np, dtype=np.int32)
        check(lib().te_level_link_get_parents(self.h, out.ctypes.data))
        return out

    def restrict(self, fine: DeviceArray, coarse: DeviceArray, stream=None):
        check(lib().te_restrict(self.h, _dt(fine), fine.ptr, coarse.ptr, fine.shape[1], stream))

    def restrict_extrapolate(self, fine: DeviceArray, coarse: DeviceArray, stream=None):
        """LinearRestrictor::restrict with extrapolate_boundary_ghosts"""
        check(lib().te_restrict_extrapolate(self.h, _dt(fine), fine.ptr, coarse.ptr, fine.shape[1], stream))

    def residual_restrict(self, filler: "GhostFiller", f: DeviceArray, u: DeviceArray, fc: DeviceArray,
                          op_kind=0, neumann=False, coeffs: DeviceArray | None = None, stream=None):
        """fc = restriction of f - A u, one launch (u's dtype is the arithmetic type)."""
        check(lib().te_residual_restrict(filler.h, self.h, op_kind, int(neumann), coeffs.ptr if coeffs else None,
                                         _dt(u), _dt(f), f.ptr, u.ptr, _dt(fc), fc.ptr, stream))

    def interpolate(self, coarse: DeviceArray, fine: DeviceArray, stream=None):
        check(lib().te_interpolate(self.h, _dt(fine), coarse.ptr, fine.ptr, fine.shape[1], stream))


class GhostFiller:
    """BiLinear / BiQuadratic / TriLinear ghost filler (te_ghost_filler)."""

    def __init__(self, domain: Domain, kind=BILINEAR, fill_type=FACES):
        self.domain = domain
        self.kind, self.fill_type = kind, fill_type
        h = C.c_void_p()
        check(lib().te_ghost_filler_create(domain.h, kind, fill_type, C.byref(h)))
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().te_ghost_filler_destroy(self.h)
            self.h = None

    def fill_ghost(self, u: DeviceArray, stream=None):
        check(lib().te_ghost_fill(self.h, _dt(u), u.ptr, u.shape[1], stream))

    def star_apply(self, u: DeviceArray, f: DeviceArray, neumann=False, stream=None):
        check(lib().te_star_apply(self.h, _dt(u), int(neumann), u.ptr, f.ptr, u.shape[1], stream))

    def varstar_apply(self, coeffs: DeviceArray, u: DeviceArray, f: DeviceArray, stream=None):
        check(lib().te_varstar_apply(self.h, _dt(u), coeffs.ptr, u.ptr, f.ptr, u.shape[1], stream))

    def residual(self, f, u, r, op_kind=0, neumann=False, coeffs=None, stream=None):
        check(lib().te_residual(self.h, _dt(u), op_kind, int(neumann), coeffs.ptr if coeffs else None, f.ptr,
                                u.ptr, r.ptr, u.shape[1], stream))

    def patch_bicgstab_smooth(self, f, u, op_kind=0, neumann=False, coeffs=None, max_it=1000, tol=1e-12,
                              continue_on_breakdown=False, iters: DeviceArray | None = None, stream=None):
        check(lib().te_patch_bicgstab_smooth(self.h, _dt(u), op_kind, int(neumann),
                                             coeffs.ptr if coeffs else None, max_it, tol,
                                             int(continue_on_breakdown), f.ptr, u.ptr,
                                             iters.ptr if iters is not None else None, stream))


class DftPlan:
    """``Poisson::DFTPatchSolver`` (te_dft_plan)."""

    def __init__(self, filler: GhostFiller, neumann_mask=0, op_kind=0, coeffs: DeviceArray | None = None):
        self.filler = filler
        self.coeffs = coeffs
        h = C.c_void_p()
        check(lib().te_dft_plan_create(filler.h, neumann_mask, op_kind, coeffs.ptr if coeffs else None,
                                       C.byref(h)))
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().te_dft_plan_destroy(self.h)
            self.h = None

    def smooth(self, f: DeviceArray, u: DeviceArray, stream=None):
        check(lib().te_dft_smooth(self.h, _dt(u), f.ptr, u.ptr, stream))

    def apply(self, f: DeviceArray, u: DeviceArray, stream=None):
        check(lib().te_dft_apply(self.h, _dt(u), f.ptr, u.ptr, stream))

    def apply_mixed(self, dtype, f: DeviceArray, out: DeviceArray, stream=None):
        """From u = 0, computing in `dtype`; f and out keep their own dtypes."""
        check(lib().te_dft_apply_mixed(self.h, dtype, _dt(f), _dt(out), f.ptr, out.ptr, stream))

    def time_variant(self, f: DeviceArray, u: DeviceArray, out: DeviceArray, variant: int, reps=20):
        """Average milliseconds of the patch solve (ghosts read from u) with a launch variant
        forced (te_tune_time_dft); u's dtype is the arithmetic type."""
        ms = C.c_double()
        check(lib().te_tune_time_dft(self.h, _dt(u), f.ptr, u.ptr, out.ptr, variant, reps, C.byref(ms)))
        return ms.value

    def smooth_interp(self, link: "LevelLink", f: DeviceArray, u: DeviceArray, uc: DeviceArray, out: DeviceArray,
                      stream=None):
        """out = smooth(f, u + P uc), one launch (u's dtype is the arithmetic type)."""
        check(lib().te_dft_smooth_interp(self.h, link.h, _dt(u), _dt(f), _dt(out), f.ptr, u.ptr, _dt(uc), uc.ptr,
                                         out.ptr, stream))


# the autotuner's table (te_tune_*): families
TUNE_DFT_GENERIC, TUNE_DFT_MFMA_2D, TUNE_C3_SWEEP = 0, 1, 2


def tune_get(family, dim, n, dtype):
    return lib().te_tune_get(family, dim, n, dtype)


def tune_set(family, dim, n, dtype, variant):
    check(lib().te_tune_set(family, dim, n, dtype, variant))


def tune_num_variants(family):
    return lib().te_tune_num_variants(family)


def tune_load(path):
    check(lib().te_tune_load(str(path).encode()))


def tune_clear():
    check(lib().te_tune_clear())


def tune_table_path():
    buf = C.create_string_buffer(4096)
    lib().te_tune_table_path(buf, len(buf))
    return buf.value.decode()


def tune_family(dim, n, use_mfma=True):
    """the tuned family of the DFT patch solve a level of this shape runs (3D 16^3: the
    compact cycle's fused sweeps)"""
    if use_mfma and dim == 3 and n == 16:
        return TUNE_C3_SWEEP
    return TUNE_DFT_MFMA_2D if (use_mfma and dim == 2 and n in (16, 32)) else TUNE_DFT_GENERIC


class Stream:
    """A HIP stream at the default priority."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().te_stream_create(C.byref(h)))
        self.h = h.value

    def sync(self):
        check(lib().te_stream_synchronize(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().te_stream_destroy(self.h)
            self.h = None


class Event:
    """A HIP event; elapsed_ms(a, b) waits for b."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().te_event_create(C.byref(h)))
        self.h = h.value

    def record(self, stream=None):
        check(lib().te_event_record(self.h, stream.h if isinstance(stream, Stream) else stream))

    @staticmethod
    def elapsed_ms(a: "Event", b: "Event") -> float:
        ms = C.c_float()
        check(lib().te_event_elapsed_ms(a.h, b.h, C.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().te_event_destroy(self.h)
            self.h = None


TIMER_NONE = 2147483647  # no domain / no patch (the reference's INT_MAX)


class Timer:
    """``Timer`` (reference ``Timer.h:28-204``) over te_timer.

    Timings without a stream use the host clock, as the reference does; with
    ``stream=`` (a Stream, a raw handle, or 0 for the null stream) they are timed by
    HIP events recorded on that stream and read when the timer is serialized or
    printed.  Errors raise TEError with the reference's messages.
    """

    def __init__(self):
        h = C.c_void_p()
        check(lib().te_timer_create(C.byref(h)))
        self.h = h.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().te_timer_destroy(self.h)
            self.h = None

    @staticmethod
    def _s(stream):
        return stream.h if isinstance(stream, Stream) else stream

    def _go(self, begin, name, domain_id, patch_id, stream):
        n = name.encode()
        if stream is None:
            fn = lib().te_timer_start if begin else lib().te_timer_stop
            check(fn(self.h, n, domain_id, patch_id))
        else:
            fn = lib().te_timer_start_on if begin else lib().te_timer_stop_on
            check(fn(self.h, n, domain_id, patch_id, self._s(stream)))

    def add_domain(self, domain_id: int, domain):
        """Timer::addDomain: the domain's patches in the reference's mesh schema.  domain:
        a Domain, or anything with its dim, ns and patches (PATCH_DTYPE records)."""
        patches = np.ascontiguousarray(domain.patches, dtype=mesh.PATCH_DTYPE)
        nsa = (C.c_int * 3)(*(list(domain.ns) + [1] * (3 - domain.dim)))
        check(lib().te_timer_add_domain(self.h, domain_id, domain.dim, nsa, len(patches), patches.ctypes.data))

    def start(self, name, stream=None):
        self._go(True, name, TIMER_NONE, TIMER_NONE, stream)

    def stop(self, name, stream=None):
        self._go(False, name, TIMER_NONE, TIMER_NONE, stream)

    def start_domain_timing(self, domain_id, name, stream=None):
        self._go(True, name, domain_id, TIMER_NONE, stream)

    def stop_domain_timing(self, domain_id, name, stream=None):
        self._go(False, name, domain_id, TIMER_NONE, stream)

    def start_patch_timing(self, patch_id, domain_id, name, stream=None):
        self._go(True, name, domain_id, patch_id, stream)

    def stop_patch_timing(self, patch_id, domain_id, name, stream=None):
        self._go(False, name, domain_id, patch_id, stream)

    def add_int_info(self, name, info: int):
        check(lib().te_timer_add_int_info(self.h, name.encode(), int(info)))

    def add_double_info(self, name, info: float):
        check(lib().te_timer_add_double_info(self.h, name.encode(), float(info)))

    def _text(self, fn, *args) -> str:
        n = C.c_size_t()
        check(fn(self.h, *args, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(fn(self.h, *args, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def dumps(self, indent: int = -1) -> str:
        """The JSON text (to_json): compact, or indented by ``indent`` spaces."""
        return self._text(lib().te_timer_to_json, indent)

    def to_json(self):
        import json

        return json.loads(self.dumps())

    def report(self) -> str:
        """operator<<: the text report."""
        return self._text(lib().te_timer_print)

    __str__ = report

    def save_to_file(self, filename: str):
        check(lib().te_timer_save(self.h, filename.encode()))


def device_count():
    n = C.c_int(0)
    rc = lib().te_device_count(C.byref(n))
    return n.value if rc == 0 else 0
