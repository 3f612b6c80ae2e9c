// Geometric multigrid V-cycle and the outer BiCGStab on the device.
//
//  - GMG::CycleBuilder / Level / VCycle (GMG/CycleBuilder.h:97-192, GMG/Level.h,
//    GMG/VCycle.h:46-70) with Cycle::restrict (GMG/Cycle.h:57-65): every level owns
//    its ghost filler, smoother plan, link to the coarser level and preallocated
//    f / u / r vectors, so applying the cycle allocates nothing and the whole
//    cycle can be captured as one hipGraph.
//  - Mixed precision: the V-cycle runs in FP32 (ghost fill, stencil, smoother,
//    restrict, interpolate) except the coarsest level, which stays FP64; the
//    finest level reads the FP64 right-hand side and its last smoothing sweep
//    writes the FP64 result directly.
//  - Iterative::BiCGStab::solve (Iterative/BiCGStab.h:111-203) in FP64 with the
//    scalars on the device; the host reads ||s|| and ||r|| once each per iteration.
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <string>
#include <vector>

#include "kernels/blas.h"
#include "kernels/cycle3d.h"
#include "kernels/dft.h"
#include "kernels/krylov.h"
#include "kernels/tail2d.h"
#include "gmg_internal.h"
#include "te_internal.h"
#include "tune.h"

using namespace te;

namespace {

#define TE_RC(x)             \
  do {                       \
    int rc_ = (x);           \
    if (rc_) return rc_;     \
  } while (0)

int zero_with_ghost(te_gmg* g, int li, void* v, int dt)
{
  return launch_vec_op(g->L[li].d, dt, VOP_SET_WITH_GHOST, v, nullptr, nullptr, 0.0, 0, 0, 1, g->s);
}

// padded elements of one patch of level li
long patch_len(const te_gmg* g, int li)
{
  const Layout L = make_layout(g->L[li].d->dim, g->L[li].d->ns, 1, 1);
  return L.ps;
}

// batch mode: the right-hand side of the finest-level patches of level li (li >= 1) from the
// cycle's input, after the restriction from level li - 1 has written the rest of its f
int batch_scatter(te_gmg* g, int li)
{
  if (!g->batch || g->bseg[li].count == 0) return TE_OK;
  const te_gmg::BatchSeg& sg = g->bseg[li];
  GLevel& Lv = g->L[li];
  const long pl = patch_len(g, li);
  return launch_convert((long)sg.count * pl, Lv.dt, (char*)Lv.f + esize(Lv.dt) * (size_t)(sg.from * pl), g->bf_dt,
                        (const char*)g->bf + esize(g->bf_dt) * (size_t)(sg.foff * pl), g->s);
}

// start (stop = false) or stop the "Total Patch Smooth" timing of level li's sweep
// (PatchSolver.h:153-168) when a timer is attached and the level is timed
int sweep_timing(te_gmg* g, int li, bool stop)
{
  if (!g->timer || g->timer_ids.empty() || g->timer_ids[li] == TE_TIMER_NONE) return TE_OK;
  const char* name = "Total Patch Smooth";
  return stop ? timer_stop_stream(g->timer, name, g->timer_ids[li], g->s)
              : timer_start_stream(g->timer, name, g->timer_ids[li], g->s);
}

int smooth_sweep(te_gmg* g, int li, const void* f, int dtf, void* u, bool uzero, void* out, int dto);

// a sweep timing requested for the finest level: its first sweep then runs in the V-cycle
bool level0_timed(const te_gmg* g) { return g->timer && !g->timer_ids.empty() && g->timer_ids[0] != TE_TIMER_NONE; }

// per-level sweep timings requested for a level of the tail: it then runs as per-level launches
bool tail_timed(const te_gmg* g)
{
  if (!g->timer || g->timer_ids.empty() || g->tail_from < 0) return false;
  for (size_t l = (size_t)g->tail_from; l < g->timer_ids.size(); l++)
    if (g->timer_ids[l] != TE_TIMER_NONE) return true;
  return false;
}

// One smoothing sweep on level li: PatchSolver::smooth (PatchSolver.h:142-170).
// uzero: u is known to be zero (fresh coarse-level vector).  out/dto: where the
// new u goes (== u except for the finest level's last mixed-precision sweep).
int smooth(te_gmg* g, int li, const void* f, int dtf, void* u, bool uzero, void* out, int dto)
{
  TE_RC(sweep_timing(g, li, false));
  TE_RC(smooth_sweep(g, li, f, dtf, u, uzero, out, dto));
  return sweep_timing(g, li, true);
}

int smooth_sweep(te_gmg* g, int li, const void* f, int dtf, void* u, bool uzero, void* out, int dto)
{
  GLevel& Lv = g->L[li];
  if (Lv.smoother == 0) {
    if (!uzero) TE_RC(launch_ghost_fill(Lv.gf, Lv.dt, u, 1, g->s));
    return launch_dft(Lv.plan, Lv.dt, dtf, dto, f, u, out, uzero, g->s);
  }
  // Iterative::PatchSolver(BiCGStab): f and u share the level dtype here
  if (uzero) TE_RC(zero_with_ghost(g, li, u, Lv.dt));
  else TE_RC(launch_ghost_fill(Lv.gf, Lv.dt, u, 1, g->s));
  double tol = Lv.inner_tol;
  if (Lv.dt == TE_F32 && tol < 1e-6) tol = 1e-6;  // an FP32 inner solve cannot reach tighter
  TE_RC(launch_patch_bicgstab(Lv.d, Lv.dt, g->o.op_kind, g->o.neumann, Lv.coeffs, f, u, Lv.inner_max_it, tol,
                              nullptr, g->d_flags, g->s));
  if (out != u) TE_RC(launch_convert(te_domain_vector_length(Lv.d, 1), dto, out, Lv.dt, u, g->s));
  return TE_OK;
}

// VCycle::visit (GMG/VCycle.h:46-70), WCycle::visit (GMG/WCycle.h:50-84) and
// FMGCycle::visit / v_visit (GMG/FMGCycle.h:43-104), one launch per step.  u_zero: this
// level's u starts at zero (the reference's fresh getZeroClone); false for the second
// coarse visit of the W- and F-cycles, which keeps the first visit's coarse u as its
// initial guess, as the reference does.  vonly: FMGCycle::v_visit (one coarse visit).
int visit(te_gmg* g, int li, const void* f, int dtf, double* u_ext, bool u_zero = true, bool vonly = false)
{
  GLevel& Lv = g->L[li];
  const int nl = (int)g->L.size();
  const bool finest = li == 0;
  const bool mixed_out = finest && Lv.dt != TE_F64;  // result must be written to u_ext in FP64
  void* u = (finest && !mixed_out) ? (void*)u_ext : Lv.u;
  if (li == nl - 1) {
    int n = g->o.coarse_sweeps;
    if (n == 0) {
      if (u_zero) TE_RC(zero_with_ghost(g, li, u, Lv.dt));
      if (mixed_out) TE_RC(launch_convert(te_domain_vector_length(Lv.d, 1), TE_F64, u_ext, Lv.dt, u, g->s));
    }
    for (int k = 0; k < n; k++) {
      bool last = k == n - 1;
      TE_RC(smooth(g, li, f, dtf, u, u_zero && k == 0, (mixed_out && last) ? (void*)u_ext : u,
                   (mixed_out && last) ? TE_F64 : Lv.dt));
    }
    return TE_OK;
  }
  for (int k = 0; k < g->o.pre_sweeps; k++) TE_RC(smooth(g, li, f, dtf, u, u_zero && k == 0, u, Lv.dt));
  if (g->o.pre_sweeps == 0 && u_zero) TE_RC(zero_with_ghost(g, li, u, Lv.dt));
  GLevel& Cv = g->L[li + 1];
  // Cycle::restrict: r = f - A u, then restrict into the coarser level's f
  auto restrict_to_coarse = [&]() -> int {
    TE_RC(launch_ghost_fill(Lv.gf, Lv.dt, u, 1, g->s));
    TE_RC(launch_star(Lv.gf, Lv.dt, dtf, Lv.dt, g->o.op_kind, g->o.neumann, Lv.coeffs, f, u, Lv.r, 1, g->s));
    return launch_restrict(Lv.link, Lv.dt, Lv.r, Cv.dt, Cv.f, 1, g->s);
  };
  TE_RC(restrict_to_coarse());
  TE_RC(batch_scatter(g, li + 1));
  TE_RC(visit(g, li + 1, Cv.f, Cv.dt, u_ext, true, false));
  TE_RC(launch_interpolate(Lv.link, Cv.dt, Cv.u, Lv.dt, u, 1, g->s));
  if (g->o.cycle_type != TE_CYCLE_V && !vonly) {
    for (int k = 0; k < g->o.mid_sweeps; k++) TE_RC(smooth(g, li, f, dtf, u, false, u, Lv.dt));
    TE_RC(restrict_to_coarse());
    TE_RC(visit(g, li + 1, Cv.f, Cv.dt, u_ext, false, g->o.cycle_type == TE_CYCLE_F));
    TE_RC(launch_interpolate(Lv.link, Cv.dt, Cv.u, Lv.dt, u, 1, g->s));
  }
  int n = g->o.post_sweeps;
  for (int k = 0; k < n; k++) {
    bool last = k == n - 1;
    TE_RC(smooth(g, li, f, dtf, u, false, (mixed_out && last) ? (void*)u_ext : u,
                 (mixed_out && last) ? TE_F64 : Lv.dt));
  }
  if (mixed_out && n == 0) TE_RC(launch_convert(te_domain_vector_length(Lv.d, 1), TE_F64, u_ext, Lv.dt, u, g->s));
  return TE_OK;
}

// VCycle::visit on the fused path: per level, one launch per smoothing sweep and one
// for Cycle::restrict (ghost fill + residual + restriction); interpolation and the
// ghost fill of the first post-smoothing sweep happen inside that sweep's launch.
// Sweeps after the first pre-smoothing one read the solution from one buffer and
// write the other; the finest level's last sweep writes the caller's vector u_ext
// (type ext_dt: FP64, or FP32 for the mixed-precision BiCGStab's own buffers, which then
// hold the same values: that sweep computes in FP32).
// pre_done: this level's first sweep (from zero) was already done by the finer level's
// launch (launch_rr_pre), its result in Lv.u.
int visit_fused(te_gmg* g, int li, const void* f, int dtf, void* u_ext, int ext_dt, bool pre_done)
{
  GLevel& Lv = g->L[li];
  const int nl = (int)g->L.size();
  const bool finest = li == 0;
  void* buf[2] = { Lv.u, Lv.u2 };
  int cur = 0;
  // one sweep from buf[cur]: gm 0 = u known zero, 1 = ghosts on the fly, 2 = with interpolation
  auto sweep = [&](int mode, bool last_of_level) -> int {
    const bool to_ext = finest && last_of_level;
    void* out = to_ext ? u_ext : buf[cur ^ 1];
    int dto = to_ext ? ext_dt : Lv.dt;
    const void* uc = nullptr;
    int dtc = TE_F64;
    if (mode == 2) {
      uc = g->L[li + 1].cur;
      dtc = g->L[li + 1].dt;
    }
    TE_RC(sweep_timing(g, li, false));
    TE_RC(launch_dft_fused(Lv.plan, Lv.dt, dtf, dto, f, buf[cur], out, mode == 0, mode == 0 ? 0 : mode, dtc,
                                uc, mode == 2 ? Lv.link : nullptr, g->s));
    TE_RC(sweep_timing(g, li, true));
    if (to_ext) {
      Lv.cur = u_ext;
    } else {
      cur ^= 1;
      Lv.cur = buf[cur];
    }
    return TE_OK;
  };
  if (li == nl - 1) {
    int n = g->o.coarse_sweeps;
    if (n == 0) {
      TE_RC(zero_with_ghost(g, li, buf[0], Lv.dt));
      Lv.cur = buf[0];
      if (finest) TE_RC(launch_convert(te_domain_vector_length(Lv.d, 1), ext_dt, u_ext, Lv.dt, buf[0], g->s));
      return TE_OK;
    }
    for (int k = pre_done ? 1 : 0; k < n; k++) TE_RC(sweep(k == 0 ? 0 : 1, k == n - 1));
    return TE_OK;
  }
  for (int k = pre_done ? 1 : 0; k < g->o.pre_sweeps; k++) TE_RC(sweep(k == 0 ? 0 : 1, false));
  if (g->o.pre_sweeps == 0) {
    TE_RC(zero_with_ghost(g, li, buf[0], Lv.dt));
    Lv.cur = buf[0];
  }
  GLevel& Cv = g->L[li + 1];
  const int next_first = li + 1 == nl - 1 ? g->o.coarse_sweeps : g->o.pre_sweeps;
  if (li + 1 == g->tail_from && !tail_timed(g)) {
    // Cycle::restrict, then the V-cycle of the levels from li + 1 down in one launch
    TE_RC(launch_resid_restrict(Lv.gf, Lv.link, g->o.op_kind, g->o.neumann, Lv.coeffs, Lv.dt, dtf, f, buf[cur],
                                Cv.dt, Cv.f, g->s));
    TE_RC(launch_dft_tail(g->tail, Cv.dt, g->L[nl - 1].dt, g->o.filler_kind, g->s));
    Cv.cur = Cv.u;
  } else if (next_first > 0 && g->rr_pre[li]) {
    // Cycle::restrict and the coarser level's first sweep (from zero) in one launch
    TE_RC(sweep_timing(g, li + 1, false));
    TE_RC(launch_rr_pre(Lv.gf, Lv.link, g->o.op_kind, g->o.neumann, Lv.coeffs, Lv.dt, dtf, f, buf[cur], Cv.plan,
                        Cv.dt, Cv.f, Cv.u, g->s));
    TE_RC(sweep_timing(g, li + 1, true));
    Cv.cur = Cv.u;
    TE_RC(visit_fused(g, li + 1, Cv.f, Cv.dt, u_ext, ext_dt, true));
  } else {
    TE_RC(launch_resid_restrict(Lv.gf, Lv.link, g->o.op_kind, g->o.neumann, Lv.coeffs, Lv.dt, dtf, f, buf[cur],
                                Cv.dt, Cv.f, g->s));
    TE_RC(batch_scatter(g, li + 1));
    TE_RC(visit_fused(g, li + 1, Cv.f, Cv.dt, u_ext, ext_dt, false));
  }
  int n = g->o.post_sweeps;
  if (n == 0) {
    // DirectInterpolator alone (GMG/DirectInterpolator.h:60-91)
    TE_RC(launch_interpolate(Lv.link, Cv.dt, Cv.cur, Lv.dt, buf[cur], 1, g->s));
    Lv.cur = buf[cur];
    if (finest) TE_RC(launch_convert(te_domain_vector_length(Lv.d, 1), ext_dt, u_ext, Lv.dt, buf[cur], g->s));
    return TE_OK;
  }
  for (int k = 0; k < n; k++) TE_RC(sweep(k == 0 ? 2 : 1, k == n - 1));
  return TE_OK;
}

// the face columns of compact buffer v (np patches, type dt), or nullptr when the solver
// does not keep them
static void* xcol_of(const te_gmg* g, const void* v, long np, int dt)
{
  return g->xcol ? c3_xcol(v, np, dt) : nullptr;
}

// VCycle::visit (GMG/VCycle.h:46-70) on compact vectors (3D, cycle3d.h): per level one launch
// per smoothing sweep and one for Cycle::restrict (residual and restriction); the
// interpolation and the ghost values of every sweep are computed inside the sweeps'
// launches.  Sweeps after the first read the solution from one buffer and write the other;
// the finest level's last sweep writes u_ext (compact, type ext_dt).  Every buffer a sweep
// writes has face columns (cycle3d.h), written by the sweep and read by the launches after it.
int visit_c3(te_gmg* g, int li, const void* f, int dtf, void* u_ext, int ext_dt)
{
  GLevel& Lv = g->L[li];
  te_gmg::C3Level& Cl = g->C3[li];
  const int nl = (int)g->L.size();
  const bool finest = li == 0;
  const long ncell = (long)Lv.d->np * kC3Cells;
  void* buf[2] = { Cl.u, Cl.u2 };
  int cur = 0;
  auto sweep = [&](int mode, bool last_of_level) -> int {
    const bool to_ext = finest && last_of_level;
    void* out = to_ext ? u_ext : buf[cur ^ 1];
    const int dto = to_ext ? ext_dt : Lv.dt;
    const void* uc = mode == 2 ? g->C3[li + 1].cur : nullptr;
    const int dtc = mode == 2 ? g->L[li + 1].dt : TE_F64;
    const long np = Lv.d->np;
    TE_RC(sweep_timing(g, li, false));
    if (mode == 2 && Cl.u3 && g->finterp && g->xcol) {
      // the boundary layer of u + P uc first, then a mode-1 sweep on it (the same values)
      TE_RC(launch_c3_face_interp(Lv.link, Lv.dt, buf[cur], xcol_of(g, buf[cur], np, Lv.dt), dtc, uc, Cl.u3,
                                  xcol_of(g, Cl.u3, np, Lv.dt), g->s));
      TE_RC(launch_c3_sweep(Lv.plan, Lv.dt, dtf, f, Cl.u3, dto, out, 1, TE_F64, nullptr, nullptr, g->s,
                            xcol_of(g, Cl.u3, np, Lv.dt), xcol_of(g, out, np, dto)));
    } else {
      TE_RC(launch_c3_sweep(Lv.plan, Lv.dt, dtf, f, buf[cur], dto, out, mode, dtc, uc, mode == 2 ? Lv.link : nullptr,
                            g->s, xcol_of(g, buf[cur], np, Lv.dt), xcol_of(g, out, np, dto)));
    }
    TE_RC(sweep_timing(g, li, true));
    if (to_ext) {
      Cl.cur = u_ext;
    } else {
      cur ^= 1;
      Cl.cur = buf[cur];
    }
    return TE_OK;
  };
  auto zero_level = [&]() -> int {
    TE_RC(launch_c3_zero(ncell + (long)Lv.d->np * kC3XCol, Lv.dt, buf[0], g->s));  // cells and face columns
    cur = 0;
    Cl.cur = buf[0];
    return TE_OK;
  };
  if (li == nl - 1) {
    const int n = g->o.coarse_sweeps;
    if (n == 0) {
      TE_RC(zero_level());
      if (finest) TE_RC(launch_convert(ncell, ext_dt, u_ext, Lv.dt, buf[0], g->s));
      return TE_OK;
    }
    for (int k = 0; k < n; k++) TE_RC(sweep(k == 0 ? 0 : 1, k == n - 1));
    return TE_OK;
  }
  for (int k = 0; k < g->o.pre_sweeps; k++) TE_RC(sweep(k == 0 ? 0 : 1, false));
  if (g->o.pre_sweeps == 0) TE_RC(zero_level());
  GLevel& Cv = g->L[li + 1];
  TE_RC(launch_c3_resid_restrict(Lv.d, Lv.link, g->o.neumann, Lv.dt, dtf, f, Cl.cur, Cv.dt, g->C3[li + 1].f, g->s,
                                 xcol_of(g, Cl.cur, Lv.d->np, Lv.dt)));
  TE_RC(visit_c3(g, li + 1, g->C3[li + 1].f, Cv.dt, u_ext, ext_dt));
  const int n = g->o.post_sweeps;
  if (n == 0) {
    // DirectInterpolator alone (GMG/DirectInterpolator.h:60-91)
    TE_RC(launch_c3_interp(Lv.link, Cv.dt, g->C3[li + 1].cur, Lv.dt, Cl.cur, buf[cur ^ 1], g->s));
    cur ^= 1;
    Cl.cur = buf[cur];
    if (finest) {
      TE_RC(launch_convert(ncell, ext_dt, u_ext, Lv.dt, buf[cur], g->s));
      Cl.cur = u_ext;
    }
    return TE_OK;
  }
  for (int k = 0; k < n; k++) TE_RC(sweep(k == 0 ? 2 : 1, k == n - 1));
  return TE_OK;
}

// Cycle::apply on the caller's padded FP64 vectors through the compact cycle
int vcycle_c3_padded(te_gmg* g, const double* f, double* u)
{
  const GLevel& L0 = g->L[0];
  TE_RC(launch_c3_pack(L0.d, L0.dt, g->C3[0].f, TE_F64, f, g->s));
  TE_RC(visit_c3(g, 0, g->C3[0].f, L0.dt, g->c_tmp, L0.dt));
  return launch_c3_unpack(L0.d, TE_F64, u, L0.dt, g->c_tmp, g->s);
}

// Cycle::apply (GMG/Cycle.h:94-98) enqueued on g->s.  dtf / udt: types of f and u (FP32
// only on the fused mixed-precision path, where the V-cycle rounds f to FP32 as it reads it
// and its last sweep computes u in FP32; see visit_fused).
// l0_done: the finest level's zero-guess sweep already ran (launch_bicg_sweep), its result
// in L[0].u.
int vcycle_enqueue_one(te_gmg* g, const void* f, void* u, int udt, int dtf, bool l0_done);

int vcycle_enqueue(te_gmg* g, const void* f, void* u, int udt = TE_F64, int dtf = TE_F64, bool l0_done = false)
{
  if (!g->batch) return vcycle_enqueue_one(g, f, u, udt, dtf, l0_done);
  g->bf = f;
  g->bf_dt = dtf;
  int rc = vcycle_enqueue_one(g, f, u, udt, dtf, false);
  g->bf = nullptr;
  TE_RC(rc);
  // the results of the finest-level patches below level 0 (level 0's last sweep wrote u)
  for (size_t li = 1; li < g->L.size(); li++) {
    const te_gmg::BatchSeg& sg = g->bseg[li];
    if (sg.count == 0) continue;
    const GLevel& Lv = g->L[li];
    const void* src = g->fused ? Lv.cur : Lv.u;
    const long pl = patch_len(g, (int)li);
    TE_RC(launch_convert((long)sg.count * pl, udt, (char*)u + esize(udt) * (size_t)(sg.foff * pl), Lv.dt,
                         (const char*)src + esize(Lv.dt) * (size_t)(sg.from * pl), g->s));
  }
  return TE_OK;
}

int vcycle_enqueue_one(te_gmg* g, const void* f, void* u, int udt, int dtf, bool l0_done)
{
  GLevel& L0 = g->L[0];
  if (g->fused) return visit_fused(g, 0, f, dtf, u, udt, l0_done);
  if (g->c3) {
    if (udt != TE_F64 || dtf != TE_F64) {
      set_error("internal: the padded entry of the 3D compact cycle takes FP64 vectors");
      return TE_ERR_UNSUPPORTED;
    }
    return vcycle_c3_padded(g, (const double*)f, (double*)u);
  }
  if (udt != TE_F64 || dtf != TE_F64) {
    set_error("internal: FP32 V-cycle input or output needs the fused path");
    return TE_ERR_UNSUPPORTED;
  }
  double* ud = (double*)u;
  if (L0.dt != TE_F64 && L0.smoother != 0) {
    // the inner-Krylov smoother works on same-dtype f and u: convert f once
    TE_RC(launch_convert(te_domain_vector_length(L0.d, 1), L0.dt, L0.f, TE_F64, f, g->s));
    return visit(g, 0, L0.f, L0.dt, ud);
  }
  return visit(g, 0, f, TE_F64, ud);
}

int vcycle_launch(te_gmg* g, const double* f, double* u)
{
  if (!g->o.use_graph || g->timer) return vcycle_enqueue(g, f, u);
  const GraphCache::Key key{ (uintptr_t)f, (uintptr_t)u, 0 };
  hipGraphExec_t it = g->vgraphs.find(key);
  if (!it) {
    hipGraph_t graph;
    TE_HIP(hipStreamBeginCapture(g->s, hipStreamCaptureModeThreadLocal));
    int rc = vcycle_enqueue(g, f, u);
    hipError_t e = hipStreamEndCapture(g->s, &graph);
    if (rc) return rc;
    if (e != hipSuccess) {
      set_error(std::string("V-cycle graph capture: ") + hipGetErrorString(e));
      return TE_ERR_HIP;
    }
    hipGraphExec_t exec;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      set_error(std::string("V-cycle graph instantiate: ") + hipGetErrorString(e));
      return TE_ERR_HIP;
    }
    g->vgraphs.put(key, exec);
    it = exec;
  }
  TE_HIP(hipGraphLaunch(it, g->s));
  return TE_OK;
}

// The solver's stream waits for the work enqueued so far on `caller` (an event recorded
// there), without blocking the host.
int join_from(te_gmg* g, hipStream_t caller)
{
  if (caller == g->s) return TE_OK;
  TE_HIP(hipEventRecord(g->ev_in, caller));
  TE_HIP(hipStreamWaitEvent(g->s, g->ev_in, 0));
  return TE_OK;
}

// `caller` waits for the work enqueued so far on the solver's stream.
int join_to(te_gmg* g, hipStream_t caller)
{
  if (caller == g->s) return TE_OK;
  TE_HIP(hipEventRecord(g->ev_out, g->s));
  TE_HIP(hipStreamWaitEvent(caller, g->ev_out, 0));
  return TE_OK;
}

// a level smooths with the inner BiCGStab and a breakdown there must be reported: the
// breakdown flag has to be read back after the cycle
bool needs_breakdown_check(const te_gmg* g)
{
  if (g->o.continue_on_breakdown) return false;
  for (const GLevel& Lv : g->L)
    if (Lv.smoother == 1) return true;
  return false;
}

int check_flags(te_gmg* g)
{
  int32_t flag = 0;
  TE_HIP(hipMemcpyAsync(&flag, g->d_flags, sizeof(int32_t), hipMemcpyDeviceToHost, g->s));
  TE_HIP(hipStreamSynchronize(g->s));
  if (flag && !g->o.continue_on_breakdown) {
    TE_HIP(hipMemsetAsync(g->d_flags, 0, sizeof(int32_t), g->s));
    set_error("BiCGStab broke down in a patch smoother (rho was 0)");
    return TE_ERR_BREAKDOWN;
  }
  return TE_OK;
}

void free_gmg(te_gmg* g)
{
  if (!g) return;
  g->vgraphs.clear();
  g->bgraphs.clear();
  for (auto& Lv : g->L) {
    if (Lv.gf) te_ghost_filler_destroy(Lv.gf);
    if (Lv.plan) te_dft_plan_destroy(Lv.plan);
    if (Lv.link) te_level_link_destroy(Lv.link);
    (void)hipFree(Lv.f);
    (void)hipFree(Lv.u);
    (void)hipFree(Lv.r);
    (void)hipFree(Lv.u2);
    (void)hipFree(Lv.coeffs32);
    (void)hipFree(Lv.coeffs64_own);
  }
  double* vs[] = { g->resid, g->rhat, g->p, g->ap, g->as, g->sv, g->ms, g->mp, g->scal, g->pa, g->pb, g->pc };
  for (double* v : vs) (void)hipFree(v);
  (void)hipFree(g->mp32);
  (void)hipFree(g->ms32);
  (void)hipFree(g->p32);
  (void)hipFree(g->s32);
  (void)hipFree(g->d_flags);
  for (auto& c : g->C3) {
    (void)hipFree(c.f);
    (void)hipFree(c.u);
    (void)hipFree(c.u2);
    (void)hipFree(c.u3);
  }
  for (void* v : { (void*)g->c_r, (void*)g->c_rhat, (void*)g->c_p, (void*)g->c_ap, (void*)g->c_as, (void*)g->c_s,
                   (void*)g->c_x,
                   g->c_mp, g->c_ms, (void*)g->c_p32, (void*)g->c_s32, (void*)g->c_rhat32, g->c_tmp })
    (void)hipFree(v);
  for (void* v : g->c_mph) (void)hipFree(v);
  for (void* v : g->c_msh) (void)hipFree(v);
  (void)hipFree(g->c_hist);
  if (g->host_scal) (void)hipHostFree(g->host_scal);
  if (g->mirror_h) (void)hipHostFree(g->mirror_h);
  if (g->ev_in) (void)hipEventDestroy(g->ev_in);
  if (g->ev_out) (void)hipEventDestroy(g->ev_out);
  if (g->s) (void)hipStreamDestroy(g->s);
  delete g;
}

}  // namespace

int te::gmg_create(const te_domain* const* levels, int num_levels, const te_gmg_options* opts,
                   const double* const* coeffs, bool batch, te_gmg** out)
{
  if (!levels || num_levels < 1 || !opts || !out) {
    set_error("te_gmg_create: bad argument");
    return TE_ERR_ARG;
  }
  if (opts->precision != TE_F64 && opts->precision != TE_F32) {
    set_error("te_gmg_create: precision must be TE_F64 or TE_F32");
    return TE_ERR_ARG;
  }
  if (opts->op_kind == 1 && !coeffs) {
    set_error("te_gmg_create: VarPoisson needs per-level coefficients");
    return TE_ERR_ARG;
  }
  for (int i = 0; i < num_levels; i++) {
    const int sk = opts->level_smoother ? opts->level_smoother[i] : opts->smoother;
    if (sk != 0 && sk != 1) {
      set_error("te_gmg_create: smoother must be 0 (DFT) or 1 (BiCGStab patch solver)");
      return TE_ERR_ARG;
    }
    const int mi = opts->level_inner_max_it ? opts->level_inner_max_it[i] : opts->inner_max_it;
    if (sk == 1 && mi < 0) {
      set_error("te_gmg_create: bad inner iteration limit");
      return TE_ERR_ARG;
    }
  }
  if (opts->cycle_type < TE_CYCLE_V || opts->cycle_type > TE_CYCLE_F) {
    set_error("te_gmg_create: Unsupported Cycle type " + std::to_string(opts->cycle_type) +
              " (0 V, 1 W, 2 F as CycleBuilder accepts)");
    return TE_ERR_ARG;
  }
  if (opts->pre_sweeps < 0 || opts->post_sweeps < 0 || opts->coarse_sweeps < 0 || opts->mid_sweeps < 0) {
    set_error("te_gmg_create: sweep counts must be >= 0");
    return TE_ERR_ARG;
  }
  auto g = new te_gmg();
  g->batch = batch;
  if (batch) g->bseg.assign(num_levels, te_gmg::BatchSeg{});
  g->o = *opts;
  g->o.level_smoother = nullptr;  // the caller's arrays are not kept (copied into the levels)
  g->o.level_inner_max_it = nullptr;
  g->o.level_inner_tol = nullptr;
  hipError_t e = hipStreamCreate(&g->s);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_out, hipEventDisableTiming);
  if (e != hipSuccess) {
    set_error(std::string("te_gmg_create: ") + hipGetErrorString(e));
    free_gmg(g);
    return TE_ERR_HIP;
  }
  g->L.resize(num_levels);
  int rc = TE_OK;
  {
    // the compact 3D path: 16^3 patches with same-level face neighbours on every level, the
    // DFT smoother everywhere, the Poisson operator, a V-cycle (TE_NO_FUSE=1 or TE_DFT_VALU=1
    // keep the one-launch-per-step cycle, for comparison)
    const char* nf = std::getenv("TE_NO_FUSE");
    const char* valu = std::getenv("TE_DFT_VALU");
    bool c3 = !batch && !(nf && nf[0] == '1') && !(valu && valu[0] == '1') && opts->cycle_type == TE_CYCLE_V &&
              opts->op_kind == 0;
    for (int i = 0; i < num_levels && c3; i++) {
      const int sk = opts->level_smoother ? opts->level_smoother[i] : opts->smoother;
      c3 = sk == 0 && c3_supported(levels[i]);
    }
    g->c3 = c3;
    const char* nx = std::getenv("TE_NO_XCOL");
    g->xcol = !(nx && nx[0] == '1');
    const char* nxr = std::getenv("TE_NO_XRFLY");
    g->xrfly = !(nxr && nxr[0] == '1');
    const char* nr32 = std::getenv("TE_NO_RHAT32");
    g->rhat32 = !(nr32 && nr32[0] == '1');
    const char* np32 = std::getenv("TE_NO_P32ONLY");
    g->p32only = !(np32 && np32[0] == '1');
    const char* nxd = std::getenv("TE_NO_XDEFER");
    g->xdefer = !(nxd && nxd[0] == '1');
    const char* nxs = std::getenv("TE_XDEFER_SLOTS");
    if (nxs) g->xslots = std::max(1, std::min(kC3XSlots, std::atoi(nxs)));
    const char* nfi = std::getenv("TE_NO_FACEINTERP");
    g->finterp = !(nfi && nfi[0] == '1');
  }
  for (int i = 0; i < num_levels && !rc; i++) {
    GLevel& Lv = g->L[i];
    Lv.d = levels[i];
    bool coarsest = i == num_levels - 1;
    Lv.dt = (opts->precision == TE_F32 && !coarsest) ? TE_F32 : TE_F64;
    Lv.smoother = opts->level_smoother ? opts->level_smoother[i] : opts->smoother;
    Lv.inner_max_it = opts->level_inner_max_it ? opts->level_inner_max_it[i] : opts->inner_max_it;
    Lv.inner_tol = opts->level_inner_tol ? opts->level_inner_tol[i] : opts->inner_tol;
    rc = te_ghost_filler_create(Lv.d, opts->filler_kind, opts->fill_type, &Lv.gf);
    if (rc) break;
    long n = te_domain_vector_length(Lv.d, 1);
    if (g->c3) {
      // compact level vectors only (the padded ones serve the one-launch-per-step cycle); the
      // solutions u, u2 carry face columns
      g->C3.resize(num_levels);
      const size_t cf = esize(Lv.dt) * (size_t)std::max<long>((long)Lv.d->np * kC3Cells, 1);
      const size_t cb = esize(Lv.dt) * (size_t)std::max<long>((long)Lv.d->np * (kC3Cells + kC3XCol), 1);
      te_gmg::C3Level& Cl = g->C3[i];
      if ((e = hipMalloc(&Cl.f, cf)) != hipSuccess || (e = hipMalloc(&Cl.u, cb)) != hipSuccess ||
          (e = hipMalloc(&Cl.u2, cb)) != hipSuccess || (e = hipMemset(Cl.f, 0, cf)) != hipSuccess ||
          (e = hipMemset(Cl.u, 0, cb)) != hipSuccess || (e = hipMemset(Cl.u2, 0, cb)) != hipSuccess ||
          (!coarsest && Lv.d->np >= kC3InterpMin &&
           ((e = hipMalloc(&Cl.u3, cb)) != hipSuccess || (e = hipMemset(Cl.u3, 0, cb)) != hipSuccess))) {
        set_error(std::string("te_gmg_create: ") + hipGetErrorString(e));
        rc = TE_ERR_HIP;
        break;
      }
    } else {
      size_t bytes = esize(Lv.dt) * (size_t)(n > 0 ? n : 1);
      if ((e = hipMalloc(&Lv.f, bytes)) != hipSuccess || (e = hipMalloc(&Lv.u, bytes)) != hipSuccess ||
          (e = hipMalloc(&Lv.r, bytes)) != hipSuccess || (e = hipMalloc(&Lv.u2, bytes)) != hipSuccess) {
        set_error(std::string("te_gmg_create: ") + hipGetErrorString(e));
        rc = TE_ERR_HIP;
        break;
      }
      (void)hipMemset(Lv.f, 0, bytes);
      (void)hipMemset(Lv.u, 0, bytes);
      (void)hipMemset(Lv.r, 0, bytes);
      (void)hipMemset(Lv.u2, 0, bytes);
    }
    if (opts->op_kind == 1) {
      if (!coeffs[i]) {
        set_error("te_gmg_create: missing coefficients for a level");
        rc = TE_ERR_ARG;
        break;
      }
      // VarPoisson::StarPatchOperator fills the coefficient ghosts once (VarPoisson/StarPatchOperator.h:70)
      if ((e = hipMalloc(&Lv.coeffs64_own, sizeof(double) * (n > 0 ? n : 1))) != hipSuccess ||
          (e = hipMemcpyAsync(Lv.coeffs64_own, coeffs[i], sizeof(double) * n, hipMemcpyDeviceToDevice, g->s)) !=
            hipSuccess) {
        set_error(std::string("te_gmg_create: ") + hipGetErrorString(e));
        rc = TE_ERR_HIP;
        break;
      }
      if ((rc = launch_ghost_fill(Lv.gf, TE_F64, Lv.coeffs64_own, 1, g->s))) break;
      Lv.coeffs64 = Lv.coeffs64_own;
      if (Lv.dt == TE_F64) {
        Lv.coeffs = Lv.coeffs64;
      } else {
        if ((e = hipMalloc(&Lv.coeffs32, sizeof(float) * (n > 0 ? n : 1))) != hipSuccess) {
          set_error(std::string("te_gmg_create: ") + hipGetErrorString(e));
          rc = TE_ERR_HIP;
          break;
        }
        rc = launch_convert(n, TE_F32, Lv.coeffs32, TE_F64, Lv.coeffs64, g->s);
        Lv.coeffs = Lv.coeffs32;
      }
      if (!rc && (e = hipStreamSynchronize(g->s)) != hipSuccess) {
        set_error(std::string("te_gmg_create: ") + hipGetErrorString(e));
        rc = TE_ERR_HIP;
      }
    }
    if (Lv.smoother == 0 && !rc)
      rc = te_dft_plan_create(Lv.gf, opts->neumann_mask, opts->op_kind, opts->op_kind == 1 ? Lv.coeffs64 : nullptr,
                              &Lv.plan);
    if (i > 0 && !rc) rc = te_level_link_create(g->L[i - 1].d, Lv.d, &g->L[i - 1].link);
  }
  if (!rc) {
    // Fused launches where every level supports them; TE_NO_FUSE=1 keeps the
    // one-kernel-per-step path (same arithmetic, for comparison and timing).
    const char* nf = std::getenv("TE_NO_FUSE");
    const bool allow = !(nf && nf[0] == '1');
    // the fused launches implement the V-cycle; W- and F-cycles run one launch per step
    bool fused = allow && opts->cycle_type == TE_CYCLE_V;
    for (int i = 0; i < num_levels && fused; i++) {
      const GLevel& Lv = g->L[i];
      fused = Lv.smoother == 0 && Lv.plan &&
              ((Lv.plan->use_mfma && dft_mfma_supported(Lv.plan)) || dft_generic_fused_supported(Lv.plan)) &&
              (i == num_levels - 1 || resid_restrict_fused_supported(Lv.gf));
    }
    g->fused = fused;
    g->fly = allow && star_dot_fly_supported(g->L[0].gf);
    g->m32 = g->fused && g->fly && g->L[0].dt == TE_F32;
    const char* nr = std::getenv("TE_NO_RRPRE");
    const bool allow_rp = fused && !batch && !(nr && nr[0] == '1');
    g->rr_pre.assign(num_levels, 0);
    // Only where the coarser level is small: there the launch and its dependent loads
    // dominate; on large levels the separate launches keep all four waves of a workgroup
    // busy in the patch solves (one wave per coarse patch solves in the fused launch).
    constexpr int kRrPreMaxCoarse = 1024;
    for (int i = 0; i + 1 < num_levels && allow_rp; i++)
      g->rr_pre[i] = (rr_pre_supported(g->L[i].gf, g->L[i + 1].plan) &&
                      te_domain_num_patches(g->L[i + 1].d) <= kRrPreMaxCoarse)
                       ? 1
                       : 0;
    // the tail: the first level below the finest with at most kTailMaxPatches patches whose
    // levels (two or more) fit the tail's LDS; Poisson, 16 x 16 patches.  TE_NO_TAIL=1
    // keeps the per-level launches (same arithmetic, for comparison).
    const char* nt = std::getenv("TE_NO_TAIL");
    if (fused && !batch && opts->op_kind == 0 && !(nt && nt[0] == '1') && levels[0]->dim == 2 && levels[0]->ns[0] == 16 &&
        levels[0]->ns[1] == 16) {
      for (int t = 1; t + 1 < num_levels; t++) {
        const int nl = num_levels - t;
        if (nl > kTailMaxLevels || te_domain_num_patches(g->L[t].d) > kTailMaxPatches) continue;
        int np[kTailMaxLevels];
        for (int l = 0; l < nl; l++) np[l] = te_domain_num_patches(g->L[t + l].d);
        TailArgs a{};
        if (tail_layout(&a, np, nl, g->L[t].dt, g->L[num_levels - 1].dt) > kTailLds) continue;
        for (int l = 0; l < nl; l++) {
          const GLevel& Lv = g->L[t + l];
          a.lv[l].dp = Lv.d->d_patches;
          a.lv[l].ring = (const int4*)Lv.d->d_ring;
          a.lv[l].parent = Lv.link ? Lv.link->d_parent : nullptr;
          a.lv[l].porth = Lv.link ? Lv.link->d_orth : nullptr;
          a.lv[l].np = np[l];
        }
        a.nl = nl;
        a.pre = opts->pre_sweeps;
        a.post = opts->post_sweeps;
        a.coarse = opts->coarse_sweeps;
        a.neumann = opts->neumann;
        a.f_in = g->L[t].f;
        a.u_out = g->L[t].u;
        const te_dft_plan* pt = g->L[t].plan;
        const te_dft_plan* pl = g->L[num_levels - 1].plan;
        const bool t64 = g->L[t].dt == TE_F64, l64 = g->L[num_levels - 1].dt == TE_F64;
        a.mats_t = t64 ? (const void*)pt->d_mats64 : (const void*)pt->d_mats32;
        a.s2_t = t64 ? (const void*)pt->d_s2_64 : (const void*)pt->d_s2_32;
        a.mats_l = l64 ? (const void*)pl->d_mats64 : (const void*)pl->d_mats32;
        a.s2_l = l64 ? (const void*)pl->d_s2_64 : (const void*)pl->d_s2_32;
        if (!a.mats_t || !a.s2_t || !a.mats_l || !a.s2_l) break;
        g->tail = a;
        g->tail_from = t;
        break;
      }
    }
    const char* nb = std::getenv("TE_NO_BSWEEP");
    g->bsweep = !batch && g->m32 && num_levels > 1 && opts->pre_sweeps >= 1 &&
                bicg_sweep_supported(g->L[0].plan) && !(nb && nb[0] == '1');
  }
  if (!rc) {
    const te_domain* d0 = levels[0];
    long n = batch ? 1 : te_domain_vector_length(d0, 1);  // a batch owns its outer vectors
    size_t bytes = sizeof(double) * (size_t)(n > 0 ? n : 1);
    double** vs[] = { &g->resid, &g->rhat, &g->p, &g->ap, &g->as, &g->sv, &g->ms, &g->mp };
    for (double** v : vs) {
      if ((e = hipMalloc(v, bytes)) != hipSuccess) break;
      (void)hipMemset(*v, 0, bytes);
    }
    if (g->m32) {
      for (float** v : { &g->mp32, &g->ms32, &g->p32, &g->s32 }) {
        if (e == hipSuccess) e = hipMalloc(v, sizeof(float) * (size_t)(n > 0 ? n : 1));
        if (e == hipSuccess) e = hipMemset(*v, 0, sizeof(float) * (size_t)(n > 0 ? n : 1));
      }
    }
    g->nblk = krylov_blocks(d0);
    if (g->c3 && e == hipSuccess) {
      const size_t nc = (size_t)std::max<long>((long)d0->np * kC3Cells, 1);
      const size_t em = esize(g->L[0].dt);
      for (double** v : { &g->c_r, &g->c_rhat, &g->c_p, &g->c_ap, &g->c_as, &g->c_s, &g->c_x }) {
        if (e == hipSuccess) e = hipMalloc(v, sizeof(double) * nc);
        if (e == hipSuccess) e = hipMemset(*v, 0, sizeof(double) * nc);
      }
      // the cycle's outputs, with face columns (the outer operator reads M p and M s)
      const size_t ncx = nc + (size_t)d0->np * kC3XCol;
      for (void** v : { &g->c_mp, &g->c_ms, &g->c_tmp }) {
        if (e == hipSuccess) e = hipMalloc(v, em * ncx);
        if (e == hipSuccess) e = hipMemset(*v, 0, em * ncx);
      }
      if (g->L[0].dt == TE_F32) {
        for (float** v : { &g->c_p32, &g->c_s32, &g->c_rhat32 }) {
          if (e == hipSuccess) e = hipMalloc(v, sizeof(float) * nc);
          if (e == hipSuccess) e = hipMemset(*v, 0, sizeof(float) * nc);
        }
        // the deferred x updates' slots 1 .. xslots - 1 (with face columns, as c_mp / c_ms)
        if (g->xdefer && g->xrfly && g->rhat32) {
          for (int j = 1; j < g->xslots && e == hipSuccess; j++)
            for (std::vector<void*>* vv : { &g->c_mph, &g->c_msh }) {
              void* v = nullptr;
              if (e == hipSuccess) e = hipMalloc(&v, em * ncx);
              if (e == hipSuccess) {
                vv->push_back(v);
                e = hipMemset(v, 0, em * ncx);
              }
            }
          if (e == hipSuccess) e = hipMalloc(&g->c_hist, sizeof(double) * 2 * kC3XSlots);
          if (e == hipSuccess) e = hipMemset(g->c_hist, 0, sizeof(double) * 2 * kC3XSlots);
        }
      }
    }
    if (e == hipSuccess) e = hipMalloc(&g->scal, sizeof(double) * 16);
    if (e == hipSuccess) e = hipMalloc(&g->pa, sizeof(double) * (g->nblk + 1));
    if (e == hipSuccess) e = hipMalloc(&g->pb, sizeof(double) * (g->nblk + 1));
    if (e == hipSuccess) e = hipMalloc(&g->pc, sizeof(double) * (g->nblk + 1));
    if (e == hipSuccess) e = hipMalloc(&g->d_flags, sizeof(int32_t) * 4);
    if (e == hipSuccess) e = hipMemset(g->d_flags, 0, sizeof(int32_t) * 4);
    if (e == hipSuccess) e = hipMemset(g->scal, 0, sizeof(double) * 16);
    if (e == hipSuccess) e = hipHostMalloc(&g->host_scal, sizeof(double) * 16, hipHostMallocDefault);
    if (e == hipSuccess)
      e = hipHostMalloc(&g->mirror_h, sizeof(double) * 16, hipHostMallocCoherent | hipHostMallocMapped);
    if (e == hipSuccess) {
      std::memset(g->mirror_h, 0, sizeof(double) * 16);
      e = hipHostGetDevicePointer((void**)&g->mirror_d, g->mirror_h, 0);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g->s);
    if (e != hipSuccess) {
      set_error(std::string("te_gmg_create: ") + hipGetErrorString(e));
      rc = TE_ERR_HIP;
    }
  }
  if (rc) {
    free_gmg(g);
    return rc;
  }
  *out = g;
  return TE_OK;
}

void te::gmg_free(te_gmg* g) { free_gmg(g); }

int te::gmg_vcycle_enqueue(te_gmg* g, const void* f, void* u, int udt, int dtf)
{
  return vcycle_enqueue(g, f, u, udt, dtf, false);
}

extern "C" {

int te_gmg_create(const te_domain* const* levels, int num_levels, const te_gmg_options* opts,
                  const double* const* coeffs, te_gmg** out)
{
  return gmg_create(levels, num_levels, opts, coeffs, false, out);
}

int te_gmg_destroy(te_gmg* g)
{
  free_gmg(g);
  return TE_OK;
}

int te_gmg_vcycle(te_gmg* g, const double* f, double* u, te_stream stream)
{
  if (!g || !f || !u) {
    set_error("te_gmg_vcycle: bad argument");
    return TE_ERR_ARG;
  }
  hipStream_t cs = (hipStream_t)stream;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (cs) TE_HIP(hipStreamIsCapturing(cs, &st));
  if (st == hipStreamCaptureStatusActive) {
    // the caller captures its stream into a graph: the cycle's launches go straight onto it
    if (needs_breakdown_check(g)) {
      set_error("te_gmg_vcycle: a cycle whose BiCGStab patch smoother stops on breakdown reads its breakdown "
                "flag back and cannot be captured (use continue_on_breakdown)");
      return TE_ERR_UNSUPPORTED;
    }
    hipStream_t own = g->s;
    g->s = cs;
    int rc = vcycle_enqueue(g, f, u);
    g->s = own;
    return rc;
  }
  TE_RC(join_from(g, cs));
  TE_RC(vcycle_launch(g, f, u));
  TE_RC(join_to(g, cs));
  if (!needs_breakdown_check(g)) return TE_OK;
  TE_HIP(hipStreamSynchronize(g->s));
  return check_flags(g);
}

// Iterative::BiCGStab::solve (Iterative/BiCGStab.h:111-203)
int te_gmg_bicgstab(te_gmg* g, const double* b, double* x, int use_precond, int max_it, double tol, int* iterations,
                    double* history, int history_cap, te_stream stream)
{
  if (!g || !b || !x) {
    set_error("te_gmg_bicgstab: bad argument");
    return TE_ERR_ARG;
  }
  TE_RC(join_from(g, (hipStream_t)stream));
  GLevel& L0 = g->L[0];
  const te_domain* d = L0.d;
  te_ghost_filler* gf = L0.gf;
  const double* cf = g->o.op_kind == 1 ? L0.coeffs64 : nullptr;  // the outer operator is FP64
  const int nb = g->nblk;
  hipStream_t s = g->s;
  double* h = g->host_scal;
  auto read_scal = [&]() -> int {
    TE_HIP(hipMemcpyAsync(h, g->scal, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, s));
    TE_HIP(hipStreamSynchronize(s));
    return TE_OK;
  };
  // M p and M s: FP64, or FP32 on the fused mixed-precision path (the same values)
  const bool m32 = use_precond && g->m32;
  const int dtm = m32 ? TE_F32 : TE_F64;
  void* mp = use_precond ? (m32 ? (void*)g->mp32 : (void*)g->mp) : (void*)g->p;
  void* ms = use_precond ? (m32 ? (void*)g->ms32 : (void*)g->ms) : (void*)g->sv;
  // One iteration, enqueued as a whole: the host reads the scalars once per iteration,
  // from the copy FIN_RHO publishes (so the next iteration is enqueued while the p update
  // still runs).
  //   mp = M p; ap = A mp (+ rhat.ap); alpha (folded by the s update's launch);
  //   s = r - alpha ap (+ s.s);
  //   ms = M s; as = A ms (+ as.s, as.as); ||s|| and the s-norm test, then omega; if the
  //   test fired: x += alpha p (mp with use_precond = 2), and the updates after it are
  //   skipped on the device (the kernels read the state word); else x, r updates; rho,
  //   beta, ||r||; p update.
  // The preconditioner and operator launches for s run before the test (the reference runs
  // them after it): when it fires their results are unused; that happens at most once per
  // solve.
  // the s and p updates also run the finest level's zero-guess sweep of the V-cycle that
  // follows them (the first iteration's comes from a separate sweep after the setup)
  const bool l0f = m32 && g->bsweep && !level0_timed(g);
  float* u0 = (float*)L0.u;
  // the compact 3D path: the same iteration on compact work vectors (cycle3d.h); x and b
  // keep the caller's padded layout.  M p and M s are of the finest level's type (the last
  // sweep computes in it), and in mixed precision the cycle's right-hand sides are FP32
  // copies of p and s written by the p and s updates.
  const int c3m = use_precond ? L0.dt : TE_F64;
  void* c3_mp = use_precond ? g->c_mp : (void*)g->c_p;
  void* c3_ms = use_precond ? g->c_ms : (void*)g->c_s;
  const bool c3_32 = use_precond && L0.dt == TE_F32;
  // the face columns of M p and M s: written when the cycle's last finest-level step is a sweep
  const bool c3_xc = use_precond && (g->L.size() == 1 ? g->o.coarse_sweeps > 0 : g->o.post_sweeps > 0);
  const void* c3_xmp = c3_xc ? xcol_of(g, c3_mp, d->np, c3m) : nullptr;
  const void* c3_xms = c3_xc ? xcol_of(g, c3_ms, d->np, c3m) : nullptr;
  // the updates form A M p from M p (M p is a buffer of its own when preconditioned)
  const bool apf = g->xrfly && use_precond;
  // mixed precision: rhat as an FP32 copy of r0.  BiCGStab uses rhat only in the inner
  // products rhat.ap and rhat.r, both formed from the same stored (widened) values, so this
  // is the same iteration with the shadow residual fl32(r0) in place of r0 (BiCGStab.h:126,132
  // sets rhat = r0; any fixed rhat with rhat.r0 != 0 serves); the recurrences stay FP64.
  const bool c3_r32 = c3_32 && g->xrfly && g->rhat32 && g->c_rhat32;
  // mixed precision: the x update deferred.  x is read only at the end of a solve, so the
  // x / r update leaves it alone and keeps alpha and omega; iteration i writes M p and M s
  // into slot i % K, and launch_c3_xcomb adds x += alpha M p + omega M s for the slots in
  // iteration order (each term as the x / r update would add it) at the end of slot K - 1's
  // iteration and after the last iteration.  The r recurrence is untouched.
  const bool xdf = c3_r32 && g->xdefer && g->c_hist != nullptr;
  // mixed precision: p kept only as its FP32 copy (the cycle's input).  p enters the
  // iteration only through M p (already applied to fl32(p)) and its own recurrence, which
  // then rounds to FP32 each iteration: a perturbation of the size the cycle's input
  // rounding already has; x, r and s stay FP64.
  const bool p32 = c3_r32 && g->p32only;
  const int K = xdf ? g->xslots : 1;
  auto slot_mp = [&](int j) -> void* { return j > 0 ? g->c_mph[j - 1] : c3_mp; };
  auto slot_ms = [&](int j) -> void* { return j > 0 ? g->c_msh[j - 1] : c3_ms; };
  auto xcomb = [&](int count, int drop_early) -> int {
    C3XComb c{};
    c.x = x;  // the caller's padded x: no compact copy of x on this path
    c.padded = 1;
    for (int j = 0; j < count; j++) {
      c.mp[j] = (const float*)slot_mp(j);
      c.ms[j] = (const float*)slot_ms(j);
    }
    c.hist = g->c_hist;
    c.scal = g->scal;
    c.count = count;
    c.drop_early = drop_early;
    return launch_c3_xcomb(d, c, s);
  };
  auto iteration_c3 = [&](int slot) -> int {
    // this iteration's M p and M s (slot 0: c3_mp, c3_ms) and their face columns
    void* smp = xdf ? slot_mp(slot) : c3_mp;
    void* sms = xdf ? slot_ms(slot) : c3_ms;
    const void* sxmp = c3_xc ? xcol_of(g, smp, d->np, c3m) : nullptr;
    const void* sxms = c3_xc ? xcol_of(g, sms, d->np, c3m) : nullptr;
    C3Bicg a{};
    a.scal = g->scal;
    a.x = g->c_x;
    a.r = g->c_r;
    a.rhat = g->c_rhat;
    a.p = g->c_p;
    a.sv = g->c_s;
    a.ap = g->c_ap;
    a.as = g->c_as;
    a.mp = smp;
    a.ms = sms;
    a.dtm = c3m;
    if (use_precond)
      TE_RC(visit_c3(g, 0, c3_32 ? (const void*)g->c_p32 : (const void*)g->c_p, c3_32 ? TE_F32 : TE_F64, smp, c3m));
    // rhat.ap (A M p itself is formed again by the updates that use it, launch_c3_upd, where
    // M p is apart from p; else stored)
    TE_RC(launch_c3_star_dot(d, g->o.neumann, c3m, smp, apf ? nullptr : g->c_ap, g->c_rhat, g->pa, nullptr, nb, s,
                             sxmp, c3_r32 ? g->c_rhat32 : nullptr));
    // s = r - alpha ap, alpha folded from the rhat.ap partials; s.s partials into pc
    a.w32 = c3_32 ? g->c_s32 : nullptr;
    a.pa = g->pc;
    a.pfold = g->pa;
    a.nfold = nb;
    if (g->xrfly) TE_RC(launch_c3_upd(BICG_S, d, g->o.neumann, a, apf, sxmp, nullptr, nb, s));
    else TE_RC(launch_c3_bicg(BICG_S, d, a, nb, s));
    if (use_precond)
      TE_RC(visit_c3(g, 0, c3_32 ? (const void*)g->c_s32 : (const void*)g->c_s, c3_32 ? TE_F32 : TE_F64, sms, c3m));
    // as.s and as.as (A M s itself is formed again by the x / r update unless xrfly is off)
    TE_RC(launch_c3_star_dot(d, g->o.neumann, c3m, sms, g->xrfly ? nullptr : g->c_as, g->c_s, g->pa, g->pb, nb, s,
                             sxms));
    TE_RC(launch_bicg_finalize(FIN_SOMEGA, g->pa, g->pb, nb, g->scal, s, 0.0, nullptr, g->pc));
    // x and r updates; after the s-norm test fired, x += alpha p (use_precond 2: alpha M p)
    a.w32 = nullptr;
    a.pfold = nullptr;
    a.nfold = 0;
    a.pa = g->pa;
    a.pb = g->pb;
    a.dir = use_precond == 2 ? (const void*)smp : (p32 ? (const void*)g->c_p32 : (const void*)g->c_p);
    a.dir_m = use_precond == 2 || p32 ? 1 : 0;
    a.rh32 = c3_r32 ? g->c_rhat32 : nullptr;
    a.hist = xdf ? g->c_hist + 2 * slot : nullptr;
    a.xpo = xdf ? x : nullptr;
    if (g->xrfly) TE_RC(launch_c3_upd(BICG_XR, d, g->o.neumann, a, apf, sxmp, sxms, nb, s));
    else TE_RC(launch_c3_bicg(BICG_XR, d, a, nb, s));
    TE_RC(launch_bicg_finalize(FIN_RHO, g->pa, g->pb, nb, g->scal, s, 0.0, g->mirror_d));
    a.dir = nullptr;
    a.rh32 = nullptr;
    a.hist = nullptr;
    if (p32) a.p = nullptr;  // p from and to c_p32 only
    a.w32 = c3_32 ? g->c_p32 : nullptr;
    if (g->xrfly) TE_RC(launch_c3_upd(BICG_P, d, g->o.neumann, a, apf, sxmp, nullptr, nb, s));
    else TE_RC(launch_c3_bicg(BICG_P, d, a, nb, s));
    // the last slot: the deferred x updates of the K iterations (less this one's when its
    // s-norm test fired: it added alpha dir to x itself)
    if (xdf && slot == K - 1) return xcomb(K, 1);
    return TE_OK;
  };
  auto iteration = [&](int slot) -> int {
    if (g->c3) return iteration_c3(slot);
    if (use_precond) TE_RC(vcycle_enqueue(g, m32 ? (const void*)g->p32 : g->p, mp, dtm, dtm, l0f));
    if (!g->fly) TE_RC(launch_ghost_fill(gf, TE_F64, mp, 1, s));
    TE_RC(launch_star_dot(gf, g->o.op_kind, g->o.neumann, cf, mp, dtm, g->ap, g->rhat, g->pa, nullptr, g->fly, s));
    // the s update folds alpha's partials itself (no FIN_ALPHA launch)
    if (l0f)
      TE_RC(launch_bicg_sweep(BICG_S, L0.plan, g->scal, g->sv, g->resid, g->ap, g->s32, u0, g->pc, nb, s, g->pa, nb));
    else
      TE_RC(launch_bicg(BICG_S, d, g->scal, g->sv, nullptr, g->resid, g->ap, nullptr, nullptr, nullptr, g->pc,
                        nullptr, s, TE_F64, m32 ? g->s32 : nullptr, nullptr, 0, g->pa, nb));
    if (use_precond) TE_RC(vcycle_enqueue(g, m32 ? (const void*)g->s32 : g->sv, ms, dtm, dtm, l0f));
    if (!g->fly) TE_RC(launch_ghost_fill(gf, TE_F64, ms, 1, s));
    TE_RC(launch_star_dot(gf, g->o.op_kind, g->o.neumann, cf, ms, dtm, g->as, g->sv, g->pa, g->pb, g->fly, s));
    // ||s|| and the s-norm test (its partials from the s update), then omega
    TE_RC(launch_bicg_finalize(FIN_SOMEGA, g->pa, g->pb, nb, g->scal, s, 0.0, nullptr, g->pc));
    // x and r updates; after the test fired, x += alpha p instead (BiCGStab.h:162-167;
    // use_precond = 2: alpha M p)
    const bool xp_m = use_precond == 2;
    TE_RC(launch_bicg(BICG_XR, d, g->scal, x, g->resid, mp, ms, g->ap, g->as, g->rhat, g->pa, g->pb, s, dtm, nullptr,
                      xp_m ? mp : (const void*)g->p, xp_m ? 1 : 0));
    TE_RC(launch_bicg_finalize(FIN_RHO, g->pa, g->pb, nb, g->scal, s, 0.0, g->mirror_d));
    if (l0f) return launch_bicg_sweep(BICG_P, L0.plan, g->scal, g->p, g->ap, g->resid, g->p32, u0, nullptr, 0, s);
    return launch_bicg(BICG_P, d, g->scal, g->p, nullptr, g->ap, g->resid, nullptr, nullptr, nullptr, nullptr,
                       nullptr, s, TE_F64, m32 ? g->p32 : nullptr);
  };
  hipGraphExec_t gxs[kC3XSlots] = {};  // one iteration graph per x-update slot
  for (int slot = 0; slot < K && g->o.use_graph && !g->timer; slot++) {
    const GraphCache::Key key{ (uintptr_t)x, (uintptr_t)use_precond + 16 * (uintptr_t)slot, 1 };
    hipGraphExec_t& gx = gxs[slot];
    gx = g->bgraphs.find(key);
    if (!gx) {
      hipGraph_t graph;
      TE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int rc = iteration(slot);
      hipError_t e = hipStreamEndCapture(s, &graph);
      if (rc) return rc;
      if (e != hipSuccess) {
        set_error(std::string("BiCGStab graph capture: ") + hipGetErrorString(e));
        return TE_ERR_HIP;
      }
      hipGraphExec_t ex;
      e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        set_error(std::string("BiCGStab graph instantiate: ") + hipGetErrorString(e));
        return TE_ERR_HIP;
      }
      g->bgraphs.put(key, ex);
      gx = ex;
    }
  }
  const bool time_its = g->timer && g->time_its;
  auto t0 = std::chrono::steady_clock::now();
  int rc = TE_OK;
  int num_its = 0;
  bool early = false;
  // resid = b - A x ; r0_norm = ||b|| ; rhat = p = resid ; rho = rhat.resid (on the compact
  // path in one launch from the padded b and x, launch_c3_init)
  const bool c3_init = g->c3 && g->xrfly;
  if (!c3_init) {
    TE_RC(launch_ghost_fill(gf, TE_F64, x, 1, s));
    TE_RC(launch_star(gf, TE_F64, TE_F64, TE_F64, g->o.op_kind, g->o.neumann, cf, b, x, g->resid, 1, s));
  }
  if (g->c3) {
    C3Bicg a{};
    a.scal = g->scal;
    a.b = b;
    a.rpad = g->resid;
    a.x = g->c_x;
    a.xpad = x;
    a.r = g->c_r;
    a.rhat = g->c_rhat;
    a.p = g->c_p;
    a.w32 = c3_32 ? g->c_p32 : nullptr;
    a.rh32 = c3_r32 ? g->c_rhat32 : nullptr;  // c3_r32 implies c3_init
    if (p32) a.p = nullptr;
    if (xdf) a.x = nullptr;  // x deferred: the updates go to the caller's x itself
    a.pr = g->pc;
    a.pa = g->pa;
    a.pb = g->pb;
    a.dtm = TE_F64;
    if (c3_init) TE_RC(launch_c3_init(d, g->o.neumann, a, nb, s));
    else TE_RC(launch_c3_bicg(BICG_INIT, d, a, nb, s));
  } else {
    TE_RC(launch_bicg(BICG_INIT, d, g->scal, g->rhat, g->p, g->resid, b, nullptr, nullptr, nullptr, g->pa, g->pb, s,
                      TE_F64, m32 ? g->p32 : nullptr));
  }
  TE_RC(launch_bicg_finalize(FIN_INIT, g->pa, g->pb, nb, g->scal, s, tol, nullptr, c3_r32 ? g->pc : nullptr));
  // the first V-cycle's zero-guess sweep on the finest level (the later ones run in the p
  // updates)
  if (l0f) TE_RC(launch_dft_mfma_fused(L0.plan, TE_F32, TE_F32, TE_F32, g->p32, u0, u0, true, 0, TE_F64, nullptr,
                                       nullptr, s));
  TE_RC(read_scal());
  double seq = h[SC_SEQ];
  // wait for the iteration with sequence number `want` to publish its scalars into h
  auto wait_published = [&](double want) -> int {
    volatile double* m = g->mirror_h;
    for (unsigned spin = 1;; spin++) {
      if (m[SC_SEQ] == want) break;
      if ((spin & 255) == 0) {
        hipError_t q = hipStreamQuery(s);
        if (q == hipSuccess && m[SC_SEQ] == want) break;
        if (q == hipSuccess) {
          set_error("BiCGStab: the iteration finished without publishing its scalars");
          return TE_ERR_HIP;
        }
        if (q != hipErrorNotReady) {
          set_error(std::string("BiCGStab iteration: ") + hipGetErrorString(q));
          return TE_ERR_HIP;
        }
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    for (int k = 0; k < SC_COUNT; k++) h[k] = m[k];
    return TE_OK;
  };
  double r0 = h[SC_R0];
  double residual = h[SC_RNORM] / r0;
  double rho = h[SC_RHO];
  if (history && history_cap > 0) history[0] = residual;
  if (r0 != 0.0) {
    while (residual > tol && num_its < max_it) {
      if (rho == 0.0) {
        set_error("BiCGStab broke down, rho was 0 on iteration " + std::to_string(num_its));
        rc = TE_ERR_BREAKDOWN;
        break;
      }
      if (time_its) TE_RC(timer_start_stream(g->timer, "Iteration", TE_TIMER_NONE, s));
      const int slot = num_its % K;
      if (gxs[slot]) TE_HIP(hipGraphLaunch(gxs[slot], s));
      else TE_RC(iteration(slot));
      if (time_its) TE_RC(timer_stop_stream(g->timer, "Iteration", TE_TIMER_NONE, s));
      seq += 1.0;
      TE_RC(wait_published(seq));
      if (h[SC_STATE] == ST_EARLY) {  // x was updated on the device
        early = true;
        break;
      }
      num_its++;
      rho = h[SC_RHO];
      residual = h[SC_RNORM] / r0;
      if (history && num_its < history_cap) history[num_its] = residual;
    }
  }
  if (xdf) {  // the deferred x updates not yet added (slot K - 1's iteration adds its group's)
    int pend = num_its % K;
    if (early && pend == K - 1) pend = 0;
    if (pend) TE_RC(xcomb(pend, 0));
  }
  if (g->c3 && !xdf) TE_RC(launch_c3_unpack(d, TE_F64, x, TE_F64, g->c_x, s));  // the compact x back
  TE_HIP(hipStreamSynchronize(s));
  auto t1 = std::chrono::steady_clock::now();
  g->last_seconds = std::chrono::duration<double>(t1 - t0).count();
  if (iterations) *iterations = num_its;
  if (rc) return rc;
  return check_flags(g);
}

// Iterative::CG::solve (Iterative/CG.h:126-194): right-preconditioned CG, ap = A (M p)
// (Solver.h, applyWithPreconditioner), x accumulates from zero and becomes M x plus the
// initial guess at the end.  The scalars go to the host every iteration (the path is not
// the headline one; it reuses the vector, operator and cycle launches).
int te_gmg_cg(te_gmg* g, const double* b, double* x, int use_precond, int max_it, double tol, int* iterations,
              double* history, int history_cap, te_stream stream)
{
  if (!g || !b || !x) {
    set_error("te_gmg_cg: bad argument");
    return TE_ERR_ARG;
  }
  TE_RC(join_from(g, (hipStream_t)stream));
  GLevel& L0 = g->L[0];
  const te_domain* d = L0.d;
  te_ghost_filler* gf = L0.gf;
  const double* cf = g->o.op_kind == 1 ? L0.coeffs64 : nullptr;
  hipStream_t s = g->s;
  double* resid = g->resid;
  double* x0 = g->rhat;  // the initial guess
  double* p = g->p;
  double* ap = g->ap;
  double* tmp = g->mp;
  auto vop = [&](int op, double* v, const double* a, double alpha) {
    return launch_vec_op(d, TE_F64, op, v, a, nullptr, alpha, 0.0, 0.0, 1, s);
  };
  // dot product (RED_DOT) or two-norm (RED_SUMSQ: the launch returns the square root)
  auto reduce = [&](int op, const double* a, const double* bv, double* out) -> int {
    TE_RC(launch_reduce(d, TE_F64, op, a, bv, 1, d->d_result, s));
    TE_HIP(hipMemcpyAsync(out, d->d_result, sizeof(double), hipMemcpyDeviceToHost, s));
    TE_HIP(hipStreamSynchronize(s));
    return TE_OK;
  };
  auto apply_a = [&](double* u, double* out) -> int {  // A.apply (PatchOperator.h:126-143)
    TE_RC(launch_ghost_fill(gf, TE_F64, u, 1, s));
    return launch_star(gf, TE_F64, TE_F64, TE_F64, g->o.op_kind, g->o.neumann, cf, nullptr, u, out, 1, s);
  };
  const bool time_its = g->timer && g->time_its;
  auto t0 = std::chrono::steady_clock::now();
  int num_its = 0, rc = TE_OK;
  // resid = b - A x (A.apply then scaleThenAdd(-1, b), as one residual launch)
  TE_RC(launch_ghost_fill(gf, TE_F64, x, 1, s));
  TE_RC(launch_star(gf, TE_F64, TE_F64, TE_F64, g->o.op_kind, g->o.neumann, cf, b, x, resid, 1, s));
  TE_RC(vop(VOP_COPY, x0, x, 0.0));    // initial_guess = x
  TE_RC(vop(VOP_SET, x, nullptr, 0.0));  // x.set(0)
  double r0 = 0.0, rho = 0.0, nr = 0.0;
  TE_RC(reduce(RED_SUMSQ, b, nullptr, &r0));  // b.twoNorm()
  TE_RC(vop(VOP_COPY, p, resid, 0.0));
  TE_RC(reduce(RED_DOT, resid, resid, &rho));
  if (r0 != 0.0) {
    TE_RC(reduce(RED_SUMSQ, resid, nullptr, &nr));
    double residual = nr / r0;
    if (history && history_cap > 0) history[0] = residual;
    while (residual > tol && num_its < max_it) {
      if (rho == 0.0) {
        set_error("CG broke down, rho was 0 on iteration " + std::to_string(num_its));
        rc = TE_ERR_BREAKDOWN;
        break;
      }
      if (time_its) TE_RC(timer_start_stream(g->timer, "Iteration", TE_TIMER_NONE, s));
      if (use_precond) {
        TE_RC(vcycle_enqueue(g, p, tmp));
        TE_RC(apply_a(tmp, ap));
      } else {
        TE_RC(apply_a(p, ap));
      }
      double pap = 0.0;
      TE_RC(reduce(RED_DOT, p, ap, &pap));
      const double alpha = rho / pap;
      TE_RC(vop(VOP_ADD_SCALED, x, p, alpha));         // x.addScaled(alpha, p)
      TE_RC(vop(VOP_ADD_SCALED, resid, ap, -alpha));   // resid.addScaled(-alpha, ap)
      double rho_new = 0.0;
      TE_RC(reduce(RED_DOT, resid, resid, &rho_new));
      const double beta = rho_new / rho;
      TE_RC(vop(VOP_SCALE_THEN_ADD, p, resid, beta));  // p.scaleThenAdd(beta, resid)
      num_its++;
      rho = rho_new;
      TE_RC(reduce(RED_SUMSQ, resid, nullptr, &nr));
      residual = nr / r0;
      if (history && num_its < history_cap) history[num_its] = residual;
      if (time_its) TE_RC(timer_stop_stream(g->timer, "Iteration", TE_TIMER_NONE, s));
    }
    if (!rc && use_precond) {
      TE_RC(vcycle_enqueue(g, x, tmp));  // Mr->apply(x, resid); x.copy(resid)
      TE_RC(vop(VOP_COPY, x, tmp, 0.0));
    }
    if (!rc) TE_RC(vop(VOP_ADD, x, x0, 0.0));  // x.add(initial_guess)
  }
  TE_HIP(hipStreamSynchronize(s));
  g->last_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (iterations) *iterations = num_its;
  if (rc) return rc;
  return check_flags(g);
}

double te_gmg_last_solve_seconds(const te_gmg* g) { return g ? g->last_seconds : -1.0; }

int te_gmg_time_kernel(te_gmg* g, int which, int reps, double* ms)
{
  if (!g || !ms || reps < 1 || which < TE_K_SWEEP_ZERO || which > TE_K_X_COMBINE) {
    set_error("te_gmg_time_kernel: bad argument");
    return TE_ERR_ARG;
  }
  if (!g->c3) {
    set_error("te_gmg_time_kernel: only the compact 3D cycle has standalone kernel timings");
    return TE_ERR_UNSUPPORTED;
  }
  const bool two = g->L.size() > 1;
  if (!two && (which == TE_K_SWEEP_INTERP || which == TE_K_RESID_RESTRICT)) {
    set_error("te_gmg_time_kernel: the transfer kernels need a coarser level");
    return TE_ERR_UNSUPPORTED;
  }
  GLevel& L0 = g->L[0];
  te_gmg::C3Level& C0 = g->C3[0];
  hipStream_t s = g->s;
  const int nb = g->nblk;
  auto launch = [&]() -> int {
    const long np = L0.d->np;
    void* xu = xcol_of(g, C0.u, np, L0.dt);
    void* xu2 = xcol_of(g, C0.u2, np, L0.dt);
    switch (which) {
      case TE_K_SWEEP_ZERO:
        return launch_c3_sweep(L0.plan, L0.dt, L0.dt, C0.f, C0.u, L0.dt, C0.u2, 0, TE_F64, nullptr, nullptr, s,
                               nullptr, xu2);
      case TE_K_SWEEP:
        return launch_c3_sweep(L0.plan, L0.dt, L0.dt, C0.f, C0.u, L0.dt, C0.u2, 1, TE_F64, nullptr, nullptr, s, xu,
                               xu2);
      case TE_K_SWEEP_INTERP:
        if (C0.u3 && g->finterp && g->xcol) {  // as the cycle runs it: boundary interpolation + mode-1 sweep
          void* xu3 = xcol_of(g, C0.u3, np, L0.dt);
          TE_RC(launch_c3_face_interp(L0.link, L0.dt, C0.u, xu, g->L[1].dt, g->C3[1].u, C0.u3, xu3, s));
          return launch_c3_sweep(L0.plan, L0.dt, L0.dt, C0.f, C0.u3, L0.dt, C0.u2, 1, TE_F64, nullptr, nullptr, s,
                                 xu3, xu2);
        }
        return launch_c3_sweep(L0.plan, L0.dt, L0.dt, C0.f, C0.u, L0.dt, C0.u2, 2, g->L[1].dt, g->C3[1].u, L0.link,
                               s, xu, xu2);
      case TE_K_RESID_RESTRICT:
        return launch_c3_resid_restrict(L0.d, L0.link, g->o.neumann, L0.dt, L0.dt, C0.f, C0.u, g->L[1].dt,
                                        g->C3[1].f, s, xu);
      case TE_K_OPERATOR_P: {
        const bool xc = g->L.size() == 1 ? g->o.coarse_sweeps > 0 : g->o.post_sweeps > 0;
        const bool r32 = L0.dt == TE_F32 && g->xrfly && g->rhat32;  // as the mixed-precision iteration
        return launch_c3_star_dot(L0.d, g->o.neumann, L0.dt, g->c_mp, g->xrfly ? nullptr : g->c_ap, g->c_rhat, g->pa,
                                  nullptr, nb, s, xc ? xcol_of(g, g->c_mp, np, L0.dt) : nullptr,
                                  r32 ? g->c_rhat32 : nullptr);
      }
      case TE_K_X_COMBINE: {  // the deferred x updates of xslots iterations
        if (!g->c_hist) {
          set_error("te_gmg_time_kernel: this solver does not defer the x update");
          return TE_ERR_UNSUPPORTED;
        }
        C3XComb c{};
        c.x = g->resid;  // a padded FP64 vector, as the solve's caller x
        c.padded = 1;
        for (int j = 0; j < g->xslots; j++) {
          c.mp[j] = (const float*)(j ? g->c_mph[j - 1] : g->c_mp);
          c.ms[j] = (const float*)(j ? g->c_msh[j - 1] : g->c_ms);
        }
        c.hist = g->c_hist;
        c.scal = g->scal;
        c.count = g->xslots;
        return launch_c3_xcomb(L0.d, c, s);
      }
      case TE_K_OPERATOR: {
        const bool xc = g->L.size() == 1 ? g->o.coarse_sweeps > 0 : g->o.post_sweeps > 0;
        return launch_c3_star_dot(L0.d, g->o.neumann, L0.dt, g->c_ms, g->xrfly ? nullptr : g->c_as, g->c_s, g->pa,
                                  g->pb, nb, s, xc ? xcol_of(g, g->c_ms, np, L0.dt) : nullptr);
      }
      default: {
        C3Bicg a{};
        a.scal = g->scal;
        a.x = g->c_x;
        a.r = g->c_r;
        a.rhat = g->c_rhat;
        a.ap = g->c_ap;
        a.as = g->c_as;
        a.mp = g->c_mp;
        a.ms = g->c_ms;
        a.sv = g->c_s;
        a.dtm = L0.dt;
        a.pa = g->pa;
        a.pb = g->pb;
        // the s and p updates as in the mixed-precision iteration (with the FP32 copies)
        int mode = BICG_XR;
        if (which == TE_K_UPDATE_S) {
          mode = BICG_S;
          a.sv = g->c_s;
          a.w32 = L0.dt == TE_F32 ? g->c_s32 : nullptr;
          a.pa = g->pc;
        } else if (which == TE_K_UPDATE_P) {
          mode = BICG_P;
          a.p = g->c_p;
          a.w32 = L0.dt == TE_F32 ? g->c_p32 : nullptr;
          if (a.w32 && g->xrfly && g->rhat32 && g->p32only) a.p = nullptr;  // as the mixed-precision solve
        }
        if (!g->xrfly) return launch_c3_bicg(mode, L0.d, a, nb, s);
        if (mode == BICG_XR && L0.dt == TE_F32 && g->rhat32) a.rh32 = g->c_rhat32;
        if (mode == BICG_XR && a.rh32 && g->c_hist) a.hist = g->c_hist;  // x deferred, as the solve runs it
        const bool xc = g->L.size() == 1 ? g->o.coarse_sweeps > 0 : g->o.post_sweeps > 0;
        return launch_c3_upd(mode, L0.d, g->o.neumann, a, true, xc ? xcol_of(g, g->c_mp, np, L0.dt) : nullptr,
                             xc ? xcol_of(g, g->c_ms, np, L0.dt) : nullptr, nb, s);
      }
    }
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  TE_HIP(hipEventCreate(&e0));
  TE_HIP(hipEventCreate(&e1));
  int rc = launch();
  if (!rc && hipEventRecord(e0, s) != hipSuccess) rc = TE_ERR_HIP;
  for (int k = 0; k < reps && !rc; k++) rc = launch();
  float t = 0.0f;
  if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&t, e0, e1) != hipSuccess)) {
    set_error("te_gmg_time_kernel: event timing failed");
    rc = TE_ERR_HIP;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc) *ms = (double)t / reps;
  return rc;
}

int te_tune_time_c3(te_gmg* g, int which, int variant, int reps, double* ms)
{
  if (!g || variant < 0 || variant >= tune_num_variants(TUNE_C3_SWEEP)) {
    set_error("te_tune_time_c3: bad argument");
    return TE_ERR_ARG;
  }
  tune_force(TUNE_C3_SWEEP, 3, kC3N, g->L[0].dt, variant);
  const int rc = te_gmg_time_kernel(g, which, reps, ms);
  tune_unforce();
  return rc;
}

int te_gmg_set_timer(te_gmg* g, te_timer* t, int iterations, const int* domain_ids)
{
  if (!g) {
    set_error("te_gmg_set_timer: bad argument");
    return TE_ERR_ARG;
  }
  std::vector<int> ids;
  if (t && domain_ids) {
    ids.assign(domain_ids, domain_ids + g->L.size());
    for (int id : ids) {
      if (id != TE_TIMER_NONE && !timer_has_domain(t, id)) {
        set_error("Domain with id " + std::to_string(id) + " was not added to timer");
        return TE_ERR_ARG;
      }
    }
  }
  g->timer = t;
  g->time_its = t && iterations;
  g->timer_ids = ids;
  return TE_OK;
}

}  // extern "C"
