// The 3D V-cycle and the outer BiCGStab's vector work on compact vectors: interior cells
// only, [patch][z][y][x], 16^3 cells per patch, for levels whose face neighbours are all on
// the same level (or a physical boundary).  The caller's vectors keep the reference's
// padded layout; the solver's own work vectors need no ghost cells, because every kernel
// here computes the ghost values it needs from the neighbours' boundary cells (the
// TriLinearGhostFiller's same-level case is a copy, TriLinearGhostFiller.cpp:33-175).
#pragma once

#include "../te_internal.h"
#include "dft.h"

namespace te {

constexpr int kC3N = 16;
constexpr int kC3Cells = kC3N * kC3N * kC3N;
// Face columns of a compact vector: the cells x = 0 and x = N - 1 of every row (z, y) of
// each patch, [patch][side][z][y], stored behind the vector's np * kC3Cells cells.  Rows
// are contiguous in x, so a neighbour's x face read from the vector itself costs a 64-byte
// (FP32) row per value; from the face columns, 16 values are one 64-byte read.  The
// kernels that write a vector with face columns write them too; a reader is given them
// only where every writer of the vector did so (nullptr: the x faces come from the cells).
constexpr int kC3XCol = 2 * kC3N * kC3N;
// levels of at least this many patches run the first post-smoothing sweep as a boundary
// interpolation (launch_c3_face_interp) and a mode-1 sweep
constexpr int kC3InterpMin = 4096;
inline void* c3_xcol(const void* v, long np, int dt)
{
  return v ? (void*)((char*)v + (size_t)np * kC3Cells * (dt == TE_F64 ? 8 : 4)) : nullptr;
}

// 3D, 16^3 cells, every face neighbour on the same level
bool c3_supported(const te_domain* d);

// One PatchSolver::smooth sweep of the DFT patch solver (Poisson/DFTPatchSolver.h:420-450)
// on compact vectors.  mode 0: u is zero (out = solve(f)); 1: the interface terms of
// modifyRHSForInternalBoundaryConditions from uin's boundary cells and the neighbours';
// 2: the same on uin + P uc (DirectInterpolator, GMG/DirectInterpolator.h:60-91; uc is the
// coarser level's compact solution of type dtc, link: this level -> the coarser one).
// out must not alias uin.  dt: arithmetic type; dtf / dto: types of f and out.  xin: uin's
// face columns (or nullptr); xout: out's face columns to write (or nullptr).
int launch_c3_sweep(const te_dft_plan* pl, int dt, int dtf, const void* f, const void* uin, int dto, void* out,
                    int mode, int dtc, const void* uc, const te_level_link* link, hipStream_t s,
                    const void* xin = nullptr, void* xout = nullptr);
// Cycle::restrict (GMG/Cycle.h:57-65): fc = restriction (GMG/LinearRestrictor.h:101-155) of
// f - A u, A the Poisson star operator (neumann: Neumann physical sides); fc compact, of the
// coarser level, type dtc.  xu: u's face columns (or nullptr).
int launch_c3_resid_restrict(const te_domain* d, const te_level_link* link, int neumann, int dt, int dtf,
                             const void* f, const void* u, int dtc, void* fc, hipStream_t s,
                             const void* xu = nullptr);
// out = A u (FP64, compact; u compact of type udt, widened as read; out may be nullptr: only
// the sums) with the per-workgroup partial sums pa = sum w out and, when pb is not null,
// pb = sum out out; nslots partial slots, those past the launch's workgroups written zero.
// xu: u's face columns (or nullptr).  w32: an FP32 w used in place of w (u FP32, pb null).
int launch_c3_star_dot(const te_domain* d, int neumann, int udt, const void* u, double* out, const double* w,
                       double* pa, double* pb, int nslots, hipStream_t s, const void* xu = nullptr,
                       const float* w32 = nullptr);
// BiCGStab vector updates (Iterative/BiCGStab.h:111-203) on compact vectors, the modes and
// scalar slots of krylov.h.  b and xpad are the caller's padded FP64 vectors; x is a compact
// FP64 copy of the caller's x that the iterations update (written back at the end).
//   BICG_INIT: r_pad (padded residual b - A x) -> r, rhat, p (+ p32), xpad -> x; sums b.b, r.r
//   BICG_S:    s = r - alpha ap (+ s32), sum s.s (alpha folded from pfold when given)
//   BICG_XR:   x += alpha mp + omega ms; r -= alpha ap + omega as; sums r.rhat, r.r; after
//              the s-norm test fired: x += alpha dir (dir of type dtm when dir_m, else FP64)
//   BICG_P:    p = r + beta (p - omega ap) (+ p32); launch_c3_upd and launch_c3_init with a.p
//              null: p kept only as its FP32 copy a.w32 (read and written in place)
struct C3Bicg {
  double* scal;
  double* x;         // compact FP64 (INIT writes it from xpad, XR updates it)
  const double* xpad;  // padded FP64 (INIT)
  const double* b;   // padded FP64 (INIT)
  const double* rpad;  // padded FP64 (INIT)
  double* r;
  double* rhat;
  double* p;
  double* sv;
  const double* ap;
  const double* as;
  const void* mp;    // type dtm
  const void* ms;
  float* w32;        // FP32 copy of the new p or s, or null
  const void* dir;   // early-exit direction (XR)
  int dir_m;
  int dtm;
  double* pa;
  double* pb;
  const double* pfold;
  int nfold;
  float* rh32;       // FP32 rhat in place of rhat, or null: INIT (launch_c3_init) writes it and
                     // the partial sums rhat.r into pr; XR (launch_c3_upd, FP32 M p and M s) reads it
  double* pr;
  double* hist;      // XR (launch_c3_upd, with rh32): x is not updated; alpha and omega go to
                     // hist[0], hist[1] for launch_c3_xcomb (null: x updated in the launch)
  double* xpo;       // XR after the s-norm test fired: x += alpha dir into this padded x instead
                     // of the compact x (the deferred path keeps no compact x; INIT: a.x null)
};
int launch_c3_bicg(int mode, const te_domain* d, const C3Bicg& a, int nslots, hipStream_t s);
// BICG_INIT from the caller's padded b and x alone: r = b - A x formed in the launch (the face
// ghosts from the same-level neighbours' cells of x or the physical condition; a.rpad unused),
// then x, r, rhat (or a.rh32), p (+ a.w32) compact and the partial sums b.b, r.r (+ rhat.r into
// a.pr with a.rh32).  One wavefront per patch.
int launch_c3_init(const te_domain* d, int neumann, const C3Bicg& a, int nslots, hipStream_t s);
// BICG_S, BICG_XR and BICG_P with the operator products formed in the launch, one wavefront
// per patch: A M s (XR) from a.ms, and with apf A M p (S, P) from a.mp (then a.ap is unused:
// the operator launches need not store either; apf needs a.mp apart from a.p, as in the
// preconditioned iteration).  XR forms r = s - omega A M s from a.sv (the same vector as
// launch_c3_bicg's r - alpha ap - omega as, rounded in another order).  xmp, xms: the face
// columns of a.mp and a.ms (or nullptr).  The early exit and partial-sum slots of
// launch_c3_bicg.
int launch_c3_upd(int mode, const te_domain* d, int neumann, const C3Bicg& a, bool apf, const void* xmp,
                  const void* xms, int nslots, hipStream_t s);
// The deferred x updates: x += alpha_j M p_j + omega_j M s_j for j < count, in that order,
// each term as the x / r update adds it (alpha_j, omega_j = hist[2 j], hist[2 j + 1]; M p_j,
// M s_j FP32 compact).  drop_early: the last slot is left out when the state word in scal
// says the s-norm test fired (that iteration added alpha dir to x itself).
constexpr int kC3XSlots = 8;
struct C3XComb {
  double* x;  // compact, or the padded layout with padded
  const float* mp[kC3XSlots];
  const float* ms[kC3XSlots];
  const double* hist;
  const double* scal;
  int count;
  int drop_early;
  int padded;
};
int launch_c3_xcomb(const te_domain* d, const C3XComb& c, hipStream_t s);
// compact <-> padded copies with element-type conversion (interior cells; the padded
// vector's ghost cells are not written)
int launch_c3_pack(const te_domain* d, int dtd, void* dst, int dts, const void* src_padded, hipStream_t s);
int launch_c3_unpack(const te_domain* d, int dtd, void* dst_padded, int dts, const void* src, hipStream_t s);
// out = u + P uc (compact, this level's type dt; uc of type dtc)
int launch_c3_interp(const te_level_link* link, int dtc, const void* uc, int dt, const void* u, void* out,
                     hipStream_t s);
// the boundary layer of u + P uc (this level's type dt, uc of type dtc) and its face columns
// into out / xout: what the interpolating sweep (mode 2) reads, for a mode-1 sweep on out;
// xu: u's face columns
int launch_c3_face_interp(const te_level_link* link, int dt, const void* u, const void* xu, int dtc, const void* uc,
                          void* out, void* xout, hipStream_t s);
// x := 0 (n elements of type dt), a kernel (no memset node in captured graphs)
int launch_c3_zero(long n, int dt, void* x, hipStream_t s);

}  // namespace te
