// Outer BiCGStab device kernels: scalar slots, update modes, launchers.
#pragma once

#include "../te_internal.h"

namespace te {

enum BicgScalar {
  SC_RHO = 0, SC_ALPHA, SC_OMEGA, SC_BETA, SC_R0, SC_SNORM, SC_RNORM, SC_TOL, SC_STATE,
  SC_SEQ,  // iterations published by FIN_RHO (never reset)
  SC_ITS,    // batches: the problem's completed iterations
  SC_MAXIT,  // batches: its iteration limit
  SC_COUNT
};
// SC_STATE: the iteration runs (ST_RUN) until the s-norm test of BiCGStab.h:162 fires
// (ST_EARLY); from then on only the early-exit update x += alpha p runs.  In a batch a
// problem whose loop of BiCGStab.h:148 has ended is ST_DONE (residual at tolerance, the
// iteration limit, r0 = 0, or the iteration after the s-norm test) or ST_BREAK (rho = 0,
// the reference's breakdown error); every update of its vectors is then skipped.
enum BicgState { ST_RUN = 0, ST_EARLY = 1, ST_DONE = 2, ST_BREAK = 3 };
enum BicgMode { BICG_INIT = 0, BICG_S, BICG_XP, BICG_XR, BICG_P };
// FIN_SOMEGA = FIN_SNORM (from the s.s partials pc) followed by FIN_OMEGA unless the test fired
enum BicgFinalize { FIN_INIT = 0, FIN_ALPHA, FIN_SNORM, FIN_OMEGA, FIN_RHO, FIN_SOMEGA };

int krylov_blocks(const te_domain* d);

// ---- batches of independent problems (te_batch) ----
// Problem q in a batched launch: its patches [poff, poff + np) of the concatenated finest
// levels, the launch's workgroups [woff, woff + nwg) and its partial-sum slots
// [soff, soff + nslot).  Each problem's patches, workgroups and slots are those of the
// same launch on that problem alone, so the sums come out with the same bits.
struct BatchProb {
  int poff, np, woff, nwg, soff, nslot;
};
// device arrays: prob[nprob], and wq[w] = the problem of workgroup w (nwg entries)
struct BatchMap {
  const BatchProb* prob = nullptr;
  const int* wq = nullptr;
  int nwg = 0;
  int nprob = 0;
};
// per-problem finalize of a batch (one workgroup per problem; scalars SC_COUNT apart).
// FIN_INIT also stores tol and max_it and decides the first state; FIN_RHO counts the
// iteration and ends the problem's loop (ST_DONE / ST_BREAK) as BiCGStab.h:148 would.
int launch_bicg_finalize_batch(int mode, const double* pa, const double* pb, const double* pc, const BatchMap& m,
                               double* scal, hipStream_t s, double tol = 0.0, int max_it = 0);
// after FIN_RHO: ctl[0] += 1; mirror[1] = problems still running, mirror[2] = problems
// broken down; then mirror[0] = ctl[0] (host-coherent memory; the host waits on it)
int launch_batch_publish(const double* scal, int nprob, double* ctl, double* mirror, hipStream_t s);
// gather (to_batch) or scatter the problems' own padded vectors: problem q's vector of
// len[q] elements at ptr[q] <-> batch vector elements [off[q], off[q] + len[q])
int launch_batch_copy(bool to_batch, double* batch, double* const* ptr, const long* off, const long* len, int nprob,
                      long maxlen, hipStream_t s);
bool star_dot_fly_supported(const te_ghost_filler* g);
// udt: type of u (FP32 only with fly: the values are widened to FP64 as they are read)
int launch_star_dot(const te_ghost_filler* g, int op_kind, int neumann, const double* coeffs, const void* u, int udt,
                    double* out, const double* w, double* pa, double* pb, bool fly, hipStream_t s);
// The maps of one batch for star_dot_2d launches: map[k] gives each problem
// min(cap[k], its slot count) workgroups, as the launch on that problem alone would run
// with an instantiation of cap[k] resident workgroups.
struct BatchStarMaps {
  static constexpr int kMax = 4;
  int n = 0;
  int cap[kMax] = {};
  BatchMap map[kMax];
};
// 2D, ghosts computed in the launch; writes `grid` partials (fused2d.hip).  With batch maps
// the launch covers every problem of the batch (grid unused).
int launch_star_dot_2d(const te_ghost_filler* g, int op_kind, int neumann, const double* coeffs, const void* u,
                       int udt, double* out, const double* w, double* pa, double* pb, int grid, hipStream_t s,
                       const BatchStarMaps* bsm = nullptr);
// the resident workgroup count of the star_dot_2d instantiation a launch with these
// arguments would run (0 when unsupported)
int star_dot_2d_cap(const te_ghost_filler* g, int op_kind, int udt, bool sq);
// dtm: type of a0 and a1 for BICG_XR (M p, M s) and of a0 for BICG_XP; FP64 for the other modes.
// w32: FP32 copy of the new p (BICG_INIT, BICG_P) or s (BICG_S), or null.
// xp_dir (BICG_XR): after the s-norm test fired, the launch does x += alpha xp_dir instead
// (xp_dir of type dtm when xp_dir_m, else FP64).
// pfold / nfold (BICG_S only): the rhat.ap partials; the launch then also does the FIN_ALPHA
// finalize (alpha = rho / their fold, stored in scal), so that finalize is not launched.
int launch_bicg(int mode, const te_domain* d, double* scal, double* v0, double* v1, const void* a0,
                const void* a1, const double* a2, const double* a3, const double* a4, double* pa, double* pb,
                hipStream_t s, int dtm = TE_F64, float* w32 = nullptr, const void* xp_dir = nullptr,
                int xp_dir_m = 0, const double* pfold = nullptr, int nfold = 0, const BatchMap* bm = nullptr);
// tol: the relative tolerance, stored by FIN_INIT for the s-norm test of FIN_SNORM.
// pc: FIN_SOMEGA's s.s partials; for FIN_INIT, when given, the rhat.r partials (rhat not r
// itself: rho from them instead of r.r).
// mirror (FIN_RHO): device pointer to host-coherent memory of SC_COUNT doubles; FIN_RHO
// copies the scalars there and then bumps SC_SEQ, so the host can read an iteration's
// outcome while the rest of the iteration is still running.
int launch_bicg_finalize(int mode, const double* pa, const double* pb, int n, double* scal, hipStream_t s,
                         double tol = 0.0, double* mirror = nullptr, const double* pc = nullptr);
int launch_patch_bicgstab(const te_domain* d, int dtype, int op_kind, int neumann, const void* coeffs, const void* f,
                          void* u, int max_it, double tol, int32_t* iters, int32_t* bd, hipStream_t s);

}  // namespace te
