/*
 * thunderegg_amd.h -- C ABI of the MI355X (gfx950) patch-parallel GMG hot path.
 *
 * Drop-in boundary for ThunderEgg's plugin/operator API on this path.  The
 * reference exposes C++ virtual interfaces only (no FFI); each entry point
 * below names the reference member function it replaces (paths relative to
 * the reference's src/ThunderEgg/).  The C++ API in include/ThunderEgg/
 * (ThunderEgg/amd.h) restores the reference's class names and method shapes on
 * top of these functions.
 *
 * Conventions
 *  - Every function returns 0 on success or a negative status; te_last_error()
 *    returns a thread-local message for the last failure.  Status -3 means a
 *    breakdown of an iterative solver (Iterative::BreakdownError).
 *  - Vectors are device pointers in the reference layout (Vector.h:125-154):
 *    patch-major, then component, then z, y, x, each axis padded by one ghost
 *    cell on both sides; patches in Domain local-index order.
 *  - dtype selects the element type: TE_F64 (double) or TE_F32 (float).
 *  - All work is enqueued on `stream` (a hipStream_t; NULL = default stream)
 *    and is asynchronous unless stated otherwise.  No entry point allocates
 *    device memory except the *_create functions and te_device_alloc.
 *  - One host thread per stream; objects are immutable after creation.
 */
#ifndef THUNDEREGG_AMD_H
#define THUNDEREGG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TE_OK 0
#define TE_ERR_ARG (-1)
#define TE_ERR_HIP (-2)
#define TE_ERR_BREAKDOWN (-3)
#define TE_ERR_UNSUPPORTED (-4)

enum te_dtype { TE_F64 = 0, TE_F32 = 1 };
/* GhostFillingType.h: Faces, Edges, Corners */
enum te_fill_type { TE_FACES = 0, TE_EDGES = 1, TE_CORNERS = 2 };
enum te_filler_kind { TE_BILINEAR = 0, TE_BIQUADRATIC = 1, TE_TRILINEAR = 2 };
/* NbrType.h:33-48 (-1 = no neighbour) */
enum te_nbr_type { TE_NBR_NONE = -1, TE_NBR_NORMAL = 0, TE_NBR_COARSE = 1, TE_NBR_FINE = 2 };

typedef void* te_stream; /* hipStream_t */

/* One patch of a Domain: PatchInfo<D> (PatchInfo.h:52-463) with its
 * Normal/Coarse/FineNbrInfo slots flattened.  Slot = Face<D,M>::sum_of_faces +
 * face value: 2D corners 0-3, sides 4-7; 3D corners 0-7, edges 8-19, sides 20-25.
 * Fine neighbours list 2^M ids in Orthant<M> order on the shared face. */
typedef struct te_patch_info {
  int32_t id, rank, parent_id, parent_rank;
  int32_t orth_on_parent; /* -1 == Orthant<D>::null() */
  int32_t refine_level;
  int32_t child_ids[8];
  int32_t child_ranks[8];
  double starts[3];
  double spacings[3];
  int32_t nbr_type[26];
  int32_t nbr_orth[26]; /* CoarseNbrInfo::orth_on_coarse, -1 otherwise */
  int32_t nbr_ids[26][4];
  int32_t nbr_ranks[26][4];
} te_patch_info;

typedef struct te_domain te_domain;
typedef struct te_ghost_filler te_ghost_filler;
typedef struct te_level_link te_level_link;
typedef struct te_dft_plan te_dft_plan;
typedef struct te_gmg te_gmg;
typedef struct te_forest te_forest;

/* ------------------------------------------------------------ runtime -- */
const char* te_last_error(void);
int te_device_count(int* out);
int te_device_alloc(size_t bytes, void** out);
int te_device_free(void* p);
int te_memcpy_h2d(void* dst, const void* src, size_t bytes);
int te_memcpy_d2h(void* dst, const void* src, size_t bytes);
int te_memcpy_d2d(void* dst, const void* src, size_t bytes, te_stream stream);
int te_memset(void* dst, int value, size_t bytes, te_stream stream);
int te_stream_synchronize(te_stream stream);
/* Create/destroy a stream at the default priority. */
int te_stream_create(te_stream* out);
int te_stream_destroy(te_stream s);
/* HIP events for timing work on a stream (hipEventRecord / hipEventElapsedTime). */
typedef void* te_event;
int te_event_create(te_event* out);
int te_event_destroy(te_event e);
int te_event_record(te_event e, te_stream s);
int te_event_elapsed_ms(te_event start, te_event stop, float* ms); /* waits for stop */

/* ------------------------------------------------------ domain / tables --
 * Domain<D>(comm, id, ns, num_ghost_cells, first, last) (Domain.h:223-248):
 * local_index = position, neighbour local indexes resolved by id (Domain.h:94-109).
 * One rank: every neighbour must be in the table.  num_ghost must be 1. */
int te_domain_create(int dim, const int* ns, int num_ghost, int num_patches, const te_patch_info* patches,
                     te_domain** out);
int te_domain_destroy(te_domain* d);
int te_domain_num_patches(const te_domain* d);
/* Number of elements of a vector with ncomp components (padded layout). */
int64_t te_domain_vector_length(const te_domain* d, int ncomp);
/* Neighbour local-index table, [num_patches][26][4], -1 where empty. */
int te_domain_get_nbr_local(const te_domain* d, int32_t* out);

/* GMG::InterLevelComm (GMG/InterLevelComm.h:125-224), one rank: every fine patch
 * has a local parent.  parent local index per fine patch. */
int te_level_link_create(const te_domain* fine, const te_domain* coarse, te_level_link** out);
int te_level_link_destroy(te_level_link* l);
int te_level_link_get_parents(const te_level_link* l, int32_t* out);

/* ------------------------------------------------------- ghost filling --
 * BiLinearGhostFiller / BiQuadraticGhostFiller (2D), TriLinearGhostFiller (3D):
 * constructors (BiLinearGhostFiller.cpp:24-27, TriLinearGhostFiller.cpp:573-585). */
int te_ghost_filler_create(const te_domain* d, int kind, int fill_type, te_ghost_filler** out);
int te_ghost_filler_destroy(te_ghost_filler* g);
/* GhostFiller::fillGhost(const Vector&) (GhostFiller.h:53; MPIGhostFiller.h:1055-1092):
 * ghost cells facing a neighbour are overwritten; interior cells are only read. */
int te_ghost_fill(const te_ghost_filler* g, int dtype, void* u, int ncomp, te_stream stream);

/* ----------------------------------------------------------- operators --
 * PatchOperator::apply (PatchOperator.h:126-143) for Poisson::StarPatchOperator
 * (Poisson/StarPatchOperator.h:75-101): fills ghosts of u, then f = A u.
 * Physical-boundary ghost values are applied inline (-u Dirichlet, +u Neumann)
 * and not stored back into u. */
int te_star_apply(const te_ghost_filler* g, int dtype, int neumann, void* u, void* f, int ncomp,
                  te_stream stream);
/* The per-patch loop of PatchOperator::apply alone (PatchOperator.h:136-142: f := 0 with
 * ghosts, then applySinglePatch on every patch): u's neighbour ghosts must already be
 * filled.  One launch of the batched stencil. */
int te_star_apply_patches(const te_ghost_filler* g, int dtype, int neumann, const void* u, void* f, int ncomp,
                          te_stream stream);
/* VarPoisson::StarPatchOperator::apply (VarPoisson/StarPatchOperator.h:61-109); coeffs
 * (same dtype, ncomp 1) must be ghost-filled once beforehand, as the reference's
 * constructor does (:70). */
int te_varstar_apply(const te_ghost_filler* g, int dtype, const void* coeffs, void* u, void* f, int ncomp,
                     te_stream stream);
/* Boundary data: one value per face cell of every side of every patch, for patch p (local
 * order) and side s (Side<D> order: west, east, south, north, bottom, top) the side's face
 * cells with the first of the other axes fastest -- the points DomainTools::GetRealCoordBound
 * gives for that side (DomainTools.h:178-210).  Values on sides with a neighbour are not read.
 * Length (doubles) for the whole domain: */
int64_t te_domain_boundary_length(const te_domain* d);
/* Poisson::StarPatchOperator::addDrichletBCToRHS (Poisson/StarPatchOperator.h:187-204):
 * on every side without a neighbour f_b += -2 g / h^2; with c != NULL the VarPoisson form
 * (VarPoisson/StarPatchOperator.h:194-220) f_b -= 2 g c / h^2.  g (Dirichlet values) and c
 * (coefficient values) are device FP64 boundary data; f a device vector (component 0). */
int te_star_add_dirichlet_bc(const te_domain* d, int dtype, void* f, const double* g, const double* c,
                             te_stream stream);
/* Poisson::StarPatchOperator::addNeumannBCToRHS (Poisson/StarPatchOperator.h:206-242): on
 * every side without a neighbour f_b += dg / h on lower sides and f_b -= dg / h on upper
 * sides; dg: device FP64 boundary data of the derivative of g along each side's axis. */
int te_star_add_neumann_bc(const te_domain* d, int dtype, void* f, const double* dg, te_stream stream);
/* Cycle::restrict residual (GMG/Cycle.h:57-65): fills ghosts of u, then r = f - A u.
 * op_kind 0 Poisson (neumann flag), 1 VarPoisson (coeffs). */
int te_residual(const te_ghost_filler* g, int dtype, int op_kind, int neumann, const void* coeffs, const void* f,
                void* u, void* r, int ncomp, te_stream stream);

/* ------------------------------------------------------------ smoothers --
 * Poisson::DFTPatchSolver (Poisson/DFTPatchSolver.h:369-450).  neumann_mask is the
 * reference's std::bitset<Side<D>::number_of> (bit s = side s).  op_kind selects the
 * modifyRHSForInternalBoundaryConditions of the wrapped operator: 0 Poisson, 1
 * VarPoisson (coeffs: ghost-filled beta, float64, ncomp 1; the plan keeps its own copy, so
 * the caller's buffer may be freed after this call returns). */
int te_dft_plan_create(const te_ghost_filler* g, unsigned neumann_mask, int op_kind, const double* coeffs,
                       te_dft_plan** out);
int te_dft_plan_destroy(te_dft_plan* p);
/* PatchSolver::smooth (PatchSolver.h:142-170): fill ghosts of u, then per patch
 * u = DFT solve of (f - interface terms from u's ghosts).  ncomp 1. */
int te_dft_smooth(const te_dft_plan* p, int dtype, const void* f, void* u, te_stream stream);
/* PatchSolver::apply (PatchSolver.h:98-134): u := 0, then per-patch solves. */
int te_dft_apply(const te_dft_plan* p, int dtype, const void* f, void* u, te_stream stream);
/* The per-patch loop of PatchSolver::smooth alone (PatchSolver.h:156-166, each patch
 * DFTPatchSolver::solveSinglePatch, Poisson/DFTPatchSolver.h:420-450): u's ghosts
 * must already be filled.  One launch of the batched patch solve. */
int te_dft_patch_solve(const te_dft_plan* p, int dtype, const void* f, void* u, te_stream stream);

/* ---- fused launches of the V-cycle (2D, even cells per side, BiLinear / BiQuadratic
 * ghosts; the DFT ones run on the matrix cores for n = 16 or 32 with Dirichlet sides, and
 * in the generic patch-solve kernel otherwise).  Each replaces a sequence of the
 * calls above by one launch that computes ghost values from the neighbours' interior
 * values instead of reading ghost cells; none of them reads or writes ghost cells.
 * dtype: arithmetic and u type; dtype_f / dtype_out / dtype_c: f, output and
 * coarse-level types (TE_F32 / TE_F64, the mixed-precision cycle's conversions). */
/* PatchSolver::apply (PatchSolver.h:98-134) from u = 0: out = per-patch solve of f. */
int te_dft_apply_mixed(const te_dft_plan* p, int dtype, int dtype_f, int dtype_out, const void* f, void* out,
                       te_stream stream);
/* The up-leg of VCycle::visit (GMG/VCycle.h:64-67): Interpolator::interpolate
 * (DirectInterpolator, GMG/DirectInterpolator.h:60-91) of the coarser level's uc into u
 * followed by PatchSolver::smooth (PatchSolver.h:142-170), as one launch writing the
 * smoothed u + P uc to out (out must not alias u; u and uc are not modified).  link:
 * this level -> the coarser level. */
int te_dft_smooth_interp(const te_dft_plan* p, const te_level_link* link, int dtype, int dtype_f, int dtype_out,
                         const void* f, const void* u, int dtype_c, const void* uc, void* out, te_stream stream);
/* Cycle::restrict (GMG/Cycle.h:57-65): r = f - A u (ghost fill, star operator) and
 * MPIRestrictor::restrict of r (GMG/LinearRestrictor.h:101-155) into the coarser
 * level's right-hand side fc, one launch; r is not stored and fc's ghost cells are not
 * written.  Even n. */
int te_residual_restrict(const te_ghost_filler* g, const te_level_link* link, int op_kind, int neumann,
                         const void* coeffs, int dtype, int dtype_f, const void* f, const void* u, int dtype_c,
                         void* fc, te_stream stream);

/* Iterative::PatchSolver<D>(BiCGStab, op) smooth (Iterative/PatchSolver.h:134-175):
 * per patch, an inner BiCGStab (Iterative/BiCGStab.h:111-203) on the patch operator
 * with internal boundary conditions, started from u.  iters (device int32[P], may be
 * NULL) receives per-patch iteration counts.  With continue_on_breakdown == 0 a
 * breakdown in any patch returns TE_ERR_BREAKDOWN (checked synchronously). */
int te_patch_bicgstab_smooth(const te_ghost_filler* g, int dtype, int op_kind, int neumann, const void* coeffs,
                             int max_it, double tol, int continue_on_breakdown, const void* f, void* u,
                             int32_t* iters, te_stream stream);

/* --------------------------------------------------- inter-level transfer --
 * MPIRestrictor::restrict + LinearRestrictor::restrictPatches (GMG/MPIRestrictor.h:55-83,
 * GMG/LinearRestrictor.h:101-191), extrapolate_boundary_ghosts = false:
 * coarse := 0 (with ghosts), then 2^D-cell averages / copies. */
int te_restrict(const te_level_link* l, int dtype, const void* fine, void* coarse, int ncomp, te_stream stream);
/* The same with extrapolate_boundary_ghosts = true (GMG/LinearRestrictor.h:49-92, 138-155):
 * the coarse ghost cells over each fine patch's exterior sides get (3 ghost - interior) / 2^D
 * of its side cells, the coarse corner ghost of its orthant (when it has no neighbour across
 * that corner) 3 ghost - interior of its corner cell, and copied patches keep their ghosts. */
int te_restrict_extrapolate(const te_level_link* l, int dtype, const void* fine, void* coarse, int ncomp,
                            te_stream stream);
/* MPIInterpolator::interpolate + DirectInterpolator (GMG/MPIInterpolator.h:79-106,
 * GMG/DirectInterpolator.h:60-91): fine += injection of coarse. */
int te_interpolate(const te_level_link* l, int dtype, const void* coarse, void* fine, int ncomp,
                   te_stream stream);

/* -------------------------------------------------------- vector algebra --
 * Vector<D> BLAS-1 (Vector.h:391-611), interior cells only. */
int te_vec_set(const te_domain* d, int dtype, void* x, double alpha, int ncomp, te_stream s);
int te_vec_set_with_ghost(const te_domain* d, int dtype, void* x, double alpha, int ncomp, te_stream s);
int te_vec_scale(const te_domain* d, int dtype, void* x, double alpha, int ncomp, te_stream s);
int te_vec_shift(const te_domain* d, int dtype, void* x, double delta, int ncomp, te_stream s);
int te_vec_copy(const te_domain* d, int dtype, void* x, const void* b, int ncomp, te_stream s);
int te_vec_add(const te_domain* d, int dtype, void* x, const void* b, int ncomp, te_stream s);
/* x += alpha a */
int te_vec_add_scaled(const te_domain* d, int dtype, void* x, double alpha, const void* a, int ncomp,
                      te_stream s);
/* x += alpha a + beta b */
int te_vec_add_scaled2(const te_domain* d, int dtype, void* x, double alpha, const void* a, double beta,
                       const void* b, int ncomp, te_stream s);
/* x = alpha x + b */
int te_vec_scale_then_add(const te_domain* d, int dtype, void* x, double alpha, const void* b, int ncomp,
                          te_stream s);
/* x = alpha x + beta b */
int te_vec_scale_then_add_scaled(const te_domain* d, int dtype, void* x, double alpha, double beta,
                                 const void* b, int ncomp, te_stream s);
/* x = alpha x + beta b + gamma c */
int te_vec_scale_then_add_scaled2(const te_domain* d, int dtype, void* x, double alpha, double beta,
                                  const void* b, double gamma, const void* c, int ncomp, te_stream s);
/* Reductions: synchronous, result in *out (Vector::dot/twoNorm/infNorm; infNorm is
 * max(0, max v), Vector.h:568-579). */
int te_vec_dot(const te_domain* d, int dtype, const void* a, const void* b, int ncomp, double* out,
               te_stream s);
int te_vec_two_norm(const te_domain* d, int dtype, const void* a, int ncomp, double* out, te_stream s);
int te_vec_inf_norm(const te_domain* d, int dtype, const void* a, int ncomp, double* out, te_stream s);
/* DomainTools::Integrate (DomainTools.h:513-536): the sum over patches and components of
 * the interior values times the patch's cell volume (product of its spacings);
 * synchronous, result in *out. */
int te_vec_integrate(const te_domain* d, int dtype, const void* u, int ncomp, double* out, te_stream s);
/* Vector::dot for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The same dot product as te_vec_dot, but its partial sums (one per patch) live in a fixed
 * device array of the library instead of the domain's reduction workspace, the result is
 * written to the caller's device double *d_out, and the call is asynchronous on `stream`;
 * it allocates nothing.  Rejects domains of more than TE_NOALLOC_MAX_PATCHES patches.
 * The array is one per GPU and shared by every call on that GPU: calls that overlap on one
 * GPU race. */
#define TE_NOALLOC_MAX_PATCHES 65536
int te_vec_dot_noalloc(const te_domain* d, int dtype, const void* a, const void* b, int ncomp, double* d_out,
                       te_stream stream);
/* Element-type conversion of a whole padded vector (dst_dtype <- src_dtype). */
int te_vec_convert(const te_domain* d, int dst_dtype, void* dst, int src_dtype, const void* src, int ncomp,
                   te_stream s);

/* ------------------------------------------------ multigrid + Krylov --
 * GMG::CycleBuilder + VCycle (GMG/CycleBuilder.h:97-192, GMG/VCycle.h:46-70) over a
 * chain of domains (finest first), with per-level device workspaces allocated here,
 * once.  precision: TE_F64 = FP64 everywhere; TE_F32 = FP32 V-cycle (ghost fill,
 * stencil, smoother, restrict/interpolate) with the coarsest-level solve in FP64. */
/* GMG::VCycle (GMG/VCycle.h:46-70), GMG::WCycle (GMG/WCycle.h:50-84), GMG::FMGCycle
 * (GMG/FMGCycle.h:43-104).  W and F run one launch per step (the fused launches are
 * V-cycle only); 0 (V) is the zero-initialised default. */
enum te_cycle_type { TE_CYCLE_V = 0, TE_CYCLE_W = 1, TE_CYCLE_F = 2 };

typedef struct te_gmg_options {
  int filler_kind;    /* te_filler_kind */
  int fill_type;      /* te_fill_type */
  int op_kind;        /* 0 Poisson star, 1 VarPoisson star */
  int neumann;        /* Poisson star: Neumann physical boundaries */
  int smoother;       /* 0 DFTPatchSolver, 1 Iterative::PatchSolver(BiCGStab) */
  unsigned neumann_mask;
  int inner_max_it;   /* inner BiCGStab (smoother 1) */
  double inner_tol;
  int continue_on_breakdown;
  int pre_sweeps, post_sweeps, coarse_sweeps; /* CycleOpts (GMG/CycleOpts.h:33-55) */
  int precision;      /* TE_F64 or TE_F32 */
  int use_graph;      /* capture the V-cycle as a hipGraph */
  int cycle_type;     /* te_cycle_type: CycleOpts::cycle_type "V" / "W" / "F" (GMG/CycleBuilder.h:182-189) */
  int mid_sweeps;     /* CycleOpts::mid_sweeps (W and F cycles) */
  /* A smoother per level, as CycleBuilder::addFinestLevel / addIntermediateLevel /
   * addCoarsestLevel take one each (GMG/CycleBuilder.h:97-160): NULL = `smoother`,
   * `inner_max_it` and `inner_tol` on every level, else num_levels entries (finest first).
   * The fused 2D launches need the DFT smoother on every level; other mixes run one
   * launch per step. */
  const int* level_smoother;
  const int* level_inner_max_it;
  const double* level_inner_tol;
} te_gmg_options;

/* coeffs: VarPoisson beta per level (device float64 pointers, ghost-filled), or NULL. */
int te_gmg_create(const te_domain* const* levels, int num_levels, const te_gmg_options* opts,
                  const double* const* coeffs, te_gmg** out);
int te_gmg_destroy(te_gmg* g);
/* Cycle::apply(f, u) (GMG/Cycle.h:94-98): u := V-cycle(f); float64 vectors of level 0.
 * Stream-ordered and asynchronous: the cycle starts after the work enqueued on `stream` and
 * `stream`'s later work waits for it (events; the launches run on the solver's own stream
 * as one cached graph).  When `stream` is being captured into a graph, the launches are
 * enqueued on it directly, so the cycle becomes part of the caller's graph.  Exception: a
 * level with the BiCGStab patch smoother and continue_on_breakdown == 0 makes the call
 * synchronous (it reads the breakdown flag back) and not capturable. */
int te_gmg_vcycle(te_gmg* g, const double* f, double* u, te_stream stream);
/* Iterative::BiCGStab::solve(A, x, b, Mr) (Iterative/BiCGStab.h:111-203), A = level-0
 * operator, Mr = this V-cycle (use_precond 1) or none (0).  x holds the initial guess.
 * The reference's early exit (||s|| small) updates x += alpha p even with a
 * preconditioner (BiCGStab.h:162-167); use_precond 2 runs the V-cycle-preconditioned
 * solve with that exit corrected to x += alpha M p, which is what reaches the
 * tolerance in the true residual.
 * history (host, may be NULL) receives the recursive relative residual per iteration
 * (entry 0 = initial), up to history_cap entries.  *iterations = the reference's
 * return value.  Returns TE_ERR_BREAKDOWN when rho == 0.
 * Synchronous, like the reference's solve: it starts after the work enqueued on `stream`
 * (event), runs on the solver's own stream, reads the residual norm back once per
 * iteration and returns when x is final.  Not capturable into a graph (each iteration is
 * itself a cached graph). */
int te_gmg_bicgstab(te_gmg* g, const double* b, double* x, int use_precond, int max_it, double tol,
                    int* iterations, double* history, int history_cap, te_stream stream);
/* Iterative::CG::solve(A, x, b, Mr) (Iterative/CG.h:126-194), right-preconditioned by
 * this cycle (use_precond 1) or plain (0); x holds the initial guess and receives
 * Mr x + guess.  Like the reference, a zero b returns 0 iterations with x set to zero.
 * history as for te_gmg_bicgstab.  Returns TE_ERR_BREAKDOWN when rho == 0.  Synchronous, as
 * te_gmg_bicgstab. */
int te_gmg_cg(te_gmg* g, const double* b, double* x, int use_precond, int max_it, double tol, int* iterations,
              double* history, int history_cap, te_stream stream);
/* Seconds spent in the last te_gmg_bicgstab / te_gmg_cg, from the first A x to convergence. */
double te_gmg_last_solve_seconds(const te_gmg* g);
/* Measurement hook, no counterpart in the reference: the average duration in milliseconds
 * of `reps` back-to-back launches of one kernel of the finest level's cycle or of the
 * solver, each launched standalone on the solver's own work buffers and stream and timed
 * with HIP events recorded on that stream (one untimed launch first).  The work buffers'
 * contents are overwritten: call it between solves.  Supported where the cycle runs on
 * compact vectors (3D levels of 16^3 patches); TE_ERR_UNSUPPORTED otherwise. */
enum te_kernel_id {
  TE_K_SWEEP_ZERO = 0,      /* patch-solve sweep from u = 0 (first pre-smoothing sweep) */
  TE_K_SWEEP = 1,           /* sweep with the interface terms from u (later sweeps) */
  TE_K_SWEEP_INTERP = 2,    /* sweep on u + P uc (first post-smoothing; on large levels the boundary
                               layer of u + P uc, then a sweep on it) */
  TE_K_RESID_RESTRICT = 3,  /* r = f - A u restricted to the coarser level */
  TE_K_OPERATOR = 4,        /* outer FP64 A (M s) with the dots as.s and as.as (A M s stored only
                               where the updates read it back: TE_NO_XRFLY=1) */
  TE_K_UPDATE_XR = 5,       /* outer x and r updates with the dots r.rhat and r.r */
  TE_K_OPERATOR_P = 6,      /* outer FP64 A (M p) with the dot rhat.ap (stored with TE_NO_XRFLY=1) */
  TE_K_UPDATE_S = 7,        /* outer s = r - alpha A M p (+ its FP32 copy) with the dot s.s */
  TE_K_UPDATE_P = 8,        /* outer p = r + beta (p - omega A M p) (+ its FP32 copy) */
  TE_K_X_COMBINE = 9        /* mixed precision: the deferred x updates of 8 iterations,
                               x += alpha_j M p_j + omega_j M s_j (TE_NO_XDEFER=1: none) */
};
int te_gmg_time_kernel(te_gmg* g, int which, int reps, double* ms);

/* ------------------------------------------- batches of independent problems --
 * Many independent 2D problems solved together.  The reference solves each with its own
 * GMG::CycleBuilder + Iterative::BiCGStab (GMG/CycleBuilder.h:97-192, GMG/VCycle.h:46-70,
 * Iterative/BiCGStab.h:111-203); a batch runs exactly those steps for every problem, with one
 * launch per cycle step and per BiCGStab step for all problems of one patch size and one
 * graph per outer iteration.  Each problem's inner products have the bits of its own
 * te_gmg_bicgstab, so its iterations and result are those of solving it alone (with the
 * cycle's one-launch-per-level forms: TE_NO_TAIL, TE_NO_RRPRE, TE_NO_BSWEEP).
 * Problem q's hierarchy is levels[first(q)] .. levels[first(q) + nlev[q] - 1], finest first
 * as te_gmg_create takes it, first(q) = nlev[0] + ... + nlev[q - 1]; every problem uses
 * `opts`, which must select the V-cycle, the Poisson operator and the DFT smoother.  The
 * domains must outlive the batch. */
typedef struct te_batch te_batch;
int te_batch_create(int nprob, const te_domain* const* levels, const int* nlev, const te_gmg_options* opts,
                    te_batch** out);
int te_batch_destroy(te_batch* b);
/* Patch-size groups of the batch (each runs its own launches, on its own stream). */
int te_batch_num_groups(const te_batch* b);
/* Cycle::apply for every problem: f[q], u[q] are problem q's own padded FP64 vectors of its
 * finest level.  Stream-ordered after `stream`; `stream` waits for the cycle. */
int te_batch_vcycle(te_batch* b, const double* const* f, double* const* u, te_stream stream);
/* Iterative::BiCGStab::solve for every problem: b[q], x[q] problem q's own padded FP64
 * vectors (x[q] the initial guess, overwritten), iterations[q] its iteration count,
 * residuals[q] (may be null) its last relative residual norm.  A problem whose rho
 * becomes 0 stops there; the call then returns TE_ERR_BREAKDOWN after all others have
 * converged (te_last_error names the first such problem).  Synchronous. */
int te_batch_bicgstab(te_batch* b, const double* const* bv, double* const* x, int use_precond, int max_it, double tol,
                      int* iterations, double* residuals, te_stream stream);

/* ------------------------------------------------------------ autotuner --
 * No counterpart in the reference.  Launch variants per patch shape: every variant of a
 * family computes each patch with the same arithmetic in the same order, only the mapping
 * of patches onto workgroups and waves differs, so the choice never changes a result.
 * thunderegg_amd/autotune.py times the variants (te_tune_time_dft) and writes the table,
 * lines of "family dim n dtype variant"; the library loads tune_table.txt beside itself (or
 * the file TE_TUNE_TABLE names; TE_TUNE_TABLE=none keeps the hand defaults, variant 0) on
 * first use. */
enum te_tune_family {
  TE_TUNE_DFT_GENERIC = 0, /* VALU DFT patch solve: patches per workgroup x threads per patch */
  TE_TUNE_DFT_MFMA_2D = 1, /* 2D matrix-core DFT patch solve (n 16, 32): workgroups per launch */
  TE_TUNE_C3_SWEEP = 2     /* 3D 16^3 fused sweeps: slices whose loads are issued together, waves per SIMD */
};
int te_tune_num_variants(int family);
int te_tune_load(const char* path);
int te_tune_clear(void);
int te_tune_set(int family, int dim, int n, int dtype, int variant);
int te_tune_get(int family, int dim, int n, int dtype);
/* the table file loaded ("" for none) into buf; returns its length */
int te_tune_table_path(char* buf, size_t cap);
/* average milliseconds of `reps` launches of plan pl's patch solve (u's ghosts read from u)
 * with variant `variant` forced, after one untimed launch; f, u, out of type dtype */
int te_tune_time_dft(const te_dft_plan* pl, int dtype, const void* f, const void* u, void* out, int variant, int reps,
                     double* ms);
/* te_gmg_time_kernel with a TE_TUNE_C3_SWEEP variant forced (the 3D compact cycle's sweeps) */
int te_tune_time_c3(te_gmg* g, int which, int variant, int reps, double* ms);

/* ------------------------------------------------------------- timings --
 * Timer (Timer.h:28-204, Timer.cpp:160-781): nested named timings, each optionally tied to
 * a domain id and a patch id (TE_TIMER_NONE = none, the reference's INT_MAX), with call
 * count, sum, min and max in seconds and named int / double infos.  te_timer_start /
 * te_timer_stop time on the host clock (Timer::start / stop, startDomainTiming,
 * startPatchTiming and their stops, with name, domain id and patch id in one call);
 * the _on variants record HIP events on a stream instead, read when the timer is
 * serialized or printed.  Errors follow the reference's RuntimeErrors: a stop that does not
 * match the innermost started timing, a domain id that was not added, an info with no
 * timing started or of the other type, adding a domain id twice.  A timer is not
 * thread-safe. */
#define TE_TIMER_NONE 2147483647
typedef struct te_timer te_timer;
int te_timer_create(te_timer** out);
int te_timer_destroy(te_timer* t);
/* Timer::addDomain(domain_id, domain) (Timer.h:71): the domain's patches (dim 2 or 3, ns
 * cells per side) serialized in the reference's mesh schema (json.cpp:161-247). */
int te_timer_add_domain(te_timer* t, int domain_id, int dim, const int* ns, int num_patches,
                        const te_patch_info* patches);
int te_timer_start(te_timer* t, const char* name, int domain_id, int patch_id);
int te_timer_stop(te_timer* t, const char* name, int domain_id, int patch_id);
int te_timer_start_on(te_timer* t, const char* name, int domain_id, int patch_id, te_stream stream);
int te_timer_stop_on(te_timer* t, const char* name, int domain_id, int patch_id, te_stream stream);
/* Timer::addIntInfo / addDoubleInfo (Timer.h:120-145): to the innermost started timing. */
int te_timer_add_int_info(te_timer* t, const char* name, int info);
int te_timer_add_double_info(te_timer* t, const char* name, double info);
/* to_json(json&, const Timer&) (Timer.cpp:700-745) as text: indent < 0 compact, else
 * pretty-printed with that many spaces; "null" for an empty timer.  *len = the text's length;
 * with buf != NULL the text and a terminating 0 are copied when cap > *len (else
 * TE_ERR_ARG). */
int te_timer_to_json(te_timer* t, int indent, char* buf, size_t cap, size_t* len);
/* operator<<(ostream&, const Timer&) (Timer.cpp:608-640): the text report, sibling timings
 * merged by name; buffer protocol of te_timer_to_json.  Fails with unfinished timings. */
int te_timer_print(te_timer* t, char* buf, size_t cap, size_t* len);
/* Timer::saveToFile (Timer.cpp:747-759): the JSON, indented by 4. */
int te_timer_save(te_timer* t, const char* filename);
/* Attach a timer to a solver (NULL detaches).  iterations != 0: te_gmg_bicgstab and
 * te_gmg_cg time each outer iteration as "Iteration" (Iterative/BiCGStab.h:145-200,
 * Iterative/CG.h:158-186).  domain_ids (num_levels entries, or NULL for none): every
 * smoothing sweep of level l is timed as the domain timing "Total Patch Smooth" of
 * domain_ids[l] (PatchSolver.h:153-168), which must have been added to the timer;
 * TE_TIMER_NONE leaves a level untimed.  Timings are recorded by events on the solver's
 * stream, with its launches enqueued one by one instead of as graphs; the arithmetic is
 * unchanged.  On the fused 2D path a sweep's launch also does the transfer fused into it
 * (interpolation into the first post-smoothing sweep, the finer level's residual and
 * restriction into the first sweep of a level of at most 1024 patches), so its timing
 * includes that work; the reference's per-patch "Single Patch Solve" timings have no
 * counterpart (a sweep solves all patches of the level in one launch).  The solver keeps
 * the pointer: detach (t = NULL) before te_timer_destroy. */
int te_gmg_set_timer(te_gmg* g, te_timer* t, int iterations, const int* domain_ids);

/* ----------------------------------------------- mesh generation (host) --
 * Quadtree / octree forest on the unit square / cube, replacing
 * P4estDomainGenerator / P8estDomainGenerator (P4estDomainGenerator.cpp:205-738):
 * Morton-ordered leaves, ids in leaf order, one-level-at-a-time coarsening of
 * complete families, neighbour linking of faces, edges and corners. */
int te_forest_create(int dim, int base_level, te_forest** out);
int te_forest_destroy(te_forest* f);
int te_forest_num_leaves(const te_forest* f);
/* levels[i], coords[i*dim + a] (integer coordinates in units of the leaf size). */
int te_forest_get_leaves(const te_forest* f, int32_t* levels, int32_t* coords);
/* Refine every leaf with flags[i] != 0 into its 2^dim children. */
int te_forest_refine(te_forest* f, const uint8_t* flags);
/* 2:1 balance across faces, edges and corners (p4est_balance with CONNECT_FULL). */
int te_forest_balance(te_forest* f);
/* Domains finest (index 0) to coarsest: number of levels, and patch counts. */
int te_forest_build_domains(te_forest* f, const int* ns, int* num_levels);
int te_forest_domain_size(const te_forest* f, int level, int* num_patches);
int te_forest_domain_patches(const te_forest* f, int level, te_patch_info* out);

#ifdef __cplusplus
}
#endif
#endif
