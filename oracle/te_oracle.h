/*
 * te_oracle.h -- FP64 CPU restatement of ThunderEgg's patch-parallel GMG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the oracle / reported CPU baseline.  The product never links it.
 *
 * It follows the reference algorithm statement by statement, in the reference's
 * own "push" order (zero ghosts -> self terms -> neighbour pushes), so that it is
 * an independent formulation from the GPU's gather ("pull") kernels.  Each
 * function cites the reference file:line it restates (paths relative to the
 * reference's src/ThunderEgg/).
 *
 * Parity pinning of each function is listed in the header comment of
 * te_oracle.c and in the tests that check it.
 */
#ifndef TE_ORACLE_H
#define TE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same binary layout as te_patch_info in include/thunderegg_amd.h (one patch
 * table format feeds both the oracle and the product).  Neighbour slots are
 * indexed Face<D,M>::sum_of_faces + face value (Face.h:67-79, PatchInfo.h:222):
 * 2D: corners 0-3, sides 4-7; 3D: corners 0-7, edges 8-19, sides 20-25. */
typedef struct ora_patch_info {
  int32_t id, rank, parent_id, parent_rank;
  int32_t orth_on_parent; /* -1 == Orthant::null() */
  int32_t refine_level;
  int32_t child_ids[8];
  int32_t child_ranks[8];
  double starts[3];
  double spacings[3];
  int32_t nbr_type[26]; /* -1 none, 0 Normal, 1 Coarse, 2 Fine (NbrType.h:33-48) */
  int32_t nbr_orth[26]; /* CoarseNbrInfo::orth_on_coarse, -1 otherwise */
  int32_t nbr_ids[26][4];
  int32_t nbr_ranks[26][4];
} ora_patch_info;

typedef struct ora_domain ora_domain;

/* Domain (Domain.h:94-248): local_index = position; neighbour local indexes by id. */
ora_domain* ora_domain_create(int dim, const int* ns, int num_ghost, int num_patches,
                              const ora_patch_info* patches);
void ora_domain_destroy(ora_domain* d);
int ora_domain_nbr_local(const ora_domain* d, int patch, int slot, int k);
/* InterLevelComm (GMG/InterLevelComm.h:125-224): parent local index per fine patch. */
int ora_domain_link_parents(ora_domain* fine, const ora_domain* coarse);
int ora_domain_parent_local(const ora_domain* fine, int patch);
long ora_domain_vec_len(const ora_domain* d, int ncomp);

/* Enumerated local ghost-fill calls (MPIGhostFiller.h:821-895), in execution
 * order.  Each record = {M, local_index, face, nbr_type, orthant, nbr_local}. */
int ora_enumerate_calls(const ora_domain* d, int fill_type, int* out, int cap);

/* Ghost fillers: kind 0 BiLinear, 1 BiQuadratic, 2 TriLinear;
 * fill_type 0 Faces, 1 Edges, 2 Corners (GhostFillingType.h).  Returns 0 or <0. */
int ora_fill_ghost(const ora_domain* d, int kind, int fill_type, double* u, int ncomp);

/* Poisson::StarPatchOperator::apply (PatchOperator.h:126-143 + Poisson/StarPatchOperator.h:75-101). */
int ora_poisson_apply(const ora_domain* d, int kind, int fill_type, int neumann,
                      double* u, double* f, int ncomp);
/* VarPoisson::StarPatchOperator::apply (VarPoisson/StarPatchOperator.h:61-109);
 * coeffs must already be ghost-filled (the reference fills them in its constructor, :70). */
int ora_varpoisson_apply(const ora_domain* d, int kind, int fill_type, const double* coeffs,
                         double* u, double* f, int ncomp);

/* Smoothers.  PatchSolver::smooth (PatchSolver.h:142-170). */
int ora_dft_smooth(const ora_domain* d, int kind, int fill_type, int op_kind,
                   const double* coeffs, unsigned neumann_mask, const double* f, double* u);
int ora_bicgstab_smooth(const ora_domain* d, int kind, int fill_type, int op_kind,
                        const double* coeffs, int neumann, int max_it, double tol,
                        int continue_on_breakdown, const double* f, double* u, int ncomp,
                        int* iters_out);

/* GMG::LinearRestrictor / MPIRestrictor::restrict (GMG/LinearRestrictor.h:101-191,
 * GMG/MPIRestrictor.h:55-83), extrapolate_boundary_ghosts = false. */
int ora_restrict(const ora_domain* fine, const ora_domain* coarse, const double* fine_v,
                 double* coarse_v, int ncomp);
/* The same with extrapolate_boundary_ghosts (GMG/LinearRestrictor.h:49-92, 138-155) when
 * extrapolate != 0: the coarse ghost cells over the fine patches' exterior sides and
 * neighbourless orthant corners get the extrapolated fine ghosts; copies include ghosts. */
int ora_restrict_ex(const ora_domain* fine, const ora_domain* coarse, const double* fine_v,
                    double* coarse_v, int ncomp, int extrapolate);
/* GMG::DirectInterpolator (GMG/DirectInterpolator.h:60-91): fine += P coarse. */
int ora_interpolate(const ora_domain* coarse, const ora_domain* fine, const double* coarse_v,
                    double* fine_v, int ncomp);

/* GMG level description for the V-cycle (GMG/Level.h, GMG/VCycle.h). */
typedef struct ora_level {
  const ora_domain* dom;
  int filler_kind;
  int fill_type;
  int op_kind;            /* 0 Poisson star, 1 VarPoisson star */
  int neumann;            /* Poisson star: Neumann physical BCs */
  const double* coeffs;   /* VarPoisson: ghost-filled beta on this level */
  int smoother_kind;      /* 0 DFTPatchSolver, 1 Iterative::PatchSolver(BiCGStab) */
  unsigned neumann_mask;  /* DFT: Neumann bitset over sides */
  int inner_max_it;       /* Iterative::PatchSolver inner BiCGStab */
  double inner_tol;
  int continue_on_breakdown;
} ora_level;

/* GMG::Cycle::apply + VCycle::visit (GMG/Cycle.h:57-98, GMG/VCycle.h:46-70), ncomp = 1. */
int ora_vcycle_apply(const ora_level* levels, int nlevels, int pre, int post, int coarse,
                     const double* f, double* u);

/* GMG::Cycle::apply with VCycle (type 0), WCycle (1, GMG/WCycle.h:50-84) or FMGCycle
 * (2, GMG/FMGCycle.h:43-104) visits; mid = CycleOpts::mid_sweeps. */
int ora_cycle_apply(const ora_level* levels, int nlevels, int type, int pre, int post, int mid, int coarse,
                    const double* f, double* u);

/* Iterative::CG::solve (Iterative/CG.h:126-194), right-preconditioned by the cycle over
 * `levels` (use_precond 1) or plain (0).  Returns iterations, -1 on breakdown (rho == 0). */
int ora_cg_solve(const ora_level* levels, int nlevels, int use_precond, int type, int pre, int post, int mid,
                 int coarse, const double* b, double* x, int max_it, double tol, double* hist, int hist_cap);

/* Iterative::BiCGStab::solve (Iterative/BiCGStab.h:111-203) with A = level 0's operator
 * and Mr = the V-cycle over `levels` (use_precond 1), or none (0).  use_precond 2 is the
 * V-cycle with the early exit updating x by alpha M p instead of the reference's alpha p.
 * Residual history (recursive ||r||/||b||, entry 0 = initial) written to hist.
 * Returns iterations, or -1 on breakdown (rho == 0). */
int ora_bicgstab_solve(const ora_level* levels, int nlevels, int use_precond, int pre, int post,
                       int coarse, const double* b, double* x, int max_it, double tol,
                       double* hist, int hist_cap);

/* Vector (Vector.h:391-611), interior only. */
double ora_vec_dot(const ora_domain* d, const double* a, const double* b, int ncomp);
double ora_vec_two_norm(const ora_domain* d, const double* a, int ncomp);
double ora_vec_inf_norm(const ora_domain* d, const double* a, int ncomp);
/* DomainTools::Integrate (DomainTools.h:513-536). */
double ora_vec_integrate(const ora_domain* d, const double* u, int ncomp);
/* StarPatchOperator::addDrichletBCToRHS / addNeumannBCToRHS with sampled boundary data */
int ora_add_bc(const ora_domain* d, double* f, const double* g, const double* c, int neumann);

/* DFTPatchSolver transform matrices (Poisson/DFTPatchSolver.h:131-196): type 0 DCT_II,
 * 1 DCT_III, 2 DCT_IV, 3 DST_II, 4 DST_III, 5 DST_IV; out[i*n+j] row-major. */
void ora_dft_matrix(int type, int n, double* out);

#ifdef __cplusplus
}
#endif
#endif
