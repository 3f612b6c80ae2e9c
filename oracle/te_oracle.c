/*
 * te_oracle.c -- FP64 CPU restatement of ThunderEgg's patch-parallel GMG path.
 *
 * TEST INFRASTRUCTURE ONLY (see te_oracle.h).  Never part of the product path.
 *
 * Written from the reference's behaviour, statement by statement, in its "push"
 * formulation: MPIGhostFiller::fillGhost zeroes neighbour-facing ghosts, adds
 * each patch's self terms, then each patch pushes (+=, or = for same-level
 * neighbours) into its neighbours' ghosts in the enumerated call order.  The
 * HIP kernels instead gather ("pull") every ghost value, so the two
 * formulations are independent and only agree if both encode the same weights.
 *
 * How each part is pinned (tests/test_oracle_pins.py; the forest and the mesh
 * tables are checked in tests/test_host.py):
 *  - tables (Domain, enumerateCalls): the reference's mesh fixtures (the JSON
 *    meshes of test/mesh_inputs); the local face calls are exactly one per
 *    (patch, side) neighbour, in patch / face / orthant order, with the table's
 *    neighbour indexes (test/MPIGhostFiller_MPI1.cpp).
 *  - ghost fillers: exactness on linear (BiLinear, TriLinear) and quadratic
 *    (BiQuadratic) functions at ghost-cell centres on the reference's refined
 *    meshes, with zeroed and with garbage ghosts, even and odd n
 *    (test/BiLinearGhostFiller_MPI1.cpp, test/BiQuadraticGhostFiller_MPI1.cpp,
 *    test/TriLinearGhostFiller_MPI1.cpp); a cubic is not reproduced (the test
 *    has teeth).
 *  - star operators: ghost fill + stencil give the exact Laplacian of quadratics
 *    away from the physical boundary on refined meshes; 2nd-order convergence on
 *    uniform meshes (test/Poisson/StarPatchOperator_MPI1.cpp,
 *    test/VarPoisson/StarPatchOperator_MPI1.cpp); VarPoisson with constant beta is
 *    beta x Poisson; the uniform-mesh matrix is symmetric negative definite.
 *  - DFT patch solver: every transform matrix against scipy.fft dst/dct
 *    (types II/III/IV); the solve is the exact inverse of the single-patch
 *    operator (2D and 3D); 2nd-order convergence of one block-Jacobi sweep
 *    (test/Poisson/DFTPatchSolver_MPI1.cpp).
 *  - restrict / interpolate: exactness on linears (test/GMG/LinearRestrictor_MPI1.cpp)
 *    and the known answer of test/GMG/DirectInterpolator_MPI1.cpp.
 *  - V-cycle: one level and one patch is the exact inverse; two levels equal the
 *    VCycle::visit sequence composed from the separately pinned smoother,
 *    operator, restrictor and interpolator.
 *  - BiCGStab: converges to the dense direct solve of the assembled operator
 *    (plain, preconditioned, corrected early exit); zero right-hand side and a
 *    converged initial guess take 0 iterations (test/Iterative/BiCGStab_MPI1.cpp).
 *  - W- and F-cycles: equal the WCycle / FMGCycle::visit sequences composed from the
 *    pinned parts on three levels; V through the general entry equals ora_vcycle_apply;
 *    W and F coincide on two levels and differ on three; both are linear in f.
 *  - CG: converges to the dense direct solve (plain, V- and W-preconditioned) on a
 *    symmetric uniform-mesh operator, the preconditioned ones in fewer iterations; a zero
 *    right-hand side leaves x at zero (the reference zeroes x before testing ||b||).
 *  - vector reductions: ghosts excluded, symmetry and bilinearity of dot, twoNorm^2 =
 *    dot(a, a), infNorm = max(0, max v).
 *  - Integrate: the midpoint rule's closed forms on every level of the fixture meshes
 *    (the integral of 1 is the meshes' box volume; products of affine functions exact;
 *    x^2 short by h^2 / 12 per unit volume of each patch); ghosts do not enter.
 *  - Neumann variants of the DFT solve: with any set of Neumann sides (2D and 3D) the
 *    solve is the exact inverse of an independently assembled dense one-patch operator
 *    (ghost = -u on Dirichlet sides, u on Neumann sides); the zero-mean solution when
 *    every side is Neumann.
 *  - inner BiCGStab patch smoother (PatchSolver.h:153-168): one sweep at a tight inner
 *    tolerance equals one DFT sweep on refined meshes (Dirichlet, Neumann, VarPoisson
 *    with beta = 1), so its interface terms and internal boundary conditions are those
 *    of the pinned DFT path.
 */
#include "te_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

enum { NB_NONE = -1, NB_NORMAL = 0, NB_COARSE = 1, NB_FINE = 2 };
enum { FT_FACES = 0, FT_EDGES = 1, FT_CORNERS = 2 };
enum { K_BILINEAR = 0, K_BIQUADRATIC = 1, K_TRILINEAR = 2 };

struct ora_domain {
  int dim;
  int ns[3];
  int g;
  int np;
  ora_patch_info* p;
  int* nbr_local;    /* np * 26 * 4 */
  int* parent_local; /* np */
  long cells_ghost;  /* per component per patch */
  long stride[3];
};

/* ---------------------------------------------------------------- faces --
 * Face<D,M> encodings (Face.h:41-584). */
static int choose_(int n, int k)
{
  int r = 1;
  for (int i = 0; i < k; i++) r = r * (n - i) / (i + 1);
  return r;
}
static int num_faces(int D, int M) { return (1 << (D - M)) * choose_(D, M); }
static int sum_of_faces(int D, int M)
{
  int s = 0;
  for (int m = 0; m < M; m++) s += num_faces(D, m);
  return s;
}
/* Face::getSides (Face.h:496-524): the D-M sides that meet in this face, in axis order. */
static int face_sides(int D, int M, int val, int* sides)
{
  if (M == D - 1) {
    sides[0] = val;
    return 1;
  }
  if (M == 0) {
    for (int i = 0; i < D; i++) sides[i] = 2 * i + ((val >> i) & 1);
    return D;
  }
  /* D == 3, M == 1: value >> 2 is the axis parallel to the edge */
  int not_axis = val >> 2, k = 0;
  for (int i = 0; i < 3; i++) {
    if (i == not_axis) continue;
    int bit = (val >> k) & 1;
    sides[k] = (i << 1) ^ bit;
    k++;
  }
  return 2;
}
static int face_free_axes(int D, int M, int val, int* axes)
{
  int sides[3];
  int nsd = face_sides(D, M, val, sides);
  int k = 0;
  for (int a = 0; a < D; a++) {
    int fixed = 0;
    for (int s = 0; s < nsd; s++)
      if ((sides[s] >> 1) == a) fixed = 1;
    if (!fixed) axes[k++] = a;
  }
  return k;
}
static int face_opposite(int D, int M, int val) { return val ^ ((1 << (D - M)) - 1); }

/* --------------------------------------------------------------- domain -- */
static int cmp_pair(const void* a, const void* b)
{
  const int* x = (const int*)a;
  const int* y = (const int*)b;
  return (x[0] > y[0]) - (x[0] < y[0]);
}
static int find_id(const int* sorted_pairs, int n, int id)
{
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    int v = sorted_pairs[2 * mid];
    if (v == id) return sorted_pairs[2 * mid + 1];
    if (v < id) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

ora_domain* ora_domain_create(int dim, const int* ns, int num_ghost, int num_patches,
                              const ora_patch_info* patches)
{
  if (dim < 2 || dim > 3 || num_ghost < 1 || num_patches < 0) return NULL;
  ora_domain* d = (ora_domain*)calloc(1, sizeof(ora_domain));
  d->dim = dim;
  d->g = num_ghost;
  d->np = num_patches;
  d->cells_ghost = 1;
  for (int a = 0; a < 3; a++) d->ns[a] = 1;
  for (int a = 0; a < dim; a++) {
    d->ns[a] = ns[a];
    d->stride[a] = d->cells_ghost;
    d->cells_ghost *= ns[a] + 2 * num_ghost;
  }
  d->p = (ora_patch_info*)malloc(sizeof(ora_patch_info) * (num_patches ? num_patches : 1));
  memcpy(d->p, patches, sizeof(ora_patch_info) * num_patches);
  d->nbr_local = (int*)malloc(sizeof(int) * 26 * 4 * (num_patches ? num_patches : 1));
  d->parent_local = (int*)malloc(sizeof(int) * (num_patches ? num_patches : 1));
  /* indexPatchesLocal (Domain.h:94-109): local index = position */
  int* pairs = (int*)malloc(sizeof(int) * 2 * (num_patches ? num_patches : 1));
  for (int i = 0; i < num_patches; i++) {
    pairs[2 * i] = patches[i].id;
    pairs[2 * i + 1] = i;
  }
  qsort(pairs, num_patches, 2 * sizeof(int), cmp_pair);
  for (int i = 0; i < num_patches; i++) {
    d->parent_local[i] = -1;
    for (int s = 0; s < 26; s++) {
      for (int k = 0; k < 4; k++) {
        int idx = -1;
        if (patches[i].nbr_type[s] != NB_NONE) idx = find_id(pairs, num_patches, patches[i].nbr_ids[s][k]);
        d->nbr_local[(i * 26 + s) * 4 + k] = idx;
      }
    }
  }
  free(pairs);
  return d;
}

void ora_domain_destroy(ora_domain* d)
{
  if (!d) return;
  free(d->p);
  free(d->nbr_local);
  free(d->parent_local);
  free(d);
}

int ora_domain_nbr_local(const ora_domain* d, int patch, int slot, int k)
{
  return d->nbr_local[(patch * 26 + slot) * 4 + k];
}

int ora_domain_link_parents(ora_domain* fine, const ora_domain* coarse)
{
  int n = coarse->np;
  int* pairs = (int*)malloc(sizeof(int) * 2 * (n ? n : 1));
  for (int i = 0; i < n; i++) {
    pairs[2 * i] = coarse->p[i].id;
    pairs[2 * i + 1] = i;
  }
  qsort(pairs, n, 2 * sizeof(int), cmp_pair);
  int bad = 0;
  for (int i = 0; i < fine->np; i++) {
    fine->parent_local[i] = find_id(pairs, n, fine->p[i].parent_id);
    if (fine->parent_local[i] < 0) bad++;
  }
  free(pairs);
  return bad ? -1 : 0;
}

int ora_domain_parent_local(const ora_domain* fine, int patch) { return fine->parent_local[patch]; }

long ora_domain_vec_len(const ora_domain* d, int ncomp) { return (long)d->np * ncomp * d->cells_ghost; }

/* -------------------------------------------------------------- helpers -- */
static inline long cell_index(const ora_domain* d, int p, int c, int ncomp, const int* x)
{
  long idx = ((long)p * ncomp + c) * d->cells_ghost;
  for (int a = 0; a < d->dim; a++) idx += (long)(x[a] + d->g) * d->stride[a];
  return idx;
}
/* Coordinates of a cell in the slice of face f: offs[k] for each fixed side k
 * (0 = first interior layer, -1 = first ghost layer; PatchView.h:199-265),
 * free[k] for each free axis in increasing order. */
static void slice_coord(const ora_domain* d, int M, int f, const int* offs, const int* freei, int* x)
{
  int sides[3], axes[3];
  int nsd = face_sides(d->dim, M, f, sides);
  int nfr = face_free_axes(d->dim, M, f, axes);
  for (int k = 0; k < nsd; k++) {
    int a = sides[k] >> 1;
    x[a] = (sides[k] & 1) ? d->ns[a] - 1 - offs[k] : offs[k];
  }
  for (int k = 0; k < nfr; k++) x[axes[k]] = freei[k];
}
static inline double* sl(const ora_domain* d, double* u, int ncomp, int p, int c, int M, int f, int off,
                         const int* freei)
{
  int offs[3] = { off, off, off };
  int x[3] = { 0, 0, 0 };
  slice_coord(d, M, f, offs, freei, x);
  return &u[cell_index(d, p, c, ncomp, x)];
}
static inline int has_nbr(const ora_domain* d, int p, int slot) { return d->p[p].nbr_type[slot] != NB_NONE; }
static inline int side_slot(int D, int s) { return sum_of_faces(D, D - 1) + s; }
static int patch_has_any_nbr(const ora_domain* d, int p)
{
  int n = sum_of_faces(d->dim, d->dim);
  for (int s = 0; s < n; s++)
    if (has_nbr(d, p, s)) return 1;
  return 0;
}

/* ------------------------------------------------------ call enumeration --
 * MPIGhostFiller::enumerateCalls (MPIGhostFiller.h:821-895), one rank:
 * per patch in local order, per face value ascending, Fine -> one call per orthant. */
typedef struct {
  int M, local, face, type, orth, nbr;
} ora_call;

static int m_included(int D, int fill_type, int M)
{
  if (M == D - 1) return 1;
  if (fill_type == FT_CORNERS) return 1;
  if (fill_type == FT_EDGES && D == 3 && M == 1) return 1;
  return 0;
}

static int build_calls(const ora_domain* d, int fill_type, ora_call* out, int cap)
{
  int D = d->dim, n = 0;
  for (int M = D - 1; M >= 0; M--) {
    if (!m_included(D, fill_type, M)) continue;
    for (int p = 0; p < d->np; p++) {
      for (int f = 0; f < num_faces(D, M); f++) {
        int slot = sum_of_faces(D, M) + f;
        int t = d->p[p].nbr_type[slot];
        if (t == NB_NONE) continue;
        int cnt = (t == NB_FINE) ? (1 << M) : 1;
        for (int i = 0; i < cnt; i++) {
          if (out && n < cap) {
            out[n].M = M;
            out[n].local = p;
            out[n].face = f;
            out[n].type = t;
            out[n].orth = (t == NB_FINE) ? i : (t == NB_COARSE ? d->p[p].nbr_orth[slot] : -1);
            out[n].nbr = d->nbr_local[(p * 26 + slot) * 4 + (t == NB_FINE ? i : 0)];
          }
          n++;
        }
      }
    }
  }
  return n;
}

int ora_enumerate_calls(const ora_domain* d, int fill_type, int* out, int cap)
{
  int n = build_calls(d, fill_type, NULL, 0);
  if (!out) return n;
  ora_call* calls = (ora_call*)malloc(sizeof(ora_call) * (n ? n : 1));
  build_calls(d, fill_type, calls, n);
  for (int i = 0; i < n && i < cap; i++) {
    out[6 * i + 0] = calls[i].M;
    out[6 * i + 1] = calls[i].local;
    out[6 * i + 2] = calls[i].face;
    out[6 * i + 3] = calls[i].type;
    out[6 * i + 4] = calls[i].orth;
    out[6 * i + 5] = calls[i].nbr;
  }
  free(calls);
  return n;
}

/* ----------------------------------------------------------- ghost fill -- */
typedef struct {
  const ora_domain* d;
  double* u;
  int nc;
  int kind;
  int ft;
} gctx;


/* zeroGhostCells (MPIGhostFiller.h:904-947): every ghost layer of each face with a nbr. */
static void zero_ghosts(const gctx* c)
{
  const ora_domain* d = c->d;
  int D = d->dim;
  for (int p = 0; p < d->np; p++) {
    for (int M = D - 1; M >= 0; M--) {
      if (!m_included(D, c->ft, M)) continue;
      for (int f = 0; f < num_faces(D, M); f++) {
        if (!has_nbr(d, p, sum_of_faces(D, M) + f)) continue;
        int axes[3];
        int nfr = face_free_axes(D, M, f, axes);
        int nfix = D - M;
        int gl = 1;
        for (int k = 0; k < nfix; k++) gl *= d->g;
        for (int go = 0; go < gl; go++) {
          int offs[3], rem = go;
          for (int k = 0; k < nfix; k++) {
            offs[k] = -1 - (rem % d->g);
            rem /= d->g;
          }
          int cnt = 1;
          for (int k = 0; k < nfr; k++) cnt *= d->ns[axes[k]];
          for (int cc = 0; cc < c->nc; cc++) {
            for (int li = 0; li < cnt; li++) {
              int freei[3], r2 = li;
              for (int k = 0; k < nfr; k++) {
                freei[k] = r2 % d->ns[axes[k]];
                r2 /= d->ns[axes[k]];
              }
              int x[3] = { 0, 0, 0 };
              slice_coord(d, M, f, offs, freei, x);
              c->u[cell_index(d, p, cc, c->nc, x)] = 0.0;
            }
          }
        }
      }
    }
  }
}

/* ---- BiLinearGhostFiller (BiLinearGhostFiller.cpp:40-388) ---- */
static void bl_nbr_side(const gctx* c, int lp, int np_, int s, int type, int orth)
{
  const ora_domain* d = c->d;
  int D = 2, M = 1, opp = face_opposite(D, M, s);
  int fa[1];
  face_free_axes(D, M, s, fa);
  int n = d->ns[fa[0]];
  int off = (orth == 1) ? n : 0;
  for (int cc = 0; cc < c->nc; cc++) {
    for (int t = 0; t < n; t++) {
      int ti = t;
      if (type == NB_NORMAL) { /* :40-49 */
        *sl(d, c->u, c->nc, np_, cc, M, opp, -1, &ti) = *sl(d, c->u, c->nc, lp, cc, M, s, 0, &ti);
      } else if (type == NB_COARSE) { /* :63-78 nbr is coarse */
        int tc = (t + off) / 2;
        *sl(d, c->u, c->nc, np_, cc, M, opp, -1, &tc) += 2.0 / 3.0 * *sl(d, c->u, c->nc, lp, cc, M, s, 0, &ti);
      } else { /* :92-107 nbr is fine */
        int tc = (t + off) / 2;
        *sl(d, c->u, c->nc, np_, cc, M, opp, -1, &ti) += 2.0 / 3.0 * *sl(d, c->u, c->nc, lp, cc, M, s, 0, &tc);
      }
    }
  }
}
static void bl_nbr_corner(const gctx* c, int lp, int np_, int f, int type)
{
  const ora_domain* d = c->d;
  int opp = face_opposite(2, 0, f);
  for (int cc = 0; cc < c->nc; cc++) {
    double v = *sl(d, c->u, c->nc, lp, cc, 0, f, 0, NULL);
    double* gp = sl(d, c->u, c->nc, np_, cc, 0, opp, -1, NULL);
    if (type == NB_NORMAL) *gp = v;             /* :167-177 */
    else if (type == NB_COARSE) *gp += 4.0 * v / 3.0; /* :190-200 */
    else *gp += 2.0 * v / 3.0;                  /* :213-223 */
  }
}
static void bl_local(const gctx* c, int p)
{
  const ora_domain* d = c->d;
  int D = 2;
  if (c->ft == FT_CORNERS) { /* :236-264 */
    for (int f = 0; f < 4; f++) {
      int t = d->p[p].nbr_type[f];
      if (t == NB_NONE || t == NB_NORMAL) continue;
      for (int cc = 0; cc < c->nc; cc++) {
        double v = *sl(d, c->u, c->nc, p, cc, 0, f, 0, NULL);
        double* gp = sl(d, c->u, c->nc, p, cc, 0, f, -1, NULL);
        if (t == NB_COARSE) *gp += v / 3.0;
        else *gp += -v / 3.0;
      }
    }
  }
  for (int s = 0; s < 4; s++) { /* :120-159 */
    int slot = side_slot(D, s);
    int t = d->p[p].nbr_type[slot];
    if (t == NB_NONE || t == NB_NORMAL) continue;
    int fa[1];
    face_free_axes(D, 1, s, fa);
    int n = d->ns[fa[0]];
    for (int cc = 0; cc < c->nc; cc++) {
      if (t == NB_COARSE) {
        int off = (d->p[p].nbr_orth[slot] == 1) ? n : 0;
        for (int ti = 0; ti < n; ti++) {
          double v = *sl(d, c->u, c->nc, p, cc, 1, s, 0, &ti);
          *sl(d, c->u, c->nc, p, cc, 1, s, -1, &ti) += 2.0 / 3.0 * v;
          int pt = ((ti + off) % 2 == 0) ? ti + 1 : ti - 1;
          *sl(d, c->u, c->nc, p, cc, 1, s, -1, &pt) += -1.0 / 3.0 * v;
        }
      } else {
        for (int ti = 0; ti < n; ti++) {
          double v = *sl(d, c->u, c->nc, p, cc, 1, s, 0, &ti);
          *sl(d, c->u, c->nc, p, cc, 1, s, -1, &ti) += -1.0 / 3.0 * v;
        }
      }
    }
  }
}

/* ---- BiQuadraticGhostFiller (BiQuadraticGhostFiller.cpp:44-470) ---- */
static void bq_nbr_side(const gctx* c, int lp, int np_, int s, int type, int orth)
{
  const ora_domain* d = c->d;
  int M = 1, opp = face_opposite(2, 1, s);
  int fa[1];
  face_free_axes(2, 1, s, fa);
  int n = d->ns[fa[0]];
#define LS(i) (*sl(d, c->u, c->nc, lp, cc, M, s, 0, &(int){ (i) }))
#define LIN(i) (*sl(d, c->u, c->nc, lp, cc, M, s, 1, &(int){ (i) }))
#define NG(i) (*sl(d, c->u, c->nc, np_, cc, M, opp, -1, &(int){ (i) }))
  for (int cc = 0; cc < c->nc; cc++) {
    if (type == NB_NORMAL) { /* :89-98 */
      for (int t = 0; t < n; t++) NG(t) = LS(t);
    } else if (type == NB_COARSE) {
      if (orth == 0) { /* :113-132 */
        for (int idx = 0; idx < n; idx++) NG(idx / 2) += LS(idx) / 3 + LIN(idx) / 5;
      } else { /* :134-155 */
        for (int idx = 0; idx < n; idx++) NG((idx + n) / 2) += LS(idx) / 3 + LIN(idx) / 5;
      }
    } else {
      if (orth == 0) { /* :170-190 */
        NG(0) += 3 * LS(0) / 4 - 3 * LS(1) / 10 + LS(2) / 12;
        NG(1) += 7 * LS(0) / 20 + 7 * LS(1) / 30 - LS(2) / 20;
        for (int idx = 2; idx < n; idx++) {
          if (idx % 2 == 0) NG(idx) += LS(idx / 2 - 1) / 12 + LS(idx / 2) / 2 - LS(idx / 2 + 1) / 20;
          else NG(idx) += -LS(idx / 2 - 1) / 20 + LS(idx / 2) / 2 + LS(idx / 2 + 1) / 12;
        }
      } else { /* :205-226 */
        for (int idx = 0; idx < n - 2; idx++) {
          int m = (idx + n) / 2;
          if ((idx + n) % 2 == 0) NG(idx) += LS(m - 1) / 12 + LS(m) / 2 - LS(m + 1) / 20;
          else NG(idx) += -LS(m - 1) / 20 + LS(m) / 2 + LS(m + 1) / 12;
        }
        NG(n - 2) += 7 * LS(n - 1) / 20 + 7 * LS(n - 2) / 30 - LS(n - 3) / 20;
        NG(n - 1) += 3 * LS(n - 1) / 4 - 3 * LS(n - 2) / 10 + LS(n - 3) / 12;
      }
    }
  }
#undef LS
#undef LIN
#undef NG
}
static void bq_corner_inner(const ora_domain* d, int f, int* x)
{
  /* getSliceOn(corner, {1,1}) */
  int offs[2] = { 1, 1 };
  slice_coord(d, 0, f, offs, NULL, x);
}
static void bq_nbr_corner(const gctx* c, int lp, int np_, int f, int type)
{
  const ora_domain* d = c->d;
  int opp = face_opposite(2, 0, f);
  for (int cc = 0; cc < c->nc; cc++) {
    double v = *sl(d, c->u, c->nc, lp, cc, 0, f, 0, NULL);
    double* gp = sl(d, c->u, c->nc, np_, cc, 0, opp, -1, NULL);
    if (type == NB_NORMAL) { /* :239-248 */
      *gp = v;
    } else if (type == NB_COARSE) { /* :258-269 */
      int x[3] = { 0, 0, 0 };
      bq_corner_inner(d, f, x);
      double vin = c->u[cell_index(d, lp, cc, c->nc, x)];
      *gp += 2 * v / 3 + 2 * vin / 5;
    } else { /* :303-313 */
      *gp += 8 * v / 15;
    }
  }
}
static void bq_local(const gctx* c, int p)
{
  const ora_domain* d = c->d;
  if (c->ft == FT_CORNERS) { /* :366-385 */
    for (int f = 0; f < 4; f++) {
      int t = d->p[p].nbr_type[f];
      if (t == NB_NONE || t == NB_NORMAL) continue;
      for (int cc = 0; cc < c->nc; cc++) {
        double v = *sl(d, c->u, c->nc, p, cc, 0, f, 0, NULL);
        double* gp = sl(d, c->u, c->nc, p, cc, 0, f, -1, NULL);
        if (t == NB_COARSE) { /* :325-334 */
          int x[3] = { 0, 0, 0 };
          bq_corner_inner(d, f, x);
          double vin = c->u[cell_index(d, p, cc, c->nc, x)];
          *gp += 2 * v / 3 - vin / 5;
        } else { /* :281-289 */
          *gp += -v / 15;
        }
      }
    }
  }
  for (int s = 0; s < 4; s++) { /* :344-364 */
    int slot = side_slot(2, s);
    int t = d->p[p].nbr_type[slot];
    if (t == NB_NONE || t == NB_NORMAL) continue;
    int fa[1];
    face_free_axes(2, 1, s, fa);
    int n = d->ns[fa[0]];
#define S0(i) (*sl(d, c->u, c->nc, p, cc, 1, s, 0, &(int){ (i) }))
#define S1(i) (*sl(d, c->u, c->nc, p, cc, 1, s, 1, &(int){ (i) }))
#define G(i) (*sl(d, c->u, c->nc, p, cc, 1, s, -1, &(int){ (i) }))
    for (int cc = 0; cc < c->nc; cc++) {
      if (t == NB_COARSE) { /* :44-56 */
        for (int idx = 0; idx < n; idx++) G(idx) += 2 * S0(idx) / 3 - S1(idx) / 5;
      } else { /* :58-81 */
        G(0) += -S0(0) / 10 + S0(1) / 15 - S0(2) / 30;
        for (int idx = 1; idx < n - 1; idx++) G(idx) += -S0(idx - 1) / 30 - S0(idx + 1) / 30;
        G(n - 1) += -S0(n - 1) / 10 + S0(n - 2) / 15 - S0(n - 3) / 30;
      }
    }
#undef S0
#undef S1
#undef G
  }
}

/* ---- TriLinearGhostFiller (TriLinearGhostFiller.cpp:33-589) ---- */
/* getOffset (:33-47): orthant bit k -> n of the k-th free axis of the side. */
static void tl_offset(const ora_domain* d, int s, int orth, int* o)
{
  int fa[2];
  face_free_axes(3, 2, s, fa);
  for (int k = 0; k < 2; k++) o[k] = ((orth >> k) & 1) ? d->ns[fa[k]] : 0;
}
static void tl_nbr_side(const gctx* c, int lp, int np_, int s, int type, int orth)
{
  const ora_domain* d = c->d;
  int M = 2, opp = face_opposite(3, 2, s);
  int fa[2];
  face_free_axes(3, 2, s, fa);
  int ni = d->ns[fa[0]], nj = d->ns[fa[1]];
  int o[2] = { 0, 0 };
  if (type != NB_NORMAL) tl_offset(d, s, orth, o);
  for (int cc = 0; cc < c->nc; cc++) {
    for (int j = 0; j < nj; j++) {
      for (int i = 0; i < ni; i++) {
        int fr[2] = { i, j };
        int frc[2] = { (i + o[0]) / 2, (j + o[1]) / 2 };
        if (type == NB_NORMAL) { /* :60-69 */
          *sl(d, c->u, c->nc, np_, cc, M, opp, -1, fr) = *sl(d, c->u, c->nc, lp, cc, M, s, 0, fr);
        } else if (type == NB_COARSE) { /* :81-98 */
          *sl(d, c->u, c->nc, np_, cc, M, opp, -1, frc) += 1.0 / 3.0 * *sl(d, c->u, c->nc, lp, cc, M, s, 0, fr);
        } else { /* :126-143 */
          *sl(d, c->u, c->nc, np_, cc, M, opp, -1, fr) += 4.0 / 6.0 * *sl(d, c->u, c->nc, lp, cc, M, s, 0, frc);
        }
      }
    }
  }
}
static void tl_nbr_edge(const gctx* c, int lp, int np_, int e, int type, int orth)
{
  const ora_domain* d = c->d;
  int M = 1, opp = face_opposite(3, 1, e);
  int fa[1];
  face_free_axes(3, 1, e, fa);
  int n = d->ns[fa[0]];
  int off = (type != NB_NORMAL && orth == 1) ? n : 0;
  for (int cc = 0; cc < c->nc; cc++) {
    for (int i = 0; i < n; i++) {
      int ii = i, ic = (i + off) / 2;
      if (type == NB_NORMAL) { /* :188-197 */
        *sl(d, c->u, c->nc, np_, cc, M, opp, -1, &ii) = *sl(d, c->u, c->nc, lp, cc, M, e, 0, &ii);
      } else if (type == NB_COARSE) { /* :209-226 */
        *sl(d, c->u, c->nc, np_, cc, M, opp, -1, &ic) += 2.0 / 3.0 * *sl(d, c->u, c->nc, lp, cc, M, e, 0, &ii);
      } else { /* :238-255 */
        *sl(d, c->u, c->nc, np_, cc, M, opp, -1, &ii) += 2.0 / 3.0 * *sl(d, c->u, c->nc, lp, cc, M, e, 0, &ic);
      }
    }
  }
}
static void tl_nbr_corner(const gctx* c, int lp, int np_, int f, int type)
{
  const ora_domain* d = c->d;
  int opp = face_opposite(3, 0, f);
  for (int cc = 0; cc < c->nc; cc++) {
    double v = *sl(d, c->u, c->nc, lp, cc, 0, f, 0, NULL);
    double* gp = sl(d, c->u, c->nc, np_, cc, 0, opp, -1, NULL);
    if (type == NB_NORMAL) *gp = v;                   /* :318-327 */
    else if (type == NB_COARSE) *gp += 4.0 * v / 3.0; /* :339-349 */
    else *gp += 2.0 * v / 3.0;                        /* :361-371 */
  }
}
static void tl_local(const gctx* c, int p)
{
  const ora_domain* d = c->d;
  if (c->ft == FT_CORNERS) { /* :383-405, :455-473 */
    for (int f = 0; f < 8; f++) {
      int t = d->p[p].nbr_type[f];
      if (t == NB_NONE || t == NB_NORMAL) continue;
      for (int cc = 0; cc < c->nc; cc++) {
        double v = *sl(d, c->u, c->nc, p, cc, 0, f, 0, NULL);
        double* gp = sl(d, c->u, c->nc, p, cc, 0, f, -1, NULL);
        if (t == NB_COARSE) *gp += v / 3.0;
        else *gp += -v / 3.0;
      }
    }
  }
  if (c->ft == FT_CORNERS || c->ft == FT_EDGES) { /* :265-305, :435-453 */
    for (int e = 0; e < 12; e++) {
      int slot = sum_of_faces(3, 1) + e;
      int t = d->p[p].nbr_type[slot];
      if (t == NB_NONE || t == NB_NORMAL) continue;
      int fa[1];
      face_free_axes(3, 1, e, fa);
      int n = d->ns[fa[0]];
      for (int cc = 0; cc < c->nc; cc++) {
        if (t == NB_COARSE) {
          int off = (d->p[p].nbr_orth[slot] == 1) ? n : 0;
          for (int i = 0; i < n; i++) {
            double v = *sl(d, c->u, c->nc, p, cc, 1, e, 0, &i);
            *sl(d, c->u, c->nc, p, cc, 1, e, -1, &i) += 2.0 / 3.0 * v;
            int pt = ((i + off) % 2 == 0) ? i + 1 : i - 1;
            *sl(d, c->u, c->nc, p, cc, 1, e, -1, &pt) += -1.0 / 3.0 * v;
          }
        } else {
          for (int i = 0; i < n; i++) {
            double v = *sl(d, c->u, c->nc, p, cc, 1, e, 0, &i);
            *sl(d, c->u, c->nc, p, cc, 1, e, -1, &i) += -1.0 / 3.0 * v;
          }
        }
      }
    }
  }
  for (int s = 0; s < 6; s++) { /* :100-175, :414-433 */
    int slot = side_slot(3, s);
    int t = d->p[p].nbr_type[slot];
    if (t == NB_NONE || t == NB_NORMAL) continue;
    int fa[2];
    face_free_axes(3, 2, s, fa);
    int ni = d->ns[fa[0]], nj = d->ns[fa[1]];
    int o[2] = { 0, 0 };
    if (t == NB_COARSE) tl_offset(d, s, d->p[p].nbr_orth[slot], o);
    for (int cc = 0; cc < c->nc; cc++) {
      for (int j = 0; j < nj; j++) {
        int oj = ((j + o[1]) % 2 == 0) ? (j + 1) : (j - 1);
        for (int i = 0; i < ni; i++) {
          int fr[2] = { i, j };
          double* gp = sl(d, c->u, c->nc, p, cc, 2, s, -1, fr);
          if (t == NB_FINE) {
            *gp -= 1.0 / 3.0 * *sl(d, c->u, c->nc, p, cc, 2, s, 0, fr);
          } else {
            int oi = ((i + o[0]) % 2 == 0) ? (i + 1) : (i - 1);
            int f1[2] = { oi, j }, f2[2] = { i, oj }, f3[2] = { oi, oj };
            *gp += 5.0 / 6.0 * *sl(d, c->u, c->nc, p, cc, 2, s, 0, fr);
            *gp -= 1.0 / 6.0 * *sl(d, c->u, c->nc, p, cc, 2, s, 0, f1);
            *gp -= 1.0 / 6.0 * *sl(d, c->u, c->nc, p, cc, 2, s, 0, f2);
            *gp -= 1.0 / 6.0 * *sl(d, c->u, c->nc, p, cc, 2, s, 0, f3);
          }
        }
      }
    }
  }
}

/* MPIGhostFiller::fillGhost (MPIGhostFiller.h:1055-1092), one rank. */
int ora_fill_ghost(const ora_domain* d, int kind, int fill_type, double* u, int ncomp)
{
  int D = d->dim;
  if (kind == K_TRILINEAR && D != 3) return -1;
  if (kind != K_TRILINEAR && D != 2) return -1;
  if (D == 2 && fill_type == FT_EDGES) return -2; /* BiLinear/BiQuadratic throw on Edges */
  if (kind == K_TRILINEAR)
    for (int a = 0; a < 3; a++)
      if (d->ns[a] % 2) return -3; /* TriLinearGhostFiller.cpp:576-584 */
  gctx c = { d, u, ncomp, kind, fill_type };
  zero_ghosts(&c);
  for (int p = 0; p < d->np; p++) {
    if (kind == K_BILINEAR) bl_local(&c, p);
    else if (kind == K_BIQUADRATIC) bq_local(&c, p);
    else tl_local(&c, p);
  }
  int ncall = build_calls(d, fill_type, NULL, 0);
  ora_call* calls = (ora_call*)malloc(sizeof(ora_call) * (ncall ? ncall : 1));
  build_calls(d, fill_type, calls, ncall);
  for (int k = 0; k < ncall; k++) {
    ora_call* q = &calls[k];
    if (q->nbr < 0) continue; /* remote neighbour: not on this rank */
    if (q->M == D - 1) {
      if (kind == K_BILINEAR) bl_nbr_side(&c, q->local, q->nbr, q->face, q->type, q->orth);
      else if (kind == K_BIQUADRATIC) bq_nbr_side(&c, q->local, q->nbr, q->face, q->type, q->orth);
      else tl_nbr_side(&c, q->local, q->nbr, q->face, q->type, q->orth);
    } else if (q->M > 0) {
      tl_nbr_edge(&c, q->local, q->nbr, q->face, q->type, q->orth);
    } else {
      if (kind == K_BILINEAR) bl_nbr_corner(&c, q->local, q->nbr, q->face, q->type);
      else if (kind == K_BIQUADRATIC) bq_nbr_corner(&c, q->local, q->nbr, q->face, q->type);
      else tl_nbr_corner(&c, q->local, q->nbr, q->face, q->type);
    }
  }
  free(calls);
  return 0;
}

/* -------------------------------------------------------- star operators -- */
typedef struct {
  const ora_domain* d;
  int kind, ft;
  int op_kind; /* 0 Poisson, 1 VarPoisson */
  int neumann;
  const double* coeffs;
} opctx;

/* enforceBoundaryConditions (Poisson/StarPatchOperator.h:115-144,
 * VarPoisson/StarPatchOperator.h:132-150): physical ghosts. */
static void enforce_bc(const opctx* o, int pi, int ai, double* u, int nc, int internal)
{
  const ora_domain* d = o->d;
  int D = d->dim, M = D - 1;
  for (int pass = internal ? 0 : 1; pass < 2; pass++) {
    /* pass 0: enforceInternalBoundaryConditions (:145-164), sides with a neighbour: ghost = -interior;
     * pass 1: enforceBoundaryConditions (:115-144), physical sides: ghost = -interior (Dirichlet)
     * or +interior (Neumann) */
    for (int s = 0; s < 2 * D; s++) {
      int hn = has_nbr(d, pi, side_slot(D, s));
      if (pass == 0 ? !hn : hn) continue;
      int fa[2];
      int nfr = face_free_axes(D, M, s, fa);
      int cnt = 1;
      for (int k = 0; k < nfr; k++) cnt *= d->ns[fa[k]];
      int neu = (pass == 1) && (o->op_kind == 0) && o->neumann;
      for (int cc = 0; cc < nc; cc++)
        for (int li = 0; li < cnt; li++) {
          int fr[2] = { li % d->ns[fa[0]], nfr > 1 ? li / d->ns[fa[0]] : 0 };
          double v = *sl(d, u, nc, ai, cc, M, s, 0, fr);
          *sl(d, u, nc, ai, cc, M, s, -1, fr) = neu ? v : -v;
        }
    }
  }
}

/* applySinglePatch (Poisson/StarPatchOperator.h:75-101, VarPoisson/StarPatchOperator.h:61-109). */
static void apply_patch(const opctx* o, int p, int ai, double* u, double* f, int nc, int internal)
{
  const ora_domain* d = o->d;
  int D = d->dim;
  enforce_bc(o, p, ai, u, nc, internal);
  double h2[3];
  for (int a = 0; a < D; a++) h2[a] = d->p[p].spacings[a] * d->p[p].spacings[a];
  int n0 = d->ns[0], n1 = d->ns[1], n2 = D == 3 ? d->ns[2] : 1;
  for (int a = 0; a < D; a++) {
    long st = d->stride[a];
    for (int cc = 0; cc < nc; cc++)
      for (int z = 0; z < n2; z++)
        for (int y = 0; y < n1; y++)
          for (int x = 0; x < n0; x++) {
            int xc[3] = { x, y, z };
            long i = cell_index(d, ai, cc, nc, xc);
            double lower = u[i - st], mid = u[i], upper = u[i + st];
            double add = (a == 0) ? 0 : 1;
            if (o->op_kind == 0) {
              f[i] = add * f[i] + (upper - 2 * mid + lower) / h2[a];
            } else {
              long ci = cell_index(d, p, 0, 1, xc);
              double cl = o->coeffs[ci - st], cm = o->coeffs[ci], cu = o->coeffs[ci + st];
              f[i] = add * f[i] + ((cu + cm) * (upper - mid) - (cl + cm) * (mid - lower)) / (2 * h2[a]);
            }
          }
  }
}

/* modifyRHSForInternalBoundaryConditions (Poisson :165-180, VarPoisson :168-186). */
static void modify_rhs(const opctx* o, int p, const double* u, double* f, int nc, int up, int fp)
{
  const ora_domain* d = o->d;
  int D = d->dim, M = D - 1;
  for (int s = 0; s < 2 * D; s++) {
    if (!has_nbr(d, p, side_slot(D, s))) continue;
    double h2 = pow(d->p[p].spacings[s >> 1], 2);
    int fa[2];
    int nfr = face_free_axes(D, M, s, fa);
    int cnt = 1;
    for (int k = 0; k < nfr; k++) cnt *= d->ns[fa[k]];
    for (int cc = 0; cc < nc; cc++)
      for (int li = 0; li < cnt; li++) {
        int fr[2] = { li % d->ns[fa[0]], nfr > 1 ? li / d->ns[fa[0]] : 0 };
        double ug = *sl(d, (double*)u, nc, up, cc, M, s, -1, fr);
        double ub = *sl(d, (double*)u, nc, up, cc, M, s, 0, fr);
        double* fb = sl(d, f, nc, fp, cc, M, s, 0, fr);
        if (o->op_kind == 0) {
          *fb -= (ug + ub) / h2;
        } else {
          double cg = *sl(d, (double*)o->coeffs, 1, p, 0, M, s, -1, fr);
          double cb = *sl(d, (double*)o->coeffs, 1, p, 0, M, s, 0, fr);
          *fb -= (ug + ub) * (cb + cg) / (2 * h2);
        }
      }
  }
}

/* PatchOperator::apply (PatchOperator.h:126-143). */
static int op_apply(const opctx* o, double* u, double* f, int nc)
{
  const ora_domain* d = o->d;
  long len = ora_domain_vec_len(d, nc);
  memset(f, 0, sizeof(double) * len);
  int rc = ora_fill_ghost(d, o->kind, o->ft, u, nc);
  if (rc) return rc;
  /* patches are independent here (each writes its own f and its own physical ghosts) */
#pragma omp parallel for schedule(static)
  for (int p = 0; p < d->np; p++) apply_patch(o, p, p, u, f, nc, 0);
  return 0;
}

int ora_poisson_apply(const ora_domain* d, int kind, int fill_type, int neumann, double* u, double* f,
                      int ncomp)
{
  opctx o = { d, kind, fill_type, 0, neumann, NULL };
  return op_apply(&o, u, f, ncomp);
}

int ora_varpoisson_apply(const ora_domain* d, int kind, int fill_type, const double* coeffs, double* u,
                         double* f, int ncomp)
{
  opctx o = { d, kind, fill_type, 1, 0, coeffs };
  return op_apply(&o, u, f, ncomp);
}

/* Poisson::StarPatchOperator::addDrichletBCToRHS (Poisson/StarPatchOperator.h:187-204),
 * VarPoisson::StarPatchOperator::addDrichletBCToRHS (VarPoisson/StarPatchOperator.h:194-220)
 * and Poisson::StarPatchOperator::addNeumannBCToRHS (:206-242), with the functions sampled
 * at the face points of every side (g: Dirichlet values or, neumann != 0, the derivative
 * along the side's axis; c: the VarPoisson coefficient there, or NULL).  Sample layout:
 * per patch, per side in order, the side's face cells with the first free axis fastest. */
int ora_add_bc(const ora_domain* d, double* f, const double* g, const double* c, int neumann)
{
  int D = d->dim, M = D - 1;
  long off = 0;
  for (int p = 0; p < d->np; p++) {
    for (int s = 0; s < 2 * D; s++) {
      int fa[2];
      int nfr = face_free_axes(D, M, s, fa);
      int cnt = 1;
      for (int k = 0; k < nfr; k++) cnt *= d->ns[fa[k]];
      if (!has_nbr(d, p, side_slot(D, s))) {
        double h = d->p[p].spacings[s >> 1];
        double h2 = pow(h, 2);
        for (int li = 0; li < cnt; li++) {
          int fr[2] = { li % d->ns[fa[0]], nfr > 1 ? li / d->ns[fa[0]] : 0 };
          double* fb = sl(d, f, 1, p, 0, M, s, 0, fr);
          double gv = g[off + li];
          if (neumann) {
            if ((s & 1) == 0) *fb += gv / h;
            else *fb -= gv / h;
          } else if (c) {
            *fb -= 2 * gv * c[off + li] / h2;
          } else {
            *fb += -2.0 * gv / h2;
          }
        }
      }
      off += cnt;
    }
  }
  return 0;
}

/* ---------------------------------------------------------- vector ops -- */
typedef struct {
  const ora_domain* d;
  int np; /* number of patches in the space (1 for single-patch solves) */
  int nc;
} vspace;

static long vs_len(const vspace* v) { return (long)v->np * v->nc * v->d->cells_ghost; }
#define FOR_INTERIOR(v, i)                                                                          \
  for (int p_ = 0; p_ < (v)->np; p_++)                                                              \
    for (int c_ = 0; c_ < (v)->nc; c_++)                                                            \
      for (int z_ = 0; z_ < ((v)->d->dim == 3 ? (v)->d->ns[2] : 1); z_++)                           \
        for (int y_ = 0; y_ < (v)->d->ns[1]; y_++)                                                  \
          for (int x_ = 0; x_ < (v)->d->ns[0]; x_++)                                                \
            for (long i = cell_index((v)->d, p_, c_, (v)->nc, (int[3]){ x_, y_, z_ }), once_ = 1; once_; \
                 once_ = 0)

static double vs_dot(const vspace* v, const double* a, const double* b)
{
  double r = 0;
  FOR_INTERIOR(v, i) r += a[i] * b[i];
  return r;
}
static double vs_norm(const vspace* v, const double* a) { return sqrt(vs_dot(v, a, a)); }

double ora_vec_dot(const ora_domain* d, const double* a, const double* b, int ncomp)
{
  vspace v = { d, d->np, ncomp };
  return vs_dot(&v, a, b);
}
double ora_vec_two_norm(const ora_domain* d, const double* a, int ncomp)
{
  vspace v = { d, d->np, ncomp };
  return vs_norm(&v, a);
}
/* DomainTools::Integrate (DomainTools.h:513-536): per patch and component, the sum of the
 * interior values, multiplied by each spacing in axis order, added over patches. */
double ora_vec_integrate(const ora_domain* d, const double* u, int ncomp)
{
  double sum = 0;
  for (int p = 0; p < d->np; p++) {
    for (int c = 0; c < ncomp; c++) {
      double ps = 0;
      for (int z = 0; z < (d->dim == 3 ? d->ns[2] : 1); z++)
        for (int y = 0; y < d->ns[1]; y++)
          for (int x = 0; x < d->ns[0]; x++) ps += u[cell_index(d, p, c, ncomp, (int[3]){ x, y, z })];
      for (int a = 0; a < d->dim; a++) ps *= d->p[p].spacings[a];
      sum += ps;
    }
  }
  return sum;
}
/* Vector::infNorm (Vector.h:568-579) is max(0, max v), not max|v|. */
double ora_vec_inf_norm(const ora_domain* d, const double* a, int ncomp)
{
  vspace v = { d, d->np, ncomp };
  double m = 0;
  FOR_INTERIOR(&v, i) m = fmax(a[i], m);
  return m;
}

/* ------------------------------------------------------------ BiCGStab --
 * Iterative::BiCGStab::solve (Iterative/BiCGStab.h:111-203). */
typedef int (*apply_fn)(void* ctx, double* in, double* out);

static int bicgstab(const vspace* V, apply_fn A, void* actx, double* x, const double* b, apply_fn Mr,
                    void* mctx, int max_it, double tol, double* hist, int hist_cap, int* breakdown, int exit_mp)
{
  long len = vs_len(V);
  double* resid = (double*)calloc(len, sizeof(double));
  double* ms = Mr ? (double*)calloc(len, sizeof(double)) : NULL;
  double* mp = Mr ? (double*)calloc(len, sizeof(double)) : NULL;
  double* rhat = (double*)malloc(sizeof(double) * len);
  double* p = (double*)malloc(sizeof(double) * len);
  double* ap = (double*)calloc(len, sizeof(double));
  double* as = (double*)calloc(len, sizeof(double));
  double* s = (double*)calloc(len, sizeof(double));
  *breakdown = 0;
  A(actx, x, resid);
  FOR_INTERIOR(V, i) resid[i] = resid[i] * -1 + b[i]; /* scaleThenAdd(-1, b) */
  double r0_norm = vs_norm(V, b);
  memcpy(rhat, resid, sizeof(double) * len);
  memcpy(p, resid, sizeof(double) * len);
  double rho = vs_dot(V, rhat, resid);
  int num_its = 0;
  double residual = vs_norm(V, resid) / r0_norm;
  if (hist && hist_cap > 0) hist[0] = residual;
  if (r0_norm == 0) goto done;
  while (residual > tol && num_its < max_it) {
    if (rho == 0) {
      *breakdown = 1;
      goto done;
    }
    if (Mr) {
      Mr(mctx, p, mp);
      A(actx, mp, ap);
    } else {
      A(actx, p, ap);
    }
    double alpha = rho / vs_dot(V, rhat, ap);
    FOR_INTERIOR(V, i) s[i] = resid[i];
    FOR_INTERIOR(V, i) s[i] += ap[i] * -alpha;
    if (vs_norm(V, s) / r0_norm <= tol) {
      /* the reference adds alpha p, not alpha M p (BiCGStab.h:162-167); exit_mp selects M p */
      const double* dirp = (exit_mp && Mr) ? mp : p;
      FOR_INTERIOR(V, i) x[i] += dirp[i] * alpha;
      break;
    }
    if (Mr) {
      Mr(mctx, s, ms);
      A(actx, ms, as);
    } else {
      A(actx, s, as);
    }
    double omega = vs_dot(V, as, s) / vs_dot(V, as, as);
    if (Mr) {
      FOR_INTERIOR(V, i) x[i] += mp[i] * alpha + ms[i] * omega;
    } else {
      FOR_INTERIOR(V, i) x[i] += p[i] * alpha + s[i] * omega;
    }
    FOR_INTERIOR(V, i) resid[i] += ap[i] * -alpha + as[i] * -omega;
    double rho_new = vs_dot(V, resid, rhat);
    double beta = rho_new * alpha / (rho * omega);
    FOR_INTERIOR(V, i) p[i] += ap[i] * -omega;
    FOR_INTERIOR(V, i) p[i] = p[i] * beta + resid[i];
    num_its++;
    rho = rho_new;
    residual = vs_norm(V, resid) / r0_norm;
    if (hist && num_its < hist_cap) hist[num_its] = residual;
  }
done:
  free(resid);
  free(ms);
  free(mp);
  free(rhat);
  free(p);
  free(ap);
  free(as);
  free(s);
  return num_its;
}

/* ------------------------------------------------------------ smoothers -- */
typedef struct {
  const opctx* o;
  int p;  /* patch whose PatchInfo drives the BCs */
  int nc;
} spctx;

/* SinglePatchOp::apply (Iterative/PatchSolver.h:76-81): a 1-patch domain view. */
static int single_patch_apply(void* ctx, double* in, double* out)
{
  spctx* s = (spctx*)ctx;
  const ora_domain* d = s->o->d;
  /* in/out are single padded patches (array index 0); BCs come from PatchInfo s->p */
  (void)d;
  apply_patch(s->o, s->p, 0, in, out, s->nc, 1);
  return 0;
}

/* DFTPatchSolver (Poisson/DFTPatchSolver.h:131-450). */
void ora_dft_matrix(int type, int n, double* m)
{
  switch (type) {
    case 0: /* DCT_II */
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) m[i * n + j] = cos(M_PI / n * (i * (j + 0.5)));
      break;
    case 1: /* DCT_III */
      for (int i = 0; i < n; i++) m[i * n] = 0.5;
      for (int j = 1; j < n; j++)
        for (int i = 0; i < n; i++) m[i * n + j] = cos(M_PI / n * ((i + 0.5) * j));
      break;
    case 2: /* DCT_IV */
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) m[i * n + j] = cos(M_PI / n * ((i + 0.5) * (j + 0.5)));
      break;
    case 3: /* DST_II */
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) m[i * n + j] = sin(M_PI / n * ((i + 1) * (j + 0.5)));
      break;
    case 4: /* DST_III */
      for (int i = 0; i < n; i += 2) m[i * n + n - 1] = 0.5;
      for (int i = 1; i < n; i += 2) m[i * n + n - 1] = -0.5;
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n - 1; j++) m[i * n + j] = sin(M_PI / n * ((i + 0.5) * (j + 1)));
      break;
    default: /* DST_IV */
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) m[i * n + j] = sin(M_PI / n * ((i + 0.5) * (j + 0.5)));
  }
}

static int neumann_on_side(const ora_domain* d, int p, int s, unsigned mask)
{
  return !has_nbr(d, p, side_slot(d->dim, s)) && ((mask >> s) & 1);
}

/* executePlan (:205-265): y = M x along each axis in turn, on an n^D interior block. */
static void transform_axes(const ora_domain* d, double* const* mats, double* a, double* tmp)
{
  int D = d->dim;
  int n[3] = { d->ns[0], d->ns[1], D == 3 ? d->ns[2] : 1 };
  long st[3] = { 1, n[0], (long)n[0] * n[1] };
  long tot = (long)n[0] * n[1] * n[2];
  for (int ax = 0; ax < D; ax++) {
    const double* m = mats[ax];
    int na = n[ax];
    for (long base = 0; base < tot; base++) {
      int coord_ax = (int)((base / st[ax]) % na);
      if (coord_ax != 0) continue;
      for (int i = 0; i < na; i++) {
        double acc = 0;
        for (int j = 0; j < na; j++) acc += m[i * na + j] * a[base + j * st[ax]];
        tmp[base + i * st[ax]] = acc;
      }
    }
    memcpy(a, tmp, sizeof(double) * tot);
  }
}

static int dft_smooth(const opctx* o, unsigned mask, const double* f, double* u)
{
  const ora_domain* d = o->d;
  int D = d->dim;
  int rc = ora_fill_ghost(d, o->kind, o->ft, u, 1); /* PatchSolver::smooth (PatchSolver.h:155) */
  if (rc) return rc;
  int n[3] = { d->ns[0], d->ns[1], D == 3 ? d->ns[2] : 1 };
  long tot = (long)n[0] * n[1] * n[2];
  double* mats[6][3];
  for (int t = 0; t < 6; t++)
    for (int a = 0; a < D; a++) {
      mats[t][a] = (double*)malloc(sizeof(double) * n[a] * n[a]);
      ora_dft_matrix(t, n[a], mats[t][a]);
    }
  int all_neumann = 1;
  for (int s = 0; s < 2 * D; s++)
    if (!((mask >> s) & 1)) all_neumann = 0;
  /* the patch solves are independent (u's ghosts were filled above and only interior
   * cells are written), so they run in parallel with per-thread work buffers */
#pragma omp parallel
  {
  double* blk = (double*)malloc(sizeof(double) * tot);
  double* tmp = (double*)malloc(sizeof(double) * tot);
  double* fcopy = (double*)malloc(sizeof(double) * d->cells_ghost);
#pragma omp for schedule(static)
  for (int p = 0; p < d->np; p++) {
    /* f_copy = f ; modifyRHS (:424-430) */
    memset(fcopy, 0, sizeof(double) * d->cells_ghost);
    FOR_INTERIOR(&((vspace){ d, 1, 1 }), i) fcopy[i] = f[(long)p * d->cells_ghost + i];
    modify_rhs(o, p, u, fcopy, 1, p, 0);
    double* fw[3];
    double* iv[3];
    double eig_scale = 1;
    for (int a = 0; a < D; a++) {
      int lo = neumann_on_side(d, p, 2 * a, mask), hi = neumann_on_side(d, p, 2 * a + 1, mask);
      int tf = (lo && hi) ? 0 : lo ? 2 : hi ? 5 : 3; /* getTransformsForPatch (:272-291) */
      int ti = (lo && hi) ? 1 : lo ? 2 : hi ? 5 : 4; /* getInverseTransformsForPatch (:294-313) */
      fw[a] = mats[tf][a];
      iv[a] = mats[ti][a];
      eig_scale *= 2.0 / n[a];
    }
    long k = 0;
    for (int z = 0; z < n[2]; z++)
      for (int y = 0; y < n[1]; y++)
        for (int x = 0; x < n[0]; x++) {
          int xc[3] = { x, y, z };
          blk[k++] = fcopy[cell_index(d, 0, 0, 1, xc)];
        }
    transform_axes(d, fw, blk, tmp);
    /* getEigenValues (:320-361), then tmp /= eig (:434-436) */
    k = 0;
    for (int z = 0; z < n[2]; z++)
      for (int y = 0; y < n[1]; y++)
        for (int x = 0; x < n[0]; x++) {
          int xc[3] = { x, y, z };
          double eig = 0;
          for (int a = 0; a < D; a++) {
            int na = n[a];
            double h = d->p[p].spacings[a];
            int lo = neumann_on_side(d, p, 2 * a, mask), hi = neumann_on_side(d, p, 2 * a + 1, mask);
            int xi = xc[a];
            double val;
            if (lo && hi) val = 4 / (h * h) * pow(sin(xi * M_PI / (2 * na)), 2);
            else if (lo || hi) val = 4 / (h * h) * pow(sin((xi + 0.5) * M_PI / (2 * na)), 2);
            else val = 4 / (h * h) * pow(sin((xi + 1) * M_PI / (2 * na)), 2);
            eig -= val;
          }
          blk[k] /= eig;
          k++;
        }
    if (all_neumann && !patch_has_any_nbr(d, p)) blk[0] = 0; /* :438-440 */
    transform_axes(d, iv, blk, tmp);
    k = 0;
    for (int z = 0; z < n[2]; z++)
      for (int y = 0; y < n[1]; y++)
        for (int x = 0; x < n[0]; x++) {
          int xc[3] = { x, y, z };
          u[cell_index(d, p, 0, 1, xc)] = blk[k++] * eig_scale; /* :442-449 */
        }
  }
  free(blk);
  free(tmp);
  free(fcopy);
  }
  for (int t = 0; t < 6; t++)
    for (int a = 0; a < D; a++) free(mats[t][a]);
  return 0;
}

int ora_dft_smooth(const ora_domain* d, int kind, int fill_type, int op_kind, const double* coeffs,
                   unsigned neumann_mask, const double* f, double* u)
{
  opctx o = { d, kind, fill_type, op_kind, 0, coeffs };
  return dft_smooth(&o, neumann_mask, f, u);
}

/* Iterative::PatchSolver::solveSinglePatch (Iterative/PatchSolver.h:134-175) within
 * PatchSolver::smooth (PatchSolver.h:142-170). */
static int bicgstab_smooth(const opctx* o, int max_it, double tol, int cont, const double* f, double* u,
                           int nc, int* iters_out)
{
  const ora_domain* d = o->d;
  int rc = ora_fill_ghost(d, o->kind, o->ft, u, nc);
  if (rc) return rc;
  long plen = (long)nc * d->cells_ghost;
  vspace V1 = { d, 1, nc };
  int failed = 0;
  /* independent patch solves (each reads the filled ghosts and writes its own interior) */
#pragma omp parallel
  {
    double* fcopy = (double*)malloc(sizeof(double) * plen);
#pragma omp for schedule(dynamic, 4)
    for (int p = 0; p < d->np; p++) {
      memset(fcopy, 0, sizeof(double) * plen);
      FOR_INTERIOR(&V1, i) fcopy[i] = f[p * plen + i];
      modify_rhs(o, p, u, fcopy, nc, p, 0);
      spctx sc = { o, p, nc };
      int bd = 0;
      int its = bicgstab(&V1, single_patch_apply, &sc, u + p * plen, fcopy, NULL, NULL, max_it, tol, NULL, 0, &bd, 0);
      if (bd && !cont) {
#pragma omp atomic write
        failed = 1;
      }
      if (iters_out) iters_out[p] = its;
    }
    free(fcopy);
  }
  return failed ? -10 : 0;
}

int ora_bicgstab_smooth(const ora_domain* d, int kind, int fill_type, int op_kind, const double* coeffs,
                        int neumann, int max_it, double tol, int cont, const double* f, double* u, int ncomp,
                        int* iters_out)
{
  opctx o = { d, kind, fill_type, op_kind, neumann, coeffs };
  return bicgstab_smooth(&o, max_it, tol, cont, f, u, ncomp, iters_out);
}

/* ------------------------------------------------- restrict / interpolate -- */
/* LinearRestrictor::extrapolateBoundaries (GMG/LinearRestrictor.h:49-92) for fine patch p
 * (orthant orth on its parent cp): on every exterior side s of the orthant, each ghost
 * cell of the fine side adds (3 ghost - interior) / 2^D to the coarse ghost cell over it
 * (the commented-out hasNbr test: whether or not s has a neighbour), the side's slices
 * walked first free axis fastest; then, when the patch has no neighbour across the corner
 * of its orthant, the coarse corner ghost gains 3 ghost - interior of the fine corner. */
static void extrapolate_boundaries(const ora_domain* fine, const ora_domain* coarse, int p, int cp, int orth,
                                   const double* fv, double* cv, int nc)
{
  const int D = fine->dim;
  int n[3] = { fine->ns[0], fine->ns[1], D == 3 ? fine->ns[2] : 1 };
  int st[3] = { 0, 0, 0 };
  for (int a = 0; a < D; a++) st[a] = ((orth >> a) & 1) ? n[a] : 0;
  for (int a = 0; a < D; a++) { /* Orthant::getExteriorSides: one side per axis, in axis order */
    const int up = (orth >> a) & 1;
    int fa[2], na = 0; /* the free axes, increasing */
    for (int b = 0; b < D; b++)
      if (b != a) fa[na++] = b;
    const int n1 = na > 0 ? n[fa[0]] : 1, n2 = na > 1 ? n[fa[1]] : 1;
    for (int cc = 0; cc < nc; cc++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          int xg[3] = { 0, 0, 0 }, xi[3] = { 0, 0, 0 }, xc[3] = { 0, 0, 0 };
          xg[a] = up ? n[a] : -1;
          xi[a] = up ? n[a] - 1 : 0;
          xc[a] = xg[a];
          const int t[2] = { i, j };
          for (int k = 0; k < na; k++) {
            xg[fa[k]] = xi[fa[k]] = t[k];
            xc[fa[k]] = (t[k] + st[fa[k]]) / 2;
          }
          const double g = fv[cell_index(fine, p, cc, nc, xg)], in = fv[cell_index(fine, p, cc, nc, xi)];
          cv[cell_index(coarse, cp, cc, nc, xc)] += (3 * g - in) / (1 << D);
        }
  }
  /* the corner of the orthant: Corner<D>(orth) has the orthant's index (slot orth) */
  if (fine->p[p].nbr_type[orth] < 0) {
    int xg[3] = { 0, 0, 0 }, xi[3] = { 0, 0, 0 };
    for (int a = 0; a < D; a++) {
      const int up = (orth >> a) & 1;
      xg[a] = up ? n[a] : -1;
      xi[a] = up ? n[a] - 1 : 0;
    }
    for (int cc = 0; cc < nc; cc++)
      cv[cell_index(coarse, cp, cc, nc, xg)] += 3 * fv[cell_index(fine, p, cc, nc, xg)] -
                                                fv[cell_index(fine, p, cc, nc, xi)];
  }
}

int ora_restrict(const ora_domain* fine, const ora_domain* coarse, const double* fv, double* cv, int nc)
{
  return ora_restrict_ex(fine, coarse, fv, cv, nc, 0);
}

int ora_restrict_ex(const ora_domain* fine, const ora_domain* coarse, const double* fv, double* cv, int nc,
                    int extrapolate)
{
  int D = fine->dim;
  memset(cv, 0, sizeof(double) * ora_domain_vec_len(coarse, nc)); /* coarse.setWithGhost(0) */
  int n[3] = { fine->ns[0], fine->ns[1], D == 3 ? fine->ns[2] : 1 };
  for (int p = 0; p < fine->np; p++)
    if (fine->parent_local[p] < 0) return -1;
  /* each coarse cell receives the values of exactly one fine patch, so the fine patches
   * are independent */
#pragma omp parallel for schedule(static)
  for (int p = 0; p < fine->np; p++) {
    int cp = fine->parent_local[p];
    int orth = fine->p[p].orth_on_parent;
    int st[3] = { 0, 0, 0 };
    if (orth >= 0)
      for (int a = 0; a < D; a++) st[a] = ((orth >> a) & 1) ? n[a] : 0;
    for (int cc = 0; cc < nc; cc++)
      for (int z = 0; z < n[2]; z++)
        for (int y = 0; y < n[1]; y++)
          for (int x = 0; x < n[0]; x++) {
            int xf[3] = { x, y, z };
            double v = fv[cell_index(fine, p, cc, nc, xf)];
            if (orth >= 0) { /* restrictToCoarserParent (:101-128) */
              int xc[3] = { (x + st[0]) / 2, (y + st[1]) / 2, (z + st[2]) / 2 };
              cv[cell_index(coarse, cp, cc, nc, xc)] += v / (1 << D);
            } else if (!extrapolate) { /* copyToParent (:138-155) */
              cv[cell_index(coarse, cp, cc, nc, xf)] += v;
            }
          }
    if (extrapolate && orth >= 0) extrapolate_boundaries(fine, coarse, p, cp, orth, fv, cv, nc);
    if (extrapolate && orth < 0) { /* copyToParent with extrapolation: every cell, ghosts included */
      long len = (long)nc * fine->cells_ghost;
      for (long k = 0; k < len; k++) cv[(long)cp * nc * coarse->cells_ghost + k] += fv[(long)p * nc * fine->cells_ghost + k];
    }
  }
  return 0;
}

int ora_interpolate(const ora_domain* coarse, const ora_domain* fine, const double* cv, double* fv, int nc)
{
  int D = fine->dim;
  int n[3] = { fine->ns[0], fine->ns[1], D == 3 ? fine->ns[2] : 1 };
  for (int p = 0; p < fine->np; p++)
    if (fine->parent_local[p] < 0) return -1;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < fine->np; p++) {
    int cp = fine->parent_local[p];
    int orth = fine->p[p].orth_on_parent;
    int st[3] = { 0, 0, 0 };
    if (orth >= 0)
      for (int a = 0; a < D; a++) st[a] = ((orth >> a) & 1) ? n[a] : 0;
    for (int cc = 0; cc < nc; cc++)
      for (int z = 0; z < n[2]; z++)
        for (int y = 0; y < n[1]; y++)
          for (int x = 0; x < n[0]; x++) {
            int xf[3] = { x, y, z };
            int xc[3] = { (x + st[0]) / 2, (y + st[1]) / 2, (z + st[2]) / 2 };
            fv[cell_index(fine, p, cc, nc, xf)] += cv[cell_index(coarse, cp, cc, nc, orth >= 0 ? xc : xf)];
          }
  }
  (void)coarse;
  return 0;
}

/* --------------------------------------------------------------- cycles -- */
typedef struct {
  const ora_level* lv;
  int nl, pre, post, coarse;
  int err;
  int type; /* 0 V (VCycle.h), 1 W (WCycle.h), 2 F (FMGCycle.h) */
  int mid;  /* CycleOpts::mid_sweeps (W and F) */
} vcctx;

static int level_smooth(const ora_level* L, const double* f, double* u)
{
  opctx o = { L->dom, L->filler_kind, L->fill_type, L->op_kind, L->neumann, L->coeffs };
  if (L->smoother_kind == 0) return dft_smooth(&o, L->neumann_mask, f, u);
  return bicgstab_smooth(&o, L->inner_max_it, L->inner_tol, L->continue_on_breakdown, f, u, 1, NULL);
}
static int level_apply(const ora_level* L, double* u, double* f)
{
  opctx o = { L->dom, L->filler_kind, L->fill_type, L->op_kind, L->neumann, L->coeffs };
  return op_apply(&o, u, f, 1);
}

/* Cycle::restrict (GMG/Cycle.h:57-65): cf = restriction of f - A u. */
static int cycle_restrict(vcctx* c, int li, const double* f, double* u, double* cf)
{
  const ora_level* L = &c->lv[li];
  long len = ora_domain_vec_len(L->dom, 1);
  double* r = (double*)calloc(len, sizeof(double));
  int rc = level_apply(L, u, r);
  if (!rc) {
    vspace V = { L->dom, L->dom->np, 1 };
    FOR_INTERIOR(&V, i) r[i] = r[i] * -1 + f[i];
    rc = ora_restrict(L->dom, c->lv[li + 1].dom, r, cf, 1);
  }
  free(r);
  return rc;
}

/* VCycle::visit (GMG/VCycle.h:46-70), WCycle::visit (GMG/WCycle.h:50-84) and
 * FMGCycle::visit / v_visit (GMG/FMGCycle.h:43-104).  vonly: FMGCycle::v_visit (one
 * coarse visit, its recursion again FMGCycle::visit).  W and F visit the coarser level a
 * second time after the mid sweeps, re-restricting the residual into the coarse f but
 * keeping the coarse u of the first visit as the initial guess, as the reference does. */
static int visit(vcctx* c, int li, const double* f, double* u, int vonly)
{
  const ora_level* L = &c->lv[li];
  int rc;
  if (li == c->nl - 1) {
    for (int i = 0; i < c->coarse; i++)
      if ((rc = level_smooth(L, f, u))) return rc;
    return 0;
  }
  for (int i = 0; i < c->pre; i++)
    if ((rc = level_smooth(L, f, u))) return rc;
  const ora_level* C = &c->lv[li + 1];
  long clen = ora_domain_vec_len(C->dom, 1);
  double* cf = (double*)calloc(clen, sizeof(double));
  double* cu = (double*)calloc(clen, sizeof(double));
  rc = cycle_restrict(c, li, f, u, cf);
  if (!rc) rc = visit(c, li + 1, cf, cu, 0);
  if (!rc) rc = ora_interpolate(C->dom, L->dom, cu, u, 1);
  if (!rc && c->type != 0 && !vonly) {
    for (int i = 0; i < c->mid && !rc; i++) rc = level_smooth(L, f, u);
    if (!rc) {
      memset(cf, 0, sizeof(double) * clen);
      rc = cycle_restrict(c, li, f, u, cf);
    }
    if (!rc) rc = visit(c, li + 1, cf, cu, c->type == 2);
    if (!rc) rc = ora_interpolate(C->dom, L->dom, cu, u, 1);
  }
  free(cf);
  free(cu);
  if (rc) return rc;
  for (int i = 0; i < c->post; i++)
    if ((rc = level_smooth(L, f, u))) return rc;
  return 0;
}

int ora_cycle_apply(const ora_level* levels, int nlevels, int type, int pre, int post, int mid, int coarse,
                    const double* f, double* u)
{
  memset(u, 0, sizeof(double) * ora_domain_vec_len(levels[0].dom, 1)); /* Cycle::apply: u.setWithGhost(0) */
  vcctx c = { levels, nlevels, pre, post, coarse, 0, type, mid };
  return visit(&c, 0, f, u, 0);
}

int ora_vcycle_apply(const ora_level* levels, int nlevels, int pre, int post, int coarse, const double* f,
                     double* u)
{
  return ora_cycle_apply(levels, nlevels, 0, pre, post, 1, coarse, f, u);
}

static int vcycle_cb(void* ctx, double* in, double* out)
{
  vcctx* c = (vcctx*)ctx;
  int rc = ora_cycle_apply(c->lv, c->nl, c->type, c->pre, c->post, c->mid, c->coarse, in, out);
  if (rc) c->err = rc;
  return rc;
}
static int level_apply_cb(void* ctx, double* in, double* out)
{
  return level_apply((const ora_level*)ctx, in, out);
}

int ora_bicgstab_solve(const ora_level* levels, int nlevels, int use_precond, int pre, int post, int coarse,
                       const double* b, double* x, int max_it, double tol, double* hist, int hist_cap)
{
  vspace V = { levels[0].dom, levels[0].dom->np, 1 };
  vcctx c = { levels, nlevels, pre, post, coarse, 0, 0, 1 };
  int bd = 0;
  int its = bicgstab(&V, level_apply_cb, (void*)&levels[0], x, b, use_precond ? vcycle_cb : NULL, &c, max_it, tol,
                     hist, hist_cap, &bd, use_precond == 2);
  if (bd) return -1;
  if (c.err) return -2;
  return its;
}

/* ------------------------------------------------------------------ CG --
 * Iterative::CG::solve (Iterative/CG.h:126-194), right-preconditioned: ap = A (Mr p)
 * (Solver.h applyWithPreconditioner), x accumulates the unpreconditioned iterate from 0,
 * and at the end x = Mr x + the initial guess. */
int ora_cg_solve(const ora_level* levels, int nlevels, int use_precond, int type, int pre, int post, int mid,
                 int coarse, const double* b, double* x, int max_it, double tol, double* hist, int hist_cap)
{
  const ora_level* L0 = &levels[0];
  vspace V = { L0->dom, L0->dom->np, 1 };
  vcctx c = { levels, nlevels, pre, post, coarse, 0, type, mid };
  long len = vs_len(&V);
  double* resid = (double*)calloc(len, sizeof(double));
  double* x0 = (double*)malloc(sizeof(double) * len);
  double* p = (double*)malloc(sizeof(double) * len);
  double* ap = (double*)calloc(len, sizeof(double));
  double* tmp = (double*)calloc(len, sizeof(double));
  int rc = 0, num_its = 0;
  level_apply(L0, x, resid);
  FOR_INTERIOR(&V, i) resid[i] = resid[i] * -1 + b[i]; /* resid.scaleThenAdd(-1, b) */
  memcpy(x0, x, sizeof(double) * len);                 /* initial_guess = x */
  FOR_INTERIOR(&V, i) x[i] = 0.0;                      /* x.set(0) */
  double r0_norm = vs_norm(&V, b);
  memcpy(p, resid, sizeof(double) * len);
  double rho = vs_dot(&V, resid, resid);
  if (r0_norm == 0) goto done;
  double residual = vs_norm(&V, resid) / r0_norm;
  if (hist && hist_cap > 0) hist[0] = residual;
  while (residual > tol && num_its < max_it) {
    if (rho == 0) {
      rc = -1;
      goto done;
    }
    if (use_precond) {
      memset(tmp, 0, sizeof(double) * len);
      if ((rc = vcycle_cb(&c, p, tmp))) goto done;
      level_apply(L0, tmp, ap);
    } else {
      level_apply(L0, p, ap);
    }
    double alpha = rho / vs_dot(&V, p, ap);
    FOR_INTERIOR(&V, i) x[i] += p[i] * alpha;
    FOR_INTERIOR(&V, i) resid[i] += ap[i] * -alpha;
    double rho_new = vs_dot(&V, resid, resid);
    double beta = rho_new / rho;
    FOR_INTERIOR(&V, i) p[i] = p[i] * beta + resid[i]; /* p.scaleThenAdd(beta, resid) */
    num_its++;
    rho = rho_new;
    residual = vs_norm(&V, resid) / r0_norm;
    if (hist && num_its < hist_cap) hist[num_its] = residual;
  }
  if (use_precond) {
    if ((rc = vcycle_cb(&c, x, resid))) goto done;
    FOR_INTERIOR(&V, i) x[i] = resid[i]; /* x.copy(resid) */
  }
  FOR_INTERIOR(&V, i) x[i] += x0[i]; /* x.add(initial_guess) */
done:
  free(resid);
  free(x0);
  free(p);
  free(ap);
  free(tmp);
  if (rc == -1) return -1;
  if (rc) return -2;
  return num_its;
}
