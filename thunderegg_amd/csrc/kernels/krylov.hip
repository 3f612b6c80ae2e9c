// Outer FP64 BiCGStab (Iterative/BiCGStab.h:111-203) as fused device kernels.
// The scalars (rho, alpha, omega, beta and the norms) live in a small device
// array; each vector kernel also emits one partial sum per workgroup for the
// next inner product, and a single-workgroup finalize kernel folds the
// partials in a fixed order and updates the scalars.  The host only reads the
// two norms that decide termination.
//
// The vector kernels are patch-blocked: a workgroup walks whole patches, and
// each thread precomputes the padded offsets of the cells it owns in a patch,
// so the per-element work is a load, the arithmetic and a store.
#include <type_traits>

#include "../te_internal.h"
#include "blas.h"
#include "krylov.h"
#include "fold.h"

namespace te {
namespace {

constexpr int kThreads = 256;
static_assert(kThreads == kFoldThreads, "the finalize folds with the Krylov block size");
constexpr int kMaxCells = 24;  // largest per-thread cell count instantiated (3D n = 16: 18^3 / 256 -> 23)

// The cells a thread owns in every patch (at most CPT; unrolled so they stay in registers).
template <int CPT>
struct CellSet {
  int n;              // cells of this thread
  int off[CPT];       // padded offset in the patch
  unsigned bnd[CPT];  // bit 2a: lower boundary on axis a, bit 2a+1: upper; bit 7: ghost cell
};

template <int CPT>
__device__ __forceinline__ void make_cells(const Layout& L, bool with_ghosts, CellSet<CPT>& cs)
{
  cs.n = 0;
  const int n0 = L.n[0], n1 = L.n[1], n2 = L.dim == 3 ? L.n[2] : 1;
  const int p0 = n0 + 2, p1 = n1 + 2, p2 = L.dim == 3 ? n2 + 2 : 1;
  const int total = with_ghosts ? p0 * p1 * p2 : n0 * n1 * n2;
#pragma unroll
  for (int k = 0; k < CPT; k++) {
    const int o = threadIdx.x + k * blockDim.x;
    cs.off[k] = 0;
    cs.bnd[k] = 0;
    if (o >= total) continue;
    int x, y, z;
    if (with_ghosts) {
      x = o % p0 - 1;
      y = (o / p0) % p1 - 1;
      z = L.dim == 3 ? o / (p0 * p1) - 1 : 0;
    } else {
      x = o % n0;
      y = (o / n0) % n1;
      z = L.dim == 3 ? o / (n0 * n1) : 0;
    }
    bool ghost = x < 0 || x >= n0 || y < 0 || y >= n1 || (L.dim == 3 && (z < 0 || z >= n2));
    unsigned char b = ghost ? 0x80 : 0;
    if (!ghost) {
      b |= (x == 0 ? 1 : 0) | (x == n0 - 1 ? 2 : 0) | (y == 0 ? 4 : 0) | (y == n1 - 1 ? 8 : 0);
      if (L.dim == 3) b |= (z == 0 ? 16 : 0) | (z == n2 - 1 ? 32 : 0);
    }
    cs.off[k] = (x + 1) + (y + 1) * L.sx + (L.dim == 3 ? (z + 1) * L.sy : 0);
    cs.bnd[k] = b;
    cs.n = k + 1;
  }
}

// out = A u (ghosts of out written 0), with optional fused sums pa = sum w*out, pb = sum out*out.
// The ghost cells of u must be filled.
template <int CPT, int D, int OP, bool WDOT, bool SQ>
__global__ void __launch_bounds__(kThreads) star_dot_kernel(const DevPatch* __restrict__ dp,
                                                            const double* __restrict__ u,
                                                            const double* __restrict__ cf, double* __restrict__ out,
                                                            const double* __restrict__ w, double* __restrict__ pa,
                                                            double* __restrict__ pb, Layout L, int neumann)
{
  CellSet<CPT> cs;
  make_cells<CPT>(L, true, cs);
  const int st[3] = { 1, L.sx, L.sy };
  const int sb = D == 2 ? 4 : 20;
  double sa = 0.0, sq = 0.0;
  for (int p = blockIdx.x; p < L.np; p += gridDim.x) {
    const DevPatch& P = dp[p];
    unsigned phys = 0;  // bit s: side s has no neighbour
#pragma unroll
    for (int s = 0; s < 2 * D; s++) phys |= (P.type[sb + s] < 0 ? 1u : 0u) << s;
    double ih2[3] = { P.inv_h2[0], P.inv_h2[1], P.inv_h2[2] };
    const int base = p * L.ps;
#pragma unroll
    for (int k = 0; k < CPT; k++) {
      if (k >= cs.n) continue;
      const int i = base + cs.off[k];
      const unsigned b = cs.bnd[k];
      if (b & 0x80) {
        out[i] = 0.0;
        continue;
      }
      double mid = u[i];
      double bc = (OP == 0 && neumann) ? mid : -mid;
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < D; a++) {
        const double lower = ((b >> (2 * a)) & 1) && ((phys >> (2 * a)) & 1) ? bc : u[i - st[a]];
        const double upper = ((b >> (2 * a + 1)) & 1) && ((phys >> (2 * a + 1)) & 1) ? bc : u[i + st[a]];
        if (OP == 0) {
          acc += (upper - 2.0 * mid + lower) * ih2[a];
        } else {
          const int ci = p * L.cs + cs.off[k];
          double cm = cf[ci], cl = cf[ci - st[a]], cu = cf[ci + st[a]];
          acc += ((cu + cm) * (upper - mid) - (cl + cm) * (mid - lower)) * (ih2[a] * 0.5);
        }
      }
      out[i] = acc;
      if (WDOT) sa += w[i] * acc;  // rhat.dot(ap) / as.dot(s): w first, as the reference
      if (SQ) sq += acc * acc;
    }
  }
  if (WDOT) {
    double r = block_reduce<RED_SUM>(sa);
    if (threadIdx.x == 0) pa[blockIdx.x] = r;
  }
  if (SQ) {
    double r = block_reduce<RED_SUM>(sq);
    if (threadIdx.x == 0) pb[blockIdx.x] = r;
  }
}

// Vector updates of one BiCGStab iteration (interior cells only).  After the s-norm test
// has fired (state ST_EARLY) only the early-exit update x += alpha dir runs: BICG_XP, or
// BICG_XR given dirv (dir of type TM when dir_m, else FP64), which otherwise updates x and r.
// TM: type of M p and M s (a0, a1 of BICG_XR; a0 of BICG_XP), widened to FP64 as read.
// w32 (BICG_INIT, BICG_S, BICG_P; may be null): an FP32 copy of the new p or s, the
// right-hand side the mixed-precision V-cycle reads (it rounds its input to FP32 anyway).
// pfold (BICG_S, may be null): the nfold rhat.ap partials; every workgroup folds them into
// alpha = rho / rhat.ap as the finalize would (same order and bits) and workgroup 0 stores
// alpha, so that no separate finalize launch is needed.
template <int CPT, int MODE, typename TM>
__global__ void __launch_bounds__(kThreads) bicg_kernel(Layout L, double* __restrict__ scal,
                                                        double* __restrict__ v0, double* __restrict__ v1,
                                                        const void* __restrict__ a0v,
                                                        const void* __restrict__ a1v,
                                                        const double* __restrict__ a2,
                                                        const double* __restrict__ a3,
                                                        const double* __restrict__ a4, double* __restrict__ pa,
                                                        double* __restrict__ pb, float* __restrict__ w32,
                                                        const void* __restrict__ dirv, int dir_m,
                                                        const double* __restrict__ pfold, int nfold, BatchMap bm)
{
  CellSet<CPT> cs;
  // this workgroup's patches p0, p0 + pstep, ... below pend and its partial slot; in a
  // batch those of its problem, whose scalars it reads
  int p0 = blockIdx.x, pstep = gridDim.x, pend = L.np, slot = blockIdx.x;
  bool lead = blockIdx.x == 0;
  if (bm.prob) {
    const int q = bm.wq[blockIdx.x];
    const BatchProb P = bm.prob[q];
    const int b = (int)blockIdx.x - P.woff;
    p0 = P.poff + b;
    pstep = P.nwg;
    pend = P.poff + P.np;
    slot = P.soff + b;
    lead = b == 0;
    scal += (long)q * SC_COUNT;
    if (pfold) pfold += P.soff;
    nfold = P.nslot;
  }
  const bool fold = MODE == BICG_S && pfold != nullptr;
  double alpha = fold ? 0.0 : scal[SC_ALPHA];
  const double omega = scal[SC_OMEGA], beta = scal[SC_BETA], rho = scal[SC_RHO];
  if (MODE != BICG_INIT) {
    const double st = scal[SC_STATE];
    if (MODE == BICG_XR && dirv && st == ST_EARLY) {
      // x += alpha dir (BiCGStab.h:162-167)
      make_cells<CPT>(L, false, cs);
      for (int p = p0; p < pend; p += pstep) {
#pragma unroll
        for (int k = 0; k < CPT; k++) {
          if (k >= cs.n) continue;
          const int i = p * L.ps + cs.off[k];
          const double dv = dir_m ? double(at((const TM*)dirv, i)) : at((const double*)dirv, i);
          at(v0, i) += dv * alpha;
        }
      }
      return;
    }
    if (MODE == BICG_XP ? st != ST_EARLY : st != ST_RUN) return;
  }
  if (fold) {  // FIN_ALPHA: alpha = rho / rhat.dot(ap)
    const double* const parts[3] = { pfold, nullptr, nullptr };
    double r[3] = { 0.0, 0.0, 0.0 };
    fold_n<1>(parts, nfold, r);
    alpha = rho / r[0];
    if (lead && threadIdx.x == 0) scal[SC_ALPHA] = alpha;
  }
  typedef typename std::conditional<MODE == BICG_XR || MODE == BICG_XP, TM, double>::type TA;
  const TA* __restrict__ a0 = (const TA*)a0v;
  const TA* __restrict__ a1 = (const TA*)a1v;
  make_cells<CPT>(L, false, cs);
  double sa = 0.0, sb = 0.0;
  for (int p = p0; p < pend; p += pstep) {
    const int base = p * L.ps;
#pragma unroll
    for (int k = 0; k < CPT; k++) {
      if (k >= cs.n) continue;
      const int i = base + cs.off[k];
      if (MODE == BICG_INIT) {
        // v0 = rhat, v1 = p, a0 = r, a1 = b: rhat = r; p = r; sums b.b and rhat.r
        double r = at(a0, i);
        at(v0, i) = r;
        at(v1, i) = r;
        if (w32) at(w32, i) = float(r);
        sa += at(a1, i) * at(a1, i);
        sb += r * r;
      } else if (MODE == BICG_S) {
        // v0 = s, a0 = r, a1 = ap: s = r; s += ap * -alpha; sum s.s
        double s = at(a0, i);
        s += at(a1, i) * -alpha;
        at(v0, i) = s;
        if (w32) at(w32, i) = float(s);
        sa += s * s;
      } else if (MODE == BICG_XP) {
        // v0 = x, a0 = p: x += p * alpha (early exit, BiCGStab.h:162-167)
        at(v0, i) += double(at(a0, i)) * alpha;
      } else if (MODE == BICG_XR) {
        // v0 = x, v1 = r, a0 = mp, a1 = ms, a2 = ap, a3 = as, a4 = rhat
        at(v0, i) += double(at(a0, i)) * alpha + double(at(a1, i)) * omega;
        double r = at(v1, i);
        r += at(a2, i) * -alpha + at(a3, i) * -omega;
        at(v1, i) = r;
        sa += r * at(a4, i);
        sb += r * r;
      } else {  // BICG_P: v0 = p, a0 = ap, a1 = r
        double pv = at(v0, i);
        pv += at(a0, i) * -omega;
        const double pn = pv * beta + at(a1, i);
        at(v0, i) = pn;
        if (w32) at(w32, i) = float(pn);
      }
    }
  }
  if (MODE == BICG_INIT || MODE == BICG_S || MODE == BICG_XR) {
    double r = block_reduce<RED_SUM>(sa);
    if (threadIdx.x == 0) pa[slot] = r;
  }
  if (MODE == BICG_INIT || MODE == BICG_XR) {
    double r = block_reduce<RED_SUM>(sb);
    if (threadIdx.x == 0) pb[slot] = r;
  }
}

// After the s-norm test has fired only FIN_RHO runs, to publish the state.
// FIN_SOMEGA: the s-norm test from pc (the s.s partials), then omega unless it fired.
__global__ void __launch_bounds__(kThreads) bicg_finalize(int mode, const double* __restrict__ pa,
                                                          const double* __restrict__ pb,
                                                          const double* __restrict__ pc, int n,
                                                          double* __restrict__ scal, double tol,
                                                          double* __restrict__ mirror)
{
  // the state and the partials are loaded together (the partial buffers are always valid
  // memory; after the s-norm test has fired their fold is simply not used)
  const double state = scal[SC_STATE];
  const double* const parts[3] = { pa, pb, pc };
  double r[3] = { 0.0, 0.0, 0.0 };
  if (mode == FIN_SOMEGA || (mode == FIN_INIT && pc)) fold_n<3>(parts, n, r);
  else if (mode == FIN_ALPHA || mode == FIN_SNORM) fold_n<1>(parts, n, r);
  else fold_n<2>(parts, n, r);
  const bool run = mode == FIN_INIT || state == ST_RUN;
  if (!run && mode != FIN_RHO) return;
  const double A = r[0], B = r[1], C = r[2];
  if (threadIdx.x != 0) return;
  switch (mode) {
    case FIN_INIT:  // pc: the rhat.r partials when rhat is not r itself (else rho = r.r)
      scal[SC_R0] = sqrt(A);
      scal[SC_RHO] = pc ? C : B;
      scal[SC_RNORM] = sqrt(B);
      scal[SC_TOL] = tol;
      scal[SC_STATE] = ST_RUN;
      break;
    case FIN_ALPHA: scal[SC_ALPHA] = scal[SC_RHO] / A; break;
    case FIN_SNORM: {
      // if (s.twoNorm() / r0_norm <= tolerance): x += alpha p and stop (BiCGStab.h:162-167)
      const double sn = sqrt(A);
      scal[SC_SNORM] = sn;
      if (sn / scal[SC_R0] <= scal[SC_TOL]) scal[SC_STATE] = ST_EARLY;
      break;
    }
    case FIN_OMEGA: scal[SC_OMEGA] = A / B; break;
    case FIN_SOMEGA: {
      const double sn = sqrt(C);
      scal[SC_SNORM] = sn;
      if (sn / scal[SC_R0] <= scal[SC_TOL]) scal[SC_STATE] = ST_EARLY;
      else scal[SC_OMEGA] = A / B;
      break;
    }
    default: {  // FIN_RHO
      if (run) {
        double rho_new = A;
        scal[SC_BETA] = rho_new * scal[SC_ALPHA] / (scal[SC_RHO] * scal[SC_OMEGA]);
        scal[SC_RHO] = rho_new;
        scal[SC_RNORM] = sqrt(B);
      }
      const double seq = scal[SC_SEQ] + 1.0;
      scal[SC_SEQ] = seq;
      if (mirror) {
        for (int k = 0; k < SC_SEQ; k++) mirror[k] = scal[k];
        __threadfence_system();  // the scalars reach the host before the sequence number
        *(volatile double*)&mirror[SC_SEQ] = seq;
      }
    }
  }
}

// bicg_finalize for every problem of a batch: workgroup q folds problem q's slots and
// updates its scalars; the loop control of BiCGStab.h:148-150 runs here (the host only
// waits until no problem runs).
__global__ void __launch_bounds__(kThreads) bicg_finalize_batch(int mode, const double* __restrict__ pa,
                                                                const double* __restrict__ pb,
                                                                const double* __restrict__ pc,
                                                                const BatchProb* __restrict__ prob,
                                                                double* __restrict__ scal_all, double tol,
                                                                int max_it)
{
  const BatchProb P = prob[blockIdx.x];
  double* __restrict__ scal = scal_all + (long)blockIdx.x * SC_COUNT;
  const double state = scal[SC_STATE];
  const double* const parts[3] = { pa + P.soff, pb ? pb + P.soff : nullptr, pc ? pc + P.soff : nullptr };
  double r[3] = { 0.0, 0.0, 0.0 };
  if (mode == FIN_SOMEGA) fold_n<3>(parts, P.nslot, r);
  else if (mode == FIN_ALPHA || mode == FIN_SNORM) fold_n<1>(parts, P.nslot, r);
  else fold_n<2>(parts, P.nslot, r);
  const bool run = mode == FIN_INIT || state == ST_RUN;
  if (!run && !(mode == FIN_RHO && state == ST_EARLY)) return;
  const double A = r[0], B = r[1], C = r[2];
  if (threadIdx.x != 0) return;
  switch (mode) {
    case FIN_INIT: {
      const double r0 = sqrt(A), rn = sqrt(B);
      scal[SC_R0] = r0;
      scal[SC_RHO] = B;
      scal[SC_RNORM] = rn;
      scal[SC_TOL] = tol;
      scal[SC_ITS] = 0.0;
      scal[SC_MAXIT] = max_it;
      // if (r0_norm == 0) no iteration; while (residual > tolerance && num_its < max_it),
      // and rho == 0 at the top of an iteration is the breakdown error
      double st = ST_RUN;
      if (r0 == 0.0 || !(rn / r0 > tol) || max_it <= 0) st = ST_DONE;
      else if (B == 0.0) st = ST_BREAK;
      scal[SC_STATE] = st;
      break;
    }
    case FIN_ALPHA: scal[SC_ALPHA] = scal[SC_RHO] / A; break;
    case FIN_SNORM: {
      const double sn = sqrt(A);
      scal[SC_SNORM] = sn;
      if (sn / scal[SC_R0] <= scal[SC_TOL]) scal[SC_STATE] = ST_EARLY;
      break;
    }
    case FIN_OMEGA: scal[SC_OMEGA] = A / B; break;
    case FIN_SOMEGA: {
      const double sn = sqrt(C);
      scal[SC_SNORM] = sn;
      if (sn / scal[SC_R0] <= scal[SC_TOL]) scal[SC_STATE] = ST_EARLY;
      else scal[SC_OMEGA] = A / B;
      break;
    }
    default: {  // FIN_RHO
      if (state == ST_EARLY) {  // x was updated; the loop ends without counting the iteration
        scal[SC_STATE] = ST_DONE;
        break;
      }
      const double rho_new = A;
      scal[SC_BETA] = rho_new * scal[SC_ALPHA] / (scal[SC_RHO] * scal[SC_OMEGA]);
      scal[SC_RHO] = rho_new;
      const double rn = sqrt(B);
      scal[SC_RNORM] = rn;
      const double its = scal[SC_ITS] + 1.0;
      scal[SC_ITS] = its;
      if (!(rn / scal[SC_R0] > scal[SC_TOL]) || its >= scal[SC_MAXIT]) scal[SC_STATE] = ST_DONE;
      else if (rho_new == 0.0) scal[SC_STATE] = ST_BREAK;
    }
  }
}

__global__ void __launch_bounds__(kThreads) batch_publish(const double* __restrict__ scal, int nprob,
                                                          double* __restrict__ ctl, double* __restrict__ mirror)
{
  __shared__ int cnt[2][kThreads / 64];
  int run = 0, brk = 0;
  for (int q = threadIdx.x; q < nprob; q += kThreads) {
    const double st = scal[(long)q * SC_COUNT + SC_STATE];
    run += (st == ST_RUN || st == ST_EARLY) ? 1 : 0;
    brk += st == ST_BREAK ? 1 : 0;
  }
  for (int off = 32; off > 0; off >>= 1) {
    run += __shfl_down(run, off, 64);
    brk += __shfl_down(brk, off, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    cnt[0][w] = run;
    cnt[1][w] = brk;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 1; k < kThreads / 64; k++) {
    run += cnt[0][k];
    brk += cnt[1][k];
  }
  // thread 0 of wave 0 holds its own wave's sums in run / brk already
  const double seq = ctl[0] + 1.0;
  ctl[0] = seq;
  mirror[1] = run;
  mirror[2] = brk;
  __threadfence_system();  // the counts reach the host before the sequence number
  *(volatile double*)&mirror[0] = seq;
}

// blockIdx.y = problem; workgroups stride over its elements
__global__ void __launch_bounds__(kThreads) batch_copy(bool to_batch, double* __restrict__ batch,
                                                       double* const* __restrict__ ptr, const long* __restrict__ off,
                                                       const long* __restrict__ len)
{
  const int q = blockIdx.y;
  const long n = len[q];
  double* own = ptr[q];
  double* bv = batch + off[q];
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads) {
    if (to_batch) bv[i] = own[i];
    else own[i] = bv[i];
  }
}

}  // namespace

// smallest instantiated per-thread cell count covering `cells`
static int pick_cpt(int cells)
{
  int need = (cells + kThreads - 1) / kThreads;
  const int opts[] = { 2, 5, 8, 16, kMaxCells };
  for (int c : opts)
    if (c >= need) return c;
  return -1;
}

int krylov_blocks(const te_domain* d)
{
  int g = d->np < 2048 ? d->np : 2048;
  return g < 1 ? 1 : g;
}

// out = A u with the dot w.out (w required) and optionally out.out (pb != null).
// fly: compute neighbour ghosts on the fly (2D, launch_star_dot_2d) instead of reading them from u.
int launch_star_dot(const te_ghost_filler* g, int op_kind, int neumann, const double* coeffs, const void* u, int udt,
                    double* out, const double* w, double* pa, double* pb, bool fly, hipStream_t s)
{
  const te_domain* d = g->d;
  Layout L = make_layout(d->dim, d->ns, 1, d->np);
  int cpt = pick_cpt(L.cs);
  if (cpt < 0 || !w) {
    set_error("fused Krylov apply: patch too large or missing dot operand");
    return TE_ERR_UNSUPPORTED;
  }
  int grid = krylov_blocks(d);
  // the partial sums cover krylov_blocks(d) workgroups whichever kernel runs
  if (fly) return launch_star_dot_2d(g, op_kind, neumann, coeffs, u, udt, out, w, pa, pb, grid, s);
  if (udt != TE_F64) {
    set_error("internal: the Krylov apply reads FP32 input only with ghosts computed in the launch");
    return TE_ERR_UNSUPPORTED;
  }
  bool sq = pb != nullptr;
#define TE_SDC(C_, D_, OP_, Q_)                                                                               \
  hipLaunchKernelGGL((star_dot_kernel<C_, D_, OP_, true, Q_>), dim3(grid), dim3(kThreads), 0, s, d->d_patches, \
                     (const double*)u, coeffs, out, w, pa, pb, L, neumann)
#define TE_SD(D_, OP_, Q_)                      \
  do {                                          \
    switch (cpt) {                              \
      case 2: TE_SDC(2, D_, OP_, Q_); break;    \
      case 5: TE_SDC(5, D_, OP_, Q_); break;    \
      case 8: TE_SDC(8, D_, OP_, Q_); break;    \
      case 16: TE_SDC(16, D_, OP_, Q_); break;  \
      default: TE_SDC(kMaxCells, D_, OP_, Q_);  \
    }                                           \
  } while (0)
#define TE_SDQ(D_, OP_)                \
  do {                                 \
    if (sq) TE_SD(D_, OP_, true);      \
    else TE_SD(D_, OP_, false);        \
  } while (0)
  if (d->dim == 2) {
    if (op_kind == 0) TE_SDQ(2, 0);
    else TE_SDQ(2, 1);
  } else {
    if (op_kind == 0) TE_SDQ(3, 0);
    else TE_SDQ(3, 1);
  }
#undef TE_SDQ
#undef TE_SD
#undef TE_SDC
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_bicg(int mode, const te_domain* d, double* scal, double* v0, double* v1, const void* a0,
                const void* a1, const double* a2, const double* a3, const double* a4, double* pa, double* pb,
                hipStream_t s, int dtm, float* w32, const void* xp_dir, int xp_dir_m, const double* pfold, int nfold,
                const BatchMap* bm)
{
  // (in a batch each problem folds its own slots, at most krylov_blocks of its patches)
  if (pfold && (mode != BICG_S || (!bm && (nfold < 1 || nfold > kFoldMax)))) {
    set_error("internal: only the s update folds partials (at most kFoldMax)");
    return TE_ERR_ARG;
  }
  if ((dtm != TE_F64 && mode != BICG_XR && mode != BICG_XP) ||
      (w32 && mode != BICG_INIT && mode != BICG_S && mode != BICG_P)) {
    set_error("internal: FP32 operands only for the x updates, FP32 copies only of p and s");
    return TE_ERR_ARG;
  }
  Layout L = make_layout(d->dim, d->ns, 1, d->np);
  int cpt = pick_cpt(d->ns[0] * d->ns[1] * (d->dim == 3 ? d->ns[2] : 1));
  if (cpt < 0) {
    set_error("patch too large for the fused Krylov kernels");
    return TE_ERR_UNSUPPORTED;
  }
  int grid = krylov_blocks(d);
  BatchMap m{};
  if (bm) {
    m = *bm;
    grid = m.nwg;
    if (grid < 1) return TE_OK;
  }
#define TE_BKC(C_, M_)                                                                                           \
  do {                                                                                                           \
    if (dtm == TE_F32)                                                                                           \
      hipLaunchKernelGGL((bicg_kernel<C_, M_, float>), dim3(grid), dim3(kThreads), 0, s, L, scal, v0, v1, a0, a1, \
                         a2, a3, a4, pa, pb, w32, xp_dir, xp_dir_m, pfold, nfold, m);                            \
    else                                                                                                         \
      hipLaunchKernelGGL((bicg_kernel<C_, M_, double>), dim3(grid), dim3(kThreads), 0, s, L, scal, v0, v1, a0,   \
                         a1, a2, a3, a4, pa, pb, w32, xp_dir, xp_dir_m, pfold, nfold, m);                        \
  } while (0)
#define TE_BK(M_)                      \
  do {                                 \
    switch (cpt) {                     \
      case 2: TE_BKC(2, M_); break;    \
      case 5: TE_BKC(5, M_); break;    \
      case 8: TE_BKC(8, M_); break;    \
      case 16: TE_BKC(16, M_); break;  \
      default: TE_BKC(kMaxCells, M_);  \
    }                                  \
  } while (0)
  switch (mode) {
    case BICG_INIT: TE_BK(BICG_INIT); break;
    case BICG_S: TE_BK(BICG_S); break;
    case BICG_XP: TE_BK(BICG_XP); break;
    case BICG_XR: TE_BK(BICG_XR); break;
    default: TE_BK(BICG_P); break;
  }
#undef TE_BK
#undef TE_BKC
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_bicg_finalize(int mode, const double* pa, const double* pb, int n, double* scal, hipStream_t s, double tol,
                         double* mirror, const double* pc)
{
  if (n > kFoldMax) {
    set_error("internal: too many partial sums for the BiCGStab finalize");
    return TE_ERR_ARG;
  }
  hipLaunchKernelGGL(bicg_finalize, dim3(1), dim3(kThreads), 0, s, mode, pa, pb, pc, n, scal, tol, mirror);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_bicg_finalize_batch(int mode, const double* pa, const double* pb, const double* pc, const BatchMap& m,
                               double* scal, hipStream_t s, double tol, int max_it)
{
  if (m.nprob < 1) return TE_OK;
  hipLaunchKernelGGL(bicg_finalize_batch, dim3(m.nprob), dim3(kThreads), 0, s, mode, pa, pb, pc, m.prob, scal, tol,
                     max_it);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_batch_publish(const double* scal, int nprob, double* ctl, double* mirror, hipStream_t s)
{
  hipLaunchKernelGGL(batch_publish, dim3(1), dim3(kThreads), 0, s, scal, nprob, ctl, mirror);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_batch_copy(bool to_batch, double* batch, double* const* ptr, const long* off, const long* len, int nprob,
                      long maxlen, hipStream_t s)
{
  if (nprob < 1 || maxlen < 1) return TE_OK;
  if (nprob > 65535) {
    set_error("internal: at most 65535 problems per batch copy");
    return TE_ERR_ARG;
  }
  long gx = (maxlen + kThreads * 4 - 1) / (kThreads * 4);
  if (gx > 256) gx = 256;
  hipLaunchKernelGGL(batch_copy, dim3((unsigned)gx, (unsigned)nprob), dim3(kThreads), 0, s, to_batch, batch, ptr, off,
                     len);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
