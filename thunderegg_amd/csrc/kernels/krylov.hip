// Outer FP64 BiCGStab (Iterative/BiCGStab.h:111-203) as fused device kernels.
// The scalars (rho, alpha, omega, beta and the norms) live in a small device
// array; each vector kernel also emits one partial sum per block for the next
// inner product, and a single-block finalize kernel folds the partials in a
// fixed order and updates the scalars.  The host only reads the two norms
// that decide termination.
#include "../te_internal.h"
#include "blas.h"
#include "krylov.h"

namespace te {
namespace {

__device__ __forceinline__ bool cell_of(const Layout& L, int idx, int* p_out, int* q_out, int* c3)
{
  int p = idx / L.ps;
  int rem = idx - p * L.ps;
  int c = rem / L.cs;
  int q = rem - c * L.cs;
  *p_out = p;
  *q_out = q;
  int x, y, z = 0;
  if (L.dim == 2) {
    y = q / L.sx;
    x = q - y * L.sx;
  } else {
    z = q / L.sy;
    int r2 = q - z * L.sy;
    y = r2 / L.sx;
    x = r2 - y * L.sx;
    z -= 1;
  }
  x -= 1;
  y -= 1;
  c3[0] = x;
  c3[1] = y;
  c3[2] = z;
  return x >= 0 && x < L.n[0] && y >= 0 && y < L.n[1] && (L.dim == 2 || (z >= 0 && z < L.n[2]));
}

// out = A u (ghosts 0), with optional fused sums  pa = sum out*w,  pb = sum out*out.
template <int D, int OP, bool WDOT, bool SQ>
__global__ void __launch_bounds__(256) star_dot_kernel(const DevPatch* __restrict__ dp, const double* __restrict__ u,
                                                       const double* __restrict__ cf, double* __restrict__ out,
                                                       const double* __restrict__ w, double* __restrict__ pa,
                                                       double* __restrict__ pb, Layout L, int neumann)
{
  long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)L.np * L.ps;
  double sa = 0.0, sq = 0.0;
  if (gid < total) {
    int idx = (int)gid, p, q, c3[3];
    if (!cell_of(L, idx, &p, &q, c3)) {
      out[idx] = 0.0;
    } else {
      const DevPatch& P = dp[p];
      double mid = u[idx];
      double bc = (OP == 0 && neumann) ? mid : -mid;
      const int st[3] = { 1, L.sx, L.sy };
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < D; a++) {
        bool lo_phys = c3[a] == 0 && P.type[(D == 2 ? 4 : 20) + 2 * a] < 0;
        bool hi_phys = c3[a] == L.n[a] - 1 && P.type[(D == 2 ? 4 : 20) + 2 * a + 1] < 0;
        double lower = lo_phys ? bc : u[idx - st[a]];
        double upper = hi_phys ? bc : u[idx + st[a]];
        if (OP == 0) {
          acc += (upper - 2.0 * mid + lower) * P.inv_h2[a];
        } else {
          double cm = cf[p * L.cs + q], cl = cf[p * L.cs + q - st[a]], cu = cf[p * L.cs + q + st[a]];
          acc += ((cu + cm) * (upper - mid) - (cl + cm) * (mid - lower)) * (P.inv_h2[a] * 0.5);
        }
      }
      out[idx] = acc;
      if (WDOT) sa = w[idx] * acc;  // rhat.dot(ap) / as.dot(s): w first, as the reference
      if (SQ) sq = acc * acc;
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

// Vector updates of one BiCGStab iteration (interior cells only).
template <int MODE>
__global__ void __launch_bounds__(256) bicg_kernel(Layout L, const double* __restrict__ scal, double* __restrict__ v0,
                                                   double* __restrict__ v1, const double* __restrict__ a0,
                                                   const double* __restrict__ a1, const double* __restrict__ a2,
                                                   const double* __restrict__ a3, const double* __restrict__ a4,
                                                   double* __restrict__ pa, double* __restrict__ pb)
{
  long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)L.np * L.ps;
  double sa = 0.0, sb = 0.0;
  int p, q, c3[3];
  if (gid < total && cell_of(L, (int)gid, &p, &q, c3)) {
    int i = (int)gid;
    if (MODE == BICG_INIT) {
      // v0 = rhat, v1 = p, a0 = r, a1 = b: rhat = r; p = r; sums b.b and rhat.r
      double r = a0[i];
      v0[i] = r;
      v1[i] = r;
      sa = a1[i] * a1[i];
      sb = r * r;
    } else if (MODE == BICG_S) {
      // v0 = s, a0 = r, a1 = ap: s = r; s += ap * -alpha; sum s.s
      double s = a0[i];
      s += a1[i] * -scal[SC_ALPHA];
      v0[i] = s;
      sa = s * s;
    } else if (MODE == BICG_XP) {
      // v0 = x, a0 = p: x += p * alpha (early exit, BiCGStab.h:162-167)
      v0[i] += a0[i] * scal[SC_ALPHA];
    } else if (MODE == BICG_XR) {
      // v0 = x, v1 = r, a0 = mp, a1 = ms, a2 = ap, a3 = as, a4 = rhat
      double alpha = scal[SC_ALPHA], omega = scal[SC_OMEGA];
      v0[i] += a0[i] * alpha + a1[i] * omega;
      double r = v1[i];
      r += a2[i] * -alpha + a3[i] * -omega;
      v1[i] = r;
      sa = r * a4[i];
      sb = r * r;
    } else {  // BICG_P: v0 = p, a0 = ap, a1 = r
      double pv = v0[i];
      pv += a0[i] * -scal[SC_OMEGA];
      v0[i] = pv * scal[SC_BETA] + a1[i];
    }
  }
  if (MODE == BICG_INIT || MODE == BICG_S || MODE == BICG_XR) {
    double r = block_reduce<RED_SUM>(sa);
    if (threadIdx.x == 0) pa[blockIdx.x] = r;
  }
  if (MODE == BICG_INIT || MODE == BICG_XR) {
    double r = block_reduce<RED_SUM>(sb);
    if (threadIdx.x == 0) pb[blockIdx.x] = r;
  }
}

__device__ double fold(const double* __restrict__ part, int n)
{
  double acc = 0.0;
  int chunk = (n + blockDim.x - 1) / blockDim.x;
  int b0 = threadIdx.x * chunk;
  for (int k = b0; k < b0 + chunk && k < n; k++) acc += part[k];
  return block_reduce<RED_SUM>(acc);
}

__global__ void __launch_bounds__(256) bicg_finalize(int mode, const double* __restrict__ pa,
                                                     const double* __restrict__ pb, int n, double* __restrict__ scal)
{
  double A = fold(pa, n);
  double B = (mode == FIN_INIT || mode == FIN_OMEGA || mode == FIN_RHO) ? fold(pb, n) : 0.0;
  if (threadIdx.x != 0) return;
  switch (mode) {
    case FIN_INIT:
      scal[SC_R0] = sqrt(A);
      scal[SC_RHO] = B;
      scal[SC_RNORM] = sqrt(B);
      break;
    case FIN_ALPHA: scal[SC_ALPHA] = scal[SC_RHO] / A; break;
    case FIN_SNORM: scal[SC_SNORM] = sqrt(A); break;
    case FIN_OMEGA: scal[SC_OMEGA] = A / B; break;
    default: {  // FIN_RHO
      double rho_new = A;
      scal[SC_BETA] = rho_new * scal[SC_ALPHA] / (scal[SC_RHO] * scal[SC_OMEGA]);
      scal[SC_RHO] = rho_new;
      scal[SC_RNORM] = sqrt(B);
    }
  }
}

}  // namespace

int krylov_blocks(const te_domain* d)
{
  Layout L = make_layout(d->dim, d->ns, 1, d->np);
  long total = (long)L.np * L.ps;
  return (int)((total + 255) / 256);
}

int launch_star_dot(const te_ghost_filler* g, int op_kind, int neumann, const double* coeffs, const double* u,
                    double* out, const double* w, double* pa, double* pb, hipStream_t s)
{
  const te_domain* d = g->d;
  Layout L = make_layout(d->dim, d->ns, 1, d->np);
  int grid = krylov_blocks(d);
  bool wd = w != nullptr, sq = pb != nullptr;
#define TE_SD(D_, OP_, W_, Q_)                                                                               \
  hipLaunchKernelGGL((star_dot_kernel<D_, OP_, W_, Q_>), dim3(grid), dim3(256), 0, s, d->d_patches, u, coeffs, \
                     out, w, pa, pb, L, neumann)
#define TE_SDW(D_, OP_)                                  \
  if (wd && sq) TE_SD(D_, OP_, true, true);              \
  else if (wd) TE_SD(D_, OP_, true, false);              \
  else if (sq) TE_SD(D_, OP_, false, true);              \
  else TE_SD(D_, OP_, false, false);
  if (d->dim == 2) {
    if (op_kind == 0) { TE_SDW(2, 0) } else { TE_SDW(2, 1) }
  } else {
    if (op_kind == 0) { TE_SDW(3, 0) } else { TE_SDW(3, 1) }
  }
#undef TE_SDW
#undef TE_SD
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_bicg(int mode, const te_domain* d, const double* scal, double* v0, double* v1, const double* a0,
                const double* a1, const double* a2, const double* a3, const double* a4, double* pa, double* pb,
                hipStream_t s)
{
  Layout L = make_layout(d->dim, d->ns, 1, d->np);
  int grid = krylov_blocks(d);
#define TE_BK(M_) \
  hipLaunchKernelGGL((bicg_kernel<M_>), dim3(grid), dim3(256), 0, s, L, scal, v0, v1, a0, a1, a2, a3, a4, pa, pb)
  switch (mode) {
    case BICG_INIT: TE_BK(BICG_INIT); break;
    case BICG_S: TE_BK(BICG_S); break;
    case BICG_XP: TE_BK(BICG_XP); break;
    case BICG_XR: TE_BK(BICG_XR); break;
    default: TE_BK(BICG_P); break;
  }
#undef TE_BK
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_bicg_finalize(int mode, const double* pa, const double* pb, int n, double* scal, hipStream_t s)
{
  hipLaunchKernelGGL(bicg_finalize, dim3(1), dim3(256), 0, s, mode, pa, pb, n, scal);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
