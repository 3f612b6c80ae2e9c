// Iterative::PatchSolver<D>(BiCGStab, op) smoothing sweep
// (Iterative/PatchSolver.h:134-175 inside PatchSolver::smooth, PatchSolver.h:142-170):
// every patch runs its own inner BiCGStab (Iterative/BiCGStab.h:111-203) on the
// single-patch operator with internal boundary conditions, started from the
// current u.  One workgroup per patch keeps the whole Krylov state on chip:
// each thread owns CPT cells of every Krylov vector in registers, the operand
// of each operator apply is staged in LDS (VarPoisson: the patch's coefficients
// too, once), and the inner products are workgroup reductions, two at a time
// where the recurrence allows (a.s with a.a, r.rhat with r.r), each over
// double-buffered slots with one barrier.  Patches iterate independently
// (per-patch termination), so a level is one launch.
#include "../te_internal.h"
#include "blas.h"

namespace te {
namespace {

constexpr int kThreads = 256;

constexpr int kWaves = kThreads / 64;

// sums over the workgroup of NV values per thread, returned to every thread; `buf` alternates
// between the two halves of `red` from call to call, so that a call's writes cannot meet the
// previous call's reads and one barrier suffices (the one before it orders the call before)
template <int NV>
__device__ __forceinline__ void wg_sums(double (&v)[NV], double (*red)[2 * kWaves], int& buf)
{
#pragma unroll
  for (int j = 0; j < NV; j++)
    for (int off = 32; off > 0; off >>= 1) v[j] += __shfl_down(v[j], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double* rb = red[buf];
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < NV; j++) rb[j * kWaves + w] = v[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; j++) {
    double r = rb[j * kWaves];
    for (int k = 1; k < kWaves; k++) r += rb[j * kWaves + k];
    v[j] = r;
  }
  buf ^= 1;
}

template <typename T, int D, int OP, int CPT>
__global__ void __launch_bounds__(kThreads) patch_bicgstab_kernel(
  const DevPatch* __restrict__ dp, const T* __restrict__ coeffs, const T* __restrict__ f, T* __restrict__ u,
  Layout L, int neumann, int max_it, double tol, int32_t* __restrict__ iters, int32_t* __restrict__ breakdown)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* pad = reinterpret_cast<T*>(smem_raw);  // padded patch of the operand being applied
  T* cfl = pad + L.cs;                      // VarPoisson: the patch's coefficients (with ghosts)
  __shared__ double red[2][2 * kWaves];
  int rbuf = 0;
  const int p = blockIdx.x;
  const DevPatch& P = dp[p];
  const int n0 = L.n[0], n1 = L.n[1], n2 = D == 3 ? L.n[2] : 1;
  const int N = n0 * n1 * n2;
  const int st[3] = { 1, L.sx, L.sy };
  const int sb = D == 2 ? 4 : 20;
  T ih2[3];
  for (int a = 0; a < 3; a++) ih2[a] = T(P.inv_h2[a]);

  // cell bookkeeping
  int gidx[CPT];   // index within the padded patch
  bool act[CPT];
  for (int k = 0; k < CPT; k++) {
    int o = threadIdx.x + k * kThreads;
    act[k] = o < N;
    int oo = act[k] ? o : 0;
    int x = oo % n0, y = (oo / n0) % n1, z = D == 3 ? oo / (n0 * n1) : 0;
    gidx[k] = lin(L, 0, 0, x, y, z);
  }
  const int base = p * L.ps;
  if (OP == 1)
    for (int k = threadIdx.x; k < L.cs; k += kThreads) cfl[k] = coeffs[base + k];

  // Operator apply on the padded LDS copy: applySinglePatchWithInternalBoundaryConditions
  // (Poisson/StarPatchOperator.h:75-101 with internal=true): every side ghost is
  // -interior, except physical Neumann sides (+interior).
  // (no barrier before the writes: a reduction, with its barrier, separates any two
  // applies, so the previous apply's reads are done)
  auto apply = [&](const T (&v)[CPT], T (&out)[CPT]) {
    for (int k = 0; k < CPT; k++)
      if (act[k]) pad[gidx[k]] = v[k];
    __syncthreads();
    for (int k = 0; k < CPT; k++) {
      if (!act[k]) continue;
      int g = gidx[k];
      int o = threadIdx.x + k * kThreads;
      int c3[3] = { o % n0, (o / n0) % n1, D == 3 ? o / (n0 * n1) : 0 };
      const int nn[3] = { n0, n1, n2 };
      T mid = pad[g];
      T acc = T(0);
#pragma unroll
      for (int a = 0; a < D; a++) {
        T lower, upper;
        if (c3[a] == 0) {
          bool neu = OP == 0 && neumann && P.type[sb + 2 * a] < 0;
          lower = neu ? mid : -mid;
        } else {
          lower = pad[g - st[a]];
        }
        if (c3[a] == nn[a] - 1) {
          bool neu = OP == 0 && neumann && P.type[sb + 2 * a + 1] < 0;
          upper = neu ? mid : -mid;
        } else {
          upper = pad[g + st[a]];
        }
        if (OP == 0) {
          acc += (upper - T(2) * mid + lower) * ih2[a];
        } else {
          T cm = cfl[g], cl = cfl[g - st[a]], cu = cfl[g + st[a]];
          acc += ((cu + cm) * (upper - mid) - (cl + cm) * (mid - lower)) * (ih2[a] * T(0.5));
        }
      }
      out[k] = acc;
    }
  };
  auto dot = [&](const T (&a)[CPT], const T (&b)[CPT]) {
    double v[1] = { 0.0 };
    for (int k = 0; k < CPT; k++)
      if (act[k]) v[0] += double(a[k]) * double(b[k]);
    wg_sums<1>(v, red, rbuf);
    return v[0];
  };
  // a.b and a.c together
  auto dot2 = [&](const T (&a)[CPT], const T (&b)[CPT], const T (&c)[CPT], double* ab, double* ac) {
    double v[2] = { 0.0, 0.0 };
    for (int k = 0; k < CPT; k++)
      if (act[k]) {
        v[0] += double(a[k]) * double(b[k]);
        v[1] += double(a[k]) * double(c[k]);
      }
    wg_sums<2>(v, red, rbuf);
    *ab = v[0];
    *ac = v[1];
  };

  T x[CPT], b[CPT], r[CPT], rhat[CPT], pv[CPT], ap[CPT], s[CPT], as[CPT];
  // f_copy = f; modifyRHSForInternalBoundaryConditions (Poisson :165-180, VarPoisson :168-186)
  for (int k = 0; k < CPT; k++) {
    x[k] = T(0);
    b[k] = T(0);
    if (!act[k]) continue;
    int g = gidx[k];
    int o = threadIdx.x + k * kThreads;
    int c3[3] = { o % n0, (o / n0) % n1, D == 3 ? o / (n0 * n1) : 0 };
    const int nn[3] = { n0, n1, n2 };
    x[k] = u[base + g];
    T v = f[base + g];
    for (int a = 0; a < D; a++) {
      for (int up = 0; up < 2; up++) {
        bool edge = up ? (c3[a] == nn[a] - 1) : (c3[a] == 0);
        if (!edge || P.type[sb + 2 * a + up] < 0) continue;
        int gg = up ? g + st[a] : g - st[a];
        if (OP == 0) v -= (u[base + gg] + u[base + g]) * ih2[a];
        else v -= (u[base + gg] + u[base + g]) * (coeffs[base + g] + coeffs[base + gg]) * (ih2[a] * T(0.5));
      }
    }
    b[k] = v;
  }
  // BiCGStab::solve
  apply(x, r);
  for (int k = 0; k < CPT; k++) r[k] = r[k] * T(-1) + b[k];
  double r0_norm = sqrt(dot(b, b));
  for (int k = 0; k < CPT; k++) {
    rhat[k] = r[k];
    pv[k] = r[k];
  }
  double rho = dot(rhat, r);
  int its = 0;
  int bd = 0;
  double residual = sqrt(dot(r, r)) / r0_norm;
  if (r0_norm != 0.0) {
    while (residual > tol && its < max_it) {
      if (rho == 0.0) {
        bd = 1;
        break;
      }
      apply(pv, ap);
      double alpha = rho / dot(rhat, ap);
      for (int k = 0; k < CPT; k++) s[k] = r[k] + ap[k] * T(-alpha);
      if (sqrt(dot(s, s)) / r0_norm <= tol) {
        for (int k = 0; k < CPT; k++) x[k] += pv[k] * T(alpha);
        break;
      }
      apply(s, as);
      double as_s, as_as;
      dot2(as, s, as, &as_s, &as_as);
      double omega = as_s / as_as;
      for (int k = 0; k < CPT; k++) {
        x[k] += pv[k] * T(alpha) + s[k] * T(omega);
        r[k] += ap[k] * T(-alpha) + as[k] * T(-omega);
      }
      double rho_new, rr;
      dot2(r, rhat, r, &rho_new, &rr);
      double beta = rho_new * alpha / (rho * omega);
      for (int k = 0; k < CPT; k++) {
        pv[k] += ap[k] * T(-omega);
        pv[k] = pv[k] * T(beta) + r[k];
      }
      its++;
      rho = rho_new;
      residual = sqrt(rr) / r0_norm;
    }
  }
  for (int k = 0; k < CPT; k++)
    if (act[k]) u[base + gidx[k]] = x[k];
  if (threadIdx.x == 0) {
    if (iters) iters[p] = its;
    if (bd) breakdown[0] = 1;
  }
}

template <typename T, int D, int OP>
int launch_cpt(const te_domain* d, int cpt, const T* coeffs, const T* f, T* u, int neumann, int max_it, double tol,
               int32_t* iters, int32_t* bd, hipStream_t s)
{
  Layout L = make_layout(d->dim, d->ns, 1, d->np);
  size_t shmem = sizeof(T) * (size_t)L.cs * (OP == 1 ? 2 : 1);
#define TE_PB(C_)                                                                                            \
  hipLaunchKernelGGL((patch_bicgstab_kernel<T, D, OP, C_>), dim3(d->np), dim3(kThreads), shmem, s, d->d_patches, \
                     coeffs, f, u, L, neumann, max_it, tol, iters, bd)
  switch (cpt) {
    case 1: TE_PB(1); break;
    case 2: TE_PB(2); break;
    case 4: TE_PB(4); break;
    case 8: TE_PB(8); break;
    default: TE_PB(16); break;
  }
#undef TE_PB
  return TE_OK;
}

}  // namespace

int launch_patch_bicgstab(const te_domain* d, int dtype, int op_kind, int neumann, const void* coeffs, const void* f,
                          void* u, int max_it, double tol, int32_t* iters, int32_t* bd, hipStream_t s)
{
  if (d->np == 0) return TE_OK;
  int N = d->ns[0] * d->ns[1] * (d->dim == 3 ? d->ns[2] : 1);
  int cpt = (N + kThreads - 1) / kThreads;
  int c = 1;
  while (c < cpt) c *= 2;
  if (c > 16) {
    set_error("patch BiCGStab smoother supports at most 4096 cells per patch");
    return TE_ERR_UNSUPPORTED;
  }
#define TE_PD(T)                                                                                                   \
  if (d->dim == 2) {                                                                                               \
    if (op_kind == 0) launch_cpt<T, 2, 0>(d, c, (const T*)coeffs, (const T*)f, (T*)u, neumann, max_it, tol, iters, bd, s); \
    else launch_cpt<T, 2, 1>(d, c, (const T*)coeffs, (const T*)f, (T*)u, neumann, max_it, tol, iters, bd, s);      \
  } else {                                                                                                         \
    if (op_kind == 0) launch_cpt<T, 3, 0>(d, c, (const T*)coeffs, (const T*)f, (T*)u, neumann, max_it, tol, iters, bd, s); \
    else launch_cpt<T, 3, 1>(d, c, (const T*)coeffs, (const T*)f, (T*)u, neumann, max_it, tol, iters, bd, s);      \
  }
  if (dtype == TE_F64) {
    TE_PD(double)
  } else {
    TE_PD(float)
  }
#undef TE_PD
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
