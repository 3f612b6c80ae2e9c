// Star stencils: Poisson (Poisson/StarPatchOperator.h:75-101) and VarPoisson
// (VarPoisson/StarPatchOperator.h:61-109), 5-point in 2D and 7-point in 3D,
// optionally fused with the residual r = f - A u of Cycle::restrict
// (GMG/Cycle.h:57-65).  One thread per padded cell: ghost cells of the output
// are written 0 (f.setWithGhost(0), PatchOperator.h:136), interior cells get
// the stencil.  Physical-boundary ghosts of u are applied inline
// (enforceBoundaryConditions, :115-144): -u (Dirichlet) or +u (Neumann).
#include "../te_internal.h"

namespace te {
namespace {

template <typename T, typename TF, typename TO, int D, int OP, bool RESID>
__global__ void __launch_bounds__(256) star_kernel(const DevPatch* __restrict__ dp, const T* __restrict__ u,
                                                   const T* __restrict__ cf, const TF* __restrict__ fin,
                                                   TO* __restrict__ out, Layout L, int neumann)
{
  long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)L.np * L.ps;
  if (gid >= total) return;
  int idx = (int)gid;
  int p = idx / L.ps;
  int rem = idx - p * L.ps;
  int c = rem / L.cs;
  int q = rem - c * L.cs;
  int x, y, z = 0;
  if (D == 2) {
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
  bool inter = x >= 0 && x < L.n[0] && y >= 0 && y < L.n[1] && (D == 2 || (z >= 0 && z < L.n[2]));
  if (!inter) {
    out[idx] = TO(0);
    return;
  }
  const DevPatch& P = dp[p];
  T mid = u[idx];
  T bc = (OP == 0 && neumann) ? mid : -mid;
  int coord[3] = { x, y, z };
  int st[3] = { 1, L.sx, L.sy };
  T acc = T(0);
#pragma unroll
  for (int a = 0; a < D; a++) {
    bool lo_phys = coord[a] == 0 && P.type[(D == 2 ? 4 : 20) + 2 * a] < 0;
    bool hi_phys = coord[a] == L.n[a] - 1 && P.type[(D == 2 ? 4 : 20) + 2 * a + 1] < 0;
    T lower = lo_phys ? bc : u[idx - st[a]];
    T upper = hi_phys ? bc : u[idx + st[a]];
    T ih2 = T(P.inv_h2[a]);
    if (OP == 0) {
      acc += (upper - T(2) * mid + lower) * ih2;
    } else {
      T cm = cf[p * L.cs + q];
      T cl = cf[p * L.cs + q - st[a]];
      T cu = cf[p * L.cs + q + st[a]];
      acc += ((cu + cm) * (upper - mid) - (cl + cm) * (mid - lower)) * (ih2 * T(0.5));
    }
  }
  if (RESID) out[idx] = TO(T(fin[idx]) - acc);
  else out[idx] = TO(acc);
}

template <typename T, typename TF, typename TO, int D, int OP, bool RESID>
void launch_one(const te_domain* d, const T* u, const T* cf, const TF* fin, TO* out, int ncomp, int neumann,
                hipStream_t s)
{
  Layout L = make_layout(d->dim, d->ns, ncomp, d->np);
  long total = (long)L.np * L.ps;
  const int B = 256;
  int grid = (int)((total + B - 1) / B);
  hipLaunchKernelGGL((star_kernel<T, TF, TO, D, OP, RESID>), dim3(grid), dim3(B), 0, s, d->d_patches, u, cf, fin,
                     out, L, neumann);
}

template <typename T, typename TF, typename TO>
void launch_dim(const te_domain* d, int op_kind, bool resid, const T* u, const T* cf, const TF* fin, TO* out,
                int ncomp, int neumann, hipStream_t s)
{
#define TE_L(D_, OP_, R_) launch_one<T, TF, TO, D_, OP_, R_>(d, u, cf, fin, out, ncomp, neumann, s)
  if (d->dim == 2) {
    if (op_kind == 0) { if (resid) TE_L(2, 0, true); else TE_L(2, 0, false); }
    else { if (resid) TE_L(2, 1, true); else TE_L(2, 1, false); }
  } else {
    if (op_kind == 0) { if (resid) TE_L(3, 0, true); else TE_L(3, 0, false); }
    else { if (resid) TE_L(3, 1, true); else TE_L(3, 1, false); }
  }
#undef TE_L
}

}  // namespace

// fin != nullptr: out = fin - A u (residual); else out = A u.  dt: type of u and coeffs,
// dtf: type of fin, dto: type of out.  Supported: (f64,f64,f64), (f32,f32,f32), (f32,f64,f32).
int launch_star(const te_ghost_filler* g, int dt, int dtf, int dto, int op_kind, int neumann, const void* coeffs,
                const void* fin, const void* u, void* out, int ncomp, hipStream_t s)
{
  const te_domain* d = g->d;
  if (d->np == 0) return TE_OK;
  bool resid = fin != nullptr;
  if (dt == TE_F64 && dtf == TE_F64 && dto == TE_F64)
    launch_dim<double, double, double>(d, op_kind, resid, (const double*)u, (const double*)coeffs,
                                       (const double*)fin, (double*)out, ncomp, neumann, s);
  else if (dt == TE_F32 && dtf == TE_F32 && dto == TE_F32)
    launch_dim<float, float, float>(d, op_kind, resid, (const float*)u, (const float*)coeffs, (const float*)fin,
                                    (float*)out, ncomp, neumann, s);
  else if (dt == TE_F32 && dtf == TE_F64 && dto == TE_F32)
    launch_dim<float, double, float>(d, op_kind, resid, (const float*)u, (const float*)coeffs,
                                     (const double*)fin, (float*)out, ncomp, neumann, s);
  else {
    set_error("unsupported dtype combination for the star stencil");
    return TE_ERR_UNSUPPORTED;
  }
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
