// Inter-level transfer.  Restriction is computed per coarse cell (a gather, so
// every coarse cell is written exactly once and no atomics are needed);
// interpolation is computed per fine cell.
//  - LinearRestrictor (GMG/LinearRestrictor.h:101-155): coarse[(i+s)/2] += fine[i]/2^D
//    for children with an orthant on the parent, coarse[i] += fine[i] for copies;
//    MPIRestrictor::restrict zeroes the coarse vector with ghosts first (:73).  With
//    extrapolate_boundary_ghosts (:49-92, 138-155) the coarse ghost cells over a child's
//    exterior sides get (3 ghost - interior) / 2^D of the child's side cells, the corner of
//    the child's orthant, when it has no neighbour there, 3 ghost - interior of the child's
//    corner cell, and copies take the ghost cells too.
//  - DirectInterpolator (GMG/DirectInterpolator.h:60-91): fine[i] += coarse[(i+s)/2].
// Children of a coarse patch are visited in increasing fine local index and fine
// cells in x-fastest order, the order in which the reference accumulates them.
#include "../te_internal.h"

namespace te {
namespace {

// the coarse ghost cell at (X, Y, Z) of patch p with extrapolate_boundary_ghosts: the sum
// over p's children in increasing local order, each term as the reference forms it
template <typename TI, int D>
__device__ TI extrapolated_ghost(const DevPatch* __restrict__ fdp, const int32_t* __restrict__ child,
                                 const int32_t* __restrict__ nchild, const int32_t* __restrict__ orth,
                                 const TI* __restrict__ fine, const Layout& L, int p, int c, int X, int Y, int Z,
                                 int rem)
{
  // 3 g - i rounded after the product, as the reference's expression (no fused multiply-add)
#pragma clang fp contract(off)
  const int cx[3] = { X, Y, Z };
  int outn = 0, bits = 0, axis = 0;
#pragma unroll
  for (int a = 0; a < D; a++) {
    if (cx[a] < 0 || cx[a] >= L.n[a]) {
      outn++;
      axis = a;
      if (cx[a] >= L.n[a]) bits |= 1 << a;
    }
  }
  TI acc = TI(0);
  const int nc = nchild[p];
  for (int k = 0; k < nc; k++) {
    const int f = child[p * 8 + k];
    const int o = orth[f];
    if (o < 0) {  // copyToParent: every cell, ghosts included
      acc += fine[f * L.ps + rem];
      continue;
    }
    if (outn == 1) {  // a side ghost: the child whose exterior side and half it lies over
      const int up = (bits >> axis) & 1;
      if (((o >> axis) & 1) != up) continue;
      int lo[2] = { 0, 0 }, hi[2] = { 0, 0 }, fa[2] = { 0, 0 }, nf = 0;
      bool mine = true;
#pragma unroll
      for (int b = 0; b < D; b++) {
        if (b == axis) continue;
        const int st = ((o >> b) & 1) ? L.n[b] : 0;
        int t0 = 2 * cx[b] - st, t1 = t0 + 1;  // the fine cells with (t + st) / 2 == cx[b]
        if (t0 < 0) t0 = 0;
        if (t1 > L.n[b] - 1) t1 = L.n[b] - 1;
        if (t0 > t1) mine = false;
        fa[nf] = b;
        lo[nf] = t0;
        hi[nf] = t1;
        nf++;
      }
      if (!mine) continue;
      const int ga = up ? L.n[axis] : -1, ia = up ? L.n[axis] - 1 : 0;
      // the side's slice, first free axis fastest
      for (int j = (D == 3 ? lo[1] : 0); j <= (D == 3 ? hi[1] : 0); j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          int xg[3] = { 0, 0, 0 }, xi[3] = { 0, 0, 0 };
          xg[axis] = ga;
          xi[axis] = ia;
          xg[fa[0]] = xi[fa[0]] = i;
          if (D == 3) xg[fa[1]] = xi[fa[1]] = j;
          const TI g = fine[lin(L, f, c, xg[0], xg[1], xg[2])], in = fine[lin(L, f, c, xi[0], xi[1], xi[2])];
          acc += (TI(3) * g - in) / TI(1 << D);
        }
    } else if (outn == D) {  // the corner of the child's orthant, when nothing lies across it
      if (o != bits || fdp[f].type[bits] >= 0) continue;
      int xg[3] = { 0, 0, 0 }, xi[3] = { 0, 0, 0 };
#pragma unroll
      for (int a = 0; a < D; a++) {
        xg[a] = ((bits >> a) & 1) ? L.n[a] : -1;
        xi[a] = ((bits >> a) & 1) ? L.n[a] - 1 : 0;
      }
      acc += TI(3) * fine[lin(L, f, c, xg[0], xg[1], xg[2])] - fine[lin(L, f, c, xi[0], xi[1], xi[2])];
    }
  }
  return acc;
}

template <typename TI, typename TO, int D, bool EX>
__global__ void __launch_bounds__(256) restrict_kernel(const int32_t* __restrict__ child,
                                                       const int32_t* __restrict__ nchild,
                                                       const int32_t* __restrict__ orth, const TI* __restrict__ fine,
                                                       TO* __restrict__ coarse, Layout L, int np_coarse,
                                                       const DevPatch* __restrict__ fdp)
{
  long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)np_coarse * L.ps;
  if (gid >= total) return;
  int idx = (int)gid;
  int p = idx / L.ps;
  int rem = idx - p * L.ps;
  int c = rem / L.cs;
  int q = rem - c * L.cs;
  int X, Y, Z = 0;
  if (D == 2) {
    Y = q / L.sx;
    X = q - Y * L.sx;
  } else {
    Z = q / L.sy;
    int r2 = q - Z * L.sy;
    Y = r2 / L.sx;
    X = r2 - Y * L.sx;
    Z -= 1;
  }
  X -= 1;
  Y -= 1;
  bool inter = X >= 0 && X < L.n[0] && Y >= 0 && Y < L.n[1] && (D == 2 || (Z >= 0 && Z < L.n[2]));
  if (!inter) {
    coarse[idx] = EX ? TO(extrapolated_ghost<TI, D>(fdp, child, nchild, orth, fine, L, p, c, X, Y, Z, rem)) : TO(0);
    return;
  }
  int nc = nchild[p];
  TI acc = TI(0);
  const TI w = TI(1.0 / (1 << D));
  for (int k = 0; k < nc; k++) {
    int f = child[p * 8 + k];
    int o = orth[f];
    if (o < 0) {  // copyToParent
      acc += fine[lin(L, f, c, X, Y, Z)];
      continue;
    }
    int s0 = (o & 1) ? L.n[0] : 0, s1 = (o & 2) ? L.n[1] : 0, s2 = (o & 4) ? L.n[2] : 0;
    int z0 = D == 3 ? 2 * Z - s2 : 0, z1 = D == 3 ? z0 + 1 : 0;
    for (int zz = z0; zz <= z1; zz++) {
      if (D == 3 && (zz < 0 || zz >= L.n[2])) continue;
      for (int yy = 2 * Y - s1; yy <= 2 * Y - s1 + 1; yy++) {
        if (yy < 0 || yy >= L.n[1]) continue;
        for (int xx = 2 * X - s0; xx <= 2 * X - s0 + 1; xx++) {
          if (xx < 0 || xx >= L.n[0]) continue;
          acc += fine[lin(L, f, c, xx, yy, zz)] * w;
        }
      }
    }
  }
  coarse[idx] = TO(acc);
}

template <typename TC, typename TF, int D>
__global__ void __launch_bounds__(256) interp_kernel(const int32_t* __restrict__ parent,
                                                     const int32_t* __restrict__ orth, const TC* __restrict__ coarse,
                                                     TF* __restrict__ fine, Layout L)
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
  if (!inter) return;
  int cp = parent[p];
  int o = orth[p];
  int ci;
  if (o < 0) {
    ci = lin(L, cp, c, x, y, z);
  } else {
    int sx0 = (o & 1) ? L.n[0] : 0, sy0 = (o & 2) ? L.n[1] : 0, sz0 = (o & 4) ? L.n[2] : 0;
    ci = lin(L, cp, c, (x + sx0) / 2, (y + sy0) / 2, D == 3 ? (z + sz0) / 2 : 0);
  }
  fine[idx] = TF(fine[idx] + TF(coarse[ci]));
}

template <typename TI, typename TO>
void restrict_t(const te_level_link* l, const TI* fine, TO* coarse, int ncomp, bool ex, hipStream_t s)
{
  const te_domain* c = l->coarse;
  Layout L = make_layout(c->dim, c->ns, ncomp, l->fine->np);
  long total = (long)c->np * L.ps;
  if (total == 0) return;
  const int B = 256;
  int grid = (int)((total + B - 1) / B);
#define TE_RK(D_, EX_)                                                                                             \
  hipLaunchKernelGGL((restrict_kernel<TI, TO, D_, EX_>), dim3(grid), dim3(B), 0, s, l->d_child, l->d_nchild,        \
                     l->d_orth, fine, coarse, L, c->np, l->fine->d_patches)
  if (c->dim == 2) {
    if (ex) TE_RK(2, true);
    else TE_RK(2, false);
  } else {
    if (ex) TE_RK(3, true);
    else TE_RK(3, false);
  }
#undef TE_RK
}

template <typename TC, typename TF>
void interp_t(const te_level_link* l, const TC* coarse, TF* fine, int ncomp, hipStream_t s)
{
  const te_domain* f = l->fine;
  Layout L = make_layout(f->dim, f->ns, ncomp, f->np);
  long total = (long)f->np * L.ps;
  if (total == 0) return;
  const int B = 256;
  int grid = (int)((total + B - 1) / B);
  if (f->dim == 2)
    hipLaunchKernelGGL((interp_kernel<TC, TF, 2>), dim3(grid), dim3(B), 0, s, l->d_parent, l->d_orth, coarse, fine,
                       L);
  else
    hipLaunchKernelGGL((interp_kernel<TC, TF, 3>), dim3(grid), dim3(B), 0, s, l->d_parent, l->d_orth, coarse, fine,
                       L);
}

}  // namespace

int launch_restrict(const te_level_link* l, int dti, const void* fine, int dto, void* coarse, int ncomp,
                    hipStream_t s, bool extrapolate)
{
  const bool ex = extrapolate;
  if (dti == TE_F64 && dto == TE_F64) restrict_t<double, double>(l, (const double*)fine, (double*)coarse, ncomp, ex, s);
  else if (dti == TE_F32 && dto == TE_F32) restrict_t<float, float>(l, (const float*)fine, (float*)coarse, ncomp, ex, s);
  else if (dti == TE_F32 && dto == TE_F64) restrict_t<float, double>(l, (const float*)fine, (double*)coarse, ncomp, ex, s);
  else restrict_t<double, float>(l, (const double*)fine, (float*)coarse, ncomp, ex, s);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_interpolate(const te_level_link* l, int dtc, const void* coarse, int dtf, void* fine, int ncomp,
                       hipStream_t s)
{
  if (dtc == TE_F64 && dtf == TE_F64) interp_t<double, double>(l, (const double*)coarse, (double*)fine, ncomp, s);
  else if (dtc == TE_F32 && dtf == TE_F32) interp_t<float, float>(l, (const float*)coarse, (float*)fine, ncomp, s);
  else if (dtc == TE_F64 && dtf == TE_F32) interp_t<double, float>(l, (const double*)coarse, (float*)fine, ncomp, s);
  else interp_t<float, double>(l, (const float*)coarse, (double*)fine, ncomp, s);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
