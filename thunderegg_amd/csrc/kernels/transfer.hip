// Inter-level transfer.  Restriction is computed per coarse cell (a gather, so
// every coarse cell is written exactly once and no atomics are needed);
// interpolation is computed per fine cell.
//  - LinearRestrictor (GMG/LinearRestrictor.h:101-155): coarse[(i+s)/2] += fine[i]/2^D
//    for children with an orthant on the parent, coarse[i] += fine[i] for copies;
//    MPIRestrictor::restrict zeroes the coarse vector with ghosts first (:73).
//  - DirectInterpolator (GMG/DirectInterpolator.h:60-91): fine[i] += coarse[(i+s)/2].
// Children of a coarse patch are visited in increasing fine local index and fine
// cells in x-fastest order, the order in which the reference accumulates them.
#include "../te_internal.h"

namespace te {
namespace {

template <typename TI, typename TO, int D>
__global__ void __launch_bounds__(256) restrict_kernel(const int32_t* __restrict__ child,
                                                       const int32_t* __restrict__ nchild,
                                                       const int32_t* __restrict__ orth, const TI* __restrict__ fine,
                                                       TO* __restrict__ coarse, Layout L, int np_coarse)
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
    coarse[idx] = TO(0);
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
void restrict_t(const te_level_link* l, const TI* fine, TO* coarse, int ncomp, hipStream_t s)
{
  const te_domain* c = l->coarse;
  Layout L = make_layout(c->dim, c->ns, ncomp, l->fine->np);
  long total = (long)c->np * L.ps;
  if (total == 0) return;
  const int B = 256;
  int grid = (int)((total + B - 1) / B);
  if (c->dim == 2)
    hipLaunchKernelGGL((restrict_kernel<TI, TO, 2>), dim3(grid), dim3(B), 0, s, l->d_child, l->d_nchild, l->d_orth,
                       fine, coarse, L, c->np);
  else
    hipLaunchKernelGGL((restrict_kernel<TI, TO, 3>), dim3(grid), dim3(B), 0, s, l->d_child, l->d_nchild, l->d_orth,
                       fine, coarse, L, c->np);
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
                    hipStream_t s)
{
  if (dti == TE_F64 && dto == TE_F64) restrict_t<double, double>(l, (const double*)fine, (double*)coarse, ncomp, s);
  else if (dti == TE_F32 && dto == TE_F32) restrict_t<float, float>(l, (const float*)fine, (float*)coarse, ncomp, s);
  else if (dti == TE_F32 && dto == TE_F64) restrict_t<float, double>(l, (const float*)fine, (double*)coarse, ncomp, s);
  else restrict_t<double, float>(l, (const double*)fine, (float*)coarse, ncomp, s);
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
