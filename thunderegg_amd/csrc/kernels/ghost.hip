// Ghost filling as a race-free gather ("pull"): every ghost cell facing a
// neighbour is computed by exactly one thread from the closed-form weights of
// the reference's push-style fillers (MPIGhostFiller.h:1055-1092 with
// BiLinearGhostFiller.cpp, BiQuadraticGhostFiller.cpp, TriLinearGhostFiller.cpp).
// Terms are summed in the order the reference accumulates them: self terms
// first (fillGhostCellsForLocalPatch runs before processLocalFills), then the
// neighbour contributions in the order their calls are enumerated.
#include "../te_internal.h"
#include "ghost2d.h"

namespace te {
namespace {

constexpr int NB_NORMAL = 0, NB_COARSE = 1;

// cell of the slice on 2D side s at depth k (k = -1 is the ghost layer), tangential t
__device__ __forceinline__ int side_cell2(const Layout& L, int p, int c, int s, int k, int t)
{
  int a = s >> 1, up = s & 1;
  int ca = up ? L.n[a] - 1 - k : k;
  return a == 0 ? lin(L, p, c, ca, t, 0) : lin(L, p, c, t, ca, 0);
}

// 3D side s: free axes are the two other axes in increasing order (i, j).
__device__ __forceinline__ int side_cell3(const Layout& L, int p, int c, int s, int k, int i, int j)
{
  int a = s >> 1, up = s & 1;
  int ca = up ? L.n[a] - 1 - k : k;
  if (a == 0) return lin(L, p, c, ca, i, j);
  if (a == 1) return lin(L, p, c, i, ca, j);
  return lin(L, p, c, i, j, ca);
}

// 2D side ghosts: one wavefront per patch walks the patch's ring of side ghost cells
// (ring_cell order: the two x faces, then the two y faces, so the y faces' ghost rows are
// written by consecutive lanes).  One 16-byte ring-table record per cell names the
// neighbour its formula reads; the closed forms are ghost2d.h's (self terms, then the
// neighbour terms, in the reference's order).  Patches are split over the XCDs
// (xcd_range), so a patch and most of its neighbours are read through one L2.
template <typename T, int KIND>
__global__ void __launch_bounds__(256) ghost_ring_2d(const int4* __restrict__ ring, T* __restrict__ u, Layout L)
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int n0 = L.n[0], n1 = L.n[1], nr = 2 * (n0 + n1);
  const XcdRange xr = xcd_range(blockIdx.x, gridDim.x, L.np, 4);
  for (int p = xr.first + wave; p < xr.end; p += xr.step) {
    for (int g = lane; g < nr; g += 64) {
      int s, t;
      const int4 r = ring[ring_index(p, g, n0, n1, &s, &t)];
      if (r.x < 0) continue;  // a physical side: the operators apply its condition themselves
      const int a = s >> 1, up = s & 1;
      const int na = a == 0 ? n0 : n1, nt = a == 0 ? n1 : n0;
      int bx, by;
      side_xy(s, 0, t, na, &bx, &by);
      const int gx = a == 0 ? (up ? n0 : -1) : bx, gy = a == 0 ? by : (up ? n1 : -1);
      for (int c = 0; c < L.ncomp; c++) {
        T* uc = u + c * L.cs;
        auto U = [&](int q, int x, int y) { return at(uc, q * L.ps + (y + 1) * L.sx + x + 1); };
        T v[4] = { T(0), T(0), T(0), T(0) };
        ghost2d_gather<KIND>(r.x, r.y, r.z, s, t, na, nt, [&](int k, int q, int x, int y) { v[k] = U(q, x, y); });
        const T G = ghost2d_combine<T, KIND>(r.x, r.y, t, nt,
                                             [&](int k, int tt) {
                                               int x, y;
                                               side_xy(s, k, tt, na, &x, &y);
                                               return U(p, x, y);
                                             },
                                             v);
        at(uc, p * L.ps + (gy + 1) * L.sx + gx + 1) = G;
      }
    }
  }
}

// 2D corners (fill type Corners).  BiLinearGhostFiller.cpp:167-264,
// BiQuadraticGhostFiller.cpp:239-334.
template <typename T, int KIND>
__global__ void ghost_corners_2d(const DevPatch* __restrict__ dp, T* __restrict__ u, Layout L)
{
  int gid = blockIdx.x * blockDim.x + threadIdx.x;
  int p = gid >> 2;
  if (p >= L.np) return;
  int cr = gid & 3;
  int type = dp[p].type[cr];
  if (type < 0) return;
  int q = dp[p].nbr[cr][0];
  int ux = cr & 1, uy = (cr >> 1) & 1;
  int nx = L.n[0], ny = L.n[1];
  for (int c = 0; c < L.ncomp; c++) {
    // interior corner cell of patch p at corner cr, and its (1,1) inner cell
    T P0 = u[lin(L, p, c, ux ? nx - 1 : 0, uy ? ny - 1 : 0, 0)];
    T P1 = u[lin(L, p, c, ux ? nx - 2 : 1, uy ? ny - 2 : 1, 0)];
    // neighbour's interior cell at the opposite corner
    T Q0 = u[lin(L, q, c, ux ? 0 : nx - 1, uy ? 0 : ny - 1, 0)];
    T Q1 = u[lin(L, q, c, ux ? 1 : nx - 2, uy ? 1 : ny - 2, 0)];
    T G;
    if (type == NB_NORMAL) {
      G = Q0;
    } else if (type == NB_COARSE) {
      if (KIND == TE_BILINEAR) G = P0 / T(3) + T(2) * Q0 / T(3);
      else G = (T(2) * P0 / T(3) - P1 / T(5)) + T(8) * Q0 / T(15);
    } else {
      if (KIND == TE_BILINEAR) G = -P0 / T(3) + T(4) * Q0 / T(3);
      else G = -P0 / T(15) + (T(2) * Q0 / T(3) + T(2) * Q1 / T(5));
    }
    u[lin(L, p, c, ux ? nx : -1, uy ? ny : -1, 0)] = G;
  }
}

// 3D faces, TriLinearGhostFiller.cpp:33-175.
template <typename T>
__global__ void __launch_bounds__(256) ghost_faces_3d(const DevPatch* __restrict__ dp, T* __restrict__ u, Layout L,
                                                      int tmax)
{
  int gid = blockIdx.x * blockDim.x + threadIdx.x;
  int per = 6 * tmax * tmax;
  int p = gid / per;
  if (p >= L.np) return;
  int r = gid - p * per;
  int s = r / (tmax * tmax);
  r -= s * tmax * tmax;
  int j = r / tmax;
  int i = r - j * tmax;
  int a = s >> 1;
  int f0 = a == 0 ? 1 : 0, f1 = a == 2 ? 1 : 2;
  int ni = L.n[f0], nj = L.n[f1];
  if (i >= ni || j >= nj) return;
  int slot = 20 + s;
  int type = dp[p].type[slot];
  if (type < 0) return;
  for (int c = 0; c < L.ncomp; c++) {
    T G;
    if (type == NB_NORMAL) {
      G = u[side_cell3(L, dp[p].nbr[slot][0], c, s ^ 1, 0, i, j)];
    } else if (type == NB_COARSE) {
      int q = dp[p].nbr[slot][0];
      int o = dp[p].orth[slot];
      int oi = (o & 1) ? ni : 0, oj = (o & 2) ? nj : 0;
      int pi = ((i + oi) % 2 == 0) ? i + 1 : i - 1;
      int pj = ((j + oj) % 2 == 0) ? j + 1 : j - 1;
      G = T(0);
      G += T(5.0 / 6.0) * u[side_cell3(L, p, c, s, 0, i, j)];
      G -= T(1.0 / 6.0) * u[side_cell3(L, p, c, s, 0, pi, j)];
      G -= T(1.0 / 6.0) * u[side_cell3(L, p, c, s, 0, i, pj)];
      G -= T(1.0 / 6.0) * u[side_cell3(L, p, c, s, 0, pi, pj)];
      G += T(4.0 / 6.0) * u[side_cell3(L, q, c, s ^ 1, 0, (i + oi) / 2, (j + oj) / 2)];
    } else {
      G = T(0);
      G -= T(1.0 / 3.0) * u[side_cell3(L, p, c, s, 0, i, j)];
      for (int o = 0; o < 4; o++) {
        int q = dp[p].nbr[slot][o];
        int oi = (o & 1) ? ni : 0, oj = (o & 2) ? nj : 0;
        for (int fj = 2 * j - oj; fj <= 2 * j - oj + 1; fj++) {
          if (fj < 0 || fj >= nj) continue;
          for (int fi = 2 * i - oi; fi <= 2 * i - oi + 1; fi++) {
            if (fi < 0 || fi >= ni) continue;
            G += T(1.0 / 3.0) * u[side_cell3(L, q, c, s ^ 1, 0, fi, fj)];
          }
        }
      }
    }
    u[side_cell3(L, p, c, s, -1, i, j)] = G;
  }
}

// 3D edge e (Face<3,1>): free axis = e >> 2; fixed sides in axis order with
// bit k of e = upper on the k-th fixed axis (Face.h:504-515).
__device__ __forceinline__ int edge_cell3(const Layout& L, int p, int c, int e, int k, int i)
{
  int fa = e >> 2;
  int x[3];
  int kk = 0;
  for (int a = 0; a < 3; a++) {
    if (a == fa) {
      x[a] = i;
    } else {
      int up = (e >> kk) & 1;
      x[a] = up ? L.n[a] - 1 - k : k;
      kk++;
    }
  }
  return lin(L, p, c, x[0], x[1], x[2]);
}

// 3D edges and corners (fill types Edges / Corners), TriLinearGhostFiller.cpp:188-405.
template <typename T>
__global__ void ghost_lowdim_3d(const DevPatch* __restrict__ dp, T* __restrict__ u, Layout L, int nmax,
                                int with_corners)
{
  int gid = blockIdx.x * blockDim.x + threadIdx.x;
  int per = 12 * nmax + 8;
  int p = gid / per;
  if (p >= L.np) return;
  int r = gid - p * per;
  if (r < 12 * nmax) {
    int e = r / nmax;
    int i = r - e * nmax;
    int n = L.n[e >> 2];
    if (i >= n) return;
    int slot = 8 + e;
    int type = dp[p].type[slot];
    if (type < 0) return;
    int eo = e ^ 3;  // opposite edge
    for (int c = 0; c < L.ncomp; c++) {
      T G;
      if (type == NB_NORMAL) {
        G = u[edge_cell3(L, dp[p].nbr[slot][0], c, eo, 0, i)];
      } else if (type == NB_COARSE) {
        int q = dp[p].nbr[slot][0];
        int off = dp[p].orth[slot] == 1 ? n : 0;
        G = T(0);
        for (int tp = i - 1; tp <= i + 1; tp++) {
          if (tp < 0 || tp >= n) continue;
          T sv = u[edge_cell3(L, p, c, e, 0, tp)];
          if (tp == i) {
            G += T(2.0 / 3.0) * sv;
          } else {
            int partner = ((tp + off) % 2 == 0) ? tp + 1 : tp - 1;
            if (partner == i) G += T(-1.0 / 3.0) * sv;
          }
        }
        G += T(2.0 / 3.0) * u[edge_cell3(L, q, c, eo, 0, (i + off) / 2)];
      } else {
        G = T(-1.0 / 3.0) * u[edge_cell3(L, p, c, e, 0, i)];
        for (int k = 0; k < 2; k++) {
          int q = dp[p].nbr[slot][k];
          int off = k * n;
          for (int fi = 2 * i - off; fi <= 2 * i - off + 1; fi++) {
            if (fi < 0 || fi >= n) continue;
            G += T(2.0 / 3.0) * u[edge_cell3(L, q, c, eo, 0, fi)];
          }
        }
      }
      u[edge_cell3(L, p, c, e, -1, i)] = G;
    }
  } else {
    if (!with_corners) return;
    int cr = r - 12 * nmax;
    int type = dp[p].type[cr];
    if (type < 0) return;
    int q = dp[p].nbr[cr][0];
    int ux = cr & 1, uy = (cr >> 1) & 1, uz = (cr >> 2) & 1;
    int nx = L.n[0], ny = L.n[1], nz = L.n[2];
    for (int c = 0; c < L.ncomp; c++) {
      T P0 = u[lin(L, p, c, ux ? nx - 1 : 0, uy ? ny - 1 : 0, uz ? nz - 1 : 0)];
      T Q0 = u[lin(L, q, c, ux ? 0 : nx - 1, uy ? 0 : ny - 1, uz ? 0 : nz - 1)];
      T G;
      if (type == NB_NORMAL) G = Q0;
      else if (type == NB_COARSE) G = P0 / T(3) + T(2) * Q0 / T(3);
      else G = -P0 / T(3) + T(4) * Q0 / T(3);
      u[lin(L, p, c, ux ? nx : -1, uy ? ny : -1, uz ? nz : -1)] = G;
    }
  }
}

template <typename T>
int ghost_fill_t(const te_ghost_filler* g, T* u, int ncomp, hipStream_t s)
{
  const te_domain* d = g->d;
  if (d->np == 0) return TE_OK;
  Layout L = make_layout(d->dim, d->ns, ncomp, d->np);
  const int B = 256;
  if (d->dim == 2) {
    // a wave per patch (4 per workgroup), at most 8192 workgroups walking the rest
    const int ngroups = (d->np + 3) / 4;
    const int grid = ngroups < 8192 ? ngroups : 8192;
    if (g->kind == TE_BILINEAR)
      hipLaunchKernelGGL((ghost_ring_2d<T, TE_BILINEAR>), dim3(grid), dim3(B), 0, s, (const int4*)d->d_ring, u, L);
    else
      hipLaunchKernelGGL((ghost_ring_2d<T, TE_BIQUADRATIC>), dim3(grid), dim3(B), 0, s, (const int4*)d->d_ring, u,
                         L);
    if (g->fill_type == TE_CORNERS) {
      int cg = (d->np * 4 + B - 1) / B;
      if (g->kind == TE_BILINEAR)
        hipLaunchKernelGGL((ghost_corners_2d<T, TE_BILINEAR>), dim3(cg), dim3(B), 0, s, d->d_patches, u, L);
      else
        hipLaunchKernelGGL((ghost_corners_2d<T, TE_BIQUADRATIC>), dim3(cg), dim3(B), 0, s, d->d_patches, u, L);
    }
  } else {
    int tmax = d->ns[0];
    for (int a = 1; a < 3; a++) tmax = d->ns[a] > tmax ? d->ns[a] : tmax;
    long total = (long)d->np * 6 * tmax * tmax;
    int grid = (int)((total + B - 1) / B);
    hipLaunchKernelGGL((ghost_faces_3d<T>), dim3(grid), dim3(B), 0, s, d->d_patches, u, L, tmax);
    if (g->fill_type != TE_FACES) {
      long tl = (long)d->np * (12 * tmax + 8);
      int lg = (int)((tl + B - 1) / B);
      hipLaunchKernelGGL((ghost_lowdim_3d<T>), dim3(lg), dim3(B), 0, s, d->d_patches, u, L, tmax,
                         g->fill_type == TE_CORNERS ? 1 : 0);
    }
  }
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace

int launch_ghost_fill(const te_ghost_filler* g, int dtype, void* u, int ncomp, hipStream_t s)
{
  if (dtype == TE_F64) return ghost_fill_t<double>(g, (double*)u, ncomp, s);
  return ghost_fill_t<float>(g, (float*)u, ncomp, s);
}

}  // namespace te
