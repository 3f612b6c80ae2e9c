// Closed-form 2D ghost values (BiLinearGhostFiller.cpp, BiQuadraticGhostFiller.cpp)
// as a gather: the value of ghost cell t on side s of patch p, computed from the
// interior values of p and of its neighbours.  `U(q, x, y)` returns the interior
// value of patch q at cell (x, y); callers plug in plain loads, or loads plus an
// interpolation correction.  Terms are summed in the reference's accumulation
// order: self terms (fillGhostCellsForLocalPatch) first, then neighbour pushes.
#pragma once

#include "../te_internal.h"

namespace te {

// interior cell of patch-side slice: side s, depth k from the side, tangential index t
__device__ __forceinline__ void side_xy(int s, int k, int t, int n, int* x, int* y)
{
  int a = s >> 1, up = s & 1;
  int ca = up ? n - 1 - k : k;
  if (a == 0) {
    *x = ca;
    *y = t;
  } else {
    *x = t;
    *y = ca;
  }
}

// n: cells along the side's axis (the depth direction); nt: cells along the side
template <typename T, int KIND, class Acc>
__device__ __forceinline__ T ghost2d(const DevPatch& P, int p, int s, int t, int n, int nt, Acc U)
{
  const int slot = 4 + s;
  const int type = P.type[slot];
  int x, y;
  auto S = [&](int q, int side, int k, int tt) -> T {
    side_xy(side, k, tt, n, &x, &y);
    return U(q, x, y);
  };
  if (type == 0) return S(P.nbr[slot][0], s ^ 1, 0, t);
  if (type == 1) {
    const int q = P.nbr[slot][0];
    const int o = P.orth[slot];
    if (KIND == TE_BILINEAR) {
      const int off = o ? nt : 0;
      T G = T(0);
      for (int tp = t - 1; tp <= t + 1; tp++) {
        if (tp < 0 || tp >= nt) continue;
        T sv = S(p, s, 0, tp);
        if (tp == t) {
          G += T(2.0 / 3.0) * sv;
        } else {
          int partner = ((tp + off) % 2 == 0) ? tp + 1 : tp - 1;
          if (partner == t) G += T(-1.0 / 3.0) * sv;
        }
      }
      return G + T(2.0 / 3.0) * S(q, s ^ 1, 0, (t + off) / 2);
    }
    T G = T(2) * S(p, s, 0, t) / T(3) - S(p, s, 1, t) / T(5);
    T add;
#define TE_CC(k) S(q, s ^ 1, 0, (k))
    if (o == 0) {
      if (t == 0) add = T(3) * TE_CC(0) / T(4) - T(3) * TE_CC(1) / T(10) + TE_CC(2) / T(12);
      else if (t == 1) add = T(7) * TE_CC(0) / T(20) + T(7) * TE_CC(1) / T(30) - TE_CC(2) / T(20);
      else if (t % 2 == 0) add = TE_CC(t / 2 - 1) / T(12) + TE_CC(t / 2) / T(2) - TE_CC(t / 2 + 1) / T(20);
      else add = -TE_CC(t / 2 - 1) / T(20) + TE_CC(t / 2) / T(2) + TE_CC(t / 2 + 1) / T(12);
    } else {
      if (t == nt - 1) add = T(3) * TE_CC(nt - 1) / T(4) - T(3) * TE_CC(nt - 2) / T(10) + TE_CC(nt - 3) / T(12);
      else if (t == nt - 2) add = T(7) * TE_CC(nt - 1) / T(20) + T(7) * TE_CC(nt - 2) / T(30) - TE_CC(nt - 3) / T(20);
      else {
        int m = (t + nt) / 2;
        if ((t + nt) % 2 == 0) add = TE_CC(m - 1) / T(12) + TE_CC(m) / T(2) - TE_CC(m + 1) / T(20);
        else add = -TE_CC(m - 1) / T(20) + TE_CC(m) / T(2) + TE_CC(m + 1) / T(12);
      }
    }
#undef TE_CC
    return G + add;
  }
  // fine neighbours
  T G;
  if (KIND == TE_BILINEAR) {
    G = T(-1.0 / 3.0) * S(p, s, 0, t);
  } else {
    if (t == 0) G = -S(p, s, 0, 0) / T(10) + S(p, s, 0, 1) / T(15) - S(p, s, 0, 2) / T(30);
    else if (t == nt - 1) G = -S(p, s, 0, nt - 1) / T(10) + S(p, s, 0, nt - 2) / T(15) - S(p, s, 0, nt - 3) / T(30);
    else G = -S(p, s, 0, t - 1) / T(30) - S(p, s, 0, t + 1) / T(30);
  }
  for (int i = 0; i < 2; i++) {
    const int q = P.nbr[slot][i];
    const int off = i * nt;
    for (int tf = 2 * t - off; tf <= 2 * t - off + 1; tf++) {
      if (tf < 0 || tf >= nt) continue;
      if (KIND == TE_BILINEAR) G += T(2.0 / 3.0) * S(q, s ^ 1, 0, tf);
      else G += S(q, s ^ 1, 0, tf) / T(3) + S(q, s ^ 1, 1, tf) / T(5);
    }
  }
  return G;
}

// ---- per-wave staging of one patch into a padded LDS tile ----
// tile[(y + 1) * (n0 + 2) + x + 1] holds cell (x, y), x in [-1, n0], y in [-1, n1].

// Side ring cell g in [0, 2 (n0 + n1)): sides 0, 1 (x faces, n1 cells each), then 2, 3.
__device__ __forceinline__ void ring_cell(int g, int n0, int n1, int* s, int* t)
{
  if (g < 2 * n1) {
    *s = g >= n1 ? 1 : 0;
    *t = g - *s * n1;
  } else {
    g -= 2 * n1;
    *s = g >= n0 ? 3 : 2;
    *t = g - (*s - 2) * n0;
  }
}

// Interior reads of the staged patch p come from the tile, other patches from G.
template <typename T, class Acc>
struct TileAcc {
  const T* tile;
  int p;
  int w;
  Acc G;
  __device__ __forceinline__ T operator()(int q, int x, int y) const
  {
    return q == p ? tile[(y + 1) * w + x + 1] : G(q, x, y);
  }
};

// Interior of patch p, coalesced along x; one wave (64 lanes) per tile.
template <typename T, class Acc>
__device__ __forceinline__ void tile_interior_2d(int p, int n0, int n1, Acc U, T* tile, int lane)
{
  const int w = n0 + 2;
  for (int q = lane; q < n0 * n1; q += 64) {
    const int y = q / n0, x = q - y * n0;
    tile[(y + 1) * w + x + 1] = U(p, x, y);
  }
}

// Side ghosts of patch p (after its interior is in the tile and visible to the wave):
// a lane per ghost cell, neighbour sides by ghost2d, physical sides as
// phys * (adjacent interior value) (the Dirichlet / Neumann image, phys = -1 / +1).
template <typename T, int KIND, class Acc>
__device__ __forceinline__ void tile_ghosts_2d(const DevPatch& P, int p, int n0, int n1, Acc U, T* tile, T phys,
                                               int lane)
{
  const int w = n0 + 2;
  TileAcc<T, Acc> A{ tile, p, w, U };
  for (int g = lane; g < 2 * (n0 + n1); g += 64) {
    int s, t;
    ring_cell(g, n0, n1, &s, &t);
    const int a = s >> 1, up = s & 1;
    const int na = a == 0 ? n0 : n1, nt = a == 0 ? n1 : n0;
    int bx, by;
    side_xy(s, 0, t, na, &bx, &by);
    const int d = up ? 1 : -1;
    const int gx = a == 0 ? bx + d : bx, gy = a == 0 ? by : by + d;
    T v;
    if (P.type[4 + s] < 0) v = phys * tile[(by + 1) * w + bx + 1];
    else v = ghost2d<T, KIND>(P, p, s, t, na, nt, A);
    tile[(gy + 1) * w + gx + 1] = v;
  }
}

}  // namespace te
