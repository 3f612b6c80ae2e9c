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

}  // namespace te
