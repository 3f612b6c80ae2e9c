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

// What a ghost formula needs to know about one side of a patch (loaded ahead of use).
struct SideInfo {
  int type;  // -1 physical boundary, 0 same level, 1 coarser, 2 finer neighbours
  int orth;  // for a coarser neighbour: which half of its side this patch touches
  int nbr0, nbr1;
};

__device__ __forceinline__ SideInfo side_info(const DevPatch& P, int s)
{
  const int slot = 4 + s;
  return SideInfo{ P.type[slot], P.orth[slot], P.nbr[slot][0], P.nbr[slot][1] };
}

// One ghost value is a sum of self terms (interior cells of the patch itself) and
// neighbour terms.  ghost2d_gather names the neighbour cells, in a fixed order, through
// ld(k, q, x, y) (k: position in the order, q: neighbour patch, (x, y): its cell);
// ghost2d_combine forms the value from the self cells S(k, tt) (depth k from side s,
// tangential index tt) and the gathered neighbour values v[k].  Split this way, a caller
// can issue the neighbour loads of a later patch early and combine them when it gets there.
// q: the neighbour the ghost reads (for finer neighbours: nbr1 when 2 t >= nt, else nbr0).
// n: cells along the side's axis (the depth direction); nt: cells along the side (even).
// The coarser-neighbour BiQuadratic stencil: three cells of the coarse side i0, i0 + di,
// i0 + 2 di and which of the reference's four weight patterns applies.
__device__ __forceinline__ void coarse_quad_cells(int o, int t, int nt, int* i0, int* di, int* cs)
{
  const int e = o == 0 ? t : nt - 1 - t;  // distance from the end of the coarse side
  if (e <= 1) {
    *i0 = o == 0 ? 0 : nt - 1;
    *di = o == 0 ? 1 : -1;
    *cs = e;
  } else {
    const int tt = o == 0 ? t : t + nt;
    *i0 = tt / 2 - 1;
    *di = 1;
    *cs = 2 + (tt & 1);
  }
}

template <int KIND, class Ld>
__device__ __forceinline__ void ghost2d_gather(int type, int orth, int q, int s, int t, int n, int nt, Ld ld)
{
  int x, y;
  const int so = s ^ 1;
  if (type == 0) {
    side_xy(so, 0, t, n, &x, &y);
    ld(0, q, x, y);
  } else if (type == 1) {
    if (KIND == TE_BILINEAR) {
      side_xy(so, 0, (t + (orth ? nt : 0)) / 2, n, &x, &y);
      ld(0, q, x, y);
    } else {
      int i0, di, cs;
      coarse_quad_cells(orth, t, nt, &i0, &di, &cs);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        side_xy(so, 0, i0 + k * di, n, &x, &y);
        ld(k, q, x, y);
      }
    }
  } else if (type == 2) {
    const int tf = 2 * t - (2 * t >= nt ? nt : 0);
#pragma unroll
    for (int j = 0; j < 2; j++) {
      if (KIND == TE_BILINEAR) {
        side_xy(so, 0, tf + j, n, &x, &y);
        ld(j, q, x, y);
      } else {
        side_xy(so, 0, tf + j, n, &x, &y);
        ld(2 * j, q, x, y);
        side_xy(so, 1, tf + j, n, &x, &y);
        ld(2 * j + 1, q, x, y);
      }
    }
  }
}

// x / c for the constant divisors of the ghost formulas: a division in FP64 (as the
// reference), a product with the rounded reciprocal in FP32, where the cycle is a
// preconditioner and the two differ by an ulp.
template <typename T>
__device__ __forceinline__ T divc(T x, int c)
{
  return sizeof(T) == 8 ? x / T(c) : x * T(1.0 / c);
}

template <typename T, int KIND, class Self>
__device__ __forceinline__ T ghost2d_combine(int type, int orth, int t, int nt, Self S, const T (&v)[4])
{
  if (type == 0) return v[0];
  if (type == 1) {
    if (KIND == TE_BILINEAR) {
      const int off = orth ? nt : 0;
      T G = T(0);
      for (int tp = t - 1; tp <= t + 1; tp++) {
        if (tp < 0 || tp >= nt) continue;
        T sv = S(0, tp);
        if (tp == t) {
          G += T(2.0 / 3.0) * sv;
        } else {
          int partner = ((tp + off) % 2 == 0) ? tp + 1 : tp - 1;
          if (partner == t) G += T(-1.0 / 3.0) * sv;
        }
      }
      return G + T(2.0 / 3.0) * v[0];
    }
    T G = divc(T(2) * S(0, t), 3) - divc(S(1, t), 5);
    int i0, di, cs;
    coarse_quad_cells(orth, t, nt, &i0, &di, &cs);
    const T c0 = v[0], c1 = v[1], c2 = v[2];
    T add;
    if (cs == 0) add = divc(T(3) * c0, 4) - divc(T(3) * c1, 10) + divc(c2, 12);
    else if (cs == 1) add = divc(T(7) * c0, 20) + divc(T(7) * c1, 30) - divc(c2, 20);
    else if (cs == 2) add = divc(c0, 12) + divc(c1, 2) - divc(c2, 20);
    else add = divc(-c0, 20) + divc(c1, 2) + divc(c2, 12);
    return G + add;
  }
  // finer neighbours (ghost t lies against fine cells 2t, 2t + 1 of neighbour 0 when
  // 2t < nt, else of neighbour 1)
  T G;
  if (KIND == TE_BILINEAR) {
    G = T(-1.0 / 3.0) * S(0, t);
  } else {
    if (t == 0 || t == nt - 1) {
      const int i0 = t == 0 ? 0 : nt - 1, di = t == 0 ? 1 : -1;
      G = divc(-S(0, i0), 10) + divc(S(0, i0 + di), 15) - divc(S(0, i0 + 2 * di), 30);
    } else {
      G = divc(-S(0, t - 1), 30) - divc(S(0, t + 1), 30);
    }
  }
#pragma unroll
  for (int j = 0; j < 2; j++) {
    if (KIND == TE_BILINEAR) G += T(2.0 / 3.0) * v[j];
    else G += divc(v[2 * j], 3) + divc(v[2 * j + 1], 5);
  }
  return G;
}

// the neighbour a ghost cell reads
__device__ __forceinline__ int ghost_nbr(const SideInfo& si, int t, int nt)
{
  return si.type == 2 && 2 * t >= nt ? si.nbr1 : si.nbr0;
}

// The value of ghost cell t on side s of patch p; `U(q, x, y)` reads every cell.
template <typename T, int KIND, class Acc>
__device__ __forceinline__ T ghost2d(const SideInfo& si, int p, int s, int t, int n, int nt, Acc U)
{
  T v[4] = { T(0), T(0), T(0), T(0) };
  ghost2d_gather<KIND>(si.type, si.orth, ghost_nbr(si, t, nt), s, t, n, nt,
                       [&](int k, int q, int x, int y) { v[k] = U(q, x, y); });
  return ghost2d_combine<T, KIND>(si.type, si.orth, t, nt,
                                  [&](int k, int tt) {
                                    int x, y;
                                    side_xy(s, k, tt, n, &x, &y);
                                    return U(p, x, y);
                                  },
                                  v);
}

template <typename T>
struct PlainAcc {  // interior value of patch q at (x, y)
  const T* u;
  Layout L;
  __device__ __forceinline__ T operator()(int q, int x, int y) const { return at(u, q * L.ps + (y + 1) * L.sx + x + 1); }
};

// (A u)(x, y) from the staged tile; ci: padded index of the cell in the coefficient vector
template <typename T, int OP>
__device__ __forceinline__ T star_tile(const T* tile, int w, int x, int y, const T* ih2, const T* cf, int ci,
                                       int sx)
{
  const int k = (y + 1) * w + x + 1;
  const T mid = tile[k];
  T acc = T(0);
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const int st = a == 0 ? 1 : w;
    const T lower = tile[k - st], upper = tile[k + st];
    if (OP == 0) {
      acc += (upper - T(2) * mid + lower) * ih2[a];
    } else {
      const int cst = a == 0 ? 1 : sx;
      const T cm = at(cf, ci), cl = at(cf, ci - cst), cu = at(cf, ci + cst);
      acc += ((cu + cm) * (upper - mid) - (cl + cm) * (mid - lower)) * (ih2[a] * T(0.5));
    }
  }
  return acc;
}

// XCD-aware work split.  Workgroups are dispatched round-robin over the 8 XCDs (block
// b runs on XCD b % 8), each with its own L2.  Giving XCD x the contiguous items
// [x N/8, (x+1) N/8) (patches in Morton order: mostly each other's neighbours), walked
// by that XCD's blocks, keeps the neighbour reads of the ghost formulas in one L2 and
// the XCDs equally loaded.  Block b covers items first, first + step, ... < end in
// units of `per` consecutive items (per = waves per block when each wave takes one).
// Grids smaller than 8 blocks fall back to plain striding.
struct XcdRange {
  int first, step, end;
};
__device__ __forceinline__ XcdRange xcd_range(int b, int G, int n, int per)
{
  if (G < 8) return XcdRange{ b * per, G * per, n };
  const int x = b & 7, i = b >> 3;
  const int gx = (G >> 3) + (x < (G & 7) ? 1 : 0);  // blocks on XCD x
  const int lo = (int)((long)x * n / 8), hi = (int)((long)(x + 1) * n / 8);
  return XcdRange{ lo + i * per, gx * per, hi };
}

// Orders the LDS accesses of one wavefront (every lane of the wave executes it): a
// wave's LDS operations are performed in issue order, so only the compiler has to be
// kept from moving memory operations across this point.  Replaces a workgroup barrier
// where each wave works on LDS of its own.
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

// Side information for ring cell lane + 64 j of a patch (j < GMAX), loaded ahead of the
// values it is combined with.
template <int GMAX>
__device__ __forceinline__ void ring_prefetch(const DevPatch& P, int n0, int n1, int lane, SideInfo (&si)[GMAX])
{
#pragma unroll
  for (int j = 0; j < GMAX; j++) {
    const int g = lane + 64 * j;
    int s = 0, t;
    if (g < 2 * (n0 + n1)) ring_cell(g, n0, n1, &s, &t);
    si[j] = side_info(P, s);
  }
}

// Side ghosts of patch p (after its interior is in the tile and visible to the wave):
// a lane per ghost cell, neighbour sides by ghost2d, physical sides as
// phys * (adjacent interior value) (the Dirichlet / Neumann image, phys = -1 / +1).
// si0: side information of the lane's first ring cell (ring_prefetch<1>); later ones are
// loaded here.  The loop is kept rolled so that one copy of the ghost formulas is live.
template <typename T, int KIND, class Acc>
__device__ __forceinline__ void tile_ghosts_2d(const DevPatch& P, const SideInfo& si0, int p, int n0, int n1, Acc U,
                                               T* tile, T phys, int lane)
{
  const int w = n0 + 2;
  TileAcc<T, Acc> A{ tile, p, w, U };
#pragma unroll 1
  for (int g = lane; g < 2 * (n0 + n1); g += 64) {
    int s, t;
    ring_cell(g, n0, n1, &s, &t);
    const SideInfo si = g == lane ? si0 : side_info(P, s);
    const int a = s >> 1, up = s & 1;
    const int na = a == 0 ? n0 : n1, nt = a == 0 ? n1 : n0;
    int bx, by;
    side_xy(s, 0, t, na, &bx, &by);
    const int d = up ? 1 : -1;
    const int gx = a == 0 ? bx + d : bx, gy = a == 0 ? by : by + d;
    T v;
    if (si.type < 0) v = phys * tile[(by + 1) * w + bx + 1];
    else v = ghost2d<T, KIND>(si, p, s, t, na, nt, A);
    tile[(gy + 1) * w + gx + 1] = v;
  }
}

// ---- ghosts from the ring table, gathered ahead ----
// Ring cell g of an n0 x n1 patch: its record in the ring table (the int4 at the
// returned index) and the side / tangential index.
__device__ __forceinline__ int ring_index(int p, int g, int n0, int n1, int* s, int* t)
{
  ring_cell(g, n0, n1, s, t);
  const int nt = (*s >> 1) == 0 ? n1 : n0;
  return (p * 4 + *s) * 2 + (2 * *t >= nt ? 1 : 0);
}

// The neighbour values ring cell g of a patch needs (record r), read from u (widened to T).
template <typename T, int KIND, typename TU>
__device__ __forceinline__ void ring_gather_plain(const int4& r, int g, int n0, int n1, const TU* __restrict__ u,
                                                  const Layout& L, T (&v)[4])
{
  int s, t;
  ring_cell(g, n0, n1, &s, &t);
  const int na = (s >> 1) == 0 ? n0 : n1, nt = (s >> 1) == 0 ? n1 : n0;
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = T(0);
  ghost2d_gather<KIND>(r.x, r.y, r.z, s, t, na, nt,
                       [&](int k, int q, int x, int y) { v[k] = T(at(u, q * L.ps + (y + 1) * L.sx + x + 1)); });
}

// The ghost of ring cell g (record r, gathered neighbour values v) written into the tile
// of the patch, whose interior is in place: ghost2d_combine with the self terms from the
// tile, or phys * (adjacent interior value) on a physical side.
template <typename T, int KIND>
__device__ __forceinline__ void tile_ghost_gathered(const int4& r, int g, int n0, int n1, T* tile, T phys,
                                                    const T (&v)[4])
{
  const int w = n0 + 2;
  int s, t;
  ring_cell(g, n0, n1, &s, &t);
  const int a = s >> 1, up = s & 1;
  const int na = a == 0 ? n0 : n1, nt = a == 0 ? n1 : n0;
  int bx, by;
  side_xy(s, 0, t, na, &bx, &by);
  const int d = up ? 1 : -1;
  const int gx = a == 0 ? bx + d : bx, gy = a == 0 ? by : by + d;
  T val;
  if (r.x < 0) {
    val = phys * tile[(by + 1) * w + bx + 1];
  } else {
    val = ghost2d_combine<T, KIND>(r.x, r.y, t, nt,
                                   [&](int k, int tt) {
                                     int x, y;
                                     side_xy(s, k, tt, na, &x, &y);
                                     return tile[(y + 1) * w + x + 1];
                                   },
                                   v);
  }
  tile[(gy + 1) * w + gx + 1] = val;
}

}  // namespace te
