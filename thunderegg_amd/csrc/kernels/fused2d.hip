// Fused 2D launches that compute ghost values inside the consumer instead of a
// separate ghost fill (MPIGhostFiller::fillGhost as a gather, GhostFiller.h):
//
//  - resid_restrict_2d: Cycle::restrict (GMG/Cycle.h:57-65) in one launch: r = f - A u
//    with the star stencil (Poisson/StarPatchOperator.h:75-101, VarPoisson :61-109),
//    restricted straight into the coarser level's right-hand side
//    (LinearRestrictor::restrictPatches, GMG/LinearRestrictor.h:101-155).  The
//    residual vector itself is never stored.
//  - star_dot_2d: the outer BiCGStab's A p / A s (Iterative/BiCGStab.h:150-170) with
//    the inner products that follow them folded in, FP64.
//
// A wavefront owns one patch at a time and stages it in a padded LDS tile: the
// interior with coalesced loads, then the side ghosts with one lane per ghost cell
// (tile_ghosts_2d), so the divergent ghost formulas and their dependent neighbour
// loads run once per patch instead of once per boundary cell.  The stencil then
// reads only the tile.
#include "../te_internal.h"
#include "blas.h"
#include "ghost2d.h"
#include "krylov.h"

namespace te {
namespace {

constexpr int kWaves = 4;

template <typename T, typename TF, typename TC, int OP, int KIND, bool P16>
__global__ void __launch_bounds__(64 * kWaves) resid_restrict_2d(const DevPatch* __restrict__ dp,
                                                                 const T* __restrict__ u, const T* __restrict__ cf,
                                                                 const TF* __restrict__ f, TC* __restrict__ fc,
                                                                 const int32_t* __restrict__ parent,
                                                                 const int32_t* __restrict__ porth, Layout L,
                                                                 int neumann, const int4* __restrict__ ring)
{
  if (P16) {  // compile-time layout constants: the address arithmetic folds
    L.n[0] = L.n[1] = 16;
    L.sx = 18;
    L.ps = L.cs = 18 * 18;
  }
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n0 = L.n[0], n1 = L.n[1], w = n0 + 2;
  const int h0 = n0 / 2, h1 = n1 / 2;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  T* tile = reinterpret_cast<T*>(smem_raw) + wave * w * (n1 + 2);
  PlainAcc<T> U{ u, L };
  const T phys = (OP == 0 && neumann) ? T(1) : T(-1);
  // each wave walks its own patches; its tile is private, so wave-level ordering suffices
  const XcdRange xr = xcd_range(blockIdx.x, gridDim.x, L.np, kWaves);
  // P16 (16 x 16 cells: ring cell = lane, 2x2 block = lane): the next patch's ring
  // records, parent, orth, u and f (the lane's 2x2 block) are loaded during the current
  // one, and the neighbour values of its ghosts as soon as the records are in
  T un[P16 ? 4 : 1], gu[4];
  TF fn[P16 ? 4 : 1];
  int on = -1, cpn = 0;
  int4 rn = make_int4(-1, 0, -1, 0);
  auto load_a = [&](int p) {
    int rs, rt;
    rn = at(ring, ring_index(p, lane, n0, n1, &rs, &rt));
    on = at(porth, p);
    cpn = at(parent, p);
#pragma unroll
    for (int k = 0; k < (P16 ? 4 : 1); k++) {
      const int q = lane + 64 * k;
      un[k] = at(u, p * L.ps + ((q >> 4) + 1) * L.sx + (q & 15) + 1);
      const int x = 2 * (lane & 7) + (k & 1), y = 2 * (lane >> 3) + (k >> 1);
      fn[k] = at(f, p * L.ps + (y + 1) * L.sx + x + 1);
    }
  };
  auto load_b = [&]() { ring_gather_plain<T, KIND>(rn, lane, n0, n1, u, L, gu); };
  if (P16 && xr.first + wave < xr.end) {
    load_a(xr.first + wave);
    load_b();
  }
  for (int p = xr.first + wave; p < xr.end; p += xr.step) {
    // everything that does not depend on the tile is loaded together with its interior
    SideInfo si[1];
    int o, cp;
    TF fcur[P16 ? 4 : 1];
    const int4 rc = rn;
    if (P16) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int q = lane + 64 * k;
        tile[((q >> 4) + 1) * w + (q & 15) + 1] = un[k];
        fcur[k] = fn[k];
      }
      o = on;
      cp = cpn;
    } else {
      tile_interior_2d<T>(p, n0, n1, U, tile, lane);
      ring_prefetch<1>(dp[p], n0, n1, lane, si);
      o = porth[p];
      cp = parent[p];
    }
    wave_sync();
    const bool has_next = P16 && p + xr.step < xr.end;
    if (has_next) load_a(p + xr.step);
    if (P16) tile_ghost_gathered<T, KIND>(rc, lane, n0, n1, tile, phys, gu);
    else tile_ghosts_2d<T, KIND>(dp[p], si[0], p, n0, n1, U, tile, phys, lane);
    if (has_next) load_b();
    wave_sync();
    {
      const DevPatch& P = dp[p];
      const T ih2[2] = { T(P.inv_h2[0]), T(P.inv_h2[1]) };
      const int ox = (o & 1) ? h0 : 0, oy = (o & 2) ? h1 : 0;
      // a lane per 2x2 block of fine cells = one coarse cell
      for (int Q = lane; Q < h0 * h1; Q += 64) {
        const int X = Q % h0, Y = Q / h0;
        T r[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int x = 2 * X + (j & 1), y = 2 * Y + (j >> 1);
          const int c = (y + 1) * L.sx + x + 1;
          const TF fv = P16 ? fcur[j] : at(f, p * L.ps + c);
          r[j] = T(fv) - star_tile<T, OP>(tile, w, x, y, ih2, cf, p * L.cs + c, L.sx);
        }
        if (o < 0) {  // copyToParent
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int x = 2 * X + (j & 1), y = 2 * Y + (j >> 1);
            at(fc, cp * L.ps + (y + 1) * L.sx + x + 1) = TC(r[j]);
          }
        } else {
          // restrictToCoarserParent: coarse[(i + s) / 2] += fine[i] / 4 over the fine
          // cells in x-fastest order
          T acc = T(0);
#pragma unroll
          for (int j = 0; j < 4; j++) acc += r[j] * T(0.25);
          at(fc, cp * L.ps + (Y + oy + 1) * L.sx + X + ox + 1) = TC(acc);
        }
      }
    }
    wave_sync();  // the tile is rewritten for the next patch
  }
}

// interior value of patch q at (x, y), widened to T
template <typename T, typename TU>
struct WideAcc {
  const TU* u;
  Layout L;
  __device__ __forceinline__ T operator()(int q, int x, int y) const { return T(at(u, q * L.ps + (y + 1) * L.sx + x + 1)); }
};

// out = A u (ghost cells of out written 0); pa[block] = sum w * out, pb[block] = sum out * out,
// and 0 in the slots from gridDim.x to nslots (the partial count the finalize folds).
// TU: type of u (FP32 values are widened to FP64 as they are read: the arithmetic is FP64).
template <int OP, int KIND, bool SQ, bool P16, typename TU>
__global__ void __launch_bounds__(64 * kWaves) star_dot_2d(const DevPatch* __restrict__ dp,
                                                           const TU* __restrict__ u,
                                                           const double* __restrict__ cf, double* __restrict__ out,
                                                           const double* __restrict__ wv, double* __restrict__ pa,
                                                           double* __restrict__ pb, Layout L, int neumann,
                                                           const int4* __restrict__ ring, int nslots, BatchMap bm)
{
  if (P16) {  // compile-time layout constants: the address arithmetic folds
    L.n[0] = L.n[1] = 16;
    L.sx = 18;
    L.ps = L.cs = 18 * 18;
  }
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n0 = L.n[0], n1 = L.n[1], w = n0 + 2;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* tile = reinterpret_cast<double*>(smem_raw) + wave * w * (n1 + 2);
  WideAcc<double, TU> U{ u, L };
  const double phys = (OP == 0 && neumann) ? 1.0 : -1.0;
  double sa = 0.0, sq = 0.0;
  // the patches of this workgroup and its partial slot; in a batch those of its problem,
  // split as the launch on that problem alone splits them
  XcdRange xr;
  int slot = blockIdx.x, zero_from = gridDim.x, zero_to = nslots;
  bool lead = blockIdx.x == 0;
  if (bm.prob) {
    const BatchProb P = bm.prob[bm.wq[blockIdx.x]];
    const int b = (int)blockIdx.x - P.woff;
    xr = xcd_range(b, P.nwg, P.np, kWaves);
    xr.first += P.poff;
    xr.end += P.poff;
    slot = P.soff + b;
    zero_from = P.soff + P.nwg;
    zero_to = P.soff + P.nslot;
    lead = b == 0;
  } else {
    xr = xcd_range(blockIdx.x, gridDim.x, L.np, kWaves);
  }
  // P16 (16 x 16 cells: ring cell = lane): the next patch's u, w and ring records are
  // loaded during the current one, and the neighbour values of its ghosts as soon as
  // the records are in
  double un[P16 ? 4 : 1], wn[P16 ? 4 : 1], wcur[P16 ? 4 : 1];
  int4 rn = make_int4(-1, 0, -1, 0);
  double gu[4];
  auto load_a = [&](int p) {
    int rs, rt;
    rn = at(ring, ring_index(p, lane, n0, n1, &rs, &rt));
#pragma unroll
    for (int k = 0; k < (P16 ? 4 : 1); k++) {
      const int q = lane + 64 * k;
      const int c = p * L.ps + ((q >> 4) + 1) * L.sx + (q & 15) + 1;
      un[k] = double(at(u, c));
      wn[k] = at(wv, c);
    }
  };
  auto load_b = [&]() { ring_gather_plain<double, KIND>(rn, lane, n0, n1, u, L, gu); };
  if (P16 && xr.first + wave < xr.end) {
    load_a(xr.first + wave);
    load_b();
  }
  for (int p = xr.first + wave; p < xr.end; p += xr.step) {
    const int base = p * L.ps;
    SideInfo si[1];
    int4 rc = rn;
    if (P16) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int q = lane + 64 * k;
        tile[((q >> 4) + 1) * w + (q & 15) + 1] = un[k];
        wcur[k] = wn[k];
      }
    } else {
      tile_interior_2d<double>(p, n0, n1, U, tile, lane);
      ring_prefetch<1>(dp[p], n0, n1, lane, si);
    }
    wave_sync();
    const bool has_next = P16 && p + xr.step < xr.end;
    if (has_next) load_a(p + xr.step);
    if (P16) tile_ghost_gathered<double, KIND>(rc, lane, n0, n1, tile, phys, gu);
    else tile_ghosts_2d<double, KIND>(dp[p], si[0], p, n0, n1, U, tile, phys, lane);
    if (has_next) load_b();
    wave_sync();
    {
      const DevPatch& P = dp[p];
      const double ih2[2] = { P.inv_h2[0], P.inv_h2[1] };
      for (int q = lane, k = 0; q < n0 * n1; q += 64, k++) {
        const int y = q / n0, x = q - y * n0;
        const int c = (y + 1) * L.sx + x + 1;
        const double acc = star_tile<double, OP>(tile, w, x, y, ih2, cf, p * L.cs + c, L.sx);
        at(out, base + c) = acc;
        const double wq = P16 ? wcur[k & 3] : at(wv, base + c);
        sa += wq * acc;  // rhat.dot(ap) / as.dot(s): w first, as the reference
        if (SQ) sq += acc * acc;
      }
      // ghost frame of out: rows -1 and n1 (with corners), then columns -1 and n0
      for (int g = lane; g < 2 * w + 2 * n1; g += 64) {
        int x, y;
        if (g < 2 * w) {
          y = g < w ? -1 : n1;
          x = (g < w ? g : g - w) - 1;
        } else {
          const int h = g - 2 * w;
          x = h < n1 ? -1 : n0;
          y = h < n1 ? h : h - n1;
        }
        at(out, base + (y + 1) * L.sx + x + 1) = 0.0;
      }
    }
    wave_sync();
  }
  double r = block_reduce<RED_SUM>(sa);
  if (threadIdx.x == 0) pa[slot] = r;
  if (SQ) {
    r = block_reduce<RED_SUM>(sq);
    if (threadIdx.x == 0) pb[slot] = r;
  }
  if (lead) {
    for (int i = zero_from + (int)threadIdx.x; i < zero_to; i += (int)blockDim.x) {
      pa[i] = 0.0;
      if (SQ) pb[i] = 0.0;
    }
  }
}

size_t tile_bytes(const te_domain* d, size_t esz) { return esz * kWaves * (d->ns[0] + 2) * (d->ns[1] + 2); }

template <typename T, typename TF, typename TC>
void launch_rr(const te_ghost_filler* g, const te_level_link* link, int op_kind, int neumann, const void* coeffs,
               const void* f, const void* u, void* fc, hipStream_t s)
{
  const te_domain* d = g->d;
  Layout L = make_layout(2, d->ns, 1, d->np);
  size_t shmem = tile_bytes(d, sizeof(T));
  const bool p16 = d->ns[0] == 16 && d->ns[1] == 16;
  int ngroups = (d->np + kWaves - 1) / kWaves;
  // 16 x 16: as many workgroups as are resident at once, so that each wave takes several
  // patches (all in one round) and its loads of the next one overlap the current one
  auto grid_of = [&](const void* k) {
    const int cap = p16 ? resident_blocks(k, 64 * kWaves, shmem) : 8192;
    return ngroups < cap ? ngroups : cap;
  };
#define TE_RR(OP_, K_, P_)                                                                                   \
  hipLaunchKernelGGL((resid_restrict_2d<T, TF, TC, OP_, K_, P_>),                                            \
                     dim3(grid_of((const void*)resid_restrict_2d<T, TF, TC, OP_, K_, P_>)), dim3(64 * kWaves), \
                     shmem, s, d->d_patches, (const T*)u, (const T*)coeffs, (const TF*)f, (TC*)fc,            \
                     link->d_parent, link->d_orth, L, neumann, (const int4*)d->d_ring)
#define TE_RRP(OP_, K_)             \
  do {                              \
    if (p16) TE_RR(OP_, K_, true);  \
    else TE_RR(OP_, K_, false);     \
  } while (0)
  if (op_kind == 0) {
    if (g->kind == TE_BILINEAR) TE_RRP(0, TE_BILINEAR);
    else TE_RRP(0, TE_BIQUADRATIC);
  } else {
    if (g->kind == TE_BILINEAR) TE_RRP(1, TE_BILINEAR);
    else TE_RRP(1, TE_BIQUADRATIC);
  }
#undef TE_RRP
#undef TE_RR
}

}  // namespace

// a tile of doubles for each of the kWaves waves must fit the 64 KiB a workgroup may use
static bool tile_fits(const te_domain* d) { return tile_bytes(d, sizeof(double)) <= 65536; }

bool resid_restrict_fused_supported(const te_ghost_filler* g)
{
  const te_domain* d = g->d;
  return d->dim == 2 && d->ns[0] % 2 == 0 && d->ns[1] % 2 == 0 && tile_fits(d) &&
         (g->kind == TE_BILINEAR || g->kind == TE_BIQUADRATIC);
}

// dt: u / coeffs / residual type, dtf: f type, dtc: coarse f type
int launch_resid_restrict(const te_ghost_filler* g, const te_level_link* link, int op_kind, int neumann,
                          const void* coeffs, int dt, int dtf, const void* f, const void* u, int dtc, void* fc,
                          hipStream_t s)
{
  if (g->d->np == 0) return TE_OK;
  if (!resid_restrict_fused_supported(g)) {
    set_error("fused residual / restriction: needs 2D, even n, BiLinear or BiQuadratic ghosts");
    return TE_ERR_UNSUPPORTED;
  }
  if (dt == TE_F64 && dtf == TE_F64 && dtc == TE_F64) launch_rr<double, double, double>(g, link, op_kind, neumann, coeffs, f, u, fc, s);
  else if (dt == TE_F32 && dtf == TE_F32 && dtc == TE_F32) launch_rr<float, float, float>(g, link, op_kind, neumann, coeffs, f, u, fc, s);
  else if (dt == TE_F32 && dtf == TE_F64 && dtc == TE_F32) launch_rr<float, double, float>(g, link, op_kind, neumann, coeffs, f, u, fc, s);
  else if (dt == TE_F32 && dtf == TE_F32 && dtc == TE_F64) launch_rr<float, float, double>(g, link, op_kind, neumann, coeffs, f, u, fc, s);
  else if (dt == TE_F32 && dtf == TE_F64 && dtc == TE_F64) launch_rr<float, double, double>(g, link, op_kind, neumann, coeffs, f, u, fc, s);
  else {
    set_error("unsupported dtype combination for the fused residual / restriction");
    return TE_ERR_UNSUPPORTED;
  }
  TE_HIP(hipGetLastError());
  return TE_OK;
}

bool star_dot_fly_supported(const te_ghost_filler* g)
{
  const te_domain* d = g->d;
  return d->dim == 2 && d->ns[0] % 2 == 0 && d->ns[1] % 2 == 0 && tile_fits(d) &&
         (g->kind == TE_BILINEAR || g->kind == TE_BIQUADRATIC);
}

// the kernel instantiation a launch runs
static const void* star_dot_2d_kernel(const te_ghost_filler* g, int op_kind, int udt, bool sq)
{
  const te_domain* d = g->d;
  const bool p16 = d->ns[0] == 16 && d->ns[1] == 16;
  const bool bq = g->kind != TE_BILINEAR;
#define TE_K(OP_, K_, Q_, P_)                                                                         \
  return udt == TE_F32 ? (const void*)star_dot_2d<OP_, K_, Q_, P_, float> \
                       : (const void*)star_dot_2d<OP_, K_, Q_, P_, double>
#define TE_KQ(OP_, K_)                          \
  do {                                          \
    if (sq && p16) TE_K(OP_, K_, true, true);   \
    if (sq) TE_K(OP_, K_, true, false);         \
    if (p16) TE_K(OP_, K_, false, true);        \
    TE_K(OP_, K_, false, false);                \
  } while (0)
  if (op_kind == 0) {
    if (!bq) TE_KQ(0, TE_BILINEAR);
    TE_KQ(0, TE_BIQUADRATIC);
  }
  if (!bq) TE_KQ(1, TE_BILINEAR);
  TE_KQ(1, TE_BIQUADRATIC);
#undef TE_KQ
#undef TE_K
}

int star_dot_2d_cap(const te_ghost_filler* g, int op_kind, int udt, bool sq)
{
  if (!star_dot_fly_supported(g)) return 0;
  return resident_blocks(star_dot_2d_kernel(g, op_kind, udt, sq), 64 * kWaves, tile_bytes(g->d, sizeof(double)));
}

int launch_star_dot_2d(const te_ghost_filler* g, int op_kind, int neumann, const double* coeffs, const void* u,
                       int udt, double* out, const double* w, double* pa, double* pb, int grid, hipStream_t s,
                       const BatchStarMaps* bsm)
{
  const te_domain* d = g->d;
  if (!star_dot_fly_supported(g)) {
    set_error("fused 2D operator: needs 2D, even n, BiLinear or BiQuadratic ghosts and a tile within 64 KiB");
    return TE_ERR_UNSUPPORTED;
  }
  Layout L = make_layout(2, d->ns, 1, d->np);
  size_t shmem = tile_bytes(d, sizeof(double));
  const bool p16 = d->ns[0] == 16 && d->ns[1] == 16;
  // as many workgroups as are resident at once (each wave takes several patches, all in one
  // round), at most the `grid` partial slots; in a batch that choice per problem (the map)
  BatchMap bm{};
  if (bsm) {
    const int cap = resident_blocks(star_dot_2d_kernel(g, op_kind, udt, pb != nullptr), 64 * kWaves, shmem);
    int k = 0;
    while (k < bsm->n && bsm->cap[k] != cap) k++;
    if (k == bsm->n) {
      set_error("internal: no batch map for this operator launch");
      return TE_ERR_ARG;
    }
    bm = bsm->map[k];
    if (bm.nwg < 1) return TE_OK;
  }
  auto fit = [&](const void* k) {
    if (bsm) return bm.nwg;
    const int r = resident_blocks(k, 64 * kWaves, shmem);
    return r < grid ? r : grid;
  };
#define TE_SD(OP_, K_, Q_, P_)                                                                                  \
  do {                                                                                                         \
    if (udt == TE_F32)                                                                                         \
      hipLaunchKernelGGL((star_dot_2d<OP_, K_, Q_, P_, float>),                                                \
                         dim3(fit((const void*)star_dot_2d<OP_, K_, Q_, P_, float>)), dim3(64 * kWaves), shmem, \
                         s, d->d_patches, (const float*)u, coeffs, out, w, pa, pb, L, neumann,                  \
                         (const int4*)d->d_ring, grid, bm);                                                     \
    else                                                                                                       \
      hipLaunchKernelGGL((star_dot_2d<OP_, K_, Q_, P_, double>),                                               \
                         dim3(fit((const void*)star_dot_2d<OP_, K_, Q_, P_, double>)), dim3(64 * kWaves),      \
                         shmem, s, d->d_patches, (const double*)u, coeffs, out, w, pa, pb, L, neumann,          \
                         (const int4*)d->d_ring, grid, bm);                                                     \
  } while (0)
#define TE_SDQ(OP_, K_)                      \
  do {                                       \
    if (pb && p16) TE_SD(OP_, K_, true, true);   \
    else if (pb) TE_SD(OP_, K_, true, false);    \
    else if (p16) TE_SD(OP_, K_, false, true);   \
    else TE_SD(OP_, K_, false, false);           \
  } while (0)
  if (op_kind == 0) {
    if (g->kind == TE_BILINEAR) TE_SDQ(0, TE_BILINEAR);
    else TE_SDQ(0, TE_BIQUADRATIC);
  } else {
    if (g->kind == TE_BILINEAR) TE_SDQ(1, TE_BILINEAR);
    else TE_SDQ(1, TE_BIQUADRATIC);
  }
#undef TE_SDQ
#undef TE_SD
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
