// DFTPatchSolver::solveSinglePatch (Poisson/DFTPatchSolver.h:420-450) on the
// matrix cores, 2D, all-Dirichlet patches, n = 16 or 32.
//
// Each wavefront solves one patch: U = S3 . ((S . F . S^T) / lambda) . S3^T * (2/n)^2,
// with S the DST-II and S3 the DST-III matrix (the reference's per-line dgemv along
// x, then along y, :205-265).  Every product is a sequence of 16x16x4 MFMAs
// (v_mfma_f32_16x16x4_f32 / v_mfma_f64_16x16x4_f64) over 16x16 tiles:
//   Y = F S^T   F from LDS in the A-operand layout, S^T fragments from LDS;
//   Z = S Y     Y stays in registers: the accumulator layout is used directly as
//               the B operand by permuting the K index to match it;
//   V = S3 Z'   (Z' = Z / lambda) the same register-resident trick;
//   U = V S3^T  V goes through LDS once to reach the A-operand layout.
// Lane layouts (checked on gfx950): A[i][k] at lane i + 16k, B[k][j] at lane j + 16k;
// accumulators: f32 lane l, reg r holds D[4(l/16) + r][l%16], f64 holds D[4r + l/16][l%16].
//
// The interface terms of modifyRHSForInternalBoundaryConditions need u's ghost
// values on sides with neighbours.  GM selects where they come from:
//   GM_MEM    read from u's ghost cells (filled by a preceding ghost fill);
//   GM_FLY    computed here from the neighbours' interior values (fused ghost fill);
//   GM_INTERP as GM_FLY on u + P uc, the DirectInterpolator correction from the
//             coarser level (GMG/DirectInterpolator.h:60-91) applied on the fly, so
//             interpolation, ghost fill and the post-smoothing sweep are one launch
//             (the result must then go to a buffer other than u).
#include "../te_internal.h"
#include "blas.h"
#include "../tune.h"
#include "dft.h"
#include "dft_wave.h"
#include "ghost2d.h"
#include "krylov.h"
#include "fold.h"

namespace te {
namespace {

constexpr int kWaves = 4;

// PF (n = 16): each wave walks several patches and loads the next one while it solves the
// current one; without PF (levels where each wave gets about one patch) a patch's loads come
// right before its solve, which needs fewer registers, so more waves fit a SIMD.
template <typename T, typename TF, typename TO, typename TC, int NT, int OP, bool UZERO, int GM, int KIND, bool PF>
__global__ void __launch_bounds__(64 * kWaves, (NT == 1 && sizeof(T) == 4) ? (PF ? (GM == 0 ? 5 : 4) : (GM == 0 ? 8 : (KIND == TE_BILINEAR ? 5 : 4))) : 1) dft_mfma_kernel(const DevPatch* __restrict__ dp,
                                                               const T* __restrict__ mats, const T* __restrict__ s2tab,
                                                               const TF* __restrict__ f, const T* __restrict__ u,
                                                               const T* __restrict__ coeffs, TO* __restrict__ out,
                                                               Layout L, FusedArgs<T, TC> fa,
                                                               const int4* __restrict__ ring)
{
  constexpr int n = 16 * NT;
  constexpr int ld = n + 1;  // padded LDS row
  // the padded layout of a one-component vector of n x n patches (L.sx, L.ps = L.cs) as
  // compile-time constants: the address arithmetic folds into shifts and adds
  constexpr int SX = n + 2, PS = SX * SX;
  L.sx = SX;  // the same constants for the accessors that take the layout
  L.ps = PS;
  L.cs = PS;
  L.n[0] = n;
  L.n[1] = n;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // the transform matrices with rows padded to n + 1 (the MFMA operand loads of a lane
  // group walk down a column: an unpadded pitch of n puts 8 lanes on one LDS bank)
  T* Sm = reinterpret_cast<T*>(smem_raw);  // DST-II, row-major n x ld
  T* S3 = Sm + n * ld;                     // DST-III
  T* s2 = S3 + n * ld;                     // sin^2((k+1) pi / 2n), k < n
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  T* X = s2 + n + wave * n * ld;  // per-wave scratch n x ld
  T* Bt = s2 + n + kWaves * n * ld + wave * 4 * n;  // per-wave interface terms of the 4 sides
  // n = 16 with ghosts computed here: the patch's own values (u, or u + P uc) are staged
  // in a padded tile first, so the ghost formulas read their self terms from LDS
  constexpr bool TILE = GM != 0 && NT == 1;
  T* Tl = s2 + n + kWaves * n * ld + kWaves * 4 * n + wave * (n + 2) * (n + 2);
  UAcc<T, TC, GM == 2> U{ u, fa.uc, fa.parent, fa.porth, L };

  for (int k = threadIdx.x; k < n * n; k += blockDim.x) {
    const int row = k / n, col = k - row * n;
    Sm[row * ld + col] = mats[(3 * 3 + 0) * kDftMatStride + k];  // DST_II, axis 0
    S3[row * ld + col] = mats[(4 * 3 + 0) * kDftMatStride + k];  // DST_III, axis 0
  }
  for (int k = threadIdx.x; k < n; k += blockDim.x) s2[k] = s2tab[(2 * 3 + 0) * kDftMatDim + k];

  // Each wave walks its own patches with private LDS (X, Bt, Tl): wave-level ordering
  // suffices inside the loop.  The one workgroup barrier orders the shared matrices; the
  // wave's first patch is staged before it so that its loads overlap the matrices'.
  SideInfo si[NT];  // ring cells lane + 64 j: side j * 64 / n + lane / n
  // ---- f_copy = f (interior) into LDS; the side information of the ghost pass is
  //      loaded at the same time
  // n = 16: the next patch's right-hand side is loaded into registers while this one is
  // solved (FPL values per lane), so a wave's loads overlap its own arithmetic
  constexpr int FPL = n * n / 64;
  constexpr bool PREF = NT == 1;
  TF fn[PREF ? FPL : 1];
  auto fload = [&](int p) {
    const int base = p * PS;
#pragma unroll
    for (int k = 0; k < (PREF ? FPL : 1); k++) {
      const int q = lane + 64 * k, y = q / n, x = q - y * n;
      fn[k] = at(f, base + (y + 1) * SX + x + 1);
    }
  };
  auto fput = [&]() {
#pragma unroll
    for (int k = 0; k < (PREF ? FPL : 1); k++) {
      const int q = lane + 64 * k, y = q / n, x = q - y * n;
      X[y * ld + x] = T(fn[k]);
    }
  };
  // with ghosts computed here (TILE): the next patch's own values and its ring records
  // (ring table: lane = ring cell, side lane / n, cell lane % n), then, once those have
  // arrived (mid-iteration), its coarse-level values and the neighbour values of every
  // lane's ghost cell (ghost2d_gather; with GM_INTERP each as u and coarse u separately)
  constexpr bool PREF_T = TILE && PREF;
  T un[PREF_T ? FPL : 1];
  TC ucn[PREF_T ? FPL : 1];
  int on = -1, parn = 0;
  int4 rn = make_int4(-1, 0, -1, 0), rc = rn;  // ring records of the next / current patch
  T gu[4];
  TC gc[4];
  T gv[4];  // the current patch's neighbour values (u, or u + P uc)
  const int rs = lane / n, rt = lane - (lane / n) * n;  // the lane's ring cell (PREF_T: 4 n = 64)
  auto tload_a = [&](int p) {
#pragma unroll
    for (int k = 0; k < (PREF_T ? FPL : 1); k++) {
      const int q = lane + 64 * k, y = q / n, x = q - y * n;
      un[k] = at(u, p * PS + (y + 1) * SX + x + 1);
    }
    if (GM == 2) {
      on = at(fa.porth, p);
      parn = at(fa.parent, p);
    }
    rn = at(ring, (p * 4 + rs) * 2 + (2 * rt >= n ? 1 : 0));
  };
  auto tload_b = [&]() {
    if (GM == 2) {
#pragma unroll
      for (int k = 0; k < (PREF_T ? FPL : 1); k++) {
        const int q = lane + 64 * k, y = q / n, x = q - y * n;
        int cx = x, cy = y;
        if (on >= 0) {
          cx = (x + ((on & 1) ? n : 0)) / 2;
          cy = (y + ((on & 2) ? n : 0)) / 2;
        }
        ucn[k] = at(fa.uc, parn * PS + (cy + 1) * SX + cx + 1);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      gu[k] = T(0);
      gc[k] = TC(0);
    }
    const int pq = rn.w;
    ghost2d_gather<KIND>(rn.x, rn.y, rn.z, rs, rt, n, n, [&](int k, int q, int x, int y) {
      gu[k] = at(u, q * PS + (y + 1) * SX + x + 1);
      if (GM == 2) {
        const int o = (pq & 7) - 1;
        int cx = x, cy = y;
        if (o >= 0) {
          cx = (x + ((o & 1) ? n : 0)) / 2;
          cy = (y + ((o & 2) ? n : 0)) / 2;
        }
        gc[k] = at(fa.uc, (pq >> 3) * PS + (cy + 1) * SX + cx + 1);
      }
    });
  };
  auto tput = [&]() {  // the values UAcc would return, staged from the registers
#pragma unroll
    for (int k = 0; k < (PREF_T ? FPL : 1); k++) {
      const int q = lane + 64 * k, y = q / n, x = q - y * n;
      Tl[(y + 1) * (n + 2) + x + 1] = GM == 2 ? T(un[k] + T(ucn[k])) : un[k];
    }
    rc = rn;
#pragma unroll
    for (int k = 0; k < 4; k++) gv[k] = GM == 2 ? T(gu[k] + T(gc[k])) : gu[k];
  };
  auto stage = [&](int p) {
    if (PREF_T) {
      fput();
      tput();
      return;
    }
    if (PREF) {
      fput();
    } else {
      const int base = p * PS;
      for (int q = lane; q < n * n; q += 64) {
        int y = q / n, x = q - y * n;
        X[y * ld + x] = T(f[base + (y + 1) * SX + x + 1]);
      }
    }
    if (TILE) tile_interior_2d<T>(p, n, n, U, Tl, lane);
    if (!UZERO) ring_prefetch<NT>(dp[p], n, n, lane, si);
  };
  const XcdRange xr = xcd_range(blockIdx.x, gridDim.x, L.np, kWaves);
  int p = xr.first + wave;
  if (PF && p < xr.end) {
    if (PREF) fload(p);
    if (PREF_T) {
      tload_a(p);
      tload_b();
    }
    stage(p);
  }
  __syncthreads();
  for (bool first = true; p < xr.end; p += xr.step, first = false) {
    if (!PF) {
      if (PREF) fload(p);
      if (PREF_T) {
        tload_a(p);
        tload_b();
      }
      stage(p);
      wave_sync();
    } else if (!first) {
      stage(p);
      wave_sync();
    }
    const bool has_next = PF && p + xr.step < xr.end;
    if (PREF && has_next) fload(p + xr.step);
    if (PREF_T && has_next) tload_a(p + xr.step);
    const int base = p * PS;
    const DevPatch& P = dp[p];
    // ---- modifyRHSForInternalBoundaryConditions: the interface term of every side cell
    //      is formed first (a lane per cell, so all its loads are in flight together),
    //      then applied west/east first, then south/north (sides in increasing order,
    //      Poisson/StarPatchOperator.h:165-180)
    if (!UZERO) {
#pragma unroll
      for (int j = 0; j < NT; j++) {  // 4 n = 64 NT ring cells
        const int q = lane + 64 * j;
        const int s = q / n, t = q - s * n;
        const int ph = s >> 1, up = s & 1;
        T term = T(0);
        if ((PREF_T ? rc.x : si[j].type) >= 0) {
          int gi, bi;  // padded index of the ghost and of the boundary cell
          int bx, by;  // boundary cell
          if (ph == 0) {
            gi = base + (t + 1) * SX + (up ? n + 1 : 0);
            bi = base + (t + 1) * SX + (up ? n : 1);
            bx = up ? n - 1 : 0;
            by = t;
          } else {
            gi = base + (up ? n + 1 : 0) * SX + t + 1;
            bi = base + (up ? n : 1) * SX + t + 1;
            bx = t;
            by = up ? n - 1 : 0;
          }
          T ug, ub;
          if (GM == 0) {
            ug = u[gi];
            ub = u[bi];
          } else if (PREF_T) {
            // self terms from the tile, neighbour terms gathered during the previous patch
            ug = ghost2d_combine<T, KIND>(rc.x, rc.y, t, n,
                                          [&](int k, int tt) {
                                            int x, y;
                                            side_xy(s, k, tt, n, &x, &y);
                                            return Tl[(y + 1) * (n + 2) + x + 1];
                                          },
                                          gv);
            ub = Tl[(by + 1) * (n + 2) + bx + 1];
          } else if (TILE) {
            TileAcc<T, UAcc<T, TC, GM == 2>> A{ Tl, p, n + 2, U };
            ug = ghost2d<T, KIND>(si[j], p, s, t, n, n, A);
            ub = Tl[(by + 1) * (n + 2) + bx + 1];
          } else {
            ug = ghost2d<T, KIND>(si[j], p, s, t, n, n, U);
            ub = U(p, bx, by);
          }
          const T ih2 = T(P.inv_h2[ph]);
          if (OP == 0) {
            term = (ug + ub) * ih2;
          } else {
            const int cg = gi - p * PS + p * PS, cb = bi - p * PS + p * PS;
            term = (ug + ub) * (coeffs[cb] + coeffs[cg]) * (ih2 * T(0.5));
          }
        }
        Bt[q] = term;
      }
      if (PREF_T && has_next) tload_b();
      wave_sync();
#pragma unroll
      for (int ph = 0; ph < 2; ph++) {
        for (int q = lane; q < 2 * n; q += 64) {
          const int up = q / n, t = q - up * n;
          const int s = 2 * ph + up;
          if (P.type[4 + s] < 0) continue;
          const int xi = ph == 0 ? t * ld + (up ? n - 1 : 0) : (up ? n - 1 : 0) * ld + t;
          X[xi] -= Bt[s * n + t];
        }
        wave_sync();
      }
    }
    const T lx = T(4.0 / (P.h[0] * P.h[0])), ly = T(4.0 / (P.h[1] * P.h[1]));
    dft_solve_wave<T, TO, NT>(X, Sm, S3, s2, lx, ly, out, base, SX, true);
    wave_sync();  // X and the tile are rewritten by the next patch
  }
}

template <typename T, typename TF, typename TO, typename TC, int NT>
void launch_nt(const te_dft_plan* pl, const TF* f, const T* u, TO* out, bool uzero, int gm, const TC* uc,
               const te_level_link* link, hipStream_t s)
{
  const te_domain* d = pl->g->d;
  Layout L = make_layout(2, d->ns, 1, d->np);
  constexpr int n = 16 * NT;
  const bool tile = gm != 0 && NT == 1;
  size_t shmem = sizeof(T) * (2 * n * (n + 1) + n + kWaves * n * (n + 1) + kWaves * 4 * n +
                              (tile ? kWaves * (n + 2) * (n + 2) : 0));
  const T* mats = (const T*)(sizeof(T) == 8 ? (const void*)pl->d_mats64 : (const void*)pl->d_mats32);
  const T* s2 = (const T*)(sizeof(T) == 8 ? (const void*)pl->d_s2_64 : (const void*)pl->d_s2_32);
  const T* cf = (const T*)(sizeof(T) == 8 ? (const void*)pl->d_coeffs64 : (const void*)pl->d_coeffs32);
  FusedArgs<T, TC> fa{ uc, link ? link->d_parent : nullptr, link ? link->d_orth : nullptr };
  // ghosts from the ring table: the link's (with the coarse parents) when interpolating
  const int4* ring = (const int4*)(gm == 2 ? link->d_ring : d->d_ring);
  const int ngroups = (d->np + kWaves - 1) / kWaves;
  const int kind = pl->g->kind;
  // the workgroup count: the autotuner's choice for this patch size and type (tune.h)
  const int var = tune_variant(TUNE_DFT_MFMA_2D, 2, n, sizeof(T) == 8 ? TE_F64 : TE_F32);
  // n = 16: when every workgroup (one patch per wave) is resident at once, one round of
  // the plain variant; else the prefetching variant (the next patch loaded during the
  // current solve) on as many workgroups as are resident at once, each wave solving
  // several patches.  n = 32: at most 4096 workgroups.
#define TE_MK1(OP_, UZ_, GM_, K_, PF_, G_)                                                                    \
  hipLaunchKernelGGL((dft_mfma_kernel<T, TF, TO, TC, NT, OP_, UZ_, GM_, K_, PF_>), dim3(G_), dim3(64 * kWaves), \
                     shmem, s, d->d_patches, mats, s2, f, u, cf, out, L, fa, ring)
#define TE_MK(OP_, UZ_, GM_, K_)                                                                            \
  do {                                                                                                      \
    if constexpr (NT == 1) {                                                                                \
      const void* plain = (const void*)dft_mfma_kernel<T, TF, TO, TC, NT, OP_, UZ_, GM_, K_, false>;        \
      if (var == MFMA_GRID_ALL ||                                                                           \
          (var == MFMA_GRID_DEFAULT && ngroups <= resident_blocks(plain, 64 * kWaves, shmem))) {            \
        TE_MK1(OP_, UZ_, GM_, K_, false, ngroups);                                                          \
      } else {                                                                                              \
        int r = resident_blocks((const void*)dft_mfma_kernel<T, TF, TO, TC, NT, OP_, UZ_, GM_, K_, true>,   \
                                64 * kWaves, shmem);                                                        \
        if (var == MFMA_GRID_HALF) r = r > 1 ? r / 2 : 1;                                                   \
        TE_MK1(OP_, UZ_, GM_, K_, true, ngroups < r ? ngroups : r);                                         \
      }                                                                                                     \
    } else {                                                                                                \
      const int cap = var == MFMA_GRID_RESIDENT ? 2048 : var == MFMA_GRID_HALF ? 8192                      \
                      : var == MFMA_GRID_ALL    ? ngroups : 4096;                                           \
      TE_MK1(OP_, UZ_, GM_, K_, false, ngroups < cap ? ngroups : cap);                                      \
    }                                                                                                       \
  } while (0)
#define TE_MKG(OP_)                                                         \
  if (uzero) TE_MK(OP_, true, 0, 0);                                        \
  else if (gm == 0) TE_MK(OP_, false, 0, 0);                                \
  else if (gm == 1 && kind == TE_BILINEAR) TE_MK(OP_, false, 1, TE_BILINEAR); \
  else if (gm == 1) TE_MK(OP_, false, 1, TE_BIQUADRATIC);                   \
  else if (kind == TE_BILINEAR) TE_MK(OP_, false, 2, TE_BILINEAR);          \
  else TE_MK(OP_, false, 2, TE_BIQUADRATIC);
  if (pl->op_kind == 0) {
    TE_MKG(0)
  } else {
    TE_MKG(1)
  }
#undef TE_MKG
#undef TE_MK
#undef TE_MK1
}

template <typename T, typename TF, typename TO, typename TC>
void launch_t(const te_dft_plan* pl, const void* f, const void* u, void* out, bool uzero, int gm, const void* uc,
              const te_level_link* link, hipStream_t s)
{
  if (pl->g->d->ns[0] == 16)
    launch_nt<T, TF, TO, TC, 1>(pl, (const TF*)f, (const T*)u, (TO*)out, uzero, gm, (const TC*)uc, link, s);
  else
    launch_nt<T, TF, TO, TC, 2>(pl, (const TF*)f, (const T*)u, (TO*)out, uzero, gm, (const TC*)uc, link, s);
}


// Cycle::restrict of a level fused with the first smoothing sweep of the next coarser
// level, which starts from u = 0 and so needs no ghost values (VCycle::visit,
// GMG/VCycle.h:46-70: the coarse level's pre-smoothing follows the restriction
// directly), 16 x 16 cells:
//   one workgroup per coarse patch P, one wave per child c of P (2^D children, or P
//   itself when it was not coarsened): c's u is staged in a tile with on-the-fly ghosts,
//   r = f - A u is formed on c's cells and restricted (2x2 averages, or a copy) into X,
//   the coarse right-hand side in LDS; then fc[P] = X (kept for the rest of the cycle)
//   and uc[P] = DFT solve of X by one wave.
// T: fine arithmetic, TF: fine f, TC: coarse level type (its f, u and DFT arithmetic).
template <typename T, typename TF, typename TC, int OP, int KIND>
__global__ void __launch_bounds__(64 * kWaves) rr_pre_kernel(const DevPatch* __restrict__ dpf,
                                                             const T* __restrict__ u, const T* __restrict__ cf,
                                                             const TF* __restrict__ f, Layout Lf, int neumann,
                                                             const int32_t* __restrict__ child,
                                                             const int32_t* __restrict__ nchild,
                                                             const int32_t* __restrict__ porth,
                                                             const DevPatch* __restrict__ dpc,
                                                             const TC* __restrict__ mats,
                                                             const TC* __restrict__ s2tab, TC* __restrict__ fc,
                                                             TC* __restrict__ uc, Layout Lc,
                                                             const int4* __restrict__ ring)
{
  constexpr int n = 16;
  constexpr int ld = n + 1;
  constexpr int w = n + 2;
  constexpr int h = n / 2;
  constexpr int SX = n + 2, PS = SX * SX;  // padded one-component layout of both levels
  Lf.sx = SX;  // the same constants for the accessors that take the layout
  Lf.ps = PS;
  Lf.cs = PS;
  Lf.n[0] = n;
  Lf.n[1] = n;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TC* Sm = reinterpret_cast<TC*>(smem_raw);  // rows padded to ld, as in dft_mfma_kernel
  TC* S3 = Sm + n * ld;
  TC* s2 = S3 + n * ld;
  TC* X = s2 + n;  // the coarse patch's right-hand side, shared by the workgroup
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  T* tile = reinterpret_cast<T*>(X + n * ld) + wave * w * w;
  PlainAcc<T> U{ u, Lf };
  const T phys = (OP == 0 && neumann) ? T(1) : T(-1);
  for (int k = threadIdx.x; k < n * n; k += blockDim.x) {
    const int row = k / n, col = k - row * n;
    Sm[row * ld + col] = mats[(3 * 3 + 0) * kDftMatStride + k];  // DST_II, axis 0
    S3[row * ld + col] = mats[(4 * 3 + 0) * kDftMatStride + k];  // DST_III, axis 0
  }
  for (int k = threadIdx.x; k < n; k += blockDim.x) s2[k] = s2tab[(2 * 3 + 0) * kDftMatDim + k];
  // (the first patch's barriers order the matrix loads before their use)
  const XcdRange xr = xcd_range(blockIdx.x, gridDim.x, Lc.np, 1);
  for (int P = xr.first; P < xr.end; P += xr.step) {
    // wave j handles child j of P (P itself when it was not coarsened: one child)
    const bool act = wave < nchild[P];
    const int c = act ? child[P * 8 + wave] : 0;
    const int Xb = lane % h, Yb = lane / h;  // a lane per 2x2 block (h * h = 64)
    int o = -1;
    int4 rr = make_int4(-1, 0, -1, 0);  // ring record of the lane's ghost cell (ring cell = lane)
    T gv[4];                            // the neighbour values of that ghost
    TF fv[4];                           // f on the lane's 2x2 block
    if (act) {
      // every global read of the child is issued before the first barrier: its ring
      // record, own values, f and orth, then the neighbour values the record names
      int rs, rt;
      rr = at(ring, ring_index(c, lane, n, n, &rs, &rt));
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int x = 2 * Xb + (q & 1), y = 2 * Yb + (q >> 1);
        fv[q] = at(f, c * PS + (y + 1) * SX + x + 1);
      }
      o = porth[c];
      tile_interior_2d<T>(c, n, n, U, tile, lane);
      ring_gather_plain<T, KIND>(rr, lane, n, n, u, Lf, gv);
    }
    __syncthreads();
    if (act) tile_ghost_gathered<T, KIND>(rr, lane, n, n, tile, phys, gv);
    __syncthreads();
    if (act) {
      const DevPatch& Pf = dpf[c];
      const T ih2[2] = { T(Pf.inv_h2[0]), T(Pf.inv_h2[1]) };
      const int ox = (o & 1) ? h : 0, oy = (o & 2) ? h : 0;
      T r[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int x = 2 * Xb + (q & 1), y = 2 * Yb + (q >> 1);
        const int ci = (y + 1) * SX + x + 1;
        r[q] = T(fv[q]) - star_tile<T, OP>(tile, w, x, y, ih2, cf, c * PS + ci, SX);
      }
      if (o < 0) {  // copyToParent
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int x = 2 * Xb + (q & 1), y = 2 * Yb + (q >> 1);
          const TC v = TC(r[q]);
          X[y * ld + x] = v;
          fc[P * PS + (y + 1) * SX + x + 1] = v;
        }
      } else {  // restrictToCoarserParent, fine cells in x-fastest order
        T acc = T(0);
#pragma unroll
        for (int q = 0; q < 4; q++) acc += r[q] * T(0.25);
        const TC v = TC(acc);
        X[(Yb + oy) * ld + Xb + ox] = v;
        fc[P * PS + (Yb + oy + 1) * SX + Xb + ox + 1] = v;
      }
    }
    __syncthreads();
    // the coarse patch solve from u = 0 by wave 0 (the others take part in its barrier)
    const DevPatch& Pc = dpc[P];
    const TC lx = TC(4.0 / (Pc.h[0] * Pc.h[0])), ly = TC(4.0 / (Pc.h[1] * Pc.h[1]));
    dft_solve_wave<TC, TC, 1>(X, Sm, S3, s2, lx, ly, uc, P * PS, SX, wave == 0);
    __syncthreads();  // X is rewritten by the next patch
  }
}

template <typename T, typename TF, typename TC>
void launch_rp(const te_ghost_filler* g, const te_level_link* link, int op_kind, int neumann, const void* coeffs,
               const void* f, const void* u, const te_dft_plan* pc, void* fc, void* uc, hipStream_t s)
{
  const te_domain* df = g->d;
  const te_domain* dc = pc->g->d;
  Layout Lf = make_layout(2, df->ns, 1, df->np);
  Layout Lc = make_layout(2, dc->ns, 1, dc->np);
  constexpr int n = 16;
  size_t shmem = sizeof(TC) * (2 * n * (n + 1) + n + n * (n + 1)) + sizeof(T) * kWaves * (n + 2) * (n + 2);
  const TC* mats = (const TC*)(sizeof(TC) == 8 ? (const void*)pc->d_mats64 : (const void*)pc->d_mats32);
  const TC* s2 = (const TC*)(sizeof(TC) == 8 ? (const void*)pc->d_s2_64 : (const void*)pc->d_s2_32);
  int grid = dc->np < 16384 ? dc->np : 16384;  // a coarse patch per workgroup
#define TE_RP(OP_, K_)                                                                                          \
  hipLaunchKernelGGL((rr_pre_kernel<T, TF, TC, OP_, K_>), dim3(grid), dim3(64 * kWaves), shmem, s, df->d_patches, \
                     (const T*)u, (const T*)coeffs, (const TF*)f, Lf, neumann, link->d_child, link->d_nchild,      \
                     link->d_orth, dc->d_patches, mats, s2, (TC*)fc, (TC*)uc, Lc, (const int4*)df->d_ring)
  if (op_kind == 0) {
    if (g->kind == TE_BILINEAR) TE_RP(0, TE_BILINEAR);
    else TE_RP(0, TE_BIQUADRATIC);
  } else {
    if (g->kind == TE_BILINEAR) TE_RP(1, TE_BILINEAR);
    else TE_RP(1, TE_BIQUADRATIC);
  }
#undef TE_RP
}

// The outer BiCGStab's s update (MODE BICG_S: s = r - alpha ap, with the s.s partials) or p
// update (BICG_P: p = (p - omega ap) beta + r) fused with the zero-guess sweep of the finest
// level (16 x 16 patches, FP32 cycle) that the next V-cycle starts with: each wave updates a
// patch in FP64, writes it and its FP32 copy, and solves the patch on the FP32 values it has
// just formed (the same values the separate sweep would read back from the copy), writing
// the level's solution buffer.  The updates are those of bicg_kernel.
// pfold (BICG_S, may be null): the rhat.ap partials of the operator launch; each workgroup
// folds them (as the finalize would, same order and bits) into alpha = rho / rhat.ap, and
// workgroup 0 stores alpha, so that no separate finalize launch is needed.
template <int MODE>
__global__ void __launch_bounds__(64 * kWaves) bicg_sweep_kernel(const DevPatch* __restrict__ dp,
                                                                 const float* __restrict__ mats,
                                                                 const float* __restrict__ s2tab,
                                                                 double* __restrict__ scal,
                                                                 double* __restrict__ v0, const double* __restrict__ a0,
                                                                 const double* __restrict__ a1, float* __restrict__ w32,
                                                                 float* __restrict__ out, double* __restrict__ pc,
                                                                 int nslots, int np, const double* __restrict__ pfold,
                                                                 int nfold)
{
  static_assert(64 * kWaves == kFoldThreads, "the in-kernel fold needs the fold's block size");
  constexpr int n = 16, ld = n + 1, SX = n + 2, PS = SX * SX;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float* Sm = reinterpret_cast<float*>(smem_raw);
  float* S3 = Sm + n * ld;
  float* s2 = S3 + n * ld;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float* X = s2 + n + wave * n * ld;
  // the scalars are loaded together with the transform matrices (the partials of alpha
  // after them: held across the staging they would cost a wave per SIMD of occupancy)
  const bool fold = MODE == BICG_S && pfold != nullptr;
  double alpha = fold ? 0.0 : scal[SC_ALPHA];
  const double omega = scal[SC_OMEGA], beta = scal[SC_BETA], rho = scal[SC_RHO];
  const double state = scal[SC_STATE];
  for (int k = threadIdx.x; k < n * n; k += blockDim.x) {
    const int row = k / n, col = k - row * n;
    Sm[row * ld + col] = mats[(3 * 3 + 0) * kDftMatStride + k];  // DST_II, axis 0
    S3[row * ld + col] = mats[(4 * 3 + 0) * kDftMatStride + k];  // DST_III, axis 0
  }
  for (int k = threadIdx.x; k < n; k += blockDim.x) s2[k] = s2tab[(2 * 3 + 0) * kDftMatDim + k];
  __syncthreads();
  const bool run = state == ST_RUN;
  if (fold) {
    const double* const parts[3] = { pfold, nullptr, nullptr };
    double r[3] = { 0.0, 0.0, 0.0 };
    fold_n<1>(parts, nfold, r);
    alpha = rho / r[0];  // FIN_ALPHA: alpha = rho / rhat.dot(ap)
    if (run && blockIdx.x == 0 && threadIdx.x == 0) scal[SC_ALPHA] = alpha;
  }
  double sa = 0.0;
  if (run) {
    const XcdRange xr = xcd_range(blockIdx.x, gridDim.x, np, kWaves);
    for (int p = xr.first + wave; p < xr.end; p += xr.step) {
      const int base = p * PS;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int q = lane + 64 * k, y = q >> 4, x = q & 15;
        const int i = base + (y + 1) * SX + x + 1;
        double v;
        if (MODE == BICG_S) {  // v0 = s, a0 = r, a1 = ap
          v = at(a0, i);
          v += at(a1, i) * -alpha;
          sa += v * v;
        } else {  // v0 = p, a0 = ap, a1 = r
          double pv = at(v0, i);
          pv += at(a0, i) * -omega;
          v = pv * beta + at(a1, i);
        }
        at(v0, i) = v;
        const float v32 = float(v);
        at(w32, i) = v32;
        X[y * ld + x] = v32;
      }
      wave_sync();
      const DevPatch& P = dp[p];
      const float lx = float(4.0 / (P.h[0] * P.h[0])), ly = float(4.0 / (P.h[1] * P.h[1]));
      dft_solve_wave<float, float, 1>(X, Sm, S3, s2, lx, ly, out, base, SX, true);
      wave_sync();  // X is rewritten by the next patch
    }
  }
  if (MODE == BICG_S) {
    const double r = block_reduce<RED_SUM>(sa);
    if (threadIdx.x == 0) pc[blockIdx.x] = r;
    if (blockIdx.x == 0)
      for (int i = (int)gridDim.x + (int)threadIdx.x; i < nslots; i += (int)blockDim.x) pc[i] = 0.0;
  }
}

}  // namespace

bool bicg_sweep_supported(const te_dft_plan* pl)
{
  const te_domain* d = pl->g->d;
  return d->dim == 2 && d->ns[0] == 16 && d->ns[1] == 16 && pl->use_mfma && pl->mask == 0 && pl->d_mats32 &&
         pl->d_s2_32;
}

int launch_bicg_sweep(int mode, const te_dft_plan* pl, double* scal, double* v0, const double* a0,
                      const double* a1, float* w32, float* out, double* pc, int nslots, hipStream_t s,
                      const double* pfold, int nfold)
{
  const te_domain* d = pl->g->d;
  if (d->np == 0) return TE_OK;
  if (!bicg_sweep_supported(pl) || (mode != BICG_S && mode != BICG_P) || !w32 || !out ||
      (mode == BICG_S && (!pc || nslots < 1)) || (pfold && (mode != BICG_S || nfold < 1 || nfold > kFoldMax))) {
    set_error("internal: fused BiCGStab update + sweep: 16 x 16 matrix-core plan, s or p update, FP32 buffers");
    return TE_ERR_ARG;
  }
  constexpr int n = 16;
  const size_t shmem = sizeof(float) * (2 * n * (n + 1) + n + kWaves * n * (n + 1));
  const int ngroups = (d->np + kWaves - 1) / kWaves;
  const void* k = mode == BICG_S ? (const void*)bicg_sweep_kernel<BICG_S> : (const void*)bicg_sweep_kernel<BICG_P>;
  int grid = resident_blocks(k, 64 * kWaves, shmem);
  if (grid > ngroups) grid = ngroups;
  if (mode == BICG_S && grid > nslots) grid = nslots;
  if (mode == BICG_S)
    hipLaunchKernelGGL((bicg_sweep_kernel<BICG_S>), dim3(grid), dim3(64 * kWaves), shmem, s, d->d_patches,
                       pl->d_mats32, pl->d_s2_32, scal, v0, a0, a1, w32, out, pc, nslots, d->np, pfold, nfold);
  else
    hipLaunchKernelGGL((bicg_sweep_kernel<BICG_P>), dim3(grid), dim3(64 * kWaves), shmem, s, d->d_patches,
                       pl->d_mats32, pl->d_s2_32, scal, v0, a0, a1, w32, out, pc, nslots, d->np, nullptr, 0);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

bool dft_mfma_supported(const te_dft_plan* pl)
{
  const te_domain* d = pl->g->d;
  return d->dim == 2 && d->ns[0] == d->ns[1] && (d->ns[0] == 16 || d->ns[0] == 32) && pl->mask == 0;
}

int launch_dft_mfma(const te_dft_plan* pl, int dt, int dtf, int dto, const void* f, const void* u, void* out,
                    bool uzero, hipStream_t s)
{
  return launch_dft_mfma_fused(pl, dt, dtf, dto, f, u, out, uzero, 0, TE_F64, nullptr, nullptr, s);
}

int launch_dft_mfma_fused(const te_dft_plan* pl, int dt, int dtf, int dto, const void* f, const void* u, void* out,
                          bool uzero, int gm, int dtc, const void* uc, const te_level_link* link, hipStream_t s)
{
  if (pl->g->d->np == 0) return TE_OK;
  if (gm == 2 && (!uc || !link)) {
    set_error("fused interpolation needs the coarser level's u and the level link");
    return TE_ERR_ARG;
  }
  const bool c64 = gm != 2 || dtc == TE_F64;
#define TE_LT(T_, TF_, TO_)                                                                      \
  do {                                                                                           \
    if (c64) launch_t<T_, TF_, TO_, double>(pl, f, u, out, uzero, gm, uc, link, s);              \
    else launch_t<T_, TF_, TO_, float>(pl, f, u, out, uzero, gm, uc, link, s);                   \
  } while (0)
  if (dt == TE_F64 && dtf == TE_F64 && dto == TE_F64) {
    if (!c64) {
      set_error("FP64 level with an FP32 coarser level is not supported");
      return TE_ERR_UNSUPPORTED;
    }
    launch_t<double, double, double, double>(pl, f, u, out, uzero, gm, uc, link, s);
  } else if (dt == TE_F32 && dtf == TE_F32 && dto == TE_F32) {
    TE_LT(float, float, float);
  } else if (dt == TE_F32 && dtf == TE_F64 && dto == TE_F32) {
    TE_LT(float, double, float);
  } else if (dt == TE_F32 && dtf == TE_F64 && dto == TE_F64) {
    TE_LT(float, double, double);
  } else {
    set_error("unsupported dtype combination for the DFT patch solver");
    return TE_ERR_UNSUPPORTED;
  }
#undef TE_LT
  TE_HIP(hipGetLastError());
  return TE_OK;
}


bool rr_pre_supported(const te_ghost_filler* gf, const te_dft_plan* pc)
{
  const te_domain* df = gf->d;
  const te_domain* dc = pc->g->d;
  return df->dim == 2 && df->ns[0] == 16 && df->ns[1] == 16 && dc->ns[0] == 16 && dc->ns[1] == 16 &&
         (gf->kind == TE_BILINEAR || gf->kind == TE_BIQUADRATIC) && pc->use_mfma && dft_mfma_supported(pc);
}

// dt: fine u / coeffs type; dtf: fine f type; dtc: coarse level type
int launch_rr_pre(const te_ghost_filler* g, const te_level_link* link, int op_kind, int neumann,
                  const void* coeffs, int dt, int dtf, const void* f, const void* u, const te_dft_plan* pc, int dtc,
                  void* fc, void* uc, hipStream_t s)
{
  if (pc->g->d->np == 0) return TE_OK;
  if (!rr_pre_supported(g, pc) || link->fine != g->d || link->coarse != pc->g->d) {
    set_error("fused restriction + coarse smoothing: 2D 16x16 patches, matrix-core plan, matching link");
    return TE_ERR_UNSUPPORTED;
  }
  if (dt == TE_F64 && dtf == TE_F64 && dtc == TE_F64) launch_rp<double, double, double>(g, link, op_kind, neumann, coeffs, f, u, pc, fc, uc, s);
  else if (dt == TE_F32 && dtf == TE_F32 && dtc == TE_F32) launch_rp<float, float, float>(g, link, op_kind, neumann, coeffs, f, u, pc, fc, uc, s);
  else if (dt == TE_F32 && dtf == TE_F64 && dtc == TE_F32) launch_rp<float, double, float>(g, link, op_kind, neumann, coeffs, f, u, pc, fc, uc, s);
  else if (dt == TE_F32 && dtf == TE_F32 && dtc == TE_F64) launch_rp<float, float, double>(g, link, op_kind, neumann, coeffs, f, u, pc, fc, uc, s);
  else if (dt == TE_F32 && dtf == TE_F64 && dtc == TE_F64) launch_rp<float, double, double>(g, link, op_kind, neumann, coeffs, f, u, pc, fc, uc, s);
  else {
    set_error("unsupported dtype combination for the fused restriction + coarse smoothing");
    return TE_ERR_UNSUPPORTED;
  }
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
