// The coarsest levels of the fused 2D V-cycle in one workgroup (VCycle::visit,
// GMG/VCycle.h:46-70, from the first tail level down): every level's f and its two
// solution buffers live in LDS; a level's patches are spread over the waves (one patch
// per wave at a time), and a workgroup barrier separates the steps.  The steps are the
// per-level launches' arithmetic, in their order and precision:
//   sweep      DFTPatchSolver::smooth (the interface terms of the ghosts computed from
//              the neighbours' values, as dft_mfma_kernel does with ghosts on the fly;
//              the first sweep of a level from u = 0 solves f alone);
//   residual   r = f - A u with the ghost formulas of resid_restrict_2d, restricted into
//              the next level's f (LinearRestrictor, GMG/LinearRestrictor.h:101-191);
//   interpolate u += P uc (DirectInterpolator, GMG/DirectInterpolator.h:60-91) into the
//              level's other buffer, which the first post-smoothing sweep then reads
//              (dft_mfma_kernel's fused interpolation forms the same values on the fly).
#include "../te_internal.h"
#include "dft.h"
#include "dft_wave.h"
#include "ghost2d.h"
#include "tail2d.h"

namespace te {
namespace {

constexpr int kN = 16, kLd = kN + 1, kSX = kN + 2, kPS = kSX * kSX;
constexpr int kScr = kN * kLd + 4 * kN + kPS;  // per-wave scratch: X, interface terms, tile

// what the steps read of a patch, copied into LDS at the start with the ring table, so
// that after the first loads every step works from LDS alone
struct TailPatch {
  double h[2];
  double ih2[2];
  int type[4];  // side neighbour types (PatchInfo slots 4..7)
  int parent, porth;
};

struct LevelLds {
  const TailPatch* meta;
  const int4* ring;
};

__device__ __forceinline__ LevelLds level_lds(const unsigned char* lds, const TailLevelArgs& lv)
{
  return LevelLds{ reinterpret_cast<const TailPatch*>(lds + lv.meta), reinterpret_cast<const int4*>(lds + lv.ringl) };
}

template <typename S>
struct Scratch {
  S* X;
  S* Bt;
  S* Tl;
};

template <typename S>
__device__ __forceinline__ Scratch<S> scratch(unsigned char* lds, int off, int wave)
{
  S* b = reinterpret_cast<S*>(lds + off) + wave * kScr;
  return Scratch<S>{ b, b + kN * kLd, b + kN * kLd + 4 * kN };
}

// one DFT sweep of patch p: out = solve(f - interface terms of u's ghosts); uzero: u = 0
template <typename S, int KIND>
__device__ void sweep_patch(int p, bool uzero, const LevelLds& lv, const S* F, const S* U, S* out,
                            const S* Sm, const S* S3, const S* s2, const Scratch<S>& w, int lane)
{
  const int base = p * kPS;
  const TailPatch& P = lv.meta[p];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int q = lane + 64 * k, y = q / kN, x = q - y * kN;
    w.X[y * kLd + x] = F[base + (y + 1) * kSX + x + 1];
  }
  if (!uzero) {
    const int rs = lane / kN, rt = lane - (lane / kN) * kN;  // ring cell = lane (4 n = 64)
    const int4 rc = at(lv.ring, (p * 4 + rs) * 2 + (2 * rt >= kN ? 1 : 0));
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int q = lane + 64 * k, y = q / kN, x = q - y * kN;
      w.Tl[(y + 1) * kSX + x + 1] = U[base + (y + 1) * kSX + x + 1];
    }
    S gv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) gv[k] = S(0);
    ghost2d_gather<KIND>(rc.x, rc.y, rc.z, rs, rt, kN, kN,
                         [&](int k, int q, int x, int y) { gv[k] = U[q * kPS + (y + 1) * kSX + x + 1]; });
    wave_sync();
    S term = S(0);
    if (rc.x >= 0) {
      const int s = rs, t = rt, ph = s >> 1, up = s & 1;
      int bx, by;
      if (ph == 0) {
        bx = up ? kN - 1 : 0;
        by = t;
      } else {
        bx = t;
        by = up ? kN - 1 : 0;
      }
      const S ug = ghost2d_combine<S, KIND>(rc.x, rc.y, t, kN,
                                            [&](int k, int tt) {
                                              int x, y;
                                              side_xy(s, k, tt, kN, &x, &y);
                                              return w.Tl[(y + 1) * kSX + x + 1];
                                            },
                                            gv);
      const S ub = w.Tl[(by + 1) * kSX + bx + 1];
      term = (ug + ub) * S(P.ih2[ph]);
    }
    w.Bt[lane] = term;
    wave_sync();
    // west / east first, then south / north (Poisson/StarPatchOperator.h:165-180)
#pragma unroll
    for (int ph = 0; ph < 2; ph++) {
      if (lane < 2 * kN) {
        const int up = lane / kN, t = lane - up * kN;
        const int s = 2 * ph + up;
        if (P.type[s] >= 0) {
          const int xi = ph == 0 ? t * kLd + (up ? kN - 1 : 0) : (up ? kN - 1 : 0) * kLd + t;
          w.X[xi] -= w.Bt[s * kN + t];
        }
      }
      wave_sync();
    }
  }
  wave_sync();
  const S lx = S(4.0 / (P.h[0] * P.h[0])), ly = S(4.0 / (P.h[1] * P.h[1]));
  dft_solve_wave<S, S, 1>(w.X, Sm, S3, s2, lx, ly, out, base, kSX, true);
  wave_sync();
}

// r = f - A u on patch p of level lv, restricted into fc (the next level's f)
template <typename S, typename SC, int KIND>
__device__ void residual_patch(int p, const LevelLds& lv, const S* F, const S* U, SC* fc, S phys,
                               const Scratch<S>& w, int lane)
{
  constexpr int h = kN / 2;
  Layout L{};
  L.n[0] = L.n[1] = kN;
  L.n[2] = 1;
  L.sx = kSX;
  L.ps = L.cs = kPS;
  int rs, rt;
  const int4 rn = at(lv.ring, ring_index(p, lane, kN, kN, &rs, &rt));
  const int o = lv.meta[p].porth, cp = lv.meta[p].parent;
  S fv[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int q = lane + 64 * k;
    w.Tl[((q >> 4) + 1) * kSX + (q & 15) + 1] = U[p * kPS + ((q >> 4) + 1) * kSX + (q & 15) + 1];
    const int x = 2 * (lane & 7) + (k & 1), y = 2 * (lane >> 3) + (k >> 1);
    fv[k] = F[p * kPS + (y + 1) * kSX + x + 1];
  }
  S gu[4];
  ring_gather_plain<S, KIND>(rn, lane, kN, kN, U, L, gu);
  wave_sync();
  tile_ghost_gathered<S, KIND>(rn, lane, kN, kN, w.Tl, phys, gu);
  wave_sync();
  const TailPatch& P = lv.meta[p];
  const S ih2[2] = { S(P.ih2[0]), S(P.ih2[1]) };
  const int ox = (o & 1) ? h : 0, oy = (o & 2) ? h : 0;
  const int X = lane & 7, Y = lane >> 3;  // a lane per 2x2 block of fine cells
  S r[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int x = 2 * X + (j & 1), y = 2 * Y + (j >> 1);
    r[j] = fv[j] - star_tile<S, 0>(w.Tl, kSX, x, y, ih2, nullptr, 0, kSX);
  }
  if (o < 0) {  // copyToParent
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int x = 2 * X + (j & 1), y = 2 * Y + (j >> 1);
      fc[cp * kPS + (y + 1) * kSX + x + 1] = SC(r[j]);
    }
  } else {  // restrictToCoarserParent, fine cells in x-fastest order
    S acc = S(0);
#pragma unroll
    for (int j = 0; j < 4; j++) acc += r[j] * S(0.25);
    fc[cp * kPS + (Y + oy + 1) * kSX + X + ox + 1] = SC(acc);
  }
  wave_sync();
}

// out = u + P uc on patch p's interior (out may be u)
template <typename S, typename SC>
__device__ void interp_patch(int p, const LevelLds& lv, const S* U, const SC* Uc, S* out, int lane)
{
  const int o = lv.meta[p].porth, cp = lv.meta[p].parent;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int q = lane + 64 * k, y = q / kN, x = q - y * kN;
    int cx = x, cy = y;
    if (o >= 0) {
      cx = (x + ((o & 1) ? kN : 0)) / 2;
      cy = (y + ((o & 2) ? kN : 0)) / 2;
    }
    const int i = p * kPS + (y + 1) * kSX + x + 1;
    out[i] = S(U[i] + S(Uc[cp * kPS + (cy + 1) * kSX + cx + 1]));
  }
}

template <typename S>
__device__ __forceinline__ S* buf(unsigned char* lds, const TailLevelArgs& lv, int which)
{
  return reinterpret_cast<S*>(lds + (which ? lv.u1 : lv.u0));
}

template <typename S>
__device__ __forceinline__ S* fbuf(unsigned char* lds, const TailLevelArgs& lv)
{
  return reinterpret_cast<S*>(lds + lv.f);
}

template <typename S>
__device__ void load_mats(unsigned char* lds, int off, const S* mats, const S* s2tab)
{
  S* Sm = reinterpret_cast<S*>(lds + off);
  S* S3 = Sm + kN * kLd;
  S* s2 = S3 + kN * kLd;
  for (int k = threadIdx.x; k < kN * kN; k += blockDim.x) {
    const int row = k / kN, col = k - row * kN;
    Sm[row * kLd + col] = mats[(3 * 3 + 0) * kDftMatStride + k];  // DST_II, axis 0
    S3[row * kLd + col] = mats[(4 * 3 + 0) * kDftMatStride + k];  // DST_III, axis 0
  }
  for (int k = threadIdx.x; k < kN; k += blockDim.x) s2[k] = s2tab[(2 * 3 + 0) * kDftMatDim + k];
}

// T: type of every tail level but the last, TL: the last level's
template <typename T, typename TL, int KIND>
__global__ void __launch_bounds__(64 * kTailWaves) dft_tail_kernel(TailArgs a)
{
  __shared__ __align__(16) unsigned char lds[kTailLds];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr bool SAME = sizeof(T) == sizeof(TL);
  const int nl = a.nl, last = nl - 1;
  load_mats<T>(lds, a.mat_t, (const T*)a.mats_t, (const T*)a.s2_t);
  if (!SAME) load_mats<TL>(lds, a.mat_l, (const TL*)a.mats_l, (const TL*)a.s2_l);
  {
    const T* fin = (const T*)a.f_in;
    T* F0 = fbuf<T>(lds, a.lv[0]);
    for (int i = threadIdx.x; i < a.lv[0].np * kPS; i += blockDim.x) F0[i] = fin[i];
  }
  for (int l = 0; l < nl; l++) {
    const TailLevelArgs& lv = a.lv[l];
    TailPatch* meta = reinterpret_cast<TailPatch*>(lds + lv.meta);
    int4* ring = reinterpret_cast<int4*>(lds + lv.ringl);
    for (int i = threadIdx.x; i < lv.np * 8; i += blockDim.x) ring[i] = lv.ring[i];
    for (int p = threadIdx.x; p < lv.np; p += blockDim.x) {
      const DevPatch& P = lv.dp[p];
      TailPatch m;
      m.h[0] = P.h[0];
      m.h[1] = P.h[1];
      m.ih2[0] = P.inv_h2[0];
      m.ih2[1] = P.inv_h2[1];
      for (int s = 0; s < 4; s++) m.type[s] = P.type[4 + s];
      m.parent = l < nl - 1 ? lv.parent[p] : 0;
      m.porth = l < nl - 1 ? lv.porth[p] : -1;
      meta[p] = m;
    }
  }
  __syncthreads();
  const T* SmT = reinterpret_cast<const T*>(lds + a.mat_t);
  const TL* SmL = reinterpret_cast<const TL*>(lds + a.mat_l);
  const Scratch<T> wt = scratch<T>(lds, a.scr_t, wave);
  const Scratch<TL> wl = scratch<TL>(lds, a.scr_l, SAME ? wave : (wave < kTailLastWaves ? wave : 0));
  const int lw = SAME ? kTailWaves : kTailLastWaves;  // waves that work on the last level
  const T physT = a.neumann ? T(1) : T(-1);
  unsigned curm = 0;  // bit l: which buffer holds level l's solution
  auto cur = [&](int l) { return (int)((curm >> l) & 1u); };

  // one sweep of level l from its current buffer into the other one
  auto sweep = [&](int l, bool uzero) {
    const TailLevelArgs& lv = a.lv[l];
    if (l == last) {
      if (wave < lw)
        for (int p = wave; p < lv.np; p += lw)
          sweep_patch<TL, KIND>(p, uzero, level_lds(lds, lv), fbuf<TL>(lds, lv), buf<TL>(lds, lv, cur(l)), buf<TL>(lds, lv, cur(l) ^ 1),
                                SmL, SmL + kN * kLd, SmL + 2 * kN * kLd, wl, lane);
    } else {
      for (int p = wave; p < lv.np; p += kTailWaves)
        sweep_patch<T, KIND>(p, uzero, level_lds(lds, lv), fbuf<T>(lds, lv), buf<T>(lds, lv, cur(l)), buf<T>(lds, lv, cur(l) ^ 1),
                             SmT, SmT + kN * kLd, SmT + 2 * kN * kLd, wt, lane);
    }
    __syncthreads();
    curm ^= 1u << l;
  };
  auto zero = [&](int l) {  // u = 0 (interior) in buffer 0
    const TailLevelArgs& lv = a.lv[l];
    for (int i = threadIdx.x; i < lv.np * kPS; i += blockDim.x) {
      if (l == last) buf<TL>(lds, lv, 0)[i] = TL(0);
      else buf<T>(lds, lv, 0)[i] = T(0);
    }
    __syncthreads();
    curm &= ~(1u << l);
  };

  // down: pre-smoothing (the first sweep from u = 0), residual, restriction
  for (int l = 0; l < nl; l++) {
    const int ns = l == last ? a.coarse : a.pre;
    if (ns == 0) zero(l);
    for (int k = 0; k < ns; k++) sweep(l, k == 0);
    if (l < last) {
      const TailLevelArgs& lv = a.lv[l];
      for (int p = wave; p < lv.np; p += kTailWaves) {
        if (l + 1 == last)
          residual_patch<T, TL, KIND>(p, level_lds(lds, lv), fbuf<T>(lds, lv), buf<T>(lds, lv, cur(l)), fbuf<TL>(lds, a.lv[l + 1]),
                                      physT, wt, lane);
        else
          residual_patch<T, T, KIND>(p, level_lds(lds, lv), fbuf<T>(lds, lv), buf<T>(lds, lv, cur(l)), fbuf<T>(lds, a.lv[l + 1]),
                                     physT, wt, lane);
      }
      __syncthreads();
    }
  }
  // up: interpolation, then post-smoothing
  for (int l = last - 1; l >= 0; l--) {
    const TailLevelArgs& lv = a.lv[l];
    const int dst = a.post == 0 ? cur(l) : cur(l) ^ 1;
    for (int p = wave; p < lv.np; p += kTailWaves) {
      if (l + 1 == last)
        interp_patch<T, TL>(p, level_lds(lds, lv), buf<T>(lds, lv, cur(l)), buf<TL>(lds, a.lv[l + 1], cur(l + 1)),
                            buf<T>(lds, lv, dst), lane);
      else
        interp_patch<T, T>(p, level_lds(lds, lv), buf<T>(lds, lv, cur(l)), buf<T>(lds, a.lv[l + 1], cur(l + 1)), buf<T>(lds, lv, dst),
                           lane);
    }
    __syncthreads();
    curm = (curm & ~(1u << l)) | ((unsigned)dst << l);
    for (int k = 0; k < a.post; k++) sweep(l, false);
  }
  // the first tail level's solution (interior) for the finer level's interpolation
  {
    const T* U0 = buf<T>(lds, a.lv[0], cur(0));
    T* uo = (T*)a.u_out;
    for (int i = threadIdx.x; i < a.lv[0].np * kN * kN; i += blockDim.x) {
      const int p = i / (kN * kN), q = i - p * (kN * kN), y = q / kN, x = q - y * kN;
      const int j = p * kPS + (y + 1) * kSX + x + 1;
      uo[j] = U0[j];
    }
  }
}

int align16(int b) { return (b + 15) & ~15; }

}  // namespace

int tail_layout(TailArgs* a, const int* np, int nl, int dt, int dtl)
{
  const int st = dt == TE_F64 ? 8 : 4, sl = dtl == TE_F64 ? 8 : 4;
  const bool same = st == sl;
  const int mats = kN * kLd * 2 + kN;
  int off = 0;
  a->mat_t = off;
  off = align16(off + mats * st);
  a->mat_l = same ? a->mat_t : off;
  if (!same) off = align16(off + mats * sl);
  a->scr_t = off;
  off = align16(off + kTailWaves * kScr * st);
  a->scr_l = same ? a->scr_t : off;
  if (!same) off = align16(off + kTailLastWaves * kScr * sl);
  for (int l = 0; l < nl; l++) {
    const int sz = l == nl - 1 ? sl : st;
    const int b = align16(np[l] * kPS * sz);
    a->lv[l].u0 = off;
    a->lv[l].u1 = off + b;
    a->lv[l].f = off + 2 * b;
    off += 3 * b;
    a->lv[l].meta = off;
    off = align16(off + np[l] * (int)sizeof(TailPatch));
    a->lv[l].ringl = off;
    off = align16(off + np[l] * 8 * (int)sizeof(int4));
  }
  return off;
}

int launch_dft_tail(const TailArgs& a, int dt, int dtl, int kind, hipStream_t s)
{
  if (a.nl < 2 || a.nl > kTailMaxLevels || a.lv[0].np > kTailMaxPatches) {
    set_error("internal: tail levels out of range");
    return TE_ERR_ARG;
  }
#define TE_TK(T_, TL_, K_) \
  hipLaunchKernelGGL((dft_tail_kernel<T_, TL_, K_>), dim3(1), dim3(64 * kTailWaves), 0, s, a)
  const bool bq = kind == TE_BIQUADRATIC;
  if (dt == TE_F32 && dtl == TE_F64) {
    if (bq) TE_TK(float, double, TE_BIQUADRATIC);
    else TE_TK(float, double, TE_BILINEAR);
  } else if (dt == TE_F32 && dtl == TE_F32) {
    if (bq) TE_TK(float, float, TE_BIQUADRATIC);
    else TE_TK(float, float, TE_BILINEAR);
  } else if (dt == TE_F64 && dtl == TE_F64) {
    if (bq) TE_TK(double, double, TE_BIQUADRATIC);
    else TE_TK(double, double, TE_BILINEAR);
  } else {
    set_error("internal: unsupported tail dtypes");
    return TE_ERR_UNSUPPORTED;
  }
#undef TE_TK
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
