// Poisson::DFTPatchSolver::solveSinglePatch (Poisson/DFTPatchSolver.h:420-450) for
// every patch of a level in one launch, generic (VALU) form for any n: a workgroup holds
// PPW patches, TPP threads each (a launch variant chosen by the autotuner's table, tune.h).
// Each patch is staged in LDS; each axis transform y = M x (executePlan, :205-265) is a
// dense n x n product over all lines of that axis; eigenvalues (:320-361) come from a
// per-axis sin^2 table.
//
// Template parameters: T = arithmetic / u type, TF = type of f, TO = type of
// the output.  UZERO: u is known to be identically zero (the first smoothing
// sweep of a freshly zeroed level), so the interface terms of
// modifyRHSForInternalBoundaryConditions vanish and u is not read.
#include "../te_internal.h"
#include "dft.h"
#include "dft_wave.h"
#include "ghost2d.h"
#include "../tune.h"

namespace te {
namespace {

// GM (2D): 0 = the ghost values of u are read from u (filled beforehand); 1 = they are
// computed here from the neighbours' cells (ghost2d, the filler's closed forms); 2 = the same
// on u + P uc (DirectInterpolator applied on the fly: uc is the coarser level's solution).
// With GM != 0, out must not alias u.
template <typename T, typename TF, typename TO, int D, int OP, bool UZERO, int GM = 0, int KIND = 0, typename TC = T>
__global__ void __launch_bounds__(256) dft_kernel(const DevPatch* __restrict__ dp, const int* __restrict__ ptype,
                                                  const T* __restrict__ mats, const T* __restrict__ s2tab,
                                                  const TF* __restrict__ f, const T* __restrict__ u,
                                                  const T* __restrict__ coeffs, TO* __restrict__ out, Layout L,
                                                  int ppw, FusedArgs<T, TC> fa)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n0 = L.n[0], n1 = L.n[1], n2 = D == 3 ? L.n[2] : 1;
  const int N = n0 * n1 * n2;
  const int nmax = n0 > n1 ? (n0 > n2 ? n0 : n2) : (n1 > n2 ? n1 : n2);
  // this thread's patch (sub-block of tpp threads) and its index t within it; the threads
  // of a patch past the last one take part in the barriers only
  const int tpp = blockDim.x / ppw, sub = threadIdx.x / tpp, t = threadIdx.x - sub * tpp;
  const bool live = blockIdx.x * ppw + sub < L.np;
  const int p = live ? blockIdx.x * ppw + sub : L.np - 1;
  T* sm = reinterpret_cast<T*>(smem_raw) + (size_t)sub * (2 * N + 2 * D * nmax * nmax);
  T* A = sm;
  T* B = sm + N;
  T* MF = sm + 2 * N;                 // forward matrices, D x nmax^2
  T* MI = MF + D * nmax * nmax;       // inverse matrices
  const int* pt = ptype + p * 8;      // fwd type per axis [0..2], eig type per axis [3..5], zero mode [6]
  const DevPatch& P = dp[p];
  const int nn[3] = { n0, n1, n2 };

  for (int a = 0; a < D; a++) {
    int na = nn[a];
    int tf = pt[a];
    int ti = tf == 0 ? 1 : (tf == 3 ? 4 : tf);  // DCT_II->DCT_III, DST_II->DST_III, IV types self-inverse
    for (int k = t; k < na * na; k += tpp) {
      MF[a * nmax * nmax + k] = mats[(tf * 3 + a) * kDftMatStride + k];
      MI[a * nmax * nmax + k] = mats[(ti * 3 + a) * kDftMatStride + k];
    }
  }
  // f_copy = f, then modifyRHSForInternalBoundaryConditions (Poisson :165-180, VarPoisson :168-186)
  for (int o = t; o < N; o += tpp) {
    int x = o % n0;
    int y = (o / n0) % n1;
    int z = D == 3 ? o / (n0 * n1) : 0;
    int gi = lin(L, p, 0, x, y, z);
    T v = T(f[gi]);
    if (!UZERO) {
      const int c3[3] = { x, y, z };
      const int st[3] = { 1, L.sx, L.sy };
      const int sbase = D == 2 ? 4 : 20;
#pragma unroll
      for (int a = 0; a < D; a++) {
#pragma unroll
        for (int up = 0; up < 2; up++) {
          bool edge = up ? (c3[a] == nn[a] - 1) : (c3[a] == 0);
          if (!edge || P.type[sbase + 2 * a + up] < 0) continue;
          int gidx = up ? gi + st[a] : gi - st[a];
          T ih2 = T(P.inv_h2[a]);
          T ug, ub;  // the ghost and the boundary cell
          if (GM == 0) {
            ug = u[gidx];
            ub = u[gi];
          } else {
            const UAcc<T, TC, GM == 2> U{ u, fa.uc, fa.parent, fa.porth, L };
            const int sd = 2 * a + up;
            ug = ghost2d<T, KIND>(side_info(P, sd), p, sd, a == 0 ? y : x, nn[a], nn[1 - a], U);
            ub = U(p, x, y);
          }
          if (OP == 0) {
            v -= (ug + ub) * ih2;
          } else {
            int cq = gi - p * L.ps;
            int cg = up ? cq + st[a] : cq - st[a];
            T cb = coeffs[p * L.cs + cq], cgv = coeffs[p * L.cs + cg];
            v -= (ug + ub) * (cb + cgv) * (ih2 * T(0.5));
          }
        }
      }
    }
    A[o] = v;
  }
  __syncthreads();
  // forward transform along each axis (DST-II / DCT-II / DCT-IV / DST-IV)
  T* src = A;
  T* dst = B;
  for (int a = 0; a < D; a++) {
    int na = nn[a];
    int sta = a == 0 ? 1 : (a == 1 ? n0 : n0 * n1);
    const T* M = MF + a * nmax * nmax;
    for (int o = t; o < N; o += tpp) {
      int k = (o / sta) % na;
      int base = o - k * sta;
      T acc = T(0);
      for (int j = 0; j < na; j++) acc += M[k * na + j] * src[base + j * sta];
      dst[o] = acc;
    }
    __syncthreads();
    T* t = src;
    src = dst;
    dst = t;
  }
  // divide by the eigenvalues
  for (int o = t; o < N; o += tpp) {
    int k3[3] = { o % n0, (o / n0) % n1, D == 3 ? o / (n0 * n1) : 0 };
    T eig = T(0);
    for (int a = 0; a < D; a++) {
      double h = P.h[a];
      eig -= T(4.0 / (h * h)) * s2tab[(pt[3 + a] * 3 + a) * kDftMatDim + k3[a]];
    }
    src[o] /= eig;
  }
  __syncthreads();
  if (pt[6] && t == 0) src[0] = T(0);  // all-Neumann patch without neighbours (:438-440)
  __syncthreads();
  for (int a = 0; a < D; a++) {
    int na = nn[a];
    int sta = a == 0 ? 1 : (a == 1 ? n0 : n0 * n1);
    const T* M = MI + a * nmax * nmax;
    for (int o = t; o < N; o += tpp) {
      int k = (o / sta) % na;
      int base = o - k * sta;
      T acc = T(0);
      for (int j = 0; j < na; j++) acc += M[k * na + j] * src[base + j * sta];
      dst[o] = acc;
    }
    __syncthreads();
    T* t = src;
    src = dst;
    dst = t;
  }
  T scale = T(1);
  for (int a = 0; a < D; a++) scale *= T(2.0) / T(nn[a]);
  for (int o = t; o < N; o += tpp) {
    int x = o % n0;
    int y = (o / n0) % n1;
    int z = D == 3 ? o / (n0 * n1) : 0;
    if (live) out[lin(L, p, 0, x, y, z)] = TO(src[o] * scale);
  }
}

// The same solve for 3D patches of 16^3 cells on the matrix cores: every axis transform is
// 16 products of 16 x 16 matrices (the lines of one slice, or of one row of slices, against
// the axis matrix), 4 per wave, each a chain of 16x16x4 MFMAs with operands read from LDS.
// The patch lives in LDS with padded pitches (rows of 17, slices of 16 * 17 + 1) so that a
// lane group's operand reads spread over the banks, in one buffer: a wave's products read
// and write only its own slices (x and y transforms) or rows (z), and each product reads
// all its operands before it writes, so the transforms run in place.  Interface terms,
// eigenvalues, the per-patch transform types and the all-Neumann zero mode are the generic
// kernel's.
constexpr int k3N = 16, k3Row = 17, k3Slice = 16 * 17 + 1, k3Cells = 16 * k3Slice, k3Mat = 16 * 17;

__device__ __forceinline__ int c3(int x, int y, int z) { return x + k3Row * y + k3Slice * z; }

// one 16 x 16 x 16 product, D[i][j] = sum_k A(i, k) B(k, j), written by put(i, j, value)
template <typename T, class GA, class GB, class Put>
__device__ __forceinline__ void mm16(GA A, GB B, Put put, int lane)
{
  typedef typename Mf<T>::V V;
  const int c = lane & 15, g = lane >> 4;
  V acc = { 0, 0, 0, 0 };
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int k = 4 * s + g;
    acc = Mf<T>::mma(A(c, k), B(k, c), acc);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) put(Mf<T>::row(r, g), c, acc[r]);
}

// dst = the transform of src along axis a with matrix M (16 x 16, rows of k3Row); dst may be src
template <typename T>
__device__ __forceinline__ void axis16(int a, const T* M, const T* src, T* dst, int wave, int lane)
{
  for (int q = wave; q < k3N; q += 4) {
    if (a == 0) {  // slice z = q: D[y][k] = sum_x X[y][x] M[k][x]
      mm16<T>([&](int y, int x) { return src[c3(x, y, q)]; }, [&](int x, int k) { return M[k * k3Row + x]; },
              [&](int y, int k, T v) { dst[c3(k, y, q)] = v; }, lane);
    } else if (a == 1) {  // slice z = q: D[k][x] = sum_y M[k][y] X[y][x]
      mm16<T>([&](int k, int y) { return M[k * k3Row + y]; }, [&](int y, int x) { return src[c3(x, y, q)]; },
              [&](int k, int x, T v) { dst[c3(x, k, q)] = v; }, lane);
    } else {  // row y = q: D[k][x] = sum_z M[k][z] X[z][x]
      mm16<T>([&](int k, int z) { return M[k * k3Row + z]; }, [&](int z, int x) { return src[c3(x, q, z)]; },
              [&](int k, int x, T v) { dst[c3(x, q, k)] = v; }, lane);
    }
  }
}

template <typename T, typename TF, typename TO, int OP, bool UZERO>
__global__ void __launch_bounds__(256) dft3_mfma_kernel(const DevPatch* __restrict__ dp, const int* __restrict__ ptype,
                                                        const T* __restrict__ mats, const T* __restrict__ s2tab,
                                                        const TF* __restrict__ f, const T* __restrict__ u,
                                                        const T* __restrict__ coeffs, TO* __restrict__ out, Layout L)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* A = reinterpret_cast<T*>(smem_raw);
  T* MF = A + k3Cells;       // forward matrices of the three axes, rows of k3Row
  T* MI = MF + 3 * k3Mat;    // inverse matrices
  const int p = blockIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int* pt = ptype + p * 8;  // fwd type per axis [0..2], eig type per axis [3..5], zero mode [6]
  const DevPatch& P = dp[p];
  for (int a = 0; a < 3; a++) {
    const int tf = pt[a];
    const int ti = tf == 0 ? 1 : (tf == 3 ? 4 : tf);  // DCT_II->DCT_III, DST_II->DST_III, IV types self-inverse
    for (int k = threadIdx.x; k < k3N * k3N; k += blockDim.x) {
      const int row = k >> 4, col = k & 15;
      MF[a * k3Mat + row * k3Row + col] = mats[(tf * 3 + a) * kDftMatStride + k];
      MI[a * k3Mat + row * k3Row + col] = mats[(ti * 3 + a) * kDftMatStride + k];
    }
  }
  // f_copy = f, then modifyRHSForInternalBoundaryConditions (Poisson :165-180, VarPoisson :168-186)
  for (int o = threadIdx.x; o < k3N * k3N * k3N; o += blockDim.x) {
    const int x = o & 15, y = (o >> 4) & 15, z = o >> 8;
    const int gi = lin(L, p, 0, x, y, z);
    T v = T(f[gi]);
    if (!UZERO) {
      const int cc[3] = { x, y, z };
      const int st[3] = { 1, L.sx, L.sy };
#pragma unroll
      for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int up = 0; up < 2; up++) {
          const bool edge = up ? (cc[a] == k3N - 1) : (cc[a] == 0);
          if (!edge || P.type[20 + 2 * a + up] < 0) continue;
          const int gidx = up ? gi + st[a] : gi - st[a];
          const T ih2 = T(P.inv_h2[a]);
          if (OP == 0) {
            v -= (u[gidx] + u[gi]) * ih2;
          } else {
            const int cq = gi - p * L.ps;
            const int cg = up ? cq + st[a] : cq - st[a];
            const T cb = coeffs[p * L.cs + cq], cgv = coeffs[p * L.cs + cg];
            v -= (u[gidx] + u[gi]) * (cb + cgv) * (ih2 * T(0.5));
          }
        }
      }
    }
    A[c3(x, y, z)] = v;
  }
  __syncthreads();
  for (int a = 0; a < 3; a++) {  // forward transform along each axis, in place
    axis16<T>(a, MF + a * k3Mat, A, A, wave, lane);
    __syncthreads();
  }
  T* src = A;
  for (int o = threadIdx.x; o < k3N * k3N * k3N; o += blockDim.x) {  // divide by the eigenvalues
    const int k3[3] = { o & 15, (o >> 4) & 15, o >> 8 };
    T eig = T(0);
    for (int a = 0; a < 3; a++) {
      const double h = P.h[a];
      eig -= T(4.0 / (h * h)) * s2tab[(pt[3 + a] * 3 + a) * kDftMatDim + k3[a]];
    }
    src[c3(k3[0], k3[1], k3[2])] /= eig;
  }
  __syncthreads();
  if (pt[6] && threadIdx.x == 0) src[0] = T(0);  // all-Neumann patch without neighbours (:438-440)
  __syncthreads();
  for (int a = 0; a < 3; a++) {
    axis16<T>(a, MI + a * k3Mat, A, A, wave, lane);
    __syncthreads();
  }
  const T scale = T(2.0) / T(16) * (T(2.0) / T(16)) * (T(2.0) / T(16));
  for (int o = threadIdx.x; o < k3N * k3N * k3N; o += blockDim.x) {
    const int x = o & 15, y = (o >> 4) & 15, z = o >> 8;
    out[lin(L, p, 0, x, y, z)] = TO(src[c3(x, y, z)] * scale);
  }
}

template <typename T, typename TF, typename TO, int D>
void launch_d(const te_dft_plan* pl, const TF* f, const T* u, TO* out, bool uzero, hipStream_t s)
{
  const te_domain* d = pl->g->d;
  Layout L = make_layout(d->dim, d->ns, 1, d->np);
  int nmax = 0;
  int N = 1;
  for (int a = 0; a < D; a++) {
    nmax = d->ns[a] > nmax ? d->ns[a] : nmax;
    N *= d->ns[a];
  }
  size_t shmem = sizeof(T) * (2 * (size_t)N + 2 * (size_t)D * nmax * nmax);
  const T* mats = (const T*)(sizeof(T) == 8 ? (const void*)pl->d_mats64 : (const void*)pl->d_mats32);
  const T* s2 = (const T*)(sizeof(T) == 8 ? (const void*)pl->d_s2_64 : (const void*)pl->d_s2_32);
  const T* cf = (const T*)(sizeof(T) == 8 ? (const void*)pl->d_coeffs64 : (const void*)pl->d_coeffs32);
  if (D == 3 && pl->use_mfma && d->ns[0] == 16 && d->ns[1] == 16 && d->ns[2] == 16) {
    const size_t sh3 = sizeof(T) * ((size_t)k3Cells + 6 * (size_t)k3Mat);
#define TE_D3(OP_, UZ_)                                                                                         \
  hipLaunchKernelGGL((dft3_mfma_kernel<T, TF, TO, OP_, UZ_>), dim3(d->np), dim3(256), sh3, s, d->d_patches,     \
                     pl->d_ptype, mats, s2, f, u, cf, out, L)
    if (pl->op_kind == 0) {
      if (uzero) TE_D3(0, true); else TE_D3(0, false);
    } else {
      if (uzero) TE_D3(1, true); else TE_D3(1, false);
    }
#undef TE_D3
    return;
  }
  // patches per workgroup and threads per patch: the autotuner's choice for this shape
  const DftGenericVariant v = dft_generic_variant(tune_variant(TUNE_DFT_GENERIC, D, d->ns[0], sizeof(T) == 8 ? TE_F64 : TE_F32));
  const int grid = (d->np + v.ppw - 1) / v.ppw;
  shmem *= (size_t)v.ppw;
  const FusedArgs<T, T> fa{ nullptr, nullptr, nullptr };
#define TE_DL(OP_, UZ_)                                                                                          \
  hipLaunchKernelGGL((dft_kernel<T, TF, TO, D, OP_, UZ_>), dim3(grid), dim3(v.ppw * v.tpp), shmem, s,          \
                     d->d_patches, pl->d_ptype, mats, s2, f, u, cf, out, L, v.ppw, fa)
  if (pl->op_kind == 0) {
    if (uzero) TE_DL(0, true); else TE_DL(0, false);
  } else {
    if (uzero) TE_DL(1, true); else TE_DL(1, false);
  }
#undef TE_DL
}

template <typename T, typename TF, typename TO>
void launch_t(const te_dft_plan* pl, const void* f, const void* u, void* out, bool uzero, hipStream_t s)
{
  if (pl->g->d->dim == 2) launch_d<T, TF, TO, 2>(pl, (const TF*)f, (const T*)u, (TO*)out, uzero, s);
  else launch_d<T, TF, TO, 3>(pl, (const TF*)f, (const T*)u, (TO*)out, uzero, s);
}

// the fused sweeps of the generic kernel, 2D: gm 0 (ghosts read from u), 1 (computed here),
// 2 (computed here on u + P uc)
template <typename T, typename TF, typename TO, typename TC>
void launch_fused_t(const te_dft_plan* pl, const void* fv, const void* uv, void* outv, bool uzero, int gm,
                    const void* ucv, const te_level_link* link, hipStream_t s)
{
  const te_domain* d = pl->g->d;
  Layout L = make_layout(2, d->ns, 1, d->np);
  const int nmax = d->ns[0] > d->ns[1] ? d->ns[0] : d->ns[1];
  const int N = d->ns[0] * d->ns[1];
  size_t shmem = sizeof(T) * (2 * (size_t)N + 4 * (size_t)nmax * nmax);
  const T* mats = (const T*)(sizeof(T) == 8 ? (const void*)pl->d_mats64 : (const void*)pl->d_mats32);
  const T* s2 = (const T*)(sizeof(T) == 8 ? (const void*)pl->d_s2_64 : (const void*)pl->d_s2_32);
  const T* cf = (const T*)(sizeof(T) == 8 ? (const void*)pl->d_coeffs64 : (const void*)pl->d_coeffs32);
  const DftGenericVariant v =
    dft_generic_variant(tune_variant(TUNE_DFT_GENERIC, 2, d->ns[0], sizeof(T) == 8 ? TE_F64 : TE_F32));
  const int grid = (d->np + v.ppw - 1) / v.ppw;
  shmem *= (size_t)v.ppw;
  const FusedArgs<T, TC> fa{ (const TC*)ucv, link ? link->d_parent : nullptr, link ? link->d_orth : nullptr };
  const TF* f = (const TF*)fv;
  const T* u = (const T*)uv;
  TO* out = (TO*)outv;
  const int kind = pl->g->kind;
#define TE_DF(OP_, UZ_, GM_, K_)                                                                                 \
  hipLaunchKernelGGL((dft_kernel<T, TF, TO, 2, OP_, UZ_, GM_, K_, TC>), dim3(grid), dim3(v.ppw * v.tpp), shmem, \
                     s, d->d_patches, pl->d_ptype, mats, s2, f, u, cf, out, L, v.ppw, fa)
#define TE_DFO(OP_)                                                        \
  do {                                                                     \
    if (uzero) TE_DF(OP_, true, 0, 0);                                     \
    else if (gm == 0) TE_DF(OP_, false, 0, 0);                             \
    else if (gm == 1 && kind == TE_BILINEAR) TE_DF(OP_, false, 1, TE_BILINEAR); \
    else if (gm == 1) TE_DF(OP_, false, 1, TE_BIQUADRATIC);                \
    else if (kind == TE_BILINEAR) TE_DF(OP_, false, 2, TE_BILINEAR);       \
    else TE_DF(OP_, false, 2, TE_BIQUADRATIC);                             \
  } while (0)
  if (pl->op_kind == 0) TE_DFO(0);
  else TE_DFO(1);
#undef TE_DFO
#undef TE_DF
}

}  // namespace

bool dft_generic_fused_supported(const te_dft_plan* pl)
{
  const te_domain* d = pl->g->d;
  return d->dim == 2 && d->ns[0] % 2 == 0 && d->ns[1] % 2 == 0 &&
         (pl->g->kind == TE_BILINEAR || pl->g->kind == TE_BIQUADRATIC);
}

int launch_dft_fused(const te_dft_plan* pl, int dt, int dtf, int dto, const void* f, const void* u, void* out,
                     bool uzero, int gm, int dtc, const void* uc, const te_level_link* link, hipStream_t s)
{
  if (pl->use_mfma && dft_mfma_supported(pl))
    return launch_dft_mfma_fused(pl, dt, dtf, dto, f, u, out, uzero, gm, dtc, uc, link, s);
  if (pl->g->d->np == 0) return TE_OK;
  if (!dft_generic_fused_supported(pl)) {
    set_error("fused DFT sweep: needs 2D, even cells per side, BiLinear or BiQuadratic ghosts");
    return TE_ERR_UNSUPPORTED;
  }
  if (gm == 2 && (!uc || !link)) {
    set_error("fused interpolation needs the coarser level's u and the level link");
    return TE_ERR_ARG;
  }
  if (gm != 0 && !uzero && out == u) {
    set_error("internal: a fused sweep that computes ghosts must not write its input");
    return TE_ERR_ARG;
  }
  const bool c64 = gm != 2 || dtc == TE_F64;
#define TE_LF(T_, TF_, TO_)                                                                      \
  do {                                                                                           \
    if (c64) launch_fused_t<T_, TF_, TO_, double>(pl, f, u, out, uzero, gm, uc, link, s);        \
    else launch_fused_t<T_, TF_, TO_, float>(pl, f, u, out, uzero, gm, uc, link, s);             \
  } while (0)
  if (dt == TE_F64 && dtf == TE_F64 && dto == TE_F64) {
    if (!c64) {
      set_error("FP64 level with an FP32 coarser level is not supported");
      return TE_ERR_UNSUPPORTED;
    }
    launch_fused_t<double, double, double, double>(pl, f, u, out, uzero, gm, uc, link, s);
  } else if (dt == TE_F32 && dtf == TE_F32 && dto == TE_F32) {
    TE_LF(float, float, float);
  } else if (dt == TE_F32 && dtf == TE_F64 && dto == TE_F32) {
    TE_LF(float, double, float);
  } else if (dt == TE_F32 && dtf == TE_F64 && dto == TE_F64) {
    TE_LF(float, double, double);
  } else if (dt == TE_F32 && dtf == TE_F32 && dto == TE_F64) {
    TE_LF(float, float, double);
  } else {
    set_error("unsupported dtype combination for the DFT patch solver");
    return TE_ERR_UNSUPPORTED;
  }
#undef TE_LF
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_dft(const te_dft_plan* pl, int dt, int dtf, int dto, const void* f, const void* u, void* out, bool uzero,
               hipStream_t s)
{
  if (pl->g->d->np == 0) return TE_OK;
  if (pl->use_mfma && dft_mfma_supported(pl)) return launch_dft_mfma(pl, dt, dtf, dto, f, u, out, uzero, s);
  if (dt == TE_F64 && dtf == TE_F64 && dto == TE_F64) launch_t<double, double, double>(pl, f, u, out, uzero, s);
  else if (dt == TE_F32 && dtf == TE_F32 && dto == TE_F32) launch_t<float, float, float>(pl, f, u, out, uzero, s);
  else if (dt == TE_F32 && dtf == TE_F64 && dto == TE_F32) launch_t<float, double, float>(pl, f, u, out, uzero, s);
  else if (dt == TE_F32 && dtf == TE_F64 && dto == TE_F64) launch_t<float, double, double>(pl, f, u, out, uzero, s);
  else {
    set_error("unsupported dtype combination for the DFT patch solver");
    return TE_ERR_UNSUPPORTED;
  }
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
