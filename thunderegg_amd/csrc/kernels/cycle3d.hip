// The 3D V-cycle and the outer BiCGStab's vector work on compact vectors (see cycle3d.h).
//
// Lane layout shared by every kernel here: a wavefront works on 16 x 16 slices of a patch;
// lane l = c + 16 g holds the four cells x = 4g .. 4g + 3 of row y = c, so a row of a slice
// is one 64-byte (FP32) or 128-byte (FP64) load by four lanes and a slice is one load
// instruction per wavefront.
//
// Patch solve (c3_sweep_kernel, one 256-thread workgroup per patch): the reference applies
// a dense n x n transform along x, then y, then z, divides by the eigenvalues and applies
// the inverse transforms (Poisson/DFTPatchSolver.h:205-265, 320-361, 420-450).  Here each
// transform is a 16 x 16 x 16 matrix product on the matrix cores (v_mfma_*_16x16x4), with
// the transform matrices held in registers:
//   per z-slice:  D1 = F Sx^T, D2 = Sy D1          (D1 stays in registers: its accumulator
//                                                   layout is the B operand of the next
//                                                   product with a matching K order)
//   per y-row:    W = Sz D2, W /= lambda, W' = Szi W (the slab passes through LDS once)
//   per z-slice:  R = Q Sxi^T, U = Syi R           (R again register-resident)
// so the 16^3 patch crosses LDS twice (after the forward x-y pass, after the z pass).
// In a product whose operands come from memory the K index of MFMA step s in lane group
// g is 4g + s (four consecutive values per lane: one vector load); in a product fed by
// the previous accumulator it is the accumulator's row of register s (Mf<T>::row).
//
// Stencils (c3_resid_restrict_kernel, c3_star_dot_kernel, one wavefront per patch): the
// wavefront walks the 16 slices keeping slices z - 1, z, z + 1 in registers; x and y
// neighbours come from other lanes (shuffles), the patch's face ghosts from the same-level
// neighbour's boundary slice or the physical boundary condition.
#include <algorithm>
#include <type_traits>

#include "../te_internal.h"
#include "blas.h"
#include "cycle3d.h"
#include "dft_wave.h"
#include "fold.h"
#include "krylov.h"
#include "../tune.h"

namespace te {
namespace {

constexpr int N = kC3N;
constexpr int NC = kC3Cells;
constexpr int kYP = 20;               // LDS pitch of a spectral row (16 + 4: 16-byte aligned rows)
constexpr int kZP = 16 * kYP + 4;     // LDS pitch of a spectral slice
constexpr int kSlab = 16 * kZP;

// Addressing: every offset here is a non-negative element index into a vector of at most
// 2 GiB (a compact vector with face columns, 4608 values per patch, is smaller than the
// padded vector of the same domain, whose size the domain checks), so the byte offset is
// formed in 32 bits (at(), te_internal.h) and a load or store addresses from the kernel
// argument's scalar base plus a 32-bit vector offset, without 64-bit vector address
// arithmetic.
template <typename V, typename TS>
__device__ __forceinline__ const V& vat(const TS* base, int off)
{
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + (unsigned)off * (unsigned)sizeof(TS));
}
template <typename V, typename TS>
__device__ __forceinline__ V& vat(TS* base, int off)
{
  return *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + (unsigned)off * (unsigned)sizeof(TS));
}

// Streaming (nontemporal) access: a load or store marked nontemporal goes to HBM without
// keeping the line in the caches (measured on this part: 7.0 TB/s for a read-only stream
// against 6.3 TB/s with plain loads, 6.2 against 5.7 TB/s for a copy).  For values that no
// other wave of the launch reads again: the own cells of a vector that is read once, never
// the cells that neighbours read as face values.  FP32 only: a lane's four FP64 values are
// two 16-byte accesses, each instruction covering every other 16 bytes of a row, and in
// that form the streaming hint measured slower (the FP64 forms below are plain accesses).
typedef float c3f4 __attribute__((ext_vector_type(4)));

// four consecutive values starting at an offset that is a multiple of 4, converted to T
// (NT: a streaming load of FP32 values)
template <bool NT = false, typename T, typename TS>
__device__ __forceinline__ void ld4(const TS* __restrict__ base, int off, T (&r)[4])
{
  if constexpr (NT && sizeof(TS) == 4) {
    const c3f4 v = __builtin_nontemporal_load(&vat<c3f4>(base, off));
    r[0] = T(v.x);
    r[1] = T(v.y);
    r[2] = T(v.z);
    r[3] = T(v.w);
  } else if constexpr (sizeof(TS) == 4) {
    const float4 v = vat<float4>(base, off);
    r[0] = T(v.x);
    r[1] = T(v.y);
    r[2] = T(v.z);
    r[3] = T(v.w);
  } else {
    const double2 a = vat<double2>(base, off);
    const double2 b = vat<double2>(base, off + 2);
    r[0] = T(a.x);
    r[1] = T(a.y);
    r[2] = T(b.x);
    r[3] = T(b.y);
  }
}

// (NT: a streaming store of FP32 values)
template <bool NT = false, typename TD, typename T>
__device__ __forceinline__ void st4(TD* __restrict__ base, int off, const T (&r)[4])
{
  if constexpr (NT && sizeof(TD) == 4) {
    __builtin_nontemporal_store(c3f4{ float(r[0]), float(r[1]), float(r[2]), float(r[3]) }, &vat<c3f4>(base, off));
  } else if constexpr (sizeof(TD) == 4) {
    vat<float4>(base, off) = make_float4(float(r[0]), float(r[1]), float(r[2]), float(r[3]));
  } else {
    vat<double2>(base, off) = make_double2(double(r[0]), double(r[1]));
    vat<double2>(base, off + 2) = make_double2(double(r[2]), double(r[3]));
  }
}

__device__ __forceinline__ int cidx(int x, int y, int z) { return x + 16 * y + 256 * z; }
// face column entry (side 0: x = 0, side 1: x = N - 1) of row (y, z) of patch p
__device__ __forceinline__ int xci(int p, int side, int y, int z) { return p * kC3XCol + 256 * side + 16 * z + y; }

// u (+ the coarser level's correction uc, INTERP) at cells of patch q, compact vectors
template <typename T, typename TC, bool INTERP>
struct C3U {
  const T* __restrict__ u;
  const TC* __restrict__ uc;
  const int32_t* __restrict__ parent;
  const int32_t* __restrict__ porth;
  // coarse cell of fine cell (x, y, z) of a patch with orthant o on its parent (o < 0: copy)
  __device__ __forceinline__ int coarse(int o, int x, int y, int z) const
  {
    if (o < 0) return cidx(x, y, z);
    return cidx((x + ((o & 1) ? N : 0)) >> 1, (y + ((o & 2) ? N : 0)) >> 1, (z + ((o & 4) ? N : 0)) >> 1);
  }
  __device__ __forceinline__ T one(int q, int x, int y, int z) const
  {
    T v = at(u, q * NC + cidx(x, y, z));
    if (INTERP) v = T(v + T(at(uc, parent[q] * NC + coarse(porth[q], x, y, z))));
    return v;
  }
  // cells x0 .. x0 + 3 of row (y, z), x0 a multiple of 4 (NT: u's cells by streaming loads)
  template <bool NT = false>
  __device__ __forceinline__ void row4(int q, int x0, int y, int z, T (&r)[4]) const
  {
    ld4<NT>(u, q * NC + cidx(x0, y, z), r);
    if (INTERP) {
      const int o = porth[q];
      const TC* b = uc + parent[q] * NC;
      if (o >= 0) {
        const int k = coarse(o, x0, y, z);
        const T a0 = T(at(b, k)), a1 = T(at(b, k + 1));
        r[0] = T(r[0] + a0);
        r[1] = T(r[1] + a0);
        r[2] = T(r[2] + a1);
        r[3] = T(r[3] + a1);
      } else {
        T a[4];
        ld4(b, cidx(x0, y, z), a);
#pragma unroll
        for (int k = 0; k < 4; k++) r[k] = T(r[k] + a[k]);
      }
    }
  }
};

// ------------------------------------------------------------------------------------------
// The patch solve.  MODE 0: u = 0 (f only); 1: interface terms from uin; 2: from uin + P uc.
// SB: slices per load batch of modes 1 / 2 (the autotuner's variants; mode 0 uses 4).
// XC: uin and out have face columns (cycle3d.h): the x faces are read from uin's, and out's
// are written.  WPE: the minimum waves per SIMD the registers are allocated for (1: no request).
// FA (modes 1, 2): the loads of f of all four slices are issued at the start, the interface
// loads per batch
template <typename T, typename TF, typename TO, typename TC, int MODE, int SB, bool XC, int WPE, bool FA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8))) c3_sweep_kernel(const DevPatch* __restrict__ dp,
                                                       const int* __restrict__ ptype, const T* __restrict__ mats,
                                                       const T* __restrict__ s2tab, const TF* __restrict__ f,
                                                       const T* __restrict__ uin, TO* __restrict__ out,
                                                       const TC* __restrict__ uc, const int32_t* __restrict__ parent,
                                                       const int32_t* __restrict__ porth, const T* __restrict__ xin,
                                                       TO* __restrict__ xout)
{
  __shared__ __align__(16) T slab[kSlab];
  typedef typename Mf<T>::V V;
  const int p = blockIdx.x;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
  const int* pt = ptype + p * 8;  // forward type per axis [0..2], eigenvalue type per axis [3..5], zero mode [6]
  int tf[3], ti[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    tf[a] = pt[a];
    ti[a] = tf[a] == 0 ? 1 : (tf[a] == 3 ? 4 : tf[a]);  // DCT-II -> DCT-III, DST-II -> DST-III, IV types self-inverse
  }
  // the transform matrices M[i][j] (y_i = sum_j M[i][j] x_j, getTransformArray :131-196) as
  // MFMA operands: lane (c, g), step s holds M[c][K] with K the product's K order
  auto mat = [&](int t, int a, int i, int j) { return mats[(t * 3 + a) * kDftMatStride + i * N + j]; };
  T SxB[4], SyA[4], SzA[4], SziA[4], SxiB[4], SyiA[4];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int k1 = 4 * g + s, k2 = Mf<T>::row(s, g);
    SxB[s] = mat(tf[0], 0, c, k1);
    SyA[s] = mat(tf[1], 1, c, k2);
    SzA[s] = mat(tf[2], 2, c, k1);
    SziA[s] = mat(ti[2], 2, c, k2);
    SxiB[s] = mat(ti[0], 0, c, k1);
    SyiA[s] = mat(ti[1], 1, c, k2);
  }
  const DevPatch& P = dp[p];
  const C3U<T, TC, MODE == 2> U{ uin, uc, parent, porth };
  // the six face neighbours (the patch itself where a side is physical: its loads are then
  // masked off) and, interpolating, their parents and orthants, read once per workgroup
  int q[6];
  bool has[6];
#pragma unroll
  for (int s = 0; s < 6; s++) {
    has[s] = P.type[20 + s] >= 0;
    q[s] = has[s] ? P.nbr[20 + s][0] : p;
  }
  int qpar[7] = { 0, 0, 0, 0, 0, 0, 0 }, qo[7] = { -1, -1, -1, -1, -1, -1, -1 };  // [6]: the patch itself
  if (MODE == 2) {
#pragma unroll
    for (int s = 0; s < 7; s++) {
      const int qq = s < 6 ? q[s] : p;
      qpar[s] = parent[qq];
      qo[s] = porth[qq];
    }
  }
  // cells x0 .. x0 + 3 of row (y, z) of patch qq (neighbour slot s, 6 = the patch), u + P uc
  // when interpolating (C3U::row4 with the parent and orthant already in registers)
  auto row = [&](int s, int qq, int x0, int y, int z, T(&r)[4]) {
    ld4(uin, qq * NC + cidx(x0, y, z), r);
    if (MODE == 2) {
      const int o = qo[s];
      const TC* b = uc + qpar[s] * NC;
      if (o >= 0) {
        const int k = U.coarse(o, x0, y, z);
        const T a0 = T(at(b, k)), a1 = T(at(b, k + 1));
        r[0] = T(r[0] + a0);
        r[1] = T(r[1] + a0);
        r[2] = T(r[2] + a1);
        r[3] = T(r[3] + a1);
      } else {
        T a[4];
        ld4(b, cidx(x0, y, z), a);
#pragma unroll
        for (int k = 0; k < 4; k++) r[k] = T(r[k] + a[k]);
      }
    }
  };
  const T ihx = T(P.inv_h2[0]), ihy = T(P.inv_h2[1]), ihz = T(P.inv_h2[2]);
  // which of this lane's cells lie on a face with a neighbour (x: the lane's first or last
  // cell, y: its row, z: its slice)
  const bool onx = (g == 0 && has[0]) || (g == 3 && has[1]);
  const bool ony = (c == 0 && has[2]) || (c == N - 1 && has[3]);
  const int sx = g == 0 ? 0 : 1, sy = c == 0 ? 2 : 3;
  // with face columns (XC) an x-face lane reads two values of its row, the own boundary cell
  // (side xs) and the neighbour's facing cell (its side 1 - xs), as one code path for both
  // x faces: the neighbour, its parent and orthant selected per lane here
  const int xs = g == 0 ? 0 : 1;
  const int qx = g == 0 ? q[0] : q[1];
  const int qxpar = g == 0 ? qpar[0] : qpar[1], qxo = g == 0 ? qo[0] : qo[1];
  // cell x = (side ? N - 1 : 0) of row (c, z) of patch qq from uin's face columns, u + P uc
  // when interpolating (par, o: qq's parent and orthant): the value row() gives for it
  auto xval = [&](int qq, int side, int par, int o, int z) -> T {
    T v = at(xin, xci(qq, side, c, z));
    if (MODE == 2) v = T(v + T(at(uc, par * NC + U.coarse(o, side ? N - 1 : 0, c, z))));
    return v;
  };

  // ---- forward transforms along x and y, one z-slice at a time; the loads of a batch of
  // SB slices (f, and for the interface terms the lane's own boundary rows and the
  // neighbours' facing rows) are all issued before its first product (more slices per
  // batch: more loads in flight, more registers, fewer waves per SIMD)
  T Fa[FA ? 4 : 1][4];
  if (FA) {
#pragma unroll
    for (int i = 0; i < (FA ? 4 : 1); i++) ld4<true>(f, p * NC + cidx(4 * g, c, w + 4 * i), Fa[i]);
  }
#pragma unroll
  for (int i0 = 0; i0 < 4; i0 += SB) {
    T F[SB][4], ub[SB][4], gx[SB][4], gy[SB][4], gz[SB][4];
    T xo[SB], xn[SB];  // XC: the x-face lane's own boundary value and the neighbour's
#pragma unroll
    for (int j = 0; j < SB; j++) {
      const int z = w + 4 * (i0 + j);
      if (FA) {
#pragma unroll
        for (int k = 0; k < 4; k++) F[j][k] = Fa[FA ? i0 + j : 0][k];
      } else {
        ld4<true>(f, p * NC + cidx(4 * g, c, z), F[j]);
      }
      if (MODE != 0) {
        const bool onz = (z == 0 && has[4]) || (z == N - 1 && has[5]);
        if (XC) {
          if (onx) {
            xo[j] = xval(p, xs, qpar[6], qo[6], z);
            xn[j] = xval(qx, 1 - xs, qxpar, qxo, z);
          }
          if (ony || onz) row(6, p, 4 * g, c, z, ub[j]);
        } else {
          if (onx || ony || onz) row(6, p, 4 * g, c, z, ub[j]);
          if (onx) row(sx, q[sx], g == 0 ? 12 : 0, c, z, gx[j]);
        }
        if (ony) row(sy, q[sy], 4 * g, c == 0 ? N - 1 : 0, z, gy[j]);
        if (onz) row(z == 0 ? 4 : 5, q[z == 0 ? 4 : 5], 4 * g, c, z == 0 ? N - 1 : 0, gz[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < SB; j++) {
      const int z = w + 4 * (i0 + j);
      if (MODE != 0) {
        // modifyRHSForInternalBoundaryConditions (Poisson/StarPatchOperator.h:165-180): on
        // every side with a neighbour, f_b -= (ghost + u_b) / h^2, sides in the reference's order
        if (g == 0 && has[0]) F[j][0] -= ((XC ? xn[j] : gx[j][3]) + (XC ? xo[j] : ub[j][0])) * ihx;
        if (g == 3 && has[1]) F[j][3] -= ((XC ? xn[j] : gx[j][0]) + (XC ? xo[j] : ub[j][3])) * ihx;
        if (ony) {
#pragma unroll
          for (int k = 0; k < 4; k++) F[j][k] -= (gy[j][k] + ub[j][k]) * ihy;
        }
        if ((z == 0 && has[4]) || (z == N - 1 && has[5])) {
#pragma unroll
          for (int k = 0; k < 4; k++) F[j][k] -= (gz[j][k] + ub[j][k]) * ihz;
        }
      }
      V d1 = { 0, 0, 0, 0 };  // D1[y][kx] = sum_x F[y][x] Sx[kx][x]
#pragma unroll
      for (int s = 0; s < 4; s++) d1 = Mf<T>::mma(F[j][s], SxB[s], d1);
      V d2 = { 0, 0, 0, 0 };  // D2[ky][kx] = sum_y Sy[ky][y] D1[y][kx]
#pragma unroll
      for (int s = 0; s < 4; s++) d2 = Mf<T>::mma(SyA[s], d1[s], d2);
#pragma unroll
      for (int r = 0; r < 4; r++) slab[z * kZP + Mf<T>::row(r, g) * kYP + c] = d2[r];
    }
  }
  __syncthreads();

  // ---- along z, one ky-row at a time: forward, eigenvalues, inverse
  const int et0 = pt[3], et1 = pt[4], et2 = pt[5];
  const T lx = T(4.0 / (P.h[0] * P.h[0])), ly = T(4.0 / (P.h[1] * P.h[1])), lz = T(4.0 / (P.h[2] * P.h[2]));
  const T s2x = s2tab[(et0 * 3 + 0) * kDftMatDim + c];
  T s2z[4];
#pragma unroll
  for (int r = 0; r < 4; r++) s2z[r] = s2tab[(et2 * 3 + 2) * kDftMatDim + Mf<T>::row(r, g)];
  const bool zero_mode = pt[6] != 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int ky = w + 4 * i;
    V b = { 0, 0, 0, 0 };  // W[kz][kx] = sum_z Sz[kz][z] D2[z][ky][kx]
#pragma unroll
    for (int s = 0; s < 4; s++) b = Mf<T>::mma(SzA[s], slab[(4 * g + s) * kZP + ky * kYP + c], b);
    const T s2y = s2tab[(et1 * 3 + 1) * kDftMatDim + ky];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      // getEigenValues (:320-361): -sum_a 4/h_a^2 sin^2(...)
      T eig = T(0);
      eig -= lx * s2x;
      eig -= ly * s2y;
      eig -= lz * s2z[r];
      if constexpr (sizeof(T) == 4) b[r] = b[r] * T(__builtin_amdgcn_rcpf(float(eig)));
      else b[r] = b[r] / eig;
      // all-Neumann patch without neighbours: the zero mode is set to 0 (:438-440)
      if (zero_mode && ky == 0 && c == 0 && Mf<T>::row(r, g) == 0) b[r] = T(0);
    }
    V b2 = { 0, 0, 0, 0 };  // W'[z][kx] = sum_kz Szi[z][kz] W[kz][kx]
#pragma unroll
    for (int s = 0; s < 4; s++) b2 = Mf<T>::mma(SziA[s], b[s], b2);
#pragma unroll
    for (int r = 0; r < 4; r++) slab[Mf<T>::row(r, g) * kZP + ky * kYP + c] = b2[r];
  }
  __syncthreads();

  // ---- inverse transforms along x and y, one z-slice at a time, scaled by prod 2/n (:442-449)
  const T scale = T(2.0 / N) * T(2.0 / N) * T(2.0 / N);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int z = w + 4 * i;
    T Q[4];
    const T* qrow = slab + z * kZP + c * kYP + 4 * g;
#pragma unroll
    for (int k = 0; k < 4; k++) Q[k] = qrow[k];
    V r1 = { 0, 0, 0, 0 };  // R[ky][x] = sum_kx Q[ky][kx] Sxi[x][kx]
#pragma unroll
    for (int s = 0; s < 4; s++) r1 = Mf<T>::mma(Q[s], SxiB[s], r1);
    V u2 = { 0, 0, 0, 0 };  // U[y][x] = sum_ky Syi[y][ky] R[ky][x]
#pragma unroll
    for (int s = 0; s < 4; s++) u2 = Mf<T>::mma(SyiA[s], r1[s], u2);
    TO v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      v[r] = TO(u2[r] * scale);
      if constexpr (sizeof(TO) == 4) __builtin_nontemporal_store(v[r], &at(out, p * NC + cidx(c, Mf<T>::row(r, g), z)));
      else at(out, p * NC + cidx(c, Mf<T>::row(r, g), z)) = v[r];
    }
    if (XC && (c == 0 || c == N - 1)) {
      if constexpr (sizeof(T) == 4) {
        st4(xout, xci(p, c == 0 ? 0 : 1, 4 * g, z), v);  // FP32 layout: rows 4g .. 4g + 3
      } else {
#pragma unroll
        for (int r = 0; r < 4; r++) at(xout, xci(p, c == 0 ? 0 : 1, Mf<T>::row(r, g), z)) = v[r];
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// The Poisson star operator (Poisson/StarPatchOperator.h:75-101) of one slice of a patch,
// wavefront-wide.  cur / prv / nxt: the lane's row in slices z, z - 1, z + 1 (with the face
// ghosts or boundary values of the z faces already in place).  bc: the physical ghost factor
// (-1 Dirichlet, +1 Neumann, :115-144).  q: the six face neighbours (-1: physical side).
template <typename T, typename TU>
__device__ __forceinline__ void star_slice(const TU* __restrict__ u, const TU* __restrict__ xu, const int (&q)[6],
                                           int lane, int c, int g, int z, T bc, T ihx, T ihy, T ihz,
                                           const T (&cur)[4], const T (&prv)[4], const T (&nxt)[4], T (&acc)[4])
{
  T xm = __shfl(cur[3], (lane - 16) & 63, 64);
  T xp = __shfl(cur[0], (lane + 16) & 63, 64);
  T ym[4], yp[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    ym[k] = __shfl(cur[k], (lane - 1) & 63, 64);
    yp[k] = __shfl(cur[k], (lane + 1) & 63, 64);
  }
  // x faces: the neighbour's facing column, from u's face columns when it has them
  if (g == 0)
    xm = q[0] >= 0 ? T(xu ? at(xu, xci(q[0], 1, c, z)) : at(u, q[0] * NC + cidx(N - 1, c, z))) : bc * cur[0];
  if (g == 3) xp = q[1] >= 0 ? T(xu ? at(xu, xci(q[1], 0, c, z)) : at(u, q[1] * NC + cidx(0, c, z))) : bc * cur[3];
  if (c == 0) {
    if (q[2] >= 0) ld4(u, q[2] * NC + cidx(4 * g, N - 1, z), ym);
    else
#pragma unroll
      for (int k = 0; k < 4; k++) ym[k] = bc * cur[k];
  }
  if (c == N - 1) {
    if (q[3] >= 0) ld4(u, q[3] * NC + cidx(4 * g, 0, z), yp);
    else
#pragma unroll
      for (int k = 0; k < 4; k++) yp[k] = bc * cur[k];
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const T mid = cur[k];
    const T lo = k == 0 ? xm : cur[k - 1];
    const T hi = k == 3 ? xp : cur[k + 1];
    T a = T(0);
    a += (hi - T(2) * mid + lo) * ihx;
    a += (yp[k] - T(2) * mid + ym[k]) * ihy;
    a += (nxt[k] - T(2) * mid + prv[k]) * ihz;
    acc[k] = a;
  }
}

// The face values star_slice reads from the neighbours for slice z, loaded ahead of their use
// (raw, of u's type): xr the facing x-face cell of lanes g = 0 (q[0]) and g = 3 (q[1]), yr the
// facing y-face row of lanes c = 0 (q[2]) and c = N - 1 (q[3]); lanes that take no face value
// (or face a physical side) load nothing.
template <typename TU>
__device__ __forceinline__ void star_faces(const TU* __restrict__ u, const TU* __restrict__ xu, const int (&q)[6],
                                           int c, int g, int z, TU& xr, TU (&yr)[4])
{
  if (g == 0 && q[0] >= 0) xr = xu ? at(xu, xci(q[0], 1, c, z)) : at(u, q[0] * NC + cidx(N - 1, c, z));
  if (g == 3 && q[1] >= 0) xr = xu ? at(xu, xci(q[1], 0, c, z)) : at(u, q[1] * NC + cidx(0, c, z));
  if (c == 0 && q[2] >= 0) ld4(u, q[2] * NC + cidx(4 * g, N - 1, z), yr);
  if (c == N - 1 && q[3] >= 0) ld4(u, q[3] * NC + cidx(4 * g, 0, z), yr);
}
// star_slice with the face values from star_faces (the same arithmetic, so the same bits)
template <typename T, typename TU>
__device__ __forceinline__ void star_slice_f(const int (&q)[6], int lane, int c, int g, T bc, T ihx, T ihy, T ihz,
                                             TU xr, const TU (&yr)[4], const T (&cur)[4], const T (&prv)[4],
                                             const T (&nxt)[4], T (&acc)[4])
{
  T xm = __shfl(cur[3], (lane - 16) & 63, 64);
  T xp = __shfl(cur[0], (lane + 16) & 63, 64);
  T ym[4], yp[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    ym[k] = __shfl(cur[k], (lane - 1) & 63, 64);
    yp[k] = __shfl(cur[k], (lane + 1) & 63, 64);
  }
  if (g == 0) xm = q[0] >= 0 ? T(xr) : bc * cur[0];
  if (g == 3) xp = q[1] >= 0 ? T(xr) : bc * cur[3];
  if (c == 0) {
#pragma unroll
    for (int k = 0; k < 4; k++) ym[k] = q[2] >= 0 ? T(yr[k]) : bc * cur[k];
  }
  if (c == N - 1) {
#pragma unroll
    for (int k = 0; k < 4; k++) yp[k] = q[3] >= 0 ? T(yr[k]) : bc * cur[k];
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const T mid = cur[k];
    const T lo = k == 0 ? xm : cur[k - 1];
    const T hi = k == 3 ? xp : cur[k + 1];
    T a = T(0);
    a += (hi - T(2) * mid + lo) * ihx;
    a += (yp[k] - T(2) * mid + ym[k]) * ihy;
    a += (nxt[k] - T(2) * mid + prv[k]) * ihz;
    acc[k] = a;
  }
}

// the lane's row of slice z - 1 at z = 0 (bottom ghost) and of slice z + 1 at z = 15 (top)
template <typename T, typename TU>
__device__ __forceinline__ void z_ghost(const TU* __restrict__ u, int qz, int zsrc, int ro, T bc, const T (&cur)[4],
                                        T (&r)[4])
{
  if (qz >= 0) {
    ld4(u, qz * NC + 256 * zsrc + ro, r);
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) r[k] = bc * cur[k];
  }
}

// Cycle::restrict: fc = restriction of f - A u (one wavefront per fine patch)
template <typename T, typename TF, typename TC, int ZS>
__global__ void __launch_bounds__(256) c3_resid_restrict_kernel(const DevPatch* __restrict__ dp,
                                                                const T* __restrict__ u, const TF* __restrict__ f,
                                                                TC* __restrict__ fc,
                                                                const int32_t* __restrict__ parent,
                                                                const int32_t* __restrict__ porth, int np, int neumann,
                                                                const T* __restrict__ xu)
{
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
  const int ro = cidx(4 * g, c, 0);
  // a wavefront per ZS slices z0 .. z0 + ZS - 1 of a patch (ZS < N on small levels, where the
  // launch is as long as one wave's chain of slice loads)
  const int unit = blockIdx.x * 4 + wv;
  const int p = unit / (N / ZS), z0 = (unit % (N / ZS)) * ZS;
  if (p >= np) return;
  const DevPatch& P = dp[p];
  int q[6];
#pragma unroll
  for (int s = 0; s < 6; s++) q[s] = P.type[20 + s] >= 0 ? P.nbr[20 + s][0] : -1;
  const T ihx = T(P.inv_h2[0]), ihy = T(P.inv_h2[1]), ihz = T(P.inv_h2[2]);
  const T bc = neumann ? T(1) : T(-1);
  const int base = p * NC;
  const int cp = parent[p], o = porth[p];
  const T wgt = T(1.0 / 8);
  T cur[4], prv[4], nxt[4];
  ld4(u, base + 256 * z0 + ro, cur);
  if (z0 == 0) z_ghost(u, q[4], N - 1, ro, bc, cur, prv);
  else ld4(u, base + 256 * (z0 - 1) + ro, prv);
  // loads one slice ahead of their use (as in c3_star_dot_kernel): u1 slice z + 1 (at
  // z = N - 1 the neighbour's slice 0), u2 the one after, f0 / f1 and the faces x0, y0 /
  // x1, y1 of slices z and z + 1
  T u1[4], u2[4], f0[4], f1[4], x0 = T(0), x1 = T(0), y0[4], y1[4];
  ld4(u, base + 256 * (z0 + 1) + ro, u1);
  ld4<true>(f, base + 256 * z0 + ro, f0);
  star_faces(u, xu, q, c, g, z0, x0, y0);
  T hold0 = T(0), hold1 = T(0);
  for (int z = z0; z < z0 + ZS; z++) {
    if (z + 2 < N && z + 2 <= z0 + ZS) ld4(u, base + 256 * (z + 2) + ro, u2);
    else if (z + 2 == N && q[5] >= 0) ld4(u, q[5] * NC + ro, u2);
    if (z + 1 < z0 + ZS) {
      ld4<true>(f, base + 256 * (z + 1) + ro, f1);
      star_faces(u, xu, q, c, g, z + 1, x1, y1);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) nxt[k] = (z < N - 1 || q[5] >= 0) ? u1[k] : bc * cur[k];
    T acc[4];
    star_slice_f(q, lane, c, g, bc, ihx, ihy, ihz, x0, y0, cur, prv, nxt, acc);
    T fv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) fv[k] = f0[k];
    T r[4];
#pragma unroll
    for (int k = 0; k < 4; k++) r[k] = fv[k] - acc[k];
    if (o < 0) {  // copyToParent (GMG/LinearRestrictor.h:138-155)
      st4<true>(fc, cp * NC + 256 * z + ro, r);
    } else {      // restrictToCoarserParent (:101-128): coarse[(i + s) / 2] += fine[i] / 2^D
      const T sx0 = r[0] * wgt + r[1] * wgt, sx1 = r[2] * wgt + r[3] * wgt;
      const T sy0 = sx0 + __shfl(sx0, (lane + 1) & 63, 64);
      const T sy1 = sx1 + __shfl(sx1, (lane + 1) & 63, 64);
      if ((z & 1) == 0) {
        hold0 = sy0;
        hold1 = sy1;
      } else if ((c & 1) == 0) {
        const int k = cp * NC + cidx(2 * g + ((o & 1) ? 8 : 0), (c >> 1) + ((o & 2) ? 8 : 0),
                                     (z >> 1) + ((o & 4) ? 8 : 0));
        fc[k] = TC(hold0 + sy0);
        fc[k + 1] = TC(hold1 + sy1);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      prv[k] = cur[k];
      cur[k] = nxt[k];
      u1[k] = u2[k];
      f0[k] = f1[k];
      y0[k] = y1[k];
    }
    x0 = x1;
  }
}

// out = A u with the partial sums w.out (and out.out); persistent wavefronts over patches
template <typename TU, bool SQ, typename TW>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(TU) == 4 ? 4 : 1, 8))) c3_star_dot_kernel(const DevPatch* __restrict__ dp, const TU* __restrict__ u,
                                                          double* __restrict__ out, const TW* __restrict__ w,
                                                          double* __restrict__ pa, double* __restrict__ pb, int np,
                                                          int nslots, int neumann, const TU* __restrict__ xu)
{
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
  const int ro = cidx(4 * g, c, 0);
  const double bc = neumann ? 1.0 : -1.0;
  double sa = 0.0, sq = 0.0;
  for (int p = blockIdx.x * 4 + wv; p < np; p += gridDim.x * 4) {
    const DevPatch& P = dp[p];
    int q[6];
#pragma unroll
    for (int s = 0; s < 6; s++) q[s] = P.type[20 + s] >= 0 ? P.nbr[20 + s][0] : -1;
    const double ihx = P.inv_h2[0], ihy = P.inv_h2[1], ihz = P.inv_h2[2];
    const int base = p * NC;
    double cur[4], prv[4], nxt[4];
    ld4(u, base + ro, cur);
    z_ghost(u, q[4], N - 1, ro, bc, cur, prv);
    // raw loads one slice ahead of their use: u1 (slice z + 1, at z = N - 1 the neighbour's
    // slice 0), w0 (w of slice z: FP64, or FP32 by a streaming load), x0 / y0 (slice z's
    // faces); each iteration issues those of the next slice before its arithmetic
    TU u1[4], u2[4], x0 = TU(0), x1 = TU(0), y0[4], y1[4];
    TW w0[4], w1[4];
    ld4(u, base + 256 + ro, u1);
    ld4<sizeof(TW) == 4>(w, base + ro, w0);
    star_faces(u, xu, q, c, g, 0, x0, y0);
    for (int z = 0; z < N; z++) {
      if (z + 2 < N) ld4(u, base + 256 * (z + 2) + ro, u2);
      else if (z + 2 == N && q[5] >= 0) ld4(u, q[5] * NC + ro, u2);
      if (z + 1 < N) {
        ld4<sizeof(TW) == 4>(w, base + 256 * (z + 1) + ro, w1);
        star_faces(u, xu, q, c, g, z + 1, x1, y1);
      }
#pragma unroll
      for (int k = 0; k < 4; k++) nxt[k] = (z < N - 1 || q[5] >= 0) ? double(u1[k]) : bc * cur[k];
      double acc[4];
      star_slice_f(q, lane, c, g, bc, ihx, ihy, ihz, x0, y0, cur, prv, nxt, acc);
      if (out) st4(out, base + 256 * z + ro, acc);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        sa += double(w0[k]) * acc[k];  // rhat.dot(ap) / as.dot(s): w first, as the reference
        if (SQ) sq += acc[k] * acc[k];
        prv[k] = cur[k];
        cur[k] = nxt[k];
        u1[k] = u2[k];
        w0[k] = w1[k];
        y0[k] = y1[k];
      }
      x0 = x1;
    }
  }
  const double ra = block_reduce<RED_SUM>(sa);
  if (threadIdx.x == 0) pa[blockIdx.x] = ra;
  if (SQ) {
    const double rb = block_reduce<RED_SUM>(sq);
    if (threadIdx.x == 0) pb[blockIdx.x] = rb;
  }
  if (blockIdx.x == 0)
    for (int k = (int)gridDim.x + threadIdx.x; k < nslots; k += blockDim.x) {
      pa[k] = 0.0;
      if (SQ) pb[k] = 0.0;
    }
}

// ------------------------------------------------------------------------------------------
// padded offset of compact cell i (16^3 patches: 18^3 padded cells)
__device__ __forceinline__ int pad_of(int i)
{
  const int p = i >> 12, x = i & 15, y = (i >> 4) & 15, z = i >> 8 & 15;
  return p * (18 * 18 * 18) + (x + 1) + 18 * (y + 1) + 324 * (z + 1);
}

// BiCGStab vector updates on compact vectors; one partial per workgroup, nslots slots
template <int MODE, typename TM>
__global__ void __launch_bounds__(256) c3_bicg_kernel(C3Bicg a, long n, int nslots)
{
  const bool fold = MODE == BICG_S && a.pfold != nullptr;
  double* __restrict__ scal = a.scal;
  double alpha = fold ? 0.0 : scal[SC_ALPHA];
  const double omega = scal[SC_OMEGA], beta = scal[SC_BETA], rho = scal[SC_RHO];
  const long stride = (long)gridDim.x * blockDim.x;
  const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (MODE != BICG_INIT) {
    const double st = scal[SC_STATE];
    if (MODE == BICG_XR && a.dir && st == ST_EARLY) {
      // x += alpha dir (BiCGStab.h:162-167)
      for (long i = i0; i < n; i += stride) {
        const double dv = a.dir_m ? double(((const TM*)a.dir)[i]) : ((const double*)a.dir)[i];
        a.x[i] += dv * alpha;
      }
      return;
    }
    if (st != ST_RUN) return;
  }
  if (fold) {  // FIN_ALPHA: alpha = rho / rhat.dot(ap)
    const double* const parts[3] = { a.pfold, nullptr, nullptr };
    double r[3] = { 0.0, 0.0, 0.0 };
    fold_n<1>(parts, a.nfold, r);
    alpha = rho / r[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[SC_ALPHA] = alpha;
  }
  double sa = 0.0, sb = 0.0;
  for (long i = i0; i < n; i += stride) {
    if (MODE == BICG_INIT) {
      // rhat = r; p = r; sums b.b and rhat.r
      const int ip = pad_of((int)i);
      const double r = a.rpad[ip], bv = a.b[ip];
      a.x[i] = a.xpad[ip];
      a.r[i] = r;
      a.rhat[i] = r;
      a.p[i] = r;
      if (a.w32) a.w32[i] = float(r);
      sa += bv * bv;
      sb += r * r;
    } else if (MODE == BICG_S) {
      // s = r; s += ap * -alpha; sum s.s
      double s = a.r[i];
      s += a.ap[i] * -alpha;
      a.sv[i] = s;
      if (a.w32) a.w32[i] = float(s);
      sa += s * s;
    } else if (MODE == BICG_XR) {
      a.x[i] += double(((const TM*)a.mp)[i]) * alpha + double(((const TM*)a.ms)[i]) * omega;
      double r = a.r[i];
      r += a.ap[i] * -alpha + a.as[i] * -omega;
      a.r[i] = r;
      sa += r * a.rhat[i];
      sb += r * r;
    } else {  // BICG_P: p = (p - omega ap) beta + r
      double pv = a.p[i];
      pv += a.ap[i] * -omega;
      const double pn = pv * beta + a.r[i];
      a.p[i] = pn;
      if (a.w32) a.w32[i] = float(pn);
    }
  }
  if (MODE == BICG_INIT || MODE == BICG_S || MODE == BICG_XR) {
    const double r = block_reduce<RED_SUM>(sa);
    if (threadIdx.x == 0) a.pa[blockIdx.x] = r;
  }
  if (MODE == BICG_INIT || MODE == BICG_XR) {
    const double r = block_reduce<RED_SUM>(sb);
    if (threadIdx.x == 0) a.pb[blockIdx.x] = r;
  }
  if ((MODE == BICG_INIT || MODE == BICG_S || MODE == BICG_XR) && blockIdx.x == 0) {
    for (int k = (int)gridDim.x + threadIdx.x; k < nslots; k += blockDim.x) {
      a.pa[k] = 0.0;
      if (MODE != BICG_S) a.pb[k] = 0.0;
    }
  }
}

// The operator A (the Poisson star, star_slice) applied to a compact vector one slice at a
// time by one wavefront, as c3_star_dot_kernel applies it: slices z - 1, z, z + 1 of the
// lane's row in registers (widened to FP64), the face ghosts from the neighbours (x faces
// from the face columns xu when given), so A u has the operator launch's bits.  init(p)
// before slice 0; step(z, acc) forms A u on slice z; next() moves on; cur holds slice z.
template <typename TU>
struct StarWalk {
  const TU* __restrict__ u;
  const TU* __restrict__ xu;
  int base, ro, lane, c, g;
  double bc;
  double cur[4], prv[4], nxt[4];
  // raw loads one slice ahead of their use (as in c3_star_dot_kernel): u1 slice z + 1 (at
  // z = N - 1 the neighbour's slice 0), u2 the one after, x0 / y0 and x1 / y1 the faces of
  // slices z and z + 1
  TU u1[4], u2[4], x0, x1, y0[4], y1[4];
  __device__ __forceinline__ void init(int p, const int (&q)[6])
  {
    base = p * NC;
    ld4(u, base + ro, cur);
    z_ghost(u, q[4], N - 1, ro, bc, cur, prv);
    ld4(u, base + 256 + ro, u1);
    x0 = TU(0);
    x1 = TU(0);
    star_faces(u, xu, q, c, g, 0, x0, y0);
  }
  __device__ __forceinline__ void step(int z, const int (&q)[6], double ihx, double ihy, double ihz, double (&acc)[4])
  {
    if (z + 2 < N) ld4(u, base + 256 * (z + 2) + ro, u2);
    else if (z + 2 == N && q[5] >= 0) ld4(u, q[5] * NC + ro, u2);
    if (z + 1 < N) star_faces(u, xu, q, c, g, z + 1, x1, y1);
#pragma unroll
    for (int k = 0; k < 4; k++) nxt[k] = (z < N - 1 || q[5] >= 0) ? double(u1[k]) : bc * cur[k];
    star_slice_f(q, lane, c, g, bc, ihx, ihy, ihz, x0, y0, cur, prv, nxt, acc);
  }
  __device__ __forceinline__ void next()
  {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      prv[k] = cur[k];
      cur[k] = nxt[k];
      u1[k] = u2[k];
      y0[k] = y1[k];
    }
    x0 = x1;
  }
};

// BiCGStab's s, x / r and p updates (Iterative/BiCGStab.h:150-203) on compact vectors with the
// operator products formed in the launch from M p and M s, as the operator launches form them
// (StarWalk: the same stencil code on the same widened values, so the same bits), so that
// A M p and A M s never go through memory and the operator launches only form dot products:
//   BICG_S:  s = r - alpha A M p (+ s32), sum s.s (alpha folded from pfold when given)
//   BICG_XR: x += alpha M p + omega M s; r = s - omega A M s; sums r.rhat, r.r.  The
//            reference forms r - alpha A M p - omega A M s (:181); with s = r - alpha A M p
//            at hand this is the same vector with one product instead of two, rounded in
//            another order.  After the s-norm test fired, x += alpha dir only (dir of type
//            TM when dir_m).
//   BICG_P:  p = r + beta (p - omega A M p) (+ p32)
// APF (s and p updates): A M p from M p here (needs M p apart from p: the preconditioned
// iteration); else read from a.ap.  One wavefront per patch, persistent over the patches;
// xmp, xms: the face columns of M p and M s (or nullptr).
// R32 (x / r update): rhat from a.rh32, its FP32 copy, widened as read.
template <int MODE, typename TM, bool APF, bool R32, bool XD>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(TM) == 4 ? (MODE == BICG_XR && !XD ? 3 : 4) : 1, 8))) c3_upd_kernel(const DevPatch* __restrict__ dp, C3Bicg a,
                                                     const TM* __restrict__ xmp, const TM* __restrict__ xms, int np,
                                                     int nslots, int neumann)
{
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
  const int ro = cidx(4 * g, c, 0);
  const bool fold = MODE == BICG_S && a.pfold != nullptr;
  double* __restrict__ scal = a.scal;
  double alpha = fold ? 0.0 : scal[SC_ALPHA];
  const double omega = scal[SC_OMEGA], beta = scal[SC_BETA], rho = scal[SC_RHO], st = scal[SC_STATE];
  if (MODE == BICG_XR && a.dir && st == ST_EARLY) {
    for (int p = blockIdx.x * 4 + wv; p < np; p += gridDim.x * 4)
      for (int z = 0; z < N; z++) {
        const int o = p * NC + 256 * z + ro;
        double dv[4], xv[4];
        if (a.dir_m) ld4((const TM*)a.dir, o, dv);
        else ld4((const double*)a.dir, o, dv);
        if (a.xpo) {  // x deferred: the caller's padded x (pad_of: the row's 4 cells are contiguous)
          const int q = pad_of(o);
#pragma unroll
          for (int k = 0; k < 4; k++) a.xpo[q + k] += dv[k] * alpha;
        } else {
          ld4(a.x, o, xv);
#pragma unroll
          for (int k = 0; k < 4; k++) xv[k] += dv[k] * alpha;
          st4(a.x, o, xv);
        }
      }
    return;
  }
  if (st != ST_RUN) return;
  if (fold) {  // FIN_ALPHA: alpha = rho / rhat.dot(ap)
    const double* const parts[3] = { a.pfold, nullptr, nullptr };
    double r[3] = { 0.0, 0.0, 0.0 };
    fold_n<1>(parts, a.nfold, r);
    alpha = rho / r[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[SC_ALPHA] = alpha;
  }
  // XD: the x update deferred (launch_c3_xcomb adds alpha M p + omega M s later): this
  // iteration's alpha and omega into its history slot
  if (XD && blockIdx.x == 0 && threadIdx.x == 0) {
    a.hist[0] = alpha;
    a.hist[1] = omega;
  }
  const double bc = neumann ? 1.0 : -1.0;
  StarWalk<TM> wp{ (const TM*)a.mp, xmp, 0, ro, lane, c, g, bc };
  StarWalk<TM> wm{ (const TM*)a.ms, xms, 0, ro, lane, c, g, bc };
  double sa = 0.0, sb = 0.0;
  for (int p = blockIdx.x * 4 + wv; p < np; p += gridDim.x * 4) {
    const DevPatch& P = dp[p];
    int q[6];
#pragma unroll
    for (int s = 0; s < 6; s++) q[s] = P.type[20 + s] >= 0 ? P.nbr[20 + s][0] : -1;
    const double ihx = P.inv_h2[0], ihy = P.inv_h2[1], ihz = P.inv_h2[2];
    constexpr bool WP = APF && MODE != BICG_XR;
    if (WP) wp.init(p, q);
    if constexpr (MODE == BICG_XR) {
      // x / r: M p, x, s, rhat of slice z + 1 loaded (in their stored types) while slice z
      // is formed, widened at their use (XPF; FP64 M p and M s: loaded at the slice, the
      // registers of a second set would cost a wave per SIMD).  XD: r only (x, M p unread)
      constexpr bool XPF = sizeof(TM) == 4;
      typedef typename std::conditional<R32, float, double>::type TH;
      wm.init(p, q);
      TM m0[4], m1[4];
      TH h0[4], h1[4];
      double x0[4], x1[4], s0[4], s1[4];
      auto ldx = [&](int z, TM(&m)[4], double(&xx)[4], double(&ss)[4], TH(&hh)[4]) {
        const int o = p * NC + 256 * z + ro;
        if (!XD) {
          ld4<true>((const TM*)a.mp, o, m);
          ld4(a.x, o, xx);
        }
        ld4(a.sv, o, ss);
        if constexpr (R32) ld4<true>(a.rh32, o, hh);
        else ld4(a.rhat, o, hh);
      };
      if (XPF) ldx(0, m0, x0, s0, h0);
      for (int z = 0; z < N; z++) {
        const int o = p * NC + 256 * z + ro;
        if (XPF && z + 1 < N) ldx(z + 1, m1, x1, s1, h1);
        double asv[4], xv[4], rv[4];
        wm.step(z, q, ihx, ihy, ihz, asv);
        if (!XPF) ldx(z, m0, x0, s0, h0);
#pragma unroll
        for (int k = 0; k < 4; k++) {
          if (!XD) {
            xv[k] = x0[k];
            xv[k] += double(m0[k]) * alpha + wm.cur[k] * omega;
          }
          double r = s0[k];
          r += asv[k] * -omega;
          rv[k] = r;
          sa += r * double(h0[k]);
          sb += r * r;
        }
        if (!XD) st4(a.x, o, xv);
        st4(a.r, o, rv);
        wm.next();
        if (XPF) {
#pragma unroll
          for (int k = 0; k < 4; k++) {
            m0[k] = m1[k];
            h0[k] = h1[k];
            x0[k] = x1[k];
            s0[k] = s1[k];
          }
        }
      }
      continue;
    }
    // S and P: the vectors of slice z (c*: S r (, ap); P p, r (, ap)), loaded one slice ahead (n*)
    double c0[4], c1[4], c2[4], c3[4], n0[4], n1[4], n2[4], n3[4];
    auto load = [&](int z, double(&s0)[4], double(&s1)[4], double(&s2)[4], double(&s3)[4]) {
      const int o = p * NC + 256 * z + ro;
      if (MODE == BICG_P && !a.p) ld4<true>(a.w32, o, s0);  // p kept only as its FP32 copy
      else ld4(MODE == BICG_S ? a.r : a.p, o, s0);
      if (MODE == BICG_P) ld4(a.r, o, s1);
      if (!WP) ld4(a.ap, o, s3);
      (void)s2;
    };
    load(0, c0, c1, c2, c3);
    for (int z = 0; z < N; z++) {
      const int o = p * NC + 256 * z + ro;
      if (z < N - 1) load(z + 1, n0, n1, n2, n3);
      double apv[4];
      if (WP) {
        wp.step(z, q, ihx, ihy, ihz, apv);
      } else {
#pragma unroll
        for (int k = 0; k < 4; k++) apv[k] = c3[k];
      }
      if (MODE == BICG_S) {  // c0: r
        double sv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          double sk = c0[k];
          sk += apv[k] * -alpha;
          sv[k] = sk;
          sa += sk * sk;
        }
        st4(a.sv, o, sv);
        if (a.w32) st4<true>(a.w32, o, sv);
      } else {  // BICG_P, c0: p, c1: r
        double pv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          double v = c0[k];
          v += apv[k] * -omega;
          pv[k] = v * beta + c1[k];
        }
        if (a.p) st4(a.p, o, pv);
        if (a.w32) st4<true>(a.w32, o, pv);
      }
      if (WP) wp.next();
#pragma unroll
      for (int k = 0; k < 4; k++) {
        c0[k] = n0[k];
        c1[k] = n1[k];
        c2[k] = n2[k];
        c3[k] = n3[k];
      }
    }
  }
  if (MODE == BICG_S || MODE == BICG_XR) {
    const double ra = block_reduce<RED_SUM>(sa);
    if (threadIdx.x == 0) a.pa[blockIdx.x] = ra;
  }
  if (MODE == BICG_XR) {
    const double rb = block_reduce<RED_SUM>(sb);
    if (threadIdx.x == 0) a.pb[blockIdx.x] = rb;
  }
  if ((MODE == BICG_S || MODE == BICG_XR) && blockIdx.x == 0)
    for (int k = (int)gridDim.x + threadIdx.x; k < nslots; k += blockDim.x) {
      a.pa[k] = 0.0;
      if (MODE == BICG_XR) a.pb[k] = 0.0;
    }
}

// BiCGStab's setup (Iterative/BiCGStab.h:118-140) from the caller's padded b and x: r = b - A x
// (the star operator with the face ghosts taken from the same-level neighbours' cells of x or
// the physical condition, star_slice's arithmetic), then the compact x, r, rhat = r, p = r
// (+ p32) and the partial sums b.b and r.r: the ghost fill, the residual on padded vectors
// and the packing in one pass.  One wavefront per patch, persistent over the patches.
__global__ void __launch_bounds__(256) c3_init_kernel(const DevPatch* __restrict__ dp, C3Bicg a, int np,
                                                      int nslots, int neumann)
{
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
  const double bc = neumann ? 1.0 : -1.0;
  const double* __restrict__ xp = a.xpad;
  // cell (x, y, z) of patch q in the padded layout (18^3 values per patch)
  auto pidx = [](int q, int x, int y, int z) { return q * (18 * 18 * 18) + (x + 1) + 18 * (y + 1) + 324 * (z + 1); };
  auto prow = [&](const double* __restrict__ v, int q, int y, int z, double (&r)[4]) {
    const int o = pidx(q, 4 * g, y, z);
#pragma unroll
    for (int k = 0; k < 4; k++) r[k] = v[o + k];
  };
  double sa = 0.0, sb = 0.0, sr = 0.0;
  for (int p = blockIdx.x * 4 + wv; p < np; p += gridDim.x * 4) {
    const DevPatch& P = dp[p];
    int q[6];
#pragma unroll
    for (int s = 0; s < 6; s++) q[s] = P.type[20 + s] >= 0 ? P.nbr[20 + s][0] : -1;
    const double ihx = P.inv_h2[0], ihy = P.inv_h2[1], ihz = P.inv_h2[2];
    double cur[4], prv[4], nxt[4];
    prow(xp, p, c, 0, cur);
    if (q[4] >= 0) prow(xp, q[4], c, N - 1, prv);
    else
#pragma unroll
      for (int k = 0; k < 4; k++) prv[k] = bc * cur[k];
    // loads one slice ahead of their use (as in c3_star_dot_kernel): u1 slice z + 1 of x (at
    // z = N - 1 the neighbour's slice 0), u2 the one after, b0 / b1 and the faces xf0, yf0 /
    // xf1, yf1 of slices z and z + 1 (lanes g = 0, 3: the facing x-face cell; c = 0, N - 1:
    // the facing y-face row)
    auto faces = [&](int z, double& xf, double (&yf)[4]) {
      if (g == 0 && q[0] >= 0) xf = xp[pidx(q[0], N - 1, c, z)];
      if (g == 3 && q[1] >= 0) xf = xp[pidx(q[1], 0, c, z)];
      if (c == 0 && q[2] >= 0) prow(xp, q[2], N - 1, z, yf);
      if (c == N - 1 && q[3] >= 0) prow(xp, q[3], 0, z, yf);
    };
    double u1[4], u2[4], b0[4], b1[4], xf0 = 0.0, xf1 = 0.0, yf0[4], yf1[4];
    prow(xp, p, c, 1, u1);
    prow(a.b, p, c, 0, b0);
    faces(0, xf0, yf0);
    for (int z = 0; z < N; z++) {
      if (z + 2 < N) prow(xp, p, c, z + 2, u2);
      else if (z + 2 == N && q[5] >= 0) prow(xp, q[5], c, 0, u2);
      if (z + 1 < N) {
        prow(a.b, p, c, z + 1, b1);
        faces(z + 1, xf1, yf1);
      }
#pragma unroll
      for (int k = 0; k < 4; k++) nxt[k] = (z < N - 1 || q[5] >= 0) ? u1[k] : bc * cur[k];
      // the operator on slice z (star_slice with the faces from the padded neighbours)
      double xm = __shfl(cur[3], (lane - 16) & 63, 64);
      double xq = __shfl(cur[0], (lane + 16) & 63, 64);
      double ym[4], yp[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        ym[k] = __shfl(cur[k], (lane - 1) & 63, 64);
        yp[k] = __shfl(cur[k], (lane + 1) & 63, 64);
      }
      if (g == 0) xm = q[0] >= 0 ? xf0 : bc * cur[0];
      if (g == 3) xq = q[1] >= 0 ? xf0 : bc * cur[3];
      if (c == 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) ym[k] = q[2] >= 0 ? yf0[k] : bc * cur[k];
      }
      if (c == N - 1) {
#pragma unroll
        for (int k = 0; k < 4; k++) yp[k] = q[3] >= 0 ? yf0[k] : bc * cur[k];
      }
      double bv[4];
#pragma unroll
      for (int k = 0; k < 4; k++) bv[k] = b0[k];
      double rv[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const double mid = cur[k];
        const double lo = k == 0 ? xm : cur[k - 1];
        const double hi = k == 3 ? xq : cur[k + 1];
        double acc = 0.0;
        acc += (hi - 2.0 * mid + lo) * ihx;
        acc += (yp[k] - 2.0 * mid + ym[k]) * ihy;
        acc += (nxt[k] - 2.0 * mid + prv[k]) * ihz;
        rv[k] = bv[k] - acc;
        sa += bv[k] * bv[k];
        sb += rv[k] * rv[k];
        if (a.rh32) sr += double(float(rv[k])) * rv[k];  // rhat.r with the FP32 rhat
      }
      const int o = p * NC + cidx(4 * g, c, z);
      if (a.x) st4(a.x, o, cur);
      st4(a.r, o, rv);
      if (a.rh32) st4<true>(a.rh32, o, rv);
      else st4(a.rhat, o, rv);
      if (a.p) st4(a.p, o, rv);
      if (a.w32) st4<true>(a.w32, o, rv);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        prv[k] = cur[k];
        cur[k] = nxt[k];
        u1[k] = u2[k];
        b0[k] = b1[k];
        yf0[k] = yf1[k];
      }
      xf0 = xf1;
    }
  }
  const double ra = block_reduce<RED_SUM>(sa);
  if (threadIdx.x == 0) a.pa[blockIdx.x] = ra;
  const double rb = block_reduce<RED_SUM>(sb);
  if (threadIdx.x == 0) a.pb[blockIdx.x] = rb;
  if (a.rh32) {
    const double rr = block_reduce<RED_SUM>(sr);
    if (threadIdx.x == 0) a.pr[blockIdx.x] = rr;
  }
  if (blockIdx.x == 0)
    for (int k = (int)gridDim.x + threadIdx.x; k < nslots; k += blockDim.x) {
      a.pa[k] = 0.0;
      a.pb[k] = 0.0;
      if (a.rh32) a.pr[k] = 0.0;
    }
}

// The boundary layer of u + P uc (DirectInterpolator, GMG/DirectInterpolator.h:60-91) and its
// face columns: the values the interpolating sweep (c3_sweep_kernel mode 2) reads from u and
// uc, written into out / xout, so that the first post-smoothing sweep can run as a mode-1
// sweep on out (the same values, so the same bits) without gathering coarse values.  Cells
// written: slices z = 0 and 15, rows y = 0 and 15, and the face columns of every row; the
// rest of out is left alone (no sweep reads it).  One wavefront per patch.
template <typename T, typename TC>
__global__ void __launch_bounds__(256) c3_face_interp_kernel(const T* __restrict__ u, const T* __restrict__ xu,
                                                             const TC* __restrict__ uc,
                                                             const int32_t* __restrict__ parent,
                                                             const int32_t* __restrict__ porth, T* __restrict__ out,
                                                             T* __restrict__ xout, int np)
{
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
  const int p = blockIdx.x * 4 + wv;
  if (p >= np) return;
  const C3U<T, TC, true> U{ u, uc, parent, porth };
  const int cp = parent[p], o = porth[p];
  const int side = g == 0 ? 0 : 1;
  for (int z = 0; z < N; z++) {
    if (z == 0 || z == N - 1 || c == 0 || c == N - 1) {
      T r[4];
      U.template row4<true>(p, 4 * g, c, z, r);
      st4<true>(out, p * NC + cidx(4 * g, c, z), r);
    }
    if (g == 0 || g == 3) {
      T v = at(xu, xci(p, side, c, z));
      v = T(v + T(at(uc, cp * NC + U.coarse(o, side ? N - 1 : 0, c, z))));
      at(xout, xci(p, side, c, z)) = v;
    }
  }
}

// launch_c3_xcomb: four cells per thread, grid-stride
__global__ void __launch_bounds__(256) c3_xcomb_kernel(C3XComb c, long n4)
{
  int cnt = c.count;
  if (c.drop_early && c.scal[SC_STATE] == ST_EARLY) cnt--;
  for (long i4 = (long)blockIdx.x * 256 + threadIdx.x; i4 < n4; i4 += (long)gridDim.x * 256) {
    const int o = (int)(4 * i4);
    const int q = c.padded ? pad_of(o) : o;  // a row's 4 cells are contiguous either way
    double xv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) xv[k] = c.x[q + k];
    for (int j = 0; j < cnt; j++) {
      double m[4], w[4];
      ld4<true>(c.mp[j], o, m);
      ld4<true>(c.ms[j], o, w);
      const double al = c.hist[2 * j], om = c.hist[2 * j + 1];
#pragma unroll
      for (int k = 0; k < 4; k++) xv[k] += m[k] * al + w[k] * om;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) c.x[q + k] = xv[k];
  }
}

template <typename TD, typename TS, bool TO_PADDED>
__global__ void __launch_bounds__(256) c3_copy_kernel(TD* __restrict__ dst, const TS* __restrict__ src, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (TO_PADDED) dst[pad_of((int)i)] = TD(src[i]);
  else dst[i] = TD(src[pad_of((int)i)]);
}

template <typename T, typename TC>
__global__ void __launch_bounds__(256) c3_interp_kernel(const int32_t* __restrict__ parent,
                                                        const int32_t* __restrict__ porth, const TC* __restrict__ uc,
                                                        const T* __restrict__ u, T* __restrict__ out, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = (int)(i >> 12), r = (int)(i & 4095);
  const int x = r & 15, y = (r >> 4) & 15, z = r >> 8;
  const C3U<T, TC, true> U{ u, uc, parent, porth };
  out[i] = U.one(p, x, y, z);
}

template <typename T>
__global__ void __launch_bounds__(256) c3_zero_kernel(T* __restrict__ x, long n)
{
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = T(0);
}

int blocks_for(long n) { return (int)((n + 255) / 256); }

}  // namespace

bool c3_supported(const te_domain* d)
{
  if (d->dim != 3 || d->ns[0] != N || d->ns[1] != N || d->ns[2] != N) return false;
  for (int p = 0; p < d->np; p++)
    for (int s = 0; s < 6; s++) {
      const int t = d->patches[p].nbr_type[20 + s];
      if (t != TE_NBR_NONE && t != TE_NBR_NORMAL) return false;
    }
  return true;
}

int launch_c3_sweep(const te_dft_plan* pl, int dt, int dtf, const void* f, const void* uin, int dto, void* out,
                    int mode, int dtc, const void* uc, const te_level_link* link, hipStream_t s, const void* xin,
                    void* xout)
{
  const te_domain* d = pl->g->d;
  if (d->np == 0) return TE_OK;
  if (mode == 2 && (!link || !uc)) {
    set_error("internal: the interpolating sweep needs the coarser level");
    return TE_ERR_ARG;
  }
  const int32_t* par = link ? link->d_parent : nullptr;
  const int32_t* por = link ? link->d_orth : nullptr;
  // slices per load batch of modes 1 / 2: the autotuner's choice, else 2 (FP32) / 1 (FP64)
  const int var = tune_variant(TUNE_C3_SWEEP, 3, N, dt);
  // levels below kC3InterpMin patches fill a fraction of the chip, and a sweep's time is its
  // wavefronts' chain of loads: there all four slices' loads go in one batch
  const int sb = d->np < kC3InterpMin ? 4 : c3_sweep_batch(var, dt);
  const int wpe = c3_sweep_waves(var);  // with face columns only (the cycle's own path)
  const bool fa = c3_sweep_f_ahead(var);
  // face columns: written when out has them; then uin (modes 1 and 2) must have them too
  const bool xc = xout != nullptr;
  if (xc && mode != 0 && !xin) {
    set_error("internal: a sweep that writes face columns needs its input's");
    return TE_ERR_ARG;
  }
#define TE_C3X(T_, TF_, TO_, TC_, M_, SB_, XC_, W_, FA_)                                                             \
  hipLaunchKernelGGL((c3_sweep_kernel<T_, TF_, TO_, TC_, M_, SB_, XC_, W_, FA_>), dim3(d->np), dim3(256), 0, s,     \
                     d->d_patches,                                                                                  \
                     pl->d_ptype, (const T_*)(sizeof(T_) == 8 ? (const void*)pl->d_mats64 : (const void*)pl->d_mats32), \
                     (const T_*)(sizeof(T_) == 8 ? (const void*)pl->d_s2_64 : (const void*)pl->d_s2_32),             \
                     (const TF_*)f, (const T_*)uin, (TO_*)out, (const TC_*)uc, par, por, (const T_*)xin,   \
                     (TO_*)xout)
#define TE_C3S(T_, TF_, TO_, TC_, M_, SB_)                    \
  do {                                                        \
    if constexpr ((M_ == 0 || SB_ < 4) && sizeof(T_) == 4) {  \
      if (xc && wpe == 5) {                                   \
        TE_C3X(T_, TF_, TO_, TC_, M_, SB_, true, 5, false);   \
        break;                                                \
      }                                                       \
    }                                                         \
    if constexpr (M_ != 0 && SB_ < 4 && sizeof(T_) == 4) {    \
      if (xc && fa) {                                         \
        if (wpe == 6) TE_C3X(T_, TF_, TO_, TC_, M_, SB_, true, 6, true); \
        else TE_C3X(T_, TF_, TO_, TC_, M_, SB_, true, 1, true); \
        break;                                                \
      }                                                       \
    }                                                         \
    if (xc) TE_C3X(T_, TF_, TO_, TC_, M_, SB_, true, 1, false); \
    else TE_C3X(T_, TF_, TO_, TC_, M_, SB_, false, 1, false);   \
  } while (0)
#define TE_C3B(T_, TF_, TO_, TC_, M_)                \
  do {                                               \
    if (sb == 1) TE_C3S(T_, TF_, TO_, TC_, M_, 1);   \
    else if (sb == 2) TE_C3S(T_, TF_, TO_, TC_, M_, 2); \
    else TE_C3S(T_, TF_, TO_, TC_, M_, 4);           \
  } while (0)
#define TE_C3M(T_, TF_, TO_, TC_)                    \
  do {                                               \
    if (mode == 0) TE_C3S(T_, TF_, TO_, TC_, 0, 4);  \
    else if (mode == 1) TE_C3B(T_, TF_, TO_, TC_, 1); \
    else TE_C3B(T_, TF_, TO_, TC_, 2);               \
  } while (0)
  if (dt == TE_F64 && dtf == TE_F64 && dto == TE_F64 && dtc == TE_F64) TE_C3M(double, double, double, double);
  else if (dt == TE_F32 && dtf == TE_F32 && dto == TE_F32 && dtc == TE_F32) TE_C3M(float, float, float, float);
  else if (dt == TE_F32 && dtf == TE_F32 && dto == TE_F32 && dtc == TE_F64) TE_C3M(float, float, float, double);
  else if (dt == TE_F32 && dtf == TE_F64 && dto == TE_F64 && dtc == TE_F32) TE_C3M(float, double, double, float);
  else if (dt == TE_F32 && dtf == TE_F64 && dto == TE_F64 && dtc == TE_F64) TE_C3M(float, double, double, double);
  else if (dt == TE_F32 && dtf == TE_F64 && dto == TE_F32 && dtc == TE_F32) TE_C3M(float, double, float, float);
  else if (dt == TE_F32 && dtf == TE_F64 && dto == TE_F32 && dtc == TE_F64) TE_C3M(float, double, float, double);
  else if (dt == TE_F32 && dtf == TE_F32 && dto == TE_F64 && dtc == TE_F32) TE_C3M(float, float, double, float);
  else if (dt == TE_F32 && dtf == TE_F32 && dto == TE_F64 && dtc == TE_F64) TE_C3M(float, float, double, double);
  else {
    set_error("unsupported dtype combination for the 3D patch solve");
    return TE_ERR_UNSUPPORTED;
  }
#undef TE_C3M
#undef TE_C3B
#undef TE_C3S
#undef TE_C3X
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_c3_resid_restrict(const te_domain* d, const te_level_link* link, int neumann, int dt, int dtf,
                             const void* f, const void* u, int dtc, void* fc, hipStream_t s, const void* xu)
{
  if (d->np == 0) return TE_OK;
  // levels of fewer than kC3InterpMin patches: four waves per patch, four slices each
  const bool split = d->np < kC3InterpMin;
  const int grid = split ? d->np : (d->np + 3) / 4;
#define TE_C3RZ(T_, TF_, TC_, ZS_)                                                                               \
  hipLaunchKernelGGL((c3_resid_restrict_kernel<T_, TF_, TC_, ZS_>), dim3(grid), dim3(256), 0, s, d->d_patches,  \
                     (const T_*)u, (const TF_*)f, (TC_*)fc, link->d_parent, link->d_orth, d->np, neumann,    \
                     (const T_*)xu)
#define TE_C3R(T_, TF_, TC_)               \
  do {                                     \
    if (split) TE_C3RZ(T_, TF_, TC_, 4);   \
    else TE_C3RZ(T_, TF_, TC_, N);         \
  } while (0)
  if (dt == TE_F64 && dtf == TE_F64 && dtc == TE_F64) TE_C3R(double, double, double);
  else if (dt == TE_F32 && dtf == TE_F32 && dtc == TE_F32) TE_C3R(float, float, float);
  else if (dt == TE_F32 && dtf == TE_F32 && dtc == TE_F64) TE_C3R(float, float, double);
  else if (dt == TE_F32 && dtf == TE_F64 && dtc == TE_F32) TE_C3R(float, double, float);
  else if (dt == TE_F32 && dtf == TE_F64 && dtc == TE_F64) TE_C3R(float, double, double);
  else {
    set_error("unsupported dtype combination for the 3D residual / restriction");
    return TE_ERR_UNSUPPORTED;
  }
#undef TE_C3R
#undef TE_C3RZ
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_c3_star_dot(const te_domain* d, int neumann, int udt, const void* u, double* out, const double* w,
                       double* pa, double* pb, int nslots, hipStream_t s, const void* xu, const float* w32)
{
  if (nslots < 1 || nslots > kFoldMax) {
    set_error("internal: bad partial count");
    return TE_ERR_ARG;
  }
  int grid = (d->np + 3) / 4;
  if (grid > nslots) grid = nslots;
  if (grid < 1) grid = 1;
#define TE_C3D(TU_, Q_, TW_, W_)                                                                                \
  hipLaunchKernelGGL((c3_star_dot_kernel<TU_, Q_, TW_>), dim3(grid), dim3(256), 0, s, d->d_patches, (const TU_*)u, \
                     out, W_, pa, pb, d->np, nslots, neumann, (const TU_*)xu)
  if (w32 && (udt != TE_F32 || pb)) {
    set_error("internal: an FP32 weight vector goes with FP32 M p and one sum");
    return TE_ERR_ARG;
  }
  if (udt == TE_F64) {
    if (pb) TE_C3D(double, true, double, w);
    else TE_C3D(double, false, double, w);
  } else if (w32) {
    TE_C3D(float, false, float, w32);
  } else {
    if (pb) TE_C3D(float, true, double, w);
    else TE_C3D(float, false, double, w);
  }
#undef TE_C3D
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_c3_bicg(int mode, const te_domain* d, const C3Bicg& a, int nslots, hipStream_t s)
{
  if (nslots < 1 || nslots > kFoldMax || (a.pfold && (mode != BICG_S || a.nfold < 1 || a.nfold > kFoldMax))) {
    set_error("internal: bad partial counts for the 3D BiCGStab updates");
    return TE_ERR_ARG;
  }
  const long n = (long)d->np * NC;
  int grid = blocks_for(n);
  if (grid > nslots) grid = nslots;
  if (grid < 1) grid = 1;
#define TE_C3B(M_)                                                                                        \
  do {                                                                                                    \
    if (a.dtm == TE_F32) hipLaunchKernelGGL((c3_bicg_kernel<M_, float>), dim3(grid), dim3(256), 0, s, a, n, nslots);  \
    else hipLaunchKernelGGL((c3_bicg_kernel<M_, double>), dim3(grid), dim3(256), 0, s, a, n, nslots);    \
  } while (0)
  switch (mode) {
    case BICG_INIT: TE_C3B(BICG_INIT); break;
    case BICG_S: TE_C3B(BICG_S); break;
    case BICG_XR: TE_C3B(BICG_XR); break;
    case BICG_P: TE_C3B(BICG_P); break;
    default:
      set_error("internal: unsupported 3D BiCGStab update");
      return TE_ERR_ARG;
  }
#undef TE_C3B
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_c3_upd(int mode, const te_domain* d, int neumann, const C3Bicg& a, bool apf, const void* xmp,
                  const void* xms, int nslots, hipStream_t s)
{
  if (nslots < 1 || nslots > kFoldMax || (a.pfold && (mode != BICG_S || a.nfold < 1 || a.nfold > kFoldMax)) ||
      (mode != BICG_S && mode != BICG_XR && mode != BICG_P) || (mode == BICG_P && !a.p && !a.w32)) {
    set_error("internal: bad arguments for the 3D BiCGStab updates");
    return TE_ERR_ARG;
  }
  int grid = (d->np + 3) / 4;
  if (grid > nslots) grid = nslots;
  if (grid < 1) grid = 1;
#define TE_C3U5(M_, TM_, APF_, R32_, XD_)                                                                        \
  hipLaunchKernelGGL((c3_upd_kernel<M_, TM_, APF_, R32_, XD_>), dim3(grid), dim3(256), 0, s, d->d_patches, a,     \
                     (const TM_*)xmp, (const TM_*)xms, d->np, nslots, neumann)
#define TE_C3U(M_, TM_, APF_) TE_C3U5(M_, TM_, APF_, false, false)
#define TE_C3UM(M_)                                      \
  do {                                                   \
    if (a.dtm == TE_F32) {                               \
      if (apf) TE_C3U(M_, float, true);                  \
      else TE_C3U(M_, float, false);                     \
    } else {                                             \
      if (apf) TE_C3U(M_, double, true);                 \
      else TE_C3U(M_, double, false);                    \
    }                                                    \
  } while (0)
  if (mode == BICG_S) {
    TE_C3UM(BICG_S);
  } else if (mode == BICG_XR) {  // forms A M s only
    if ((a.rh32 || a.hist) && (a.dtm != TE_F32 || !a.rh32)) {
      set_error("internal: the FP32 rhat and the deferred x update go with FP32 M p and M s");
      return TE_ERR_ARG;
    }
    if (a.hist) TE_C3U5(BICG_XR, float, false, true, true);
    else if (a.rh32) TE_C3U5(BICG_XR, float, false, true, false);
    else if (a.dtm == TE_F32) TE_C3U(BICG_XR, float, false);
    else TE_C3U(BICG_XR, double, false);
  } else {
    TE_C3UM(BICG_P);
  }
#undef TE_C3UM
#undef TE_C3U
#undef TE_C3U5
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_c3_init(const te_domain* d, int neumann, const C3Bicg& a, int nslots, hipStream_t s)
{
  if (nslots < 1 || nslots > kFoldMax) {
    set_error("internal: bad partial count");
    return TE_ERR_ARG;
  }
  int grid = (d->np + 3) / 4;
  if (grid > nslots) grid = nslots;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(c3_init_kernel, dim3(grid), dim3(256), 0, s, d->d_patches, a, d->np, nslots, neumann);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_c3_face_interp(const te_level_link* link, int dt, const void* u, const void* xu, int dtc, const void* uc,
                          void* out, void* xout, hipStream_t s)
{
  const int np = link->fine->np;
  if (np == 0) return TE_OK;
  if (!u || !xu || !uc || !out || !xout) {
    set_error("internal: the boundary interpolation needs u, uc and face columns");
    return TE_ERR_ARG;
  }
  const int grid = (np + 3) / 4;
#define TE_C3F(T_, TC_)                                                                                       \
  hipLaunchKernelGGL((c3_face_interp_kernel<T_, TC_>), dim3(grid), dim3(256), 0, s, (const T_*)u, (const T_*)xu, \
                     (const TC_*)uc, link->d_parent, link->d_orth, (T_*)out, (T_*)xout, np)
  if (dt == TE_F64 && dtc == TE_F64) TE_C3F(double, double);
  else if (dt == TE_F32 && dtc == TE_F64) TE_C3F(float, double);
  else if (dt == TE_F32 && dtc == TE_F32) TE_C3F(float, float);
  else {
    set_error("unsupported dtype combination for the 3D boundary interpolation");
    return TE_ERR_UNSUPPORTED;
  }
#undef TE_C3F
  TE_HIP(hipGetLastError());
  return TE_OK;
}

template <bool TO_PADDED>
static int copy_t(const te_domain* d, int dtd, void* dst, int dts, const void* src, hipStream_t s)
{
  const long n = (long)d->np * NC;
  if (n == 0) return TE_OK;
  const int grid = blocks_for(n);
#define TE_C3C(TD_, TS_) \
  hipLaunchKernelGGL((c3_copy_kernel<TD_, TS_, TO_PADDED>), dim3(grid), dim3(256), 0, s, (TD_*)dst, (const TS_*)src, n)
  if (dtd == TE_F64 && dts == TE_F64) TE_C3C(double, double);
  else if (dtd == TE_F64 && dts == TE_F32) TE_C3C(double, float);
  else if (dtd == TE_F32 && dts == TE_F64) TE_C3C(float, double);
  else TE_C3C(float, float);
#undef TE_C3C
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_c3_xcomb(const te_domain* d, const C3XComb& c, hipStream_t s)
{
  if (c.count < 0 || c.count > kC3XSlots) {
    set_error("internal: bad count of deferred x updates");
    return TE_ERR_ARG;
  }
  const long n4 = (long)d->np * NC / 4;
  if (n4 == 0 || c.count == 0) return TE_OK;
  const int grid = (int)std::min<long>((n4 + 255) / 256, 8192);
  hipLaunchKernelGGL(c3_xcomb_kernel, dim3(grid), dim3(256), 0, s, c, n4);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_c3_pack(const te_domain* d, int dtd, void* dst, int dts, const void* src_padded, hipStream_t s)
{
  return copy_t<false>(d, dtd, dst, dts, src_padded, s);
}

int launch_c3_unpack(const te_domain* d, int dtd, void* dst_padded, int dts, const void* src, hipStream_t s)
{
  return copy_t<true>(d, dtd, dst_padded, dts, src, s);
}

int launch_c3_interp(const te_level_link* link, int dtc, const void* uc, int dt, const void* u, void* out,
                     hipStream_t s)
{
  const long n = (long)link->fine->np * NC;
  if (n == 0) return TE_OK;
  const int grid = blocks_for(n);
#define TE_C3I(T_, TC_)                                                                                          \
  hipLaunchKernelGGL((c3_interp_kernel<T_, TC_>), dim3(grid), dim3(256), 0, s, link->d_parent, link->d_orth,     \
                     (const TC_*)uc, (const T_*)u, (T_*)out, n)
  if (dt == TE_F64 && dtc == TE_F64) TE_C3I(double, double);
  else if (dt == TE_F32 && dtc == TE_F64) TE_C3I(float, double);
  else if (dt == TE_F32 && dtc == TE_F32) TE_C3I(float, float);
  else {
    set_error("unsupported dtype combination for the 3D interpolation");
    return TE_ERR_UNSUPPORTED;
  }
#undef TE_C3I
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_c3_zero(long n, int dt, void* x, hipStream_t s)
{
  if (n == 0) return TE_OK;
  if (dt == TE_F64) hipLaunchKernelGGL((c3_zero_kernel<double>), dim3(blocks_for(n)), dim3(256), 0, s, (double*)x, n);
  else hipLaunchKernelGGL((c3_zero_kernel<float>), dim3(blocks_for(n)), dim3(256), 0, s, (float*)x, n);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
