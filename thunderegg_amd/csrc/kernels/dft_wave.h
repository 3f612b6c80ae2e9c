// Building blocks of the matrix-core patch solve shared by the DFT launches: the MFMA
// operand layouts, the per-wavefront 2D patch solve (dft_solve_wave) and the accessor of
// u (+ the coarser level's correction) used by the on-the-fly ghost formulas.
#pragma once

#include "../te_internal.h"
#include "ghost2d.h"

namespace te {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mf;
template <>
struct Mf<float> {
  typedef floatx4 V;
  static __device__ __forceinline__ V mma(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  // row of accumulator register r in lane group g
  static __device__ __forceinline__ int row(int r, int g) { return 4 * g + r; }
};
template <>
struct Mf<double> {
  typedef doublex4 V;
  static __device__ __forceinline__ V mma(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ __forceinline__ int row(int r, int g) { return 4 * r + g; }
};

// u (+ injection of the coarser level's uc) at interior cell (x, y) of patch q
template <typename T, typename TC, bool INTERP>
struct UAcc {
  const T* u;
  const TC* uc;
  const int32_t* parent;
  const int32_t* porth;
  Layout L;
  __device__ __forceinline__ T operator()(int q, int x, int y) const
  {
    T v = u[q * L.ps + (y + 1) * L.sx + x + 1];
    if (INTERP) {
      int o = porth[q];
      int cx = x, cy = y;
      if (o >= 0) {
        cx = (x + ((o & 1) ? L.n[0] : 0)) / 2;
        cy = (y + ((o & 2) ? L.n[1] : 0)) / 2;
      }
      v = T(v + T(uc[parent[q] * L.ps + (cy + 1) * L.sx + cx + 1]));
    }
    return v;
  }
};

template <typename T, typename TC>
struct FusedArgs {
  const TC* uc;
  const int32_t* parent;
  const int32_t* porth;
};

// The patch solve of one wavefront: X (LDS, n x ld, the modified right-hand side) ->
// out[base + padded (x, y)] (interior only), lx = 4 / hx^2, ly = 4 / hy^2.  X is
// overwritten (the wave's own LDS; no workgroup barrier).  active = false: the wave
// skips the arithmetic.
template <typename T, typename TO, int NT>
__device__ __forceinline__ void dft_solve_wave(T* X, const T* __restrict__ Sm, const T* __restrict__ S3,
                                               const T* __restrict__ s2, T lx, T ly, TO* __restrict__ out,
                                               int base, int sx, bool active)
{
  constexpr int n = 16 * NT;
  constexpr int ld = n + 1;
  typedef typename Mf<T>::V V;
  const int lane = threadIdx.x & 63;
  const int c = lane & 15;
  const int g = lane >> 4;
  if (active) {
    // ---- Y = F S^T : Y[i][j] = sum_k F[i][k] S[j][k]
    V Y[NT][NT];
#pragma unroll
    for (int I = 0; I < NT; I++)
#pragma unroll
      for (int J = 0; J < NT; J++) {
        V acc = { 0, 0, 0, 0 };
#pragma unroll
        for (int K = 0; K < NT; K++)
#pragma unroll
          for (int s = 0; s < 4; s++) {
            int k = 16 * K + 4 * s + g;
            acc = Mf<T>::mma(X[(16 * I + c) * ld + k], Sm[(16 * J + c) * ld + k], acc);
          }
        Y[I][J] = acc;
      }
    // ---- Z = S Y : the B operand comes from Y's registers (K index = row of register s)
    V Z[NT][NT];
#pragma unroll
    for (int I = 0; I < NT; I++)
#pragma unroll
      for (int J = 0; J < NT; J++) {
        V acc = { 0, 0, 0, 0 };
#pragma unroll
        for (int K = 0; K < NT; K++)
#pragma unroll
          for (int s = 0; s < 4; s++) {
            int k = 16 * K + Mf<T>::row(s, g);
            acc = Mf<T>::mma(Sm[(16 * I + c) * ld + k], Y[K][J][s], acc);
          }
        Z[I][J] = acc;
      }
    // ---- divide by the eigenvalues -4/hx^2 s2[kx] - 4/hy^2 s2[ky] (getEigenValues, :320-361)
    {
#pragma unroll
      for (int I = 0; I < NT; I++)
#pragma unroll
        for (int J = 0; J < NT; J++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            int ky = 16 * I + Mf<T>::row(r, g), kx = 16 * J + c;
            T eig = T(0);
            eig -= lx * s2[kx];
            eig -= ly * s2[ky];
            // FP32: the hardware reciprocal (1 ulp) instead of a correctly rounded division
            if (sizeof(T) == 4) Z[I][J][r] = Z[I][J][r] * T(__builtin_amdgcn_rcpf(float(eig)));
            else Z[I][J][r] = Z[I][J][r] / eig;
          }
    }
    // ---- V = S3 Z'
#pragma unroll
    for (int I = 0; I < NT; I++)
#pragma unroll
      for (int J = 0; J < NT; J++) {
        V acc = { 0, 0, 0, 0 };
#pragma unroll
        for (int K = 0; K < NT; K++)
#pragma unroll
          for (int s = 0; s < 4; s++) {
            int k = 16 * K + Mf<T>::row(s, g);
            acc = Mf<T>::mma(S3[(16 * I + c) * ld + k], Z[K][J][s], acc);
          }
        Y[I][J] = acc;  // reuse Y for V
      }
    // V -> LDS (accumulator layout), then U = V S3^T with V in the A layout
#pragma unroll
    for (int I = 0; I < NT; I++)
#pragma unroll
      for (int J = 0; J < NT; J++)
#pragma unroll
        for (int r = 0; r < 4; r++) X[(16 * I + Mf<T>::row(r, g)) * ld + 16 * J + c] = Y[I][J][r];
  }
  wave_sync();  // V in LDS, read back in the A-operand layout
  if (active) {
    const T scale = T(4.0 / double(n * n));  // prod over axes of 2/n (:442-449)
#pragma unroll
    for (int I = 0; I < NT; I++)
#pragma unroll
      for (int J = 0; J < NT; J++) {
        V acc = { 0, 0, 0, 0 };
#pragma unroll
        for (int K = 0; K < NT; K++)
#pragma unroll
          for (int s = 0; s < 4; s++) {
            int k = 16 * K + 4 * s + g;
            acc = Mf<T>::mma(X[(16 * I + c) * ld + k], S3[(16 * J + c) * ld + k], acc);
          }
#pragma unroll
        for (int r = 0; r < 4; r++) {
          int y = 16 * I + Mf<T>::row(r, g), x = 16 * J + c;
          at(out, base + (y + 1) * sx + x + 1) = TO(acc[r] * scale);
        }
      }
  }
}

}  // namespace
}  // namespace te
