// Fixed-order folds of the per-workgroup partial sums of the Krylov kernels by one
// 256-thread workgroup (device code only).  Thread t sums partials t, t + 256, ... in that
// order, then the block reduction: up to three arrays, every load issued before the first
// add (n <= kFoldMax, the Krylov grid cap), one block reduction pass for all.  Any
// workgroup of 256 threads that folds the same partials gets the same bits.
#pragma once

#include <hip/hip_runtime.h>

namespace te {

constexpr int kFoldThreads = 256;
constexpr int kFoldPer = 8;
constexpr int kFoldMax = kFoldThreads * kFoldPer;
static_assert(kFoldMax >= 2048, "the Krylov kernels write at most 2048 partials");

// the partials a thread folds, held in registers between the loads and the reduction
template <int NA>
struct FoldRegs {
  double v[NA][kFoldPer];
};

template <int NA>
__device__ __forceinline__ void fold_load(const double* const (&part)[3], int n, FoldRegs<NA>& f)
{
#pragma unroll
  for (int a = 0; a < NA; a++)
#pragma unroll
    for (int j = 0; j < kFoldPer; j++) {
      const int k = threadIdx.x + kFoldThreads * j;
      f.v[a][j] = k < n ? part[a][k] : 0.0;
    }
}

// blockDim.x must be kFoldThreads; out[a] is valid in every thread
template <int NA>
__device__ __forceinline__ void fold_reduce(const FoldRegs<NA>& f, double (&out)[3])
{
  __shared__ double wsum[3][kFoldThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < NA; a++) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < kFoldPer; j++) acc += f.v[a][j];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) wsum[a][w] = acc;
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < NA; a++) {
    double r = wsum[a][0];
    for (int k = 1; k < kFoldThreads / 64; k++) r += wsum[a][k];
    out[a] = r;
  }
}

template <int NA>
__device__ void fold_n(const double* const (&part)[3], int n, double (&out)[3])
{
  FoldRegs<NA> f;
  fold_load<NA>(part, n, f);
  fold_reduce<NA>(f, out);
}

}  // namespace te
