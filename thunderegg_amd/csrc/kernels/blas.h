// Shared pieces of the vector kernels: operation codes, block reductions and
// launcher declarations.
#pragma once

#include "../te_internal.h"

namespace te {

enum VecOp {
  VOP_SET = 0,
  VOP_SET_WITH_GHOST,
  VOP_SCALE,
  VOP_SHIFT,
  VOP_COPY,
  VOP_ADD,
  VOP_ADD_SCALED,
  VOP_ADD_SCALED2,
  VOP_SCALE_THEN_ADD,
  VOP_SCALE_THEN_ADD_SCALED,
  VOP_SCALE_THEN_ADD_SCALED2,
};

// RED_INTEGRAL: sum of the interior values times their patch's cell volume
enum RedOp { RED_SUM = 0, RED_DOT = 0, RED_SUMSQ = 1, RED_MAX = 2, RED_INTEGRAL = 3 };

constexpr int kMaxPartials = 1024;

template <int OP>
__device__ __forceinline__ double red_combine(double a, double b)
{
  return OP == RED_MAX ? fmax(a, b) : a + b;
}

// Reduction of one double per thread over a block of up to 512 threads
// (wavefront shuffles, then one LDS slot per wavefront); result valid in thread 0.
template <int OP>
__device__ __forceinline__ double block_reduce(double v)
{
  __shared__ double wsum[8];
  for (int off = 32; off > 0; off >>= 1) v = red_combine<OP>(v, __shfl_down(v, off, 64));
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    r = wsum[0];
    for (int k = 1; k < nw; k++) r = red_combine<OP>(r, wsum[k]);
  }
  __syncthreads();
  return r;
}

int reduce_grid(const Layout& L);
int launch_vec_op(const te_domain* d, int dtype, int op, void* x, const void* a, const void* b, double alpha,
                  double beta, double gamma, int ncomp, hipStream_t s);
int launch_reduce(const te_domain* d, int dtype, int op, const void* a, const void* b, int ncomp, double* d_out,
                  hipStream_t s);
int launch_convert(long n, int dtd, void* dst, int dts, const void* src, hipStream_t s);

}  // namespace te
