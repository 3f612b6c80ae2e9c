// Vector<D> BLAS-1 over interior cells (Vector.h:391-611) and deterministic
// two-pass reductions: a fixed grid writes one partial per block (wavefront
// shuffles, then LDS), a single-block pass folds the partials in a fixed order.
#include "../te_internal.h"
#include "blas.h"

namespace te {
namespace {

__device__ __forceinline__ bool interior_of(const Layout& L, int idx, int* q_out)
{
  int p = idx / L.ps;
  int rem = idx - p * L.ps;
  int c = rem / L.cs;
  int q = rem - c * L.cs;
  *q_out = q;
  int x, y, z = 0;
  if (L.dim == 2) {
    y = q / L.sx;
    x = q - y * L.sx;
  } else {
    z = q / L.sy;
    int r2 = q - z * L.sy;
    y = r2 / L.sx;
    x = r2 - y * L.sx;
    z -= 1;
  }
  x -= 1;
  y -= 1;
  return x >= 0 && x < L.n[0] && y >= 0 && y < L.n[1] && (L.dim == 2 || (z >= 0 && z < L.n[2]));
}

template <typename T>
__global__ void __launch_bounds__(256) vec_op_kernel(int op, Layout L, T* __restrict__ x, const T* __restrict__ a,
                                                     const T* __restrict__ b, double alpha, double beta,
                                                     double gamma)
{
  long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)L.np * L.ps;
  if (gid >= total) return;
  int i = (int)gid;
  int q;
  bool in = interior_of(L, i, &q);
  if (op == VOP_SET_WITH_GHOST) {
    x[i] = T(alpha);
    return;
  }
  if (!in) return;
  switch (op) {
    case VOP_SET: x[i] = T(alpha); break;
    case VOP_SCALE: x[i] *= T(alpha); break;
    case VOP_SHIFT: x[i] += T(alpha); break;
    case VOP_COPY: x[i] = a[i]; break;
    case VOP_ADD: x[i] += a[i]; break;
    case VOP_ADD_SCALED: x[i] += a[i] * T(alpha); break;
    case VOP_ADD_SCALED2: x[i] += a[i] * T(alpha) + b[i] * T(beta); break;
    case VOP_SCALE_THEN_ADD: x[i] = x[i] * T(alpha) + a[i]; break;
    case VOP_SCALE_THEN_ADD_SCALED: x[i] = x[i] * T(alpha) + a[i] * T(beta); break;
    case VOP_SCALE_THEN_ADD_SCALED2: x[i] = x[i] * T(alpha) + a[i] * T(beta) + b[i] * T(gamma); break;
    default: break;
  }
}

template <typename TD, typename TS>
__global__ void __launch_bounds__(256) convert_kernel(long n, TD* __restrict__ d, const TS* __restrict__ s)
{
  long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid < n) d[gid] = TD(s[gid]);
}

}  // namespace

namespace {

template <typename T, int OP>
__global__ void __launch_bounds__(256) reduce_partial_kernel(Layout L, const T* __restrict__ a,
                                                             const T* __restrict__ b, const DevPatch* __restrict__ dp,
                                                             double* __restrict__ partial)
{
  long total = (long)L.np * L.ps;
  double acc = 0.0;  // max starts at 0: Vector::infNorm is max(0, max v)
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    int q;
    if (!interior_of(L, (int)i, &q)) continue;
    double va = double(a[i]);
    if (OP == RED_DOT) {
      acc += va * double(b[i]);
    } else if (OP == RED_SUMSQ) {
      acc += va * va;
    } else if (OP == RED_INTEGRAL) {
      const double* h = dp[(int)i / L.ps].h;  // h[2] = 1 in 2D
      acc += va * (h[0] * h[1] * h[2]);
    } else {
      acc = fmax(acc, va);
    }
  }
  double r = block_reduce<OP>(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

template <int OP>
__global__ void __launch_bounds__(256) reduce_final_kernel(const double* __restrict__ partial, int n,
                                                           double* __restrict__ out, int sqrt_out)
{
  double acc = 0.0;
  // each thread folds a contiguous chunk so the order is fixed
  int chunk = (n + blockDim.x - 1) / blockDim.x;
  int b0 = threadIdx.x * chunk;
  for (int k = b0; k < b0 + chunk && k < n; k++) acc = red_combine<OP>(acc, partial[k]);
  double r = block_reduce<OP>(acc);
  if (threadIdx.x == 0) out[0] = sqrt_out ? sqrt(r) : r;
}

}  // namespace

int reduce_grid(const Layout& L)
{
  long total = (long)L.np * L.ps;
  long g = (total + 255) / 256;
  if (g > kMaxPartials) g = kMaxPartials;
  if (g < 1) g = 1;
  return (int)g;
}

int launch_vec_op(const te_domain* d, int dtype, int op, void* x, const void* a, const void* b, double alpha,
                  double beta, double gamma, int ncomp, hipStream_t s)
{
  if (d->np == 0) return TE_OK;
  Layout L = make_layout(d->dim, d->ns, ncomp, d->np);
  long total = (long)L.np * L.ps;
  int grid = (int)((total + 255) / 256);
  if (dtype == TE_F64)
    hipLaunchKernelGGL((vec_op_kernel<double>), dim3(grid), dim3(256), 0, s, op, L, (double*)x, (const double*)a,
                       (const double*)b, alpha, beta, gamma);
  else
    hipLaunchKernelGGL((vec_op_kernel<float>), dim3(grid), dim3(256), 0, s, op, L, (float*)x, (const float*)a,
                       (const float*)b, alpha, beta, gamma);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_reduce(const te_domain* d, int dtype, int op, const void* a, const void* b, int ncomp, double* d_out,
                  hipStream_t s)
{
  Layout L = make_layout(d->dim, d->ns, ncomp, d->np);
  int g = reduce_grid(L);
  double* part = d->d_scratch;
#define TE_RP(T, OP_)                                                                                     \
  hipLaunchKernelGGL((reduce_partial_kernel<T, OP_>), dim3(g), dim3(256), 0, s, L, (const T*)a, (const T*)b, \
                     d->d_patches, part)
  if (dtype == TE_F64) {
    if (op == RED_DOT) TE_RP(double, RED_DOT);
    else if (op == RED_SUMSQ) TE_RP(double, RED_SUMSQ);
    else if (op == RED_INTEGRAL) TE_RP(double, RED_INTEGRAL);
    else TE_RP(double, RED_MAX);
  } else {
    if (op == RED_DOT) TE_RP(float, RED_DOT);
    else if (op == RED_SUMSQ) TE_RP(float, RED_SUMSQ);
    else if (op == RED_INTEGRAL) TE_RP(float, RED_INTEGRAL);
    else TE_RP(float, RED_MAX);
  }
#undef TE_RP
  if (op == RED_MAX)
    hipLaunchKernelGGL((reduce_final_kernel<RED_MAX>), dim3(1), dim3(256), 0, s, part, g, d_out, 0);
  else
    hipLaunchKernelGGL((reduce_final_kernel<RED_SUM>), dim3(1), dim3(256), 0, s, part, g, d_out,
                       op == RED_SUMSQ ? 1 : 0);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

int launch_convert(long n, int dtd, void* dst, int dts, const void* src, hipStream_t s)
{
  if (n == 0) return TE_OK;
  int grid = (int)((n + 255) / 256);
  if (dtd == TE_F64 && dts == TE_F32)
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid), dim3(256), 0, s, n, (double*)dst,
                       (const float*)src);
  else if (dtd == TE_F32 && dts == TE_F64)
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid), dim3(256), 0, s, n, (float*)dst,
                       (const double*)src);
  else if (dtd == TE_F64)
    hipLaunchKernelGGL((convert_kernel<double, double>), dim3(grid), dim3(256), 0, s, n, (double*)dst,
                       (const double*)src);
  else
    hipLaunchKernelGGL((convert_kernel<float, float>), dim3(grid), dim3(256), 0, s, n, (float*)dst,
                       (const float*)src);
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
