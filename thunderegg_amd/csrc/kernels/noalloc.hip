// te_vec_dot_noalloc: Vector::dot (Vector.h:582-595) for callers that cannot allocate
// device memory.  Its partial sums live in `scratch`, a file-scope device array of fixed
// size (one slot per patch, so at most kNoAllocMaxPatches patches), instead of the
// domain's library-allocated reduction workspace; the result goes to the caller's device
// memory on the caller's stream.  Each GPU holds its own copy of `scratch` until the
// process ends, at its own address: the host looks that address up once per device
// (hipGetSymbolAddress, with the device current) and passes it to the kernels.  Every
// call on one GPU shares that GPU's array, so calls that overlap on one GPU race.
#include <mutex>

#include "../te_internal.h"
#include "blas.h"

namespace te {
namespace {

constexpr int kNoAllocMaxPatches = TE_NOALLOC_MAX_PATCHES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocMaxPatches];

// one workgroup per patch: the patch's interior cells over all components, in a fixed
// order, summed into part[p]
template <typename T>
__global__ void __launch_bounds__(256) patch_dot_kernel(Layout L, const T* __restrict__ a, const T* __restrict__ b,
                                                        double* __restrict__ part)
{
  const int p = blockIdx.x;
  const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
  const int cells = L.ncomp * nz * ny * nx;
  const long base = (long)p * L.ps;
  double acc = 0.0;
  for (int k = threadIdx.x; k < cells; k += blockDim.x) {
    const int x = k % nx;
    int r = k / nx;
    const int y = r % ny;
    r /= ny;
    const int z = r % nz;
    const int c = r / nz;
    const long i = base + (long)c * L.cs + (x + 1) + (long)(y + 1) * L.sx + (L.dim == 3 ? (long)(z + 1) * L.sy : 0);
    acc += double(a[i]) * double(b[i]);
  }
  const double s = block_reduce<RED_DOT>(acc);
  if (threadIdx.x == 0) part[p] = s;
}

// one workgroup: the per-patch partials folded in patch order (contiguous chunks per
// thread, then the block reduction), written to *out
__global__ void __launch_bounds__(256) fold_kernel(const double* __restrict__ part, int n, double* __restrict__ out)
{
  const int chunk = (n + (int)blockDim.x - 1) / (int)blockDim.x;
  const int b0 = (int)threadIdx.x * chunk;
  double acc = 0.0;
  for (int k = b0; k < b0 + chunk && k < n; k++) acc += part[k];
  const double s = block_reduce<RED_DOT>(acc);
  if (threadIdx.x == 0) out[0] = s;
}

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

// device address of the current device's `scratch`, looked up on the first call per device
int scratch_address(double** out)
{
  int dev = 0;
  TE_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("te_vec_dot_noalloc: device index out of range");
    return TE_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    TE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("te_vec_dot_noalloc: no device address for the scratch array");
      return TE_ERR_HIP;
    }
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return TE_OK;
}

}  // namespace
}  // namespace te

using namespace te;

extern "C" int te_vec_dot_noalloc(const te_domain* d, int dtype, const void* a, const void* b, int ncomp,
                                  double* d_out, te_stream stream)
{
  if (!d || !a || !b || !d_out || ncomp < 1 || ncomp > 8 || (dtype != TE_F64 && dtype != TE_F32)) {
    set_error("te_vec_dot_noalloc: bad argument");
    return TE_ERR_ARG;
  }
  if (d->np > kNoAllocMaxPatches) {
    set_error("te_vec_dot_noalloc: the domain has " + std::to_string(d->np) + " patches, more than the " +
              std::to_string(kNoAllocMaxPatches) + " its fixed scratch array holds");
    return TE_ERR_ARG;
  }
  double* part = nullptr;
  int rc = scratch_address(&part);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const Layout L = make_layout(d->dim, d->ns, ncomp, d->np);
  if (d->np > 0) {
    if (dtype == TE_F64)
      hipLaunchKernelGGL((patch_dot_kernel<double>), dim3(d->np), dim3(256), 0, s, L, (const double*)a,
                         (const double*)b, part);
    else
      hipLaunchKernelGGL((patch_dot_kernel<float>), dim3(d->np), dim3(256), 0, s, L, (const float*)a,
                         (const float*)b, part);
  }
  hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(256), 0, s, (const double*)part, d->np, d_out);
  TE_HIP(hipGetLastError());
  return TE_OK;
}
