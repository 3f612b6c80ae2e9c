// Boundary data added to the right-hand side on physical sides:
//  - Poisson::StarPatchOperator::addDrichletBCToRHS (Poisson/StarPatchOperator.h:187-204):
//    f_b += -2 g / h^2, and its VarPoisson form (VarPoisson/StarPatchOperator.h:194-220):
//    f_b -= 2 g c / h^2 with c the coefficient function at the face point;
//  - Poisson::StarPatchOperator::addNeumannBCToRHS (:206-242): f_b += dg / h on lower
//    sides, -= on upper sides, dg the derivative of g along the side's axis.
// g, c and dg are sampled by the caller at the face points of every side
// (DomainTools::GetRealCoordBound) in the boundary-data layout of te_domain_boundary_length.
// One workgroup per patch; the sides are done one after another, in the reference's order,
// because a corner cell belongs to several of them.
#include "../te_internal.h"

namespace te {
namespace {

template <typename T, bool NEUMANN>
__global__ void __launch_bounds__(256) bc_rhs_kernel(const DevPatch* __restrict__ dp, T* __restrict__ f,
                                                     const double* __restrict__ g, const double* __restrict__ c,
                                                     Layout L, int stride)
{
  const int p = blockIdx.x;
  const DevPatch& P = dp[p];
  const int D = L.dim;
  const int sb = D == 2 ? 4 : 20;
  int off = p * stride;
  for (int s = 0; s < 2 * D; s++) {
    const int a = s >> 1;
    const int f0 = a == 0 ? 1 : 0;
    const int f1 = a == 2 ? 1 : 2;  // 3D only
    const int n0 = L.n[f0], n1 = D == 3 ? L.n[f1] : 1;
    const int count = n0 * n1;
    if (P.type[sb + s] < 0) {
      const double h = P.h[a];
      for (int k = threadIdx.x; k < count; k += blockDim.x) {
        int x[3] = { 0, 0, 0 };
        x[a] = (s & 1) ? L.n[a] - 1 : 0;
        x[f0] = k % n0;
        if (D == 3) x[f1] = k / n0;
        const int i = lin(L, p, 0, x[0], x[1], x[2]);
        const double gv = g[off + k];
        double v = double(f[i]);
        if (NEUMANN) {
          if (s & 1) v -= gv / h;
          else v += gv / h;
        } else if (c) {
          v -= 2 * gv * c[off + k] / (h * h);
        } else {
          v += -2.0 * gv / (h * h);
        }
        f[i] = T(v);
      }
    }
    off += count;
    __syncthreads();
  }
}

}  // namespace

long boundary_stride(const te_domain* d)
{
  long s = 0;
  for (int a = 0; a < d->dim; a++) {
    long cnt = 1;
    for (int b = 0; b < d->dim; b++)
      if (b != a) cnt *= d->ns[b];
    s += 2 * cnt;
  }
  return s;
}

int launch_bc_rhs(const te_domain* d, int dtype, void* f, const double* g, const double* c, bool neumann,
                  hipStream_t s)
{
  if (d->np == 0) return TE_OK;
  const Layout L = make_layout(d->dim, d->ns, 1, d->np);
  const int stride = (int)boundary_stride(d);
#define TE_BC(T_, N_)                                                                                        \
  hipLaunchKernelGGL((bc_rhs_kernel<T_, N_>), dim3(d->np), dim3(256), 0, s, d->d_patches, (T_*)f, g, c, L, stride)
  if (dtype == TE_F64) {
    if (neumann) TE_BC(double, true);
    else TE_BC(double, false);
  } else {
    if (neumann) TE_BC(float, true);
    else TE_BC(float, false);
  }
#undef TE_BC
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
