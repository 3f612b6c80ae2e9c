// Cycle::restrict (GMG/Cycle.h:57-65) in one launch, 2D, even n: for every fine
// patch, its ghost values are computed on the fly from its neighbours (the fill
// of MPIGhostFiller::fillGhost, as a gather), r = f - A u is evaluated with the
// star stencil (Poisson/StarPatchOperator.h:75-101, VarPoisson :61-109), and r is
// restricted straight into the coarser level's right-hand side
// (LinearRestrictor::restrictPatches, GMG/LinearRestrictor.h:101-155): a child
// patch writes its quadrant of the parent (2x2 averages), an uncoarsened patch
// copies.  The residual vector itself is never stored.  Each coarse interior
// cell is written by exactly one fine patch; the coarse ghost cells are left
// as they are (zero from allocation; nothing reads the ghosts of a right-hand side).
#include "../te_internal.h"
#include "ghost2d.h"

namespace te {
namespace {

constexpr int kWaves = 4;

template <typename T>
struct PlainAcc {
  const T* u;
  Layout L;
  __device__ __forceinline__ T operator()(int q, int x, int y) const { return u[q * L.ps + (y + 1) * L.sx + x + 1]; }
};

template <typename T, typename TF, typename TC, int OP, int KIND>
__global__ void __launch_bounds__(64 * kWaves) resid_restrict_2d(const DevPatch* __restrict__ dp,
                                                                 const T* __restrict__ u, const T* __restrict__ cf,
                                                                 const TF* __restrict__ f, TC* __restrict__ fc,
                                                                 const int32_t* __restrict__ parent,
                                                                 const int32_t* __restrict__ porth, Layout L,
                                                                 int neumann)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n0 = L.n[0], n1 = L.n[1];
  const int N = n0 * n1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  T* R = reinterpret_cast<T*>(smem_raw) + wave * N;
  PlainAcc<T> U{ u, L };
  const int ngroups = (L.np + kWaves - 1) / kWaves;
  for (int grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int p = grp * kWaves + wave;
    const bool active = p < L.np;
    const int o = active ? porth[p] : -1;
    const int cp = active ? parent[p] : 0;
    if (active) {
      const DevPatch& P = dp[p];
      const T ih2[2] = { T(P.inv_h2[0]), T(P.inv_h2[1]) };
      for (int q = lane; q < N; q += 64) {
        const int x = q % n0, y = q / n0;
        const int i = p * L.ps + (y + 1) * L.sx + x + 1;
        const T mid = u[i];
        const T bc = (OP == 0 && neumann) ? mid : -mid;
        T acc = T(0);
#pragma unroll
        for (int a = 0; a < 2; a++) {
          const int ca = a == 0 ? x : y, na = a == 0 ? n0 : n1, t = a == 0 ? y : x;
          const int st = a == 0 ? 1 : L.sx;
          T lower, upper;
          if (ca == 0) lower = P.type[4 + 2 * a] < 0 ? bc : ghost2d<T, KIND>(P, p, 2 * a, t, na, a == 0 ? n1 : n0, U);
          else lower = u[i - st];
          if (ca == na - 1)
            upper = P.type[5 + 2 * a] < 0 ? bc : ghost2d<T, KIND>(P, p, 2 * a + 1, t, na, a == 0 ? n1 : n0, U);
          else upper = u[i + st];
          if (OP == 0) {
            acc += (upper - T(2) * mid + lower) * ih2[a];
          } else {
            const int ci = p * L.cs + (y + 1) * L.sx + x + 1;
            const T cm = cf[ci], cl = cf[ci - st], cu = cf[ci + st];
            acc += ((cu + cm) * (upper - mid) - (cl + cm) * (mid - lower)) * (ih2[a] * T(0.5));
          }
        }
        const T r = T(f[i]) - acc;
        if (o < 0) fc[cp * L.ps + (y + 1) * L.sx + x + 1] = TC(r);  // copyToParent
        else R[q] = r;
      }
    }
    __syncthreads();
    if (active && o >= 0) {
      // restrictToCoarserParent: coarse[(i + s) / 2] += fine[i] / 4, fine cells in x-fastest order
      const int h0 = n0 / 2, h1 = n1 / 2;
      const int ox = (o & 1) ? h0 : 0, oy = (o & 2) ? h1 : 0;
      const T w = T(0.25);
      for (int q = lane; q < h0 * h1; q += 64) {
        const int X = q % h0, Y = q / h0;
        T acc = T(0);
        acc += R[(2 * Y) * n0 + 2 * X] * w;
        acc += R[(2 * Y) * n0 + 2 * X + 1] * w;
        acc += R[(2 * Y + 1) * n0 + 2 * X] * w;
        acc += R[(2 * Y + 1) * n0 + 2 * X + 1] * w;
        fc[cp * L.ps + (Y + oy + 1) * L.sx + X + ox + 1] = TC(acc);
      }
    }
    __syncthreads();
  }
}

template <typename T, typename TF, typename TC>
void launch_rr(const te_ghost_filler* g, const te_level_link* link, int op_kind, int neumann, const void* coeffs,
               const void* f, const void* u, void* fc, hipStream_t s)
{
  const te_domain* d = g->d;
  Layout L = make_layout(2, d->ns, 1, d->np);
  size_t shmem = sizeof(T) * kWaves * d->ns[0] * d->ns[1];
  int ngroups = (d->np + kWaves - 1) / kWaves;
  int grid = ngroups < 8192 ? ngroups : 8192;
#define TE_RR(OP_, K_)                                                                                      \
  hipLaunchKernelGGL((resid_restrict_2d<T, TF, TC, OP_, K_>), dim3(grid), dim3(64 * kWaves), shmem, s,     \
                     d->d_patches, (const T*)u, (const T*)coeffs, (const TF*)f, (TC*)fc, link->d_parent,   \
                     link->d_orth, L, neumann)
  if (op_kind == 0) {
    if (g->kind == TE_BILINEAR) TE_RR(0, TE_BILINEAR);
    else TE_RR(0, TE_BIQUADRATIC);
  } else {
    if (g->kind == TE_BILINEAR) TE_RR(1, TE_BILINEAR);
    else TE_RR(1, TE_BIQUADRATIC);
  }
#undef TE_RR
}

}  // namespace

bool resid_restrict_fused_supported(const te_ghost_filler* g)
{
  const te_domain* d = g->d;
  return d->dim == 2 && d->ns[0] % 2 == 0 && d->ns[1] % 2 == 0 && d->ns[0] * d->ns[1] <= 4096 &&
         (g->kind == TE_BILINEAR || g->kind == TE_BIQUADRATIC);
}

// dt: u / coeffs / residual type, dtf: f type, dtc: coarse f type
int launch_resid_restrict(const te_ghost_filler* g, const te_level_link* link, int op_kind, int neumann,
                          const void* coeffs, int dt, int dtf, const void* f, const void* u, int dtc, void* fc,
                          hipStream_t s)
{
  if (g->d->np == 0) return TE_OK;
  if (dt == TE_F64 && dtf == TE_F64 && dtc == TE_F64) launch_rr<double, double, double>(g, link, op_kind, neumann, coeffs, f, u, fc, s);
  else if (dt == TE_F32 && dtf == TE_F32 && dtc == TE_F32) launch_rr<float, float, float>(g, link, op_kind, neumann, coeffs, f, u, fc, s);
  else if (dt == TE_F32 && dtf == TE_F64 && dtc == TE_F32) launch_rr<float, double, float>(g, link, op_kind, neumann, coeffs, f, u, fc, s);
  else if (dt == TE_F32 && dtf == TE_F32 && dtc == TE_F64) launch_rr<float, float, double>(g, link, op_kind, neumann, coeffs, f, u, fc, s);
  else if (dt == TE_F32 && dtf == TE_F64 && dtc == TE_F64) launch_rr<float, double, double>(g, link, op_kind, neumann, coeffs, f, u, fc, s);
  else {
    set_error("unsupported dtype combination for the fused residual / restriction");
    return TE_ERR_UNSUPPORTED;
  }
  TE_HIP(hipGetLastError());
  return TE_OK;
}

}  // namespace te
