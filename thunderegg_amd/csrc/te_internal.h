// Internal declarations shared by the host code and the HIP kernels of
// libthunderegg_amd.  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/thunderegg_amd.h"

namespace te {

void set_error(const std::string& msg);

#define TE_HIP(call)                                                                                 \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess) {                                                                          \
      ::te::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                            \
      return TE_ERR_HIP;                                                                             \
    }                                                                                                \
  } while (0)

// Per-patch neighbour table on the device.  Slots as in te_patch_info.
struct DevPatch {
  int type[26];
  int orth[26];
  int nbr[26][4];
  double h[3];
  double inv_h2[3];
  int any_nbr;
  int pad;
};

// Vector layout: [patch][comp][z][y][x], each axis padded by one ghost cell.
struct Layout {
  int dim;
  int n[3];   // cells per axis (n[2] = 1 in 2D)
  int ncomp;
  int sx;     // stride of y  (= n0 + 2)
  int sy;     // stride of z  (= (n0 + 2)(n1 + 2))
  int cs;     // component stride (padded cells of one component)
  int ps;     // patch stride (= cs * ncomp)
  int np;     // number of patches
};

inline Layout make_layout(int dim, const int* n, int ncomp, int np)
{
  Layout L{};
  L.dim = dim;
  L.n[0] = n[0];
  L.n[1] = n[1];
  L.n[2] = dim == 3 ? n[2] : 1;
  L.ncomp = ncomp;
  L.sx = L.n[0] + 2;
  L.sy = (L.n[0] + 2) * (L.n[1] + 2);
  L.cs = dim == 3 ? L.sy * (L.n[2] + 2) : L.sy;
  L.ps = L.cs * ncomp;
  L.np = np;
  return L;
}

// base[idx] for idx >= 0 with the byte offset formed in 32 bits (vectors and tables stay
// below 2 GiB, checked when domains are created), so a load or store can address from the
// kernel argument's scalar base plus a 32-bit vector offset instead of a 64-bit vector address.
template <typename T>
__device__ __forceinline__ const T& at(const T* base, int idx)
{
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (unsigned)idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T& at(T* base, int idx)
{
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (unsigned)idx * (unsigned)sizeof(T));
}

__host__ __device__ inline int lin(const Layout& L, int p, int c, int x, int y, int z)
{
  return p * L.ps + c * L.cs + (x + 1) + (y + 1) * L.sx + (L.dim == 3 ? (z + 1) * L.sy : 0);
}

}  // namespace te

struct te_domain {
  int dim;
  int ns[3];
  int np;
  std::vector<te_patch_info> patches;
  std::vector<int32_t> nbr_local;  // np * 26 * 4
  te::DevPatch* d_patches = nullptr;
  int any_coarse_fine = 0;         // some face/edge/corner is a coarse/fine interface
  double* d_scratch = nullptr;     // reduction partials
  unsigned* d_counter = nullptr;   // last-block counter for reductions
  double* d_result = nullptr;      // reduction results
  int32_t* d_flags = nullptr;      // breakdown flags
  int scratch_len = 0;
  int32_t* d_ring = nullptr;       // 2D: ring table (te::ring_table), pq fields 0
};

struct te_ghost_filler {
  const te_domain* d;
  int kind;
  int fill_type;
};

struct te_level_link {
  const te_domain* fine;
  const te_domain* coarse;
  std::vector<int32_t> parent;      // per fine patch
  int32_t* d_parent = nullptr;      // per fine patch
  int32_t* d_orth = nullptr;        // per fine patch orth_on_parent (-1 copy)
  // per coarse patch: children (fine local indexes) in increasing local-index order
  int32_t* d_child = nullptr;       // [np_coarse][8]
  int32_t* d_nchild = nullptr;      // [np_coarse]
  int32_t* d_ring = nullptr;        // 2D: ring table of the fine level with the pq fields set
};

namespace te {

// 2D ring table: for patch p, side s and half h of that side (h = 1 for ghost cells
// t >= n / 2), the int4 record at [(p * 4 + s) * 2 + h] = {type, orth, q, pq} of what a
// ghost cell there reads: the side's neighbour type and orth_on_coarse, the neighbour
// patch q (for finer neighbours the one facing that half; -1 on a physical side) and,
// when the table belongs to a level link, q's parent * 8 + orth_on_parent(q) + 1 in the
// coarser level (0 otherwise).  One 16-byte load gives a lane everything it needs to
// issue the neighbour reads of its ghost cell.
std::vector<int32_t> ring_table(const te_domain* d, const int32_t* parent, const int32_t* orth_on_parent);

// compute units of the current device (queried once per device)
int cu_count();
// workgroups of `block` threads and `shmem` dynamic LDS bytes of `kernel` resident at once
// on the whole current device (occupancy query, cached)
int resident_blocks(const void* kernel, int block, size_t shmem);

// Timings on a stream for the solvers (timer.cpp): a domain timing (domain_id, or INT_MAX
// for none) started / stopped by events on s.  A null timer makes both no-ops.
int timer_start_stream(te_timer* t, const char* name, int domain_id, hipStream_t s);
int timer_stop_stream(te_timer* t, const char* name, int domain_id, hipStream_t s);
bool timer_has_domain(const te_timer* t, int domain_id);

// ---- launchers (defined in kernels/*.hip) ----
int launch_ghost_fill(const te_ghost_filler* g, int dtype, void* u, int ncomp, hipStream_t s);
int launch_star(const te_ghost_filler* g, int dt, int dtf, int dto, int op_kind, int neumann, const void* coeffs,
                const void* fin, const void* u, void* out, int ncomp, hipStream_t s);
// extrapolate: LinearRestrictor's extrapolate_boundary_ghosts
int launch_restrict(const te_level_link* l, int dtype_in, const void* fine, int dtype_out, void* coarse, int ncomp,
                    hipStream_t s, bool extrapolate = false);
int launch_interpolate(const te_level_link* l, int dtype_c, const void* coarse, int dtype_f, void* fine, int ncomp,
                       hipStream_t s);
// boundary data (kernels/bc.hip): values per patch, sum over sides of the face cell counts
long boundary_stride(const te_domain* d);
int launch_bc_rhs(const te_domain* d, int dtype, void* f, const double* g, const double* c, bool neumann,
                  hipStream_t s);

}  // namespace te
