// extern "C" entry points of libthunderegg_amd: tables, ghost filling,
// operators, smoothers, transfers and vector algebra.  Each function checks
// its arguments on the host before anything is enqueued.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "kernels/blas.h"
#include "kernels/dft.h"
#include "kernels/krylov.h"
#include "te_internal.h"

namespace te {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

int cu_count()
{
  static std::atomic<int> cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int v = cus[dev].load(std::memory_order_relaxed);
  if (v == 0) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cus[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

int resident_blocks(const void* kernel, int block, size_t shmem)
{
  static std::mutex mu;
  static std::map<std::tuple<int, const void*, int, size_t>, int> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  auto key = std::make_tuple(dev, kernel, block, shmem);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, shmem) != hipSuccess || per_cu < 1)
    per_cu = 1;
  const int v = per_cu * cu_count();
  std::lock_guard<std::mutex> lk(mu);
  cache.emplace(key, v);
  return v;
}
}  // namespace te

using namespace te;

#define TE_ARG(cond, msg)            \
  do {                               \
    if (!(cond)) {                   \
      te::set_error(msg);            \
      return TE_ERR_ARG;             \
    }                                \
  } while (0)

static inline hipStream_t S(te_stream s) { return (hipStream_t)s; }

std::vector<int32_t> te::ring_table(const te_domain* d, const int32_t* parent, const int32_t* orth_on_parent)
{
  std::vector<int32_t> r((size_t)d->np * 4 * 2 * 4, 0);
  for (int p = 0; p < d->np; p++)
    for (int s = 0; s < 4; s++) {
      const int sl = 4 + s;  // 2D face slots follow the four corner slots
      const te_patch_info& P = d->patches[p];
      const int type = P.nbr_type[sl];
      for (int h = 0; h < 2; h++) {
        int32_t* e = &r[(((size_t)p * 4 + s) * 2 + h) * 4];
        const int q = type < 0 ? -1 : d->nbr_local[((size_t)p * 26 + sl) * 4 + (type == TE_NBR_FINE ? h : 0)];
        e[0] = type;
        e[1] = P.nbr_orth[sl];
        e[2] = q;
        e[3] = (parent && q >= 0) ? parent[q] * 8 + orth_on_parent[q] + 1 : 0;
      }
    }
  return r;
}

static hipError_t upload_ring(const std::vector<int32_t>& r, int32_t** out)
{
  hipError_t e = hipMalloc(out, sizeof(int32_t) * std::max<size_t>(r.size(), 4));
  if (e == hipSuccess && !r.empty()) e = hipMemcpy(*out, r.data(), sizeof(int32_t) * r.size(), hipMemcpyHostToDevice);
  return e;
}

extern "C" {

const char* te_last_error(void) { return te::g_err.c_str(); }

int te_device_count(int* out)
{
  TE_HIP(hipGetDeviceCount(out));
  return TE_OK;
}
int te_device_alloc(size_t bytes, void** out)
{
  TE_ARG(out, "null output");
  TE_HIP(hipMalloc(out, bytes ? bytes : 1));
  return TE_OK;
}
int te_device_free(void* p)
{
  if (p) TE_HIP(hipFree(p));
  return TE_OK;
}
int te_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
  TE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return TE_OK;
}
int te_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
  TE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return TE_OK;
}
int te_memcpy_d2d(void* dst, const void* src, size_t bytes, te_stream s)
{
  TE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(s)));
  return TE_OK;
}
int te_memset(void* dst, int value, size_t bytes, te_stream s)
{
  TE_HIP(hipMemsetAsync(dst, value, bytes, S(s)));
  return TE_OK;
}
int te_stream_synchronize(te_stream s)
{
  TE_HIP(hipStreamSynchronize(S(s)));
  return TE_OK;
}
int te_stream_create(te_stream* out)
{
  hipStream_t s;
  TE_HIP(hipStreamCreate(&s));
  *out = (te_stream)s;
  return TE_OK;
}
int te_stream_destroy(te_stream s)
{
  TE_HIP(hipStreamDestroy(S(s)));
  return TE_OK;
}
int te_event_create(te_event* out)
{
  TE_ARG(out, "null output");
  hipEvent_t e;
  TE_HIP(hipEventCreate(&e));
  *out = (te_event)e;
  return TE_OK;
}
int te_event_destroy(te_event e)
{
  TE_HIP(hipEventDestroy((hipEvent_t)e));
  return TE_OK;
}
int te_event_record(te_event e, te_stream s)
{
  TE_HIP(hipEventRecord((hipEvent_t)e, S(s)));
  return TE_OK;
}
int te_event_elapsed_ms(te_event a, te_event b, float* ms)
{
  TE_ARG(ms, "null output");
  TE_HIP(hipEventSynchronize((hipEvent_t)b));
  TE_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  return TE_OK;
}

// ------------------------------------------------------------------ domain --
static int face_count(int dim) { return dim == 2 ? 8 : 26; }


int te_domain_create(int dim, const int* ns, int num_ghost, int num_patches, const te_patch_info* patches,
                     te_domain** out)
{
  TE_ARG(out, "null output");
  TE_ARG(dim == 2 || dim == 3, "dim must be 2 or 3");
  TE_ARG(num_ghost == 1, "only one ghost layer is supported");
  TE_ARG(num_patches >= 0 && (num_patches == 0 || patches), "bad patch table");
  for (int a = 0; a < dim; a++) TE_ARG(ns[a] >= 1 && ns[a] <= 1024, "bad cells per side");
  long cells = 1;
  for (int a = 0; a < dim; a++) cells *= ns[a] + 2;
  TE_ARG((long)num_patches * cells * 8 < (1L << 31), "domain too large for 32-bit indexing (8 components)");
  auto d = new te_domain();
  d->dim = dim;
  d->ns[0] = ns[0];
  d->ns[1] = ns[1];
  d->ns[2] = dim == 3 ? ns[2] : 1;
  d->np = num_patches;
  d->patches.assign(patches, patches + num_patches);
  std::unordered_map<int, int> id2local;
  id2local.reserve(num_patches * 2);
  for (int i = 0; i < num_patches; i++) {
    if (!id2local.emplace(patches[i].id, i).second) {
      delete d;
      set_error("duplicate patch id " + std::to_string(patches[i].id));
      return TE_ERR_ARG;
    }
  }
  d->nbr_local.assign((size_t)num_patches * 26 * 4, -1);
  std::vector<DevPatch> hp(num_patches);
  int nf = face_count(dim);
  for (int i = 0; i < num_patches; i++) {
    DevPatch& P = hp[i];
    std::memset(&P, 0, sizeof(P));
    P.any_nbr = 0;
    for (int a = 0; a < 3; a++) {
      P.h[a] = a < dim ? patches[i].spacings[a] : 1.0;
      P.inv_h2[a] = 1.0 / (P.h[a] * P.h[a]);
    }
    for (int sl = 0; sl < 26; sl++) {
      int t = sl < nf ? patches[i].nbr_type[sl] : -1;
      P.type[sl] = t;
      P.orth[sl] = patches[i].nbr_orth[sl];
      for (int k = 0; k < 4; k++) P.nbr[sl][k] = -1;
      if (t < 0) continue;
      if (t > 2) {
        delete d;
        set_error("bad neighbour type");
        return TE_ERR_ARG;
      }
      P.any_nbr = 1;
      int m = dim == 2 ? (sl < 4 ? 0 : 1) : (sl < 8 ? 0 : (sl < 20 ? 1 : 2));
      int cnt = t == TE_NBR_FINE ? (1 << m) : 1;
      if (t != TE_NBR_NORMAL) d->any_coarse_fine = 1;
      if (t == TE_NBR_COARSE && m > 0 && (P.orth[sl] < 0 || P.orth[sl] >= (1 << m))) {
        delete d;
        set_error("coarse neighbour without a valid orth_on_coarse");
        return TE_ERR_ARG;
      }
      for (int k = 0; k < cnt; k++) {
        auto it = id2local.find(patches[i].nbr_ids[sl][k]);
        if (it == id2local.end()) {
          delete d;
          set_error("neighbour id " + std::to_string(patches[i].nbr_ids[sl][k]) + " of patch " +
                    std::to_string(patches[i].id) + " is not in the domain (one rank only)");
          return TE_ERR_ARG;
        }
        P.nbr[sl][k] = it->second;
        d->nbr_local[((size_t)i * 26 + sl) * 4 + k] = it->second;
      }
    }
  }
  hipError_t e = hipSuccess;
  if (num_patches) {
    e = hipMalloc(&d->d_patches, sizeof(DevPatch) * num_patches);
    if (e == hipSuccess) e = hipMemcpy(d->d_patches, hp.data(), sizeof(DevPatch) * num_patches, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = hipMalloc(&d->d_scratch, sizeof(double) * kMaxPartials);
  if (e == hipSuccess) e = hipMalloc(&d->d_result, sizeof(double) * 16);
  if (e == hipSuccess) e = hipMalloc(&d->d_flags, sizeof(int32_t) * 4);
  if (e == hipSuccess && dim == 2) e = upload_ring(ring_table(d, nullptr, nullptr), &d->d_ring);
  if (e != hipSuccess) {
    set_error(std::string("te_domain_create: ") + hipGetErrorString(e));
    te_domain_destroy(d);
    return TE_ERR_HIP;
  }
  *out = d;
  return TE_OK;
}

int te_domain_destroy(te_domain* d)
{
  if (!d) return TE_OK;
  if (d->d_patches) (void)hipFree(d->d_patches);
  if (d->d_scratch) (void)hipFree(d->d_scratch);
  if (d->d_result) (void)hipFree(d->d_result);
  if (d->d_flags) (void)hipFree(d->d_flags);
  if (d->d_ring) (void)hipFree(d->d_ring);
  delete d;
  return TE_OK;
}

int te_domain_num_patches(const te_domain* d) { return d ? d->np : -1; }

int64_t te_domain_vector_length(const te_domain* d, int ncomp)
{
  Layout L = make_layout(d->dim, d->ns, ncomp, d->np);
  return (int64_t)L.np * L.ps;
}

int te_domain_get_nbr_local(const te_domain* d, int32_t* out)
{
  TE_ARG(d && out, "null argument");
  std::memcpy(out, d->nbr_local.data(), sizeof(int32_t) * d->nbr_local.size());
  return TE_OK;
}

// -------------------------------------------------------------- level link --
int te_level_link_create(const te_domain* fine, const te_domain* coarse, te_level_link** out)
{
  TE_ARG(fine && coarse && out, "null argument");
  TE_ARG(fine->dim == coarse->dim, "dimension mismatch");
  for (int a = 0; a < fine->dim; a++) TE_ARG(fine->ns[a] == coarse->ns[a], "patch size mismatch");
  std::unordered_map<int, int> cid;
  for (int i = 0; i < coarse->np; i++) cid[coarse->patches[i].id] = i;
  auto l = new te_level_link();
  l->fine = fine;
  l->coarse = coarse;
  l->parent.resize(fine->np);
  std::vector<int32_t> orth(fine->np), child((size_t)coarse->np * 8, -1), nchild(coarse->np, 0);
  for (int i = 0; i < fine->np; i++) {
    auto it = cid.find(fine->patches[i].parent_id);
    if (it == cid.end()) {
      delete l;
      set_error("fine patch " + std::to_string(fine->patches[i].id) + " has no parent in the coarser domain");
      return TE_ERR_ARG;
    }
    int cp = it->second;
    l->parent[i] = cp;
    orth[i] = fine->patches[i].orth_on_parent;
    if (nchild[cp] >= (1 << fine->dim)) {
      delete l;
      set_error("too many children for coarse patch " + std::to_string(coarse->patches[cp].id));
      return TE_ERR_ARG;
    }
    child[(size_t)cp * 8 + nchild[cp]++] = i;  // fine local order = increasing
  }
  hipError_t e = hipSuccess;
  size_t nf = std::max(fine->np, 1), nc = std::max(coarse->np, 1);
  e = hipMalloc(&l->d_parent, sizeof(int32_t) * nf);
  if (e == hipSuccess) e = hipMalloc(&l->d_orth, sizeof(int32_t) * nf);
  if (e == hipSuccess) e = hipMalloc(&l->d_child, sizeof(int32_t) * nc * 8);
  if (e == hipSuccess) e = hipMalloc(&l->d_nchild, sizeof(int32_t) * nc);
  if (e == hipSuccess && fine->np) e = hipMemcpy(l->d_parent, l->parent.data(), sizeof(int32_t) * fine->np, hipMemcpyHostToDevice);
  if (e == hipSuccess && fine->np) e = hipMemcpy(l->d_orth, orth.data(), sizeof(int32_t) * fine->np, hipMemcpyHostToDevice);
  if (e == hipSuccess && coarse->np) e = hipMemcpy(l->d_child, child.data(), sizeof(int32_t) * coarse->np * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess && coarse->np) e = hipMemcpy(l->d_nchild, nchild.data(), sizeof(int32_t) * coarse->np, hipMemcpyHostToDevice);
  if (e == hipSuccess && fine->dim == 2) e = upload_ring(ring_table(fine, l->parent.data(), orth.data()), &l->d_ring);
  if (e != hipSuccess) {
    set_error(std::string("te_level_link_create: ") + hipGetErrorString(e));
    te_level_link_destroy(l);
    return TE_ERR_HIP;
  }
  *out = l;
  return TE_OK;
}

int te_level_link_destroy(te_level_link* l)
{
  if (!l) return TE_OK;
  (void)hipFree(l->d_parent);
  (void)hipFree(l->d_orth);
  (void)hipFree(l->d_child);
  (void)hipFree(l->d_nchild);
  (void)hipFree(l->d_ring);
  delete l;
  return TE_OK;
}

int te_level_link_get_parents(const te_level_link* l, int32_t* out)
{
  TE_ARG(l && out, "null argument");
  std::memcpy(out, l->parent.data(), sizeof(int32_t) * l->parent.size());
  return TE_OK;
}

// ------------------------------------------------------------ ghost filler --
int te_ghost_filler_create(const te_domain* d, int kind, int fill_type, te_ghost_filler** out)
{
  TE_ARG(d && out, "null argument");
  TE_ARG(fill_type >= TE_FACES && fill_type <= TE_CORNERS, "bad fill type");
  if (d->dim == 2) {
    TE_ARG(kind == TE_BILINEAR || kind == TE_BIQUADRATIC, "2D domains take BiLinear or BiQuadratic filling");
    TE_ARG(fill_type != TE_EDGES, "2D ghost fillers support Faces and Corners (GhostFillingType::Edges is 3D)");
    if (kind == TE_BIQUADRATIC && d->any_coarse_fine)
      TE_ARG(d->ns[0] >= 4 && d->ns[1] >= 4, "BiQuadraticGhostFiller needs n >= 4 at coarse/fine faces");
  } else {
    TE_ARG(kind == TE_TRILINEAR, "3D domains take TriLinear filling");
    for (int a = 0; a < 3; a++)
      TE_ARG(d->ns[a] % 2 == 0, "TriLinearGhostFiller only supports even number patch sizes");
  }
  *out = new te_ghost_filler{ d, kind, fill_type };
  return TE_OK;
}

int te_ghost_filler_destroy(te_ghost_filler* g)
{
  delete g;
  return TE_OK;
}

int te_ghost_fill(const te_ghost_filler* g, int dtype, void* u, int ncomp, te_stream s)
{
  TE_ARG(g && u && ncomp >= 1 && ncomp <= 8, "bad argument");
  TE_ARG(dtype == TE_F64 || dtype == TE_F32, "bad dtype");
  return launch_ghost_fill(g, dtype, u, ncomp, S(s));
}

// --------------------------------------------------------------- operators --
int te_star_apply(const te_ghost_filler* g, int dtype, int neumann, void* u, void* f, int ncomp, te_stream s)
{
  TE_ARG(g && u && f && ncomp >= 1 && ncomp <= 8, "bad argument");
  TE_ARG(u != f, "u and f must be distinct vectors");
  int rc = launch_ghost_fill(g, dtype, u, ncomp, S(s));
  if (rc) return rc;
  return launch_star(g, dtype, dtype, dtype, 0, neumann, nullptr, nullptr, u, f, ncomp, S(s));
}

int te_star_apply_patches(const te_ghost_filler* g, int dtype, int neumann, const void* u, void* f, int ncomp,
                          te_stream s)
{
  TE_ARG(g && u && f && ncomp >= 1 && ncomp <= 8, "bad argument");
  TE_ARG(u != f, "u and f must be distinct vectors");
  return launch_star(g, dtype, dtype, dtype, 0, neumann, nullptr, nullptr, u, f, ncomp, S(s));
}

int te_varstar_apply(const te_ghost_filler* g, int dtype, const void* coeffs, void* u, void* f, int ncomp,
                     te_stream s)
{
  TE_ARG(g && coeffs && u && f && ncomp >= 1 && ncomp <= 8, "bad argument");
  TE_ARG(u != f, "u and f must be distinct vectors");
  int rc = launch_ghost_fill(g, dtype, u, ncomp, S(s));
  if (rc) return rc;
  return launch_star(g, dtype, dtype, dtype, 1, 0, coeffs, nullptr, u, f, ncomp, S(s));
}

int64_t te_domain_boundary_length(const te_domain* d) { return d ? (int64_t)d->np * boundary_stride(d) : -1; }

int te_star_add_dirichlet_bc(const te_domain* d, int dtype, void* f, const double* g, const double* c, te_stream s)
{
  TE_ARG(d && f && g, "bad argument");
  TE_ARG(dtype == TE_F64 || dtype == TE_F32, "bad dtype");
  return launch_bc_rhs(d, dtype, f, g, c, false, S(s));
}

int te_star_add_neumann_bc(const te_domain* d, int dtype, void* f, const double* dg, te_stream s)
{
  TE_ARG(d && f && dg, "bad argument");
  TE_ARG(dtype == TE_F64 || dtype == TE_F32, "bad dtype");
  return launch_bc_rhs(d, dtype, f, dg, nullptr, true, S(s));
}

int te_residual(const te_ghost_filler* g, int dtype, int op_kind, int neumann, const void* coeffs, const void* f,
                void* u, void* r, int ncomp, te_stream s)
{
  TE_ARG(g && f && u && r && ncomp >= 1 && ncomp <= 8, "bad argument");
  TE_ARG(op_kind == 0 || coeffs, "VarPoisson needs coefficients");
  TE_ARG(r != u, "r and u must be distinct vectors");
  int rc = launch_ghost_fill(g, dtype, u, ncomp, S(s));
  if (rc) return rc;
  return launch_star(g, dtype, dtype, dtype, op_kind, neumann, coeffs, f, u, r, ncomp, S(s));
}

// ------------------------------------------------------------------- DFT ----
// Transform matrices of DFTPatchSolver::getTransformArray (Poisson/DFTPatchSolver.h:131-196),
// row-major M[i][j]: y_i = sum_j M[i][j] x_j.
static void dft_matrix(int type, int n, double* m)
{
  const double pi = M_PI;
  switch (type) {
    case 0:
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) m[i * n + j] = std::cos(pi / n * (i * (j + 0.5)));
      break;
    case 1:
      for (int i = 0; i < n; i++) {
        m[i * n] = 0.5;
        for (int j = 1; j < n; j++) m[i * n + j] = std::cos(pi / n * ((i + 0.5) * j));
      }
      break;
    case 2:
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) m[i * n + j] = std::cos(pi / n * ((i + 0.5) * (j + 0.5)));
      break;
    case 3:
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) m[i * n + j] = std::sin(pi / n * ((i + 1) * (j + 0.5)));
      break;
    case 4:
      for (int i = 0; i < n; i++) {
        for (int j = 0; j < n - 1; j++) m[i * n + j] = std::sin(pi / n * ((i + 0.5) * (j + 1)));
        m[i * n + n - 1] = (i % 2 == 0) ? 0.5 : -0.5;
      }
      break;
    default:
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) m[i * n + j] = std::sin(pi / n * ((i + 0.5) * (j + 0.5)));
  }
}

int te_dft_plan_create(const te_ghost_filler* g, unsigned neumann_mask, int op_kind, const double* coeffs,
                       te_dft_plan** out)
{
  TE_ARG(g && out, "null argument");
  TE_ARG(op_kind == 0 || coeffs, "VarPoisson needs coefficients");
  const te_domain* d = g->d;
  int D = d->dim;
  int N = 1;
  for (int a = 0; a < D; a++) {
    TE_ARG(d->ns[a] >= 1 && d->ns[a] <= kDftMatDim, "DFT patch solver supports n <= 64");
    N *= d->ns[a];
  }
  TE_ARG(N <= 4096, "DFT patch solver supports at most 4096 cells per patch");
  auto pl = new te_dft_plan();
  pl->g = g;
  pl->mask = neumann_mask;
  pl->op_kind = op_kind;
  const char* valu = std::getenv("TE_DFT_VALU");  // force the generic kernel (testing)
  pl->use_mfma = (valu && valu[0] == '1') ? 0 : 1;
  std::vector<double> mats((size_t)6 * 3 * kDftMatStride, 0.0), s2((size_t)3 * 3 * kDftMatDim, 0.0);
  for (int t = 0; t < 6; t++)
    for (int a = 0; a < D; a++) dft_matrix(t, d->ns[a], &mats[((size_t)t * 3 + a) * kDftMatStride]);
  // sin^2 tables of getEigenValues (:320-361): 0 both Neumann, 1 one Neumann, 2 Dirichlet
  for (int a = 0; a < D; a++) {
    int n = d->ns[a];
    for (int k = 0; k < n; k++) {
      s2[(0 * 3 + a) * kDftMatDim + k] = std::pow(std::sin(k * M_PI / (2 * n)), 2);
      s2[(1 * 3 + a) * kDftMatDim + k] = std::pow(std::sin((k + 0.5) * M_PI / (2 * n)), 2);
      s2[(2 * 3 + a) * kDftMatDim + k] = std::pow(std::sin((k + 1) * M_PI / (2 * n)), 2);
    }
  }
  bool all_neumann = true;
  for (int s = 0; s < 2 * D; s++)
    if (!((neumann_mask >> s) & 1)) all_neumann = false;
  std::vector<int> pt((size_t)std::max(d->np, 1) * 8, 0);
  int sb = D == 2 ? 4 : 20;
  for (int p = 0; p < d->np; p++) {
    bool any = false;
    for (int sl = 0; sl < face_count(D); sl++)
      if (d->patches[p].nbr_type[sl] >= 0) any = true;
    for (int a = 0; a < D; a++) {
      bool lo = d->patches[p].nbr_type[sb + 2 * a] < 0 && ((neumann_mask >> (2 * a)) & 1);
      bool hi = d->patches[p].nbr_type[sb + 2 * a + 1] < 0 && ((neumann_mask >> (2 * a + 1)) & 1);
      pt[p * 8 + a] = (lo && hi) ? 0 : lo ? 2 : hi ? 5 : 3;  // getTransformsForPatch (:272-291)
      pt[p * 8 + 3 + a] = (lo && hi) ? 0 : (lo || hi) ? 1 : 2;
    }
    pt[p * 8 + 6] = (all_neumann && !any) ? 1 : 0;
  }
  std::vector<float> mats32(mats.begin(), mats.end()), s2_32(s2.begin(), s2.end());
  hipError_t e = hipMalloc(&pl->d_mats64, sizeof(double) * mats.size());
  if (e == hipSuccess) e = hipMalloc(&pl->d_mats32, sizeof(float) * mats.size());
  if (e == hipSuccess) e = hipMalloc(&pl->d_s2_64, sizeof(double) * s2.size());
  if (e == hipSuccess) e = hipMalloc(&pl->d_s2_32, sizeof(float) * s2.size());
  if (e == hipSuccess) e = hipMalloc(&pl->d_ptype, sizeof(int) * pt.size());
  if (e == hipSuccess) e = hipMemcpy(pl->d_mats64, mats.data(), sizeof(double) * mats.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(pl->d_mats32, mats32.data(), sizeof(float) * mats.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(pl->d_s2_64, s2.data(), sizeof(double) * s2.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(pl->d_s2_32, s2_32.data(), sizeof(float) * s2.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(pl->d_ptype, pt.data(), sizeof(int) * pt.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && op_kind == 1) {
    // the plan keeps its own copies: the caller's coefficient vector may be freed first
    long n = te_domain_vector_length(d, 1);
    e = hipMalloc(&pl->d_coeffs64, sizeof(double) * std::max(n, 1L));
    if (e == hipSuccess) e = hipMalloc(&pl->d_coeffs32, sizeof(float) * std::max(n, 1L));
    if (e == hipSuccess && n) e = hipMemcpy(pl->d_coeffs64, coeffs, sizeof(double) * n, hipMemcpyDeviceToDevice);
    if (e == hipSuccess && launch_convert(n, TE_F32, pl->d_coeffs32, TE_F64, coeffs, 0) != TE_OK) e = hipErrorUnknown;
    if (e == hipSuccess) e = hipDeviceSynchronize();
  }
  if (e != hipSuccess) {
    set_error(std::string("te_dft_plan_create: ") + hipGetErrorString(e));
    te_dft_plan_destroy(pl);
    return TE_ERR_HIP;
  }
  *out = pl;
  return TE_OK;
}

int te_dft_plan_destroy(te_dft_plan* p)
{
  if (!p) return TE_OK;
  (void)hipFree(p->d_mats64);
  (void)hipFree(p->d_mats32);
  (void)hipFree(p->d_s2_64);
  (void)hipFree(p->d_s2_32);
  (void)hipFree(p->d_ptype);
  if (p->d_coeffs64) (void)hipFree(p->d_coeffs64);
  if (p->d_coeffs32) (void)hipFree(p->d_coeffs32);
  delete p;
  return TE_OK;
}

int te_dft_smooth(const te_dft_plan* p, int dtype, const void* f, void* u, te_stream s)
{
  TE_ARG(p && f && u, "null argument");
  TE_ARG(dtype == TE_F64 || dtype == TE_F32, "bad dtype");
  int rc = launch_ghost_fill(p->g, dtype, u, 1, S(s));
  if (rc) return rc;
  return launch_dft(p, dtype, dtype, dtype, f, u, u, false, S(s));
}

int te_dft_patch_solve(const te_dft_plan* p, int dtype, const void* f, void* u, te_stream s)
{
  TE_ARG(p && f && u, "null argument");
  TE_ARG(dtype == TE_F64 || dtype == TE_F32, "bad dtype");
  return launch_dft(p, dtype, dtype, dtype, f, u, u, false, S(s));
}

int te_dft_apply(const te_dft_plan* p, int dtype, const void* f, void* u, te_stream s)
{
  TE_ARG(p && f && u, "null argument");
  int rc = launch_vec_op(p->g->d, dtype, VOP_SET_WITH_GHOST, u, nullptr, nullptr, 0.0, 0, 0, 1, S(s));
  if (rc) return rc;
  return launch_dft(p, dtype, dtype, dtype, f, u, u, true, S(s));
}

static bool dtype_ok(int t) { return t == TE_F64 || t == TE_F32; }

int te_dft_apply_mixed(const te_dft_plan* p, int dtype, int dtype_f, int dtype_out, const void* f, void* out,
                       te_stream s)
{
  TE_ARG(p && f && out, "null argument");
  TE_ARG(dtype_ok(dtype) && dtype_ok(dtype_f) && dtype_ok(dtype_out), "bad dtype");
  TE_ARG((p->use_mfma && dft_mfma_supported(p)) || dft_generic_fused_supported(p),
         "fused DFT launches need a 2D plan with even cells per side and BiLinear or BiQuadratic ghosts");
  return launch_dft_fused(p, dtype, dtype_f, dtype_out, f, out, out, true, 0, TE_F64, nullptr, nullptr, S(s));
}

int te_dft_smooth_interp(const te_dft_plan* p, const te_level_link* l, int dtype, int dtype_f, int dtype_out,
                         const void* f, const void* u, int dtype_c, const void* uc, void* out, te_stream s)
{
  TE_ARG(p && l && f && u && uc && out, "null argument");
  TE_ARG(out != u, "out must not alias u");
  TE_ARG(dtype_ok(dtype) && dtype_ok(dtype_f) && dtype_ok(dtype_out) && dtype_ok(dtype_c), "bad dtype");
  TE_ARG(l->fine == p->g->d, "the link's fine level is not the plan's domain");
  TE_ARG((p->use_mfma && dft_mfma_supported(p)) || dft_generic_fused_supported(p),
         "fused DFT launches need a 2D plan with even cells per side and BiLinear or BiQuadratic ghosts");
  TE_ARG(p->g->kind == TE_BILINEAR || p->g->kind == TE_BIQUADRATIC, "fused ghosts: BiLinear or BiQuadratic");
  return launch_dft_fused(p, dtype, dtype_f, dtype_out, f, u, out, false, 2, dtype_c, uc, l, S(s));
}

int te_residual_restrict(const te_ghost_filler* g, const te_level_link* l, int op_kind, int neumann,
                         const void* coeffs, int dtype, int dtype_f, const void* f, const void* u, int dtype_c,
                         void* fc, te_stream s)
{
  TE_ARG(g && l && f && u && fc, "null argument");
  TE_ARG(op_kind == 0 || coeffs, "VarPoisson needs coefficients");
  TE_ARG(dtype_ok(dtype) && dtype_ok(dtype_f) && dtype_ok(dtype_c), "bad dtype");
  TE_ARG(l->fine == g->d, "the link's fine level is not the filler's domain");
  TE_ARG(resid_restrict_fused_supported(g), "fused residual / restriction: 2D, even n, BiLinear or BiQuadratic");
  return launch_resid_restrict(g, l, op_kind, neumann, coeffs, dtype, dtype_f, f, u, dtype_c, fc, S(s));
}

int te_patch_bicgstab_smooth(const te_ghost_filler* g, int dtype, int op_kind, int neumann, const void* coeffs,
                             int max_it, double tol, int continue_on_breakdown, const void* f, void* u,
                             int32_t* iters, te_stream s)
{
  TE_ARG(g && f && u, "null argument");
  TE_ARG(dtype == TE_F64 || dtype == TE_F32, "bad dtype");
  TE_ARG(op_kind == 0 || coeffs, "VarPoisson needs coefficients");
  TE_ARG(max_it >= 0, "bad iteration limit");
  int rc = launch_ghost_fill(g, dtype, u, 1, S(s));
  if (rc) return rc;
  int32_t* flag = g->d->d_flags;
  TE_HIP(hipMemsetAsync(flag, 0, sizeof(int32_t), S(s)));
  rc = launch_patch_bicgstab(g->d, dtype, op_kind, neumann, coeffs, f, u, max_it, tol, iters, flag, S(s));
  if (rc) return rc;
  int32_t h = 0;
  TE_HIP(hipMemcpyAsync(&h, flag, sizeof(int32_t), hipMemcpyDeviceToHost, S(s)));
  TE_HIP(hipStreamSynchronize(S(s)));
  if (h && !continue_on_breakdown) {
    set_error("BiCGStab broke down in a patch solve (rho was 0)");
    return TE_ERR_BREAKDOWN;
  }
  return TE_OK;
}

// ---------------------------------------------------------------- transfer --
int te_restrict(const te_level_link* l, int dtype, const void* fine, void* coarse, int ncomp, te_stream s)
{
  TE_ARG(l && fine && coarse && ncomp >= 1 && ncomp <= 8, "bad argument");
  return launch_restrict(l, dtype, fine, dtype, coarse, ncomp, S(s));
}

int te_restrict_extrapolate(const te_level_link* l, int dtype, const void* fine, void* coarse, int ncomp,
                            te_stream s)
{
  TE_ARG(l && fine && coarse && ncomp >= 1 && ncomp <= 8, "bad argument");
  return launch_restrict(l, dtype, fine, dtype, coarse, ncomp, S(s), true);
}

int te_interpolate(const te_level_link* l, int dtype, const void* coarse, void* fine, int ncomp, te_stream s)
{
  TE_ARG(l && fine && coarse && ncomp >= 1 && ncomp <= 8, "bad argument");
  return launch_interpolate(l, dtype, coarse, dtype, fine, ncomp, S(s));
}

// ------------------------------------------------------------------ vectors --
#define TE_VEC_CHECK() TE_ARG(d && x && ncomp >= 1 && ncomp <= 8 && (dtype == TE_F64 || dtype == TE_F32), "bad argument")

int te_vec_set(const te_domain* d, int dtype, void* x, double alpha, int ncomp, te_stream s)
{
  TE_VEC_CHECK();
  return launch_vec_op(d, dtype, VOP_SET, x, nullptr, nullptr, alpha, 0, 0, ncomp, S(s));
}
int te_vec_set_with_ghost(const te_domain* d, int dtype, void* x, double alpha, int ncomp, te_stream s)
{
  TE_VEC_CHECK();
  return launch_vec_op(d, dtype, VOP_SET_WITH_GHOST, x, nullptr, nullptr, alpha, 0, 0, ncomp, S(s));
}
int te_vec_scale(const te_domain* d, int dtype, void* x, double alpha, int ncomp, te_stream s)
{
  TE_VEC_CHECK();
  return launch_vec_op(d, dtype, VOP_SCALE, x, nullptr, nullptr, alpha, 0, 0, ncomp, S(s));
}
int te_vec_shift(const te_domain* d, int dtype, void* x, double delta, int ncomp, te_stream s)
{
  TE_VEC_CHECK();
  return launch_vec_op(d, dtype, VOP_SHIFT, x, nullptr, nullptr, delta, 0, 0, ncomp, S(s));
}
int te_vec_copy(const te_domain* d, int dtype, void* x, const void* b, int ncomp, te_stream s)
{
  TE_VEC_CHECK();
  return launch_vec_op(d, dtype, VOP_COPY, x, b, nullptr, 0, 0, 0, ncomp, S(s));
}
int te_vec_add(const te_domain* d, int dtype, void* x, const void* b, int ncomp, te_stream s)
{
  TE_VEC_CHECK();
  return launch_vec_op(d, dtype, VOP_ADD, x, b, nullptr, 0, 0, 0, ncomp, S(s));
}
int te_vec_add_scaled(const te_domain* d, int dtype, void* x, double alpha, const void* a, int ncomp, te_stream s)
{
  TE_VEC_CHECK();
  return launch_vec_op(d, dtype, VOP_ADD_SCALED, x, a, nullptr, alpha, 0, 0, ncomp, S(s));
}
int te_vec_add_scaled2(const te_domain* d, int dtype, void* x, double alpha, const void* a, double beta,
                       const void* b, int ncomp, te_stream s)
{
  TE_VEC_CHECK();
  return launch_vec_op(d, dtype, VOP_ADD_SCALED2, x, a, b, alpha, beta, 0, ncomp, S(s));
}
int te_vec_scale_then_add(const te_domain* d, int dtype, void* x, double alpha, const void* b, int ncomp,
                          te_stream s)
{
  TE_VEC_CHECK();
  return launch_vec_op(d, dtype, VOP_SCALE_THEN_ADD, x, b, nullptr, alpha, 0, 0, ncomp, S(s));
}
int te_vec_scale_then_add_scaled(const te_domain* d, int dtype, void* x, double alpha, double beta, const void* b,
                                 int ncomp, te_stream s)
{
  TE_VEC_CHECK();
  return launch_vec_op(d, dtype, VOP_SCALE_THEN_ADD_SCALED, x, b, nullptr, alpha, beta, 0, ncomp, S(s));
}
int te_vec_scale_then_add_scaled2(const te_domain* d, int dtype, void* x, double alpha, double beta, const void* b,
                                  double gamma, const void* c, int ncomp, te_stream s)
{
  TE_VEC_CHECK();
  return launch_vec_op(d, dtype, VOP_SCALE_THEN_ADD_SCALED2, x, b, c, alpha, beta, gamma, ncomp, S(s));
}

static int reduce_sync(const te_domain* d, int dtype, int op, const void* a, const void* b, int ncomp, double* out,
                       te_stream s)
{
  TE_ARG(d && a && out && ncomp >= 1 && ncomp <= 8, "bad argument");
  if (d->np == 0) {
    *out = 0.0;
    return TE_OK;
  }
  int rc = launch_reduce(d, dtype, op, a, b, ncomp, d->d_result, S(s));
  if (rc) return rc;
  TE_HIP(hipMemcpyAsync(out, d->d_result, sizeof(double), hipMemcpyDeviceToHost, S(s)));
  TE_HIP(hipStreamSynchronize(S(s)));
  return TE_OK;
}

int te_vec_dot(const te_domain* d, int dtype, const void* a, const void* b, int ncomp, double* out, te_stream s)
{
  TE_ARG(b, "null argument");
  return reduce_sync(d, dtype, RED_DOT, a, b, ncomp, out, s);
}
int te_vec_two_norm(const te_domain* d, int dtype, const void* a, int ncomp, double* out, te_stream s)
{
  return reduce_sync(d, dtype, RED_SUMSQ, a, nullptr, ncomp, out, s);
}
int te_vec_inf_norm(const te_domain* d, int dtype, const void* a, int ncomp, double* out, te_stream s)
{
  return reduce_sync(d, dtype, RED_MAX, a, nullptr, ncomp, out, s);
}
int te_vec_integrate(const te_domain* d, int dtype, const void* u, int ncomp, double* out, te_stream s)
{
  return reduce_sync(d, dtype, RED_INTEGRAL, u, nullptr, ncomp, out, s);
}
int te_vec_convert(const te_domain* d, int dst_dtype, void* dst, int src_dtype, const void* src, int ncomp,
                   te_stream s)
{
  TE_ARG(d && dst && src && ncomp >= 1 && ncomp <= 8, "bad argument");
  return launch_convert(te_domain_vector_length(d, ncomp), dst_dtype, dst, src_dtype, src, S(s));
}

}  // extern "C"
