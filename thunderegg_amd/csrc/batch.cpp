// Many independent problems solved together: one launch per cycle step and per BiCGStab
// step for all problems of a patch size, one graph per outer iteration.
//
// The reference solves each problem with its own GMG::CycleBuilder and
// Iterative::BiCGStab (GMG/CycleBuilder.h:97-192, GMG/VCycle.h:46-70,
// Iterative/BiCGStab.h:111-203); a batch gives every problem exactly those steps.
//
//  - Problems of one patch size form a group.  The group's levels are unions of its
//    problems' levels aligned at the coarsest: union j holds level nlev - 1 - j of every
//    problem with more than j levels (patch ids offset per problem so they stay unique), so
//    each patch of union j + 1 has its parent in union j and every cycle step is one launch
//    over a union.  A problem's finest level sits in the union of its depth; the cycle
//    scatters that level's right-hand side in after the restriction into the union has
//    run and gathers its result at the end (te_gmg batch mode, gmg_internal.h).
//  - The outer BiCGStab runs on the concatenation of the problems' finest levels with one
//    scalar set per problem.  Each problem's inner products are folded from partial sums
//    written by the workgroups, over the patches, that its own solve would use (BatchMap),
//    so they have the same bits; the loop control of BiCGStab.h:148-150 runs on the device
//    per problem (ST_DONE, ST_BREAK), and the host only waits until no problem runs.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "gmg_internal.h"
#include "kernels/blas.h"
#include "kernels/dft.h"
#include "kernels/krylov.h"
#include "te_internal.h"

using namespace te;

namespace {

#define TE_RC(x)             \
  do {                       \
    int rc_ = (x);           \
    if (rc_) return rc_;     \
  } while (0)

struct Group {
  int ns[2] = { 0, 0 };
  std::vector<int> probs;     // the caller's problem indexes, deepest first (stable)
  std::vector<int> nlev;      // their level counts
  std::vector<te_domain*> U;  // unions, finest first (the te_gmg's levels)
  te_domain* F = nullptr;     // the problems' finest levels, concatenated in group order
  te_ghost_filler* Fgf = nullptr;
  te_gmg* g = nullptr;
  hipStream_t s = nullptr;    // g's stream
  long flen = 0;              // elements of a batch vector
  double *b = nullptr, *x = nullptr, *resid = nullptr, *rhat = nullptr, *p = nullptr, *ap = nullptr, *as = nullptr,
         *sv = nullptr, *mp = nullptr, *ms = nullptr;
  float *p32 = nullptr, *s32 = nullptr, *mp32 = nullptr, *ms32 = nullptr;
  double *scal = nullptr, *pa = nullptr, *pb = nullptr, *pc = nullptr, *ctl = nullptr;
  bool m32 = false;
  BatchMap bmap;        // vector kernels: a workgroup per slot
  BatchStarMaps smaps;  // operator launches
  double** d_bptr = nullptr;  // the problems' own vectors (group order) for the copies
  double** d_xptr = nullptr;
  long* d_off = nullptr;
  long* d_len = nullptr;
  long maxlen = 0;
  std::vector<void*> dev;  // device allocations owned by the group
  GraphCache graphs;       // one BiCGStab iteration per use_precond
  double* mirror_h = nullptr;  // host-coherent [seq, running, broken]
  double* mirror_d = nullptr;
  double seq = 0.0;
  double running = 0.0;
};

template <typename T>
hipError_t dev_alloc(Group& G, T** p, size_t count)
{
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (count > 0 ? count : 1));
  if (e == hipSuccess) {
    G.dev.push_back(*p);
    e = hipMemset(*p, 0, sizeof(T) * (count > 0 ? count : 1));
  }
  return e;
}

template <typename T>
hipError_t dev_upload(Group& G, T** p, const std::vector<T>& v)
{
  hipError_t e = dev_alloc(G, p, v.size());
  if (e == hipSuccess && !v.empty()) e = hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
  return e;
}

void free_group(Group& G)
{
  G.graphs.clear();
  if (G.g) gmg_free(G.g);
  if (G.Fgf) te_ghost_filler_destroy(G.Fgf);
  if (G.F) te_domain_destroy(G.F);
  for (te_domain* d : G.U) te_domain_destroy(d);
  for (void* p : G.dev) (void)hipFree(p);
  if (G.mirror_h) (void)hipHostFree(G.mirror_h);
  G = Group{};
}

// the largest patch id a problem's levels use (ids, parents, children, neighbours)
int max_id(const te_domain* const* lv, int n)
{
  int m = 0;
  for (int l = 0; l < n; l++)
    for (const te_patch_info& P : lv[l]->patches) {
      m = std::max(m, std::max(P.id, P.parent_id));
      for (int k = 0; k < 8; k++) m = std::max(m, P.child_ids[k]);
      for (int s = 0; s < 26; s++)
        for (int k = 0; k < 4; k++) m = std::max(m, P.nbr_ids[s][k]);
    }
  return m;
}

void append_offset(std::vector<te_patch_info>& out, const te_domain* d, int off)
{
  for (te_patch_info P : d->patches) {
    P.id += off;
    if (P.parent_id >= 0) P.parent_id += off;
    for (int k = 0; k < 8; k++)
      if (P.child_ids[k] >= 0) P.child_ids[k] += off;
    for (int s = 0; s < 26; s++)
      for (int k = 0; k < 4; k++)
        if (P.nbr_ids[s][k] >= 0) P.nbr_ids[s][k] += off;
    out.push_back(P);
  }
}

// the problems' maps: problem k (group order, np[k] finest patches) gets nslot[k] =
// krylov_blocks of its finest level partial slots and nwg(k) workgroups
BatchMap make_map(Group& G, const std::vector<int>& np, const std::vector<int>& nslot, int cap, hipError_t* e)
{
  std::vector<BatchProb> prob(np.size());
  std::vector<int> wq;
  int poff = 0, woff = 0, soff = 0;
  for (size_t k = 0; k < np.size(); k++) {
    BatchProb& P = prob[k];
    P.poff = poff;
    P.np = np[k];
    P.woff = woff;
    P.nwg = cap > 0 ? std::min(cap, nslot[k]) : nslot[k];
    P.soff = soff;
    P.nslot = nslot[k];
    for (int w = 0; w < P.nwg; w++) wq.push_back((int)k);
    poff += np[k];
    woff += P.nwg;
    soff += nslot[k];
  }
  BatchMap m{};
  m.nwg = woff;
  m.nprob = (int)np.size();
  BatchProb* dprob = nullptr;
  int* dwq = nullptr;
  *e = dev_upload(G, &dprob, prob);
  if (*e == hipSuccess) *e = dev_upload(G, &dwq, wq);
  m.prob = dprob;
  m.wq = dwq;
  return m;
}

int build_group(Group& G, const std::vector<const te_domain* const*>& lv, const te_gmg_options* o)
{
  const int ng = (int)G.probs.size();
  int J = 0;
  for (int k = 0; k < ng; k++) J = std::max(J, G.nlev[k] - 1);
  std::vector<int> idoff(ng);
  {
    long off = 0;
    for (int k = 0; k < ng; k++) {
      idoff[k] = (int)off;
      off += (long)max_id(lv[k], G.nlev[k]) + 1;
      if (off >= (1L << 31)) {
        set_error("te_batch_create: too many patch ids in one batch");
        return TE_ERR_ARG;
      }
    }
  }
  const int nsz[3] = { G.ns[0], G.ns[1], 1 };
  // unions U_j (coarse index j), stored finest first: G.U[J - j]
  G.U.assign(J + 1, nullptr);
  std::vector<te_gmg::BatchSeg> seg(J + 1);
  std::vector<int> fpatches(ng);
  for (int k = 0; k < ng; k++) fpatches[k] = lv[k][0]->np;
  for (int j = 0; j <= J; j++) {
    std::vector<te_patch_info> pi;
    te_gmg::BatchSeg sg;
    int foff = 0;
    bool in_seg = false;
    for (int k = 0; k < ng; k++) {
      if (G.nlev[k] == j + 1 && !in_seg) {
        sg.from = (int)pi.size();
        sg.foff = foff;
        in_seg = true;
      }
      if (G.nlev[k] > j) append_offset(pi, lv[k][G.nlev[k] - 1 - j], idoff[k]);
      if (G.nlev[k] == j + 1) sg.count += fpatches[k];
      if (!in_seg) foff += fpatches[k];
    }
    seg[J - j] = sg;
    TE_RC(te_domain_create(2, nsz, 1, (int)pi.size(), pi.data(), &G.U[J - j]));
  }
  {
    std::vector<te_patch_info> pi;
    for (int k = 0; k < ng; k++) append_offset(pi, lv[k][0], idoff[k]);
    TE_RC(te_domain_create(2, nsz, 1, (int)pi.size(), pi.data(), &G.F));
  }
  TE_RC(te_ghost_filler_create(G.F, o->filler_kind, o->fill_type, &G.Fgf));
  if (!star_dot_fly_supported(G.Fgf)) {
    set_error("te_batch_create: the batched operator needs even cells per side and a patch tile within 64 KiB");
    return TE_ERR_UNSUPPORTED;
  }
  std::vector<const te_domain*> levels(G.U.begin(), G.U.end());
  TE_RC(gmg_create(levels.data(), J + 1, o, nullptr, true, &G.g));
  G.g->bseg = seg;
  G.s = G.g->s;
  G.m32 = G.g->fused && G.g->L[0].dt == TE_F32;
  // vectors over F
  const long pl = (long)(G.ns[0] + 2) * (G.ns[1] + 2);
  G.flen = (long)G.F->np * pl;
  hipError_t e = hipSuccess;
  for (double** v : { &G.b, &G.x, &G.resid, &G.rhat, &G.p, &G.ap, &G.as, &G.sv, &G.mp, &G.ms })
    if (e == hipSuccess) e = dev_alloc(G, v, (size_t)G.flen);
  if (G.m32)
    for (float** v : { &G.p32, &G.s32, &G.mp32, &G.ms32 })
      if (e == hipSuccess) e = dev_alloc(G, v, (size_t)G.flen);
  // slots and maps
  std::vector<int> nslot(ng);
  int ns_total = 0;
  for (int k = 0; k < ng; k++) {
    nslot[k] = krylov_blocks(lv[k][0]);
    ns_total += nslot[k];
  }
  if (e == hipSuccess) G.bmap = make_map(G, fpatches, nslot, 0, &e);
  for (int udt : { TE_F64, TE_F32 })
    for (bool sq : { false, true }) {
      const int cap = star_dot_2d_cap(G.Fgf, o->op_kind, udt, sq);
      bool have = false;
      for (int k = 0; k < G.smaps.n; k++) have = have || G.smaps.cap[k] == cap;
      if (have || e != hipSuccess) continue;
      if (G.smaps.n == BatchStarMaps::kMax) {
        set_error("internal: too many operator instantiations for the batch maps");
        return TE_ERR_UNSUPPORTED;
      }
      G.smaps.cap[G.smaps.n] = cap;
      G.smaps.map[G.smaps.n] = make_map(G, fpatches, nslot, cap, &e);
      G.smaps.n++;
    }
  if (e == hipSuccess) e = dev_alloc(G, &G.scal, (size_t)ng * SC_COUNT);
  if (e == hipSuccess) e = dev_alloc(G, &G.pa, (size_t)ns_total);
  if (e == hipSuccess) e = dev_alloc(G, &G.pb, (size_t)ns_total);
  if (e == hipSuccess) e = dev_alloc(G, &G.pc, (size_t)ns_total);
  if (e == hipSuccess) e = dev_alloc(G, &G.ctl, 4);
  // copy tables
  std::vector<long> off(ng), len(ng);
  long o_ = 0;
  for (int k = 0; k < ng; k++) {
    off[k] = o_;
    len[k] = (long)fpatches[k] * pl;
    o_ += len[k];
    G.maxlen = std::max(G.maxlen, len[k]);
  }
  if (e == hipSuccess) e = dev_upload(G, &G.d_off, off);
  if (e == hipSuccess) e = dev_upload(G, &G.d_len, len);
  if (e == hipSuccess) e = dev_alloc(G, &G.d_bptr, (size_t)ng);
  if (e == hipSuccess) e = dev_alloc(G, &G.d_xptr, (size_t)ng);
  if (e == hipSuccess) e = hipHostMalloc(&G.mirror_h, sizeof(double) * 4, hipHostMallocCoherent | hipHostMallocMapped);
  if (e == hipSuccess) {
    std::memset(G.mirror_h, 0, sizeof(double) * 4);
    e = hipHostGetDevicePointer((void**)&G.mirror_d, G.mirror_h, 0);
  }
  if (e != hipSuccess) {
    set_error(std::string("te_batch_create: ") + hipGetErrorString(e));
    return TE_ERR_HIP;
  }
  return TE_OK;
}

// wait until the group's publish with sequence number G.seq has reached the host
int wait_published(Group& G)
{
  volatile double* m = G.mirror_h;
  for (unsigned spin = 1;; spin++) {
    if (m[0] == G.seq) break;
    if ((spin & 255) == 0) {
      hipError_t q = hipStreamQuery(G.s);
      if (q == hipSuccess && m[0] == G.seq) break;
      if (q == hipSuccess) {
        set_error("batched BiCGStab: an iteration finished without publishing its state");
        return TE_ERR_HIP;
      }
      if (q != hipErrorNotReady) {
        set_error(std::string("batched BiCGStab iteration: ") + hipGetErrorString(q));
        return TE_ERR_HIP;
      }
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  G.running = m[1];
  return TE_OK;
}

}  // namespace

struct te_batch {
  int nprob = 0;
  te_gmg_options o{};
  std::vector<Group> grp;
  std::vector<std::pair<int, int>> where;  // problem -> (group, index in the group)
  hipEvent_t ev = nullptr;
};

namespace {

void free_batch(te_batch* B)
{
  if (!B) return;
  for (Group& G : B->grp) free_group(G);
  if (B->ev) (void)hipEventDestroy(B->ev);
  delete B;
}

// every group's stream waits for the work enqueued so far on `caller`
int join_from(te_batch* B, hipStream_t caller)
{
  TE_HIP(hipEventRecord(B->ev, caller));
  for (Group& G : B->grp) TE_HIP(hipStreamWaitEvent(G.s, B->ev, 0));
  return TE_OK;
}

// `caller` waits for every group's stream
int join_to(te_batch* B, hipStream_t caller)
{
  for (Group& G : B->grp) {
    if (caller == G.s) continue;
    TE_HIP(hipEventRecord(B->ev, G.s));
    TE_HIP(hipStreamWaitEvent(caller, B->ev, 0));
  }
  return TE_OK;
}

int upload_ptrs(Group& G, double** dst, const double* const* src)
{
  std::vector<double*> h(G.probs.size());
  for (size_t k = 0; k < G.probs.size(); k++) h[k] = const_cast<double*>(src[G.probs[k]]);
  TE_HIP(hipMemcpy(dst, h.data(), sizeof(double*) * h.size(), hipMemcpyHostToDevice));
  return TE_OK;
}

// one BiCGStab iteration of every problem of the group (the sequence of te_gmg_bicgstab)
int iteration(Group& G, int use_precond)
{
  const te_gmg_options& o = G.g->o;
  hipStream_t s = G.s;
  const bool m32 = use_precond && G.m32;
  const int dtm = m32 ? TE_F32 : TE_F64;
  void* mp = use_precond ? (m32 ? (void*)G.mp32 : (void*)G.mp) : (void*)G.p;
  void* ms = use_precond ? (m32 ? (void*)G.ms32 : (void*)G.ms) : (void*)G.sv;
  if (use_precond) TE_RC(gmg_vcycle_enqueue(G.g, m32 ? (const void*)G.p32 : G.p, mp, dtm, dtm));
  TE_RC(launch_star_dot_2d(G.Fgf, o.op_kind, o.neumann, nullptr, mp, dtm, G.ap, G.rhat, G.pa, nullptr, 0, s,
                           &G.smaps));
  TE_RC(launch_bicg(BICG_S, G.F, G.scal, G.sv, nullptr, G.resid, G.ap, nullptr, nullptr, nullptr, G.pc, nullptr, s,
                    TE_F64, m32 ? G.s32 : nullptr, nullptr, 0, G.pa, 0, &G.bmap));
  if (use_precond) TE_RC(gmg_vcycle_enqueue(G.g, m32 ? (const void*)G.s32 : G.sv, ms, dtm, dtm));
  TE_RC(launch_star_dot_2d(G.Fgf, o.op_kind, o.neumann, nullptr, ms, dtm, G.as, G.sv, G.pa, G.pb, 0, s, &G.smaps));
  TE_RC(launch_bicg_finalize_batch(FIN_SOMEGA, G.pa, G.pb, G.pc, G.bmap, G.scal, s));
  const bool xp_m = use_precond == 2;
  TE_RC(launch_bicg(BICG_XR, G.F, G.scal, G.x, G.resid, mp, ms, G.ap, G.as, G.rhat, G.pa, G.pb, s, dtm, nullptr,
                    xp_m ? mp : (const void*)G.p, xp_m ? 1 : 0, nullptr, 0, &G.bmap));
  TE_RC(launch_bicg_finalize_batch(FIN_RHO, G.pa, G.pb, nullptr, G.bmap, G.scal, s));
  TE_RC(launch_bicg(BICG_P, G.F, G.scal, G.p, nullptr, G.ap, G.resid, nullptr, nullptr, nullptr, nullptr, nullptr, s,
                    TE_F64, m32 ? G.p32 : nullptr, nullptr, 0, nullptr, 0, &G.bmap));
  return launch_batch_publish(G.scal, (int)G.probs.size(), G.ctl, G.mirror_d, s);
}

int iteration_graph(Group& G, int use_precond, hipGraphExec_t* out)
{
  const GraphCache::Key key{ (uintptr_t)use_precond, 0, 2 };
  *out = G.graphs.find(key);
  if (*out) return TE_OK;
  hipGraph_t graph;
  TE_HIP(hipStreamBeginCapture(G.s, hipStreamCaptureModeThreadLocal));
  int rc = iteration(G, use_precond);
  hipError_t e = hipStreamEndCapture(G.s, &graph);
  if (rc) return rc;
  if (e != hipSuccess) {
    set_error(std::string("batched BiCGStab graph capture: ") + hipGetErrorString(e));
    return TE_ERR_HIP;
  }
  hipGraphExec_t ex;
  e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    set_error(std::string("batched BiCGStab graph instantiate: ") + hipGetErrorString(e));
    return TE_ERR_HIP;
  }
  G.graphs.put(key, ex);
  *out = ex;
  return TE_OK;
}

}  // namespace

extern "C" {

int te_batch_create(int nprob, const te_domain* const* levels, const int* nlev, const te_gmg_options* opts,
                    te_batch** out)
{
  if (nprob < 1 || !levels || !nlev || !opts || !out) {
    set_error("te_batch_create: bad argument");
    return TE_ERR_ARG;
  }
  if (opts->cycle_type != TE_CYCLE_V || opts->op_kind != 0) {
    set_error("te_batch_create: batches run the V-cycle with the Poisson operator");
    return TE_ERR_UNSUPPORTED;
  }
  if (opts->smoother != 0 || opts->level_smoother) {
    set_error("te_batch_create: batches smooth with the DFT patch solver on every level");
    return TE_ERR_UNSUPPORTED;
  }
  auto B = new te_batch();
  B->nprob = nprob;
  B->o = *opts;
  B->o.level_smoother = nullptr;
  B->o.level_inner_max_it = nullptr;
  B->o.level_inner_tol = nullptr;
  B->where.assign(nprob, { -1, -1 });
  // problem q's levels start at levels[first[q]]
  std::vector<long> first(nprob);
  long acc = 0;
  for (int q = 0; q < nprob; q++) {
    if (nlev[q] < 1) {
      free_batch(B);
      set_error("te_batch_create: every problem needs at least one level");
      return TE_ERR_ARG;
    }
    first[q] = acc;
    acc += nlev[q];
    for (int l = 0; l < nlev[q]; l++) {
      const te_domain* d = levels[first[q] + l];
      if (!d || d->dim != 2 || d->ns[0] != levels[first[q]]->ns[0] || d->ns[1] != levels[first[q]]->ns[1]) {
        free_batch(B);
        set_error("te_batch_create: problem " + std::to_string(q) +
                  ": the levels must be 2D domains of one patch size");
        return TE_ERR_ARG;
      }
    }
  }
  // groups by patch size, in order of first appearance; problems deepest first (stable)
  std::vector<std::pair<int, int>> sizes;
  for (int q = 0; q < nprob; q++) {
    const te_domain* d = levels[first[q]];
    std::pair<int, int> key{ d->ns[0], d->ns[1] };
    if (std::find(sizes.begin(), sizes.end(), key) == sizes.end()) sizes.push_back(key);
  }
  B->grp.resize(sizes.size());
  int rc = TE_OK;
  for (size_t gi = 0; gi < sizes.size() && !rc; gi++) {
    Group& G = B->grp[gi];
    G.ns[0] = sizes[gi].first;
    G.ns[1] = sizes[gi].second;
    for (int q = 0; q < nprob; q++) {
      const te_domain* d = levels[first[q]];
      if (d->ns[0] == G.ns[0] && d->ns[1] == G.ns[1]) G.probs.push_back(q);
    }
    std::stable_sort(G.probs.begin(), G.probs.end(), [&](int a, int b) { return nlev[a] > nlev[b]; });
    std::vector<const te_domain* const*> lv;
    for (size_t k = 0; k < G.probs.size(); k++) {
      const int q = G.probs[k];
      G.nlev.push_back(nlev[q]);
      lv.push_back(levels + first[q]);
      B->where[q] = { (int)gi, (int)k };
    }
    rc = build_group(G, lv, &B->o);
  }
  if (!rc) {
    hipError_t e = hipEventCreateWithFlags(&B->ev, hipEventDisableTiming);
    if (e != hipSuccess) {
      set_error(std::string("te_batch_create: ") + hipGetErrorString(e));
      rc = TE_ERR_HIP;
    }
  }
  if (rc) {
    free_batch(B);
    return rc;
  }
  *out = B;
  return TE_OK;
}

int te_batch_destroy(te_batch* b)
{
  free_batch(b);
  return TE_OK;
}

int te_batch_vcycle(te_batch* B, const double* const* f, double* const* u, te_stream stream)
{
  if (!B || !f || !u) {
    set_error("te_batch_vcycle: bad argument");
    return TE_ERR_ARG;
  }
  for (Group& G : B->grp) {
    TE_RC(upload_ptrs(G, G.d_bptr, f));
    TE_RC(upload_ptrs(G, G.d_xptr, u));
  }
  TE_RC(join_from(B, (hipStream_t)stream));
  for (Group& G : B->grp) {
    const int ng = (int)G.probs.size();
    TE_RC(launch_batch_copy(true, G.b, G.d_bptr, G.d_off, G.d_len, ng, G.maxlen, G.s));
    TE_RC(gmg_vcycle_enqueue(G.g, G.b, G.x, TE_F64, TE_F64));
    TE_RC(launch_batch_copy(false, G.x, G.d_xptr, G.d_off, G.d_len, ng, G.maxlen, G.s));
  }
  return join_to(B, (hipStream_t)stream);
}

int te_batch_bicgstab(te_batch* B, const double* const* b, double* const* x, int use_precond, int max_it, double tol,
                      int* iterations, double* residuals, te_stream stream)
{
  if (!B || !b || !x) {
    set_error("te_batch_bicgstab: bad argument");
    return TE_ERR_ARG;
  }
  for (Group& G : B->grp) {
    TE_RC(upload_ptrs(G, G.d_bptr, b));
    TE_RC(upload_ptrs(G, G.d_xptr, x));
  }
  TE_RC(join_from(B, (hipStream_t)stream));
  // resid = b - A x; r0 = ||b||; rhat = p = resid; rho = rhat.resid; the first states
  for (Group& G : B->grp) {
    const int ng = (int)G.probs.size();
    const bool m32 = use_precond && G.m32;
    TE_RC(launch_batch_copy(true, G.b, G.d_bptr, G.d_off, G.d_len, ng, G.maxlen, G.s));
    TE_RC(launch_batch_copy(true, G.x, G.d_xptr, G.d_off, G.d_len, ng, G.maxlen, G.s));
    TE_RC(launch_ghost_fill(G.Fgf, TE_F64, G.x, 1, G.s));
    TE_RC(launch_star(G.Fgf, TE_F64, TE_F64, TE_F64, B->o.op_kind, B->o.neumann, nullptr, G.b, G.x, G.resid, 1,
                      G.s));
    TE_RC(launch_bicg(BICG_INIT, G.F, G.scal, G.rhat, G.p, G.resid, G.b, nullptr, nullptr, nullptr, G.pa, G.pb, G.s,
                      TE_F64, m32 ? G.p32 : nullptr, nullptr, 0, nullptr, 0, &G.bmap));
    TE_RC(launch_bicg_finalize_batch(FIN_INIT, G.pa, G.pb, nullptr, G.bmap, G.scal, G.s, tol, max_it));
    TE_RC(launch_batch_publish(G.scal, ng, G.ctl, G.mirror_d, G.s));
    G.seq += 1.0;
  }
  for (Group& G : B->grp) TE_RC(wait_published(G));
  // every problem's loop ends within max_it iterations (ST_DONE at the limit)
  for (int it = 0; it <= max_it; it++) {
    bool any = false;
    for (Group& G : B->grp) {
      if (G.running == 0.0) continue;
      hipGraphExec_t gx = nullptr;
      if (B->o.use_graph) TE_RC(iteration_graph(G, use_precond, &gx));
      if (gx) TE_HIP(hipGraphLaunch(gx, G.s));
      else TE_RC(iteration(G, use_precond));
      G.seq += 1.0;
      any = true;
    }
    if (!any) break;
    for (Group& G : B->grp)
      if (G.running != 0.0) TE_RC(wait_published(G));
  }
  std::vector<std::vector<double>> sc(B->grp.size());
  for (size_t gi = 0; gi < B->grp.size(); gi++) {
    Group& G = B->grp[gi];
    const int ng = (int)G.probs.size();
    TE_RC(launch_batch_copy(false, G.x, G.d_xptr, G.d_off, G.d_len, ng, G.maxlen, G.s));
    sc[gi].resize((size_t)ng * SC_COUNT);
    TE_HIP(hipMemcpyAsync(sc[gi].data(), G.scal, sizeof(double) * sc[gi].size(), hipMemcpyDeviceToHost, G.s));
  }
  TE_RC(join_to(B, (hipStream_t)stream));
  for (Group& G : B->grp) TE_HIP(hipStreamSynchronize(G.s));
  std::string broken;
  for (int q = 0; q < B->nprob; q++) {
    const auto w = B->where[q];
    const double* S = sc[w.first].data() + (size_t)w.second * SC_COUNT;
    if (iterations) iterations[q] = (int)S[SC_ITS];
    if (residuals) residuals[q] = S[SC_R0] != 0.0 ? S[SC_RNORM] / S[SC_R0] : 0.0;
    if (S[SC_STATE] == ST_BREAK && broken.empty())
      broken = "BiCGStab broke down, rho was 0 on iteration " + std::to_string((int)S[SC_ITS]) + " of problem " +
               std::to_string(q);
  }
  if (!broken.empty()) {
    set_error(broken);
    return TE_ERR_BREAKDOWN;
  }
  return TE_OK;
}

int te_batch_num_groups(const te_batch* B) { return B ? (int)B->grp.size() : -1; }

}  // extern "C"
