// Internal view of the device GMG solver (gmg.cpp) for the code that drives it: the
// per-level state, the graph cache and the solver object.  Not part of the public C ABI.
#pragma once

#include <array>
#include <cstdint>
#include <utility>
#include <vector>

#include "kernels/tail2d.h"
#include "te_internal.h"

namespace te {

struct GLevel {
  const te_domain* d = nullptr;
  te_ghost_filler* gf = nullptr;
  te_dft_plan* plan = nullptr;
  te_level_link* link = nullptr;  // this level -> next coarser level
  int dt = TE_F64;
  void* f = nullptr;
  void* u = nullptr;
  void* r = nullptr;
  void* u2 = nullptr;  // second solution buffer (fused path ping-pongs between u and u2)
  void* cur = nullptr; // buffer holding this level's solution after its visit
  const void* coeffs = nullptr;      // in this level's dtype
  const double* coeffs64 = nullptr;  // FP64 coefficients (ghost-filled)
  float* coeffs32 = nullptr;         // owned when dt == F32 and VarPoisson
  double* coeffs64_own = nullptr;    // ghost-filled copy of the caller's coefficients
  int smoother = 0;                  // 0 DFTPatchSolver, 1 Iterative::PatchSolver(BiCGStab)
  int inner_max_it = 1000;           // inner BiCGStab settings (smoother 1)
  double inner_tol = 1e-12;
};

inline size_t esize(int dt) { return dt == TE_F64 ? 8 : 4; }

// Instantiated graphs keyed by the buffers they were captured with, most recently used
// first; at most kCap are kept (callers that pass fresh buffers every call would otherwise
// accumulate one per call).
struct GraphCache {
  static constexpr size_t kCap = 24;  // a solve's iteration graphs: one per x-update slot
  using Key = std::array<uintptr_t, 3>;
  std::vector<std::pair<Key, hipGraphExec_t>> e;
  hipGraphExec_t find(const Key& k)
  {
    for (size_t i = 0; i < e.size(); i++)
      if (e[i].first == k) {
        auto hit = e[i];
        e.erase(e.begin() + (long)i);
        e.insert(e.begin(), hit);
        return hit.second;
      }
    return nullptr;
  }
  void put(const Key& k, hipGraphExec_t x)
  {
    if (e.size() >= kCap) {
      (void)hipGraphExecDestroy(e.back().second);
      e.pop_back();
    }
    e.insert(e.begin(), { k, x });
  }
  void clear()
  {
    for (auto& kv : e) (void)hipGraphExecDestroy(kv.second);
    e.clear();
  }
};

}  // namespace te

struct te_gmg {
  te_gmg_options o{};
  std::vector<te::GLevel> L;
  hipStream_t s = nullptr;
  double *resid = nullptr, *rhat = nullptr, *p = nullptr, *ap = nullptr, *as = nullptr, *sv = nullptr,
         *ms = nullptr, *mp = nullptr;
  double *scal = nullptr, *pa = nullptr, *pb = nullptr, *pc = nullptr;  // pc: the s.s partials
  double* host_scal = nullptr;  // pinned
  // host-coherent copy of the scalars written by the device at the end of each BiCGStab
  // iteration (host and device addresses)
  double* mirror_h = nullptr;
  double* mirror_d = nullptr;
  int nblk = 0;
  int32_t* d_flags = nullptr;
  // the V-cycle per (f, u); one BiCGStab iteration per (x, use_precond) (b is not read
  // inside the iteration)
  te::GraphCache vgraphs, bgraphs;
  double last_seconds = 0.0;
  bool fused = false;  // 2D matrix-core path: ghost fills fused into the stencil / smoother launches
  std::vector<char> rr_pre;  // per level: restriction fused with the coarser level's first sweep
  bool fly = false;    // outer operator computes ghosts on the fly
  // mixed precision on the fused path: M p and M s are FP32 values (the finest level's last
  // sweep computes in FP32), so the BiCGStab keeps them in FP32 buffers
  // and hands it FP32 copies of p and s as right-hand sides (written by the p and s updates)
  bool m32 = false;
  float* mp32 = nullptr;
  float* ms32 = nullptr;
  float* p32 = nullptr;
  float* s32 = nullptr;
  // te_gmg_set_timer: "Iteration" timings (time_its) and per-level "Total Patch Smooth"
  // domain timings (timer_ids[level], TE_TIMER_NONE = untimed), events on s
  te_timer* timer = nullptr;
  int time_its = 0;
  std::vector<int> timer_ids;
  // fused path: the coarsest levels from tail_from down in one launch (launch_dft_tail),
  // -1 when none qualify
  int tail_from = -1;
  te::TailArgs tail{};
  // fused mixed-precision path: the BiCGStab s and p updates run the finest level's
  // zero-guess sweep too (launch_bicg_sweep); TE_NO_BSWEEP=1 keeps them apart
  bool bsweep = false;
  // stream ordering with the caller's stream (no host blocking)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // 3D hierarchies of 16^3 patches whose face neighbours are all on the same level: the
  // cycle and the BiCGStab work on compact vectors (interior cells only, cycle3d.h)
  bool c3 = false;
  // the compact path keeps face columns behind its solution buffers (cycle3d.h);
  // TE_NO_XCOL=1 reads the x faces from the cells instead (same values)
  bool xcol = true;
  // the compact path's s, x / r and p updates form A M p and A M s themselves
  // (launch_c3_upd), so the operator launches store nothing, and the setup forms b - A x
  // from the padded b and x in one launch (launch_c3_init); TE_NO_XRFLY=1 stores the
  // products and reads them back, and sets up through the padded residual
  bool xrfly = true;
  // mixed precision on the compact path (FP32 cycle, xrfly): BiCGStab's shadow residual rhat
  // is kept as an FP32 copy of r0 (c_rhat32); TE_NO_RHAT32=1 keeps it in FP64
  bool rhat32 = true;
  // mixed precision on the compact path (with rhat32): the x / r update leaves x alone and
  // keeps alpha and omega; M p and M s of kC3XSlots iterations rotate through slots (slot 0:
  // c_mp / c_ms, the others c_mph / c_msh), and launch_c3_xcomb adds the slots' terms to x
  // every xslots iterations and at the end of a solve.  TE_NO_XDEFER=1: x updated in the
  // x / r update; TE_XDEFER_SLOTS=k (1..8): k slots
  // mixed precision on the compact path (with rhat32): the direction p is kept only as its
  // FP32 copy c_p32, which the cycle reads anyway; TE_NO_P32ONLY=1 keeps an FP64 p as well
  bool p32only = true;
  bool xdefer = true;
  int xslots = 8;
  std::vector<void*> c_mph, c_msh;
  double* c_hist = nullptr;  // alpha, omega per slot
  struct C3Level {
    void* f = nullptr;
    void* u = nullptr;
    void* u2 = nullptr;
    // levels of at least kC3InterpMin patches below a coarser one: the boundary layer of
    // u + P uc for the first post-smoothing sweep (launch_c3_face_interp), or nullptr
    void* u3 = nullptr;
    void* cur = nullptr;  // the buffer holding this level's solution after its visit
  };
  // TE_NO_FACEINTERP=1: the first post-smoothing sweep gathers u + P uc itself (mode 2)
  bool finterp = true;
  std::vector<C3Level> C3;
  double *c_r = nullptr, *c_rhat = nullptr, *c_p = nullptr, *c_ap = nullptr, *c_as = nullptr, *c_s = nullptr;
  double* c_x = nullptr;  // compact copy of the caller's x during a solve
  void* c_mp = nullptr;  // M p, M s: the finest level's type (FP32 in mixed precision)
  void* c_ms = nullptr;
  float* c_p32 = nullptr;  // FP32 copies of p and s (mixed precision: the cycle's input)
  float* c_s32 = nullptr;
  float* c_rhat32 = nullptr;  // rhat as FP32 (rhat32)
  void* c_tmp = nullptr;   // the cycle's output for padded callers
  // Batch mode (te_batch): the levels are unions of many independent problems' levels,
  // aligned at their coarsest level, so a problem's finest level can be any of them.  On
  // level l the patches [bseg[l].from, bseg[l].from + bseg[l].count) are finest-level
  // patches: their right-hand side is the cycle's input and their result its output, at
  // patch bseg[l].foff of the batch vectors (level 0 is the start of those vectors).  No
  // cycle step couples two problems, so each problem gets its own V-cycle.
  struct BatchSeg {
    int from = 0, count = 0, foff = 0;
  };
  bool batch = false;
  std::vector<BatchSeg> bseg;
  const void* bf = nullptr;  // the cycle's input while it is enqueued (batch layout), its type
  int bf_dt = TE_F64;
};


namespace te {

// te_gmg_create; batch: the levels are a batch's unions (te_batch), which keeps every
// launch that assumes one problem per level (the LDS-resident tail, the fused restriction
// + coarse sweep, the sweep fused into the BiCGStab updates, the compact 3D path) off, and
// leaves the outer BiCGStab vectors unallocated (the batch owns its own)
int gmg_create(const te_domain* const* levels, int num_levels, const te_gmg_options* opts,
               const double* const* coeffs, bool batch, te_gmg** out);
void gmg_free(te_gmg* g);
// Cycle::apply enqueued on g->s; f / u of types dtf / udt (FP32 only on the fused
// mixed-precision path).  In batch mode f and u are batch vectors (bseg).
int gmg_vcycle_enqueue(te_gmg* g, const void* f, void* u, int udt, int dtf);

}  // namespace te
