// The coarsest levels of the fused 2D V-cycle (16 x 16 patches, Poisson, DFT smoother) in
// one single-workgroup launch: from the first level with at most kTailMaxPatches patches
// down to the coarsest and back, every level's right-hand side and solution kept in LDS.
// Each step is the arithmetic of the per-level launches it replaces (the same sweeps,
// ghost formulas, residual, restriction and interpolation, in the same order and
// precision), without their global-memory round trips.
#pragma once

#include "../te_internal.h"

namespace te {

constexpr int kTailWaves = 8;       // waves of the workgroup
constexpr int kTailLastWaves = 4;   // waves that sweep the last level when its type differs
constexpr int kTailMaxLevels = 6;
constexpr int kTailMaxPatches = 16; // patches of the tail's first level (two per wave)
constexpr int kTailLds = 158 * 1024;

struct TailLevelArgs {
  const DevPatch* dp;      // patch table
  const int4* ring;        // ring table (te::ring_table, pq fields 0)
  const int32_t* parent;   // per patch: parent in the next tail level (unused on the last)
  const int32_t* porth;    // per patch: orth_on_parent, -1 for a copy
  int np;
  int u0, u1, f;           // LDS byte offsets of the two solution buffers and of f
  int meta, ringl;         // LDS byte offsets of the per-patch records and of the ring table
};

struct TailArgs {
  TailLevelArgs lv[kTailMaxLevels];
  int nl;                  // levels in the tail (2..kTailMaxLevels)
  int pre, post, coarse;   // sweeps (CycleOpts)
  int neumann;             // Poisson: Neumann physical boundaries in the residual
  const void* f_in;        // right-hand side of the first tail level (global, its type)
  void* u_out;             // solution of the first tail level (global, its type)
  const void* mats_t;      // DFT tables of the levels' type (te_dft_plan layout)
  const void* s2_t;
  const void* mats_l;      // and of the last level's type
  const void* s2_l;
  int mat_t, mat_l;        // LDS byte offsets of the transform matrices
  int scr_t, scr_l;        // LDS byte offsets of the per-wave scratch
};

// LDS bytes of a tail over levels with the given patch counts; dt: type of all but the
// last level, dtl: the last level's (TE_F32 / TE_F64).  Fills the offsets of `a`.
int tail_layout(TailArgs* a, const int* np, int nl, int dt, int dtl);

// dt / dtl as in tail_layout; kind: TE_BILINEAR or TE_BIQUADRATIC
int launch_dft_tail(const TailArgs& a, int dt, int dtl, int kind, hipStream_t s);

}  // namespace te
