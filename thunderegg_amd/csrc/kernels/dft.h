// DFT patch-solver plan shared by the host code and the kernels.
#pragma once

#include "../te_internal.h"

namespace te {
constexpr int kDftMatDim = 64;                      // largest supported n
constexpr int kDftMatStride = kDftMatDim * kDftMatDim;
}  // namespace te

struct te_dft_plan {
  const te_ghost_filler* g = nullptr;
  unsigned mask = 0;
  int op_kind = 0;
  int* d_ptype = nullptr;       // per patch: fwd type per axis [0..2], eig type [3..5], zero mode [6]
  double* d_mats64 = nullptr;   // [6 types][3 axes][kDftMatStride]
  float* d_mats32 = nullptr;
  double* d_s2_64 = nullptr;    // [3 eig types][3 axes][kDftMatDim]
  float* d_s2_32 = nullptr;
  double* d_coeffs64 = nullptr;  // VarPoisson beta: plan-owned copy of the caller's values
  float* d_coeffs32 = nullptr;   // plan-owned float copy
  int use_mfma = 1;                    // matrix-core kernel where supported (2D, n 16/32, Dirichlet)
};

namespace te {
int launch_dft(const te_dft_plan* pl, int dt, int dtf, int dto, const void* f, const void* u, void* out, bool uzero,
               hipStream_t s);
bool dft_mfma_supported(const te_dft_plan* pl);
int launch_dft_mfma(const te_dft_plan* pl, int dt, int dtf, int dto, const void* f, const void* u, void* out,
                    bool uzero, hipStream_t s);
// gm: 0 ghosts read from u, 1 ghosts computed on the fly, 2 on the fly from u + P uc
// (uc: the coarser level's solution of dtype dtc, link: fine -> coarse).
int launch_dft_mfma_fused(const te_dft_plan* pl, int dt, int dtf, int dto, const void* f, const void* u, void* out,
                          bool uzero, int gm, int dtc, const void* uc, const te_level_link* link, hipStream_t s);
// The same fused sweeps for any 2D plan with even cells per side: the matrix-core kernel
// where it applies, else the generic kernel with the ghosts (and the interpolation) computed
// in the launch.
bool dft_generic_fused_supported(const te_dft_plan* pl);
int launch_dft_fused(const te_dft_plan* pl, int dt, int dtf, int dto, const void* f, const void* u, void* out,
                     bool uzero, int gm, int dtc, const void* uc, const te_level_link* link, hipStream_t s);
bool resid_restrict_fused_supported(const te_ghost_filler* g);
int launch_resid_restrict(const te_ghost_filler* g, const te_level_link* link, int op_kind, int neumann,
                          const void* coeffs, int dt, int dtf, const void* f, const void* u, int dtc, void* fc,
                          hipStream_t s);
// Cycle::restrict of the level of g fused with the zero-guess sweep of the next coarser
// level (plan pc): writes the coarse f (fc) and the coarse u (uc); 2D, 16x16 patches.
bool rr_pre_supported(const te_ghost_filler* g, const te_dft_plan* pc);
int launch_rr_pre(const te_ghost_filler* g, const te_level_link* link, int op_kind, int neumann,
                  const void* coeffs, int dt, int dtf, const void* f, const void* u, const te_dft_plan* pc, int dtc,
                  void* fc, void* uc, hipStream_t s);
// BiCGStab's s update (mode BICG_S, with the s.s partials in pc: nslots slots, those past
// the launch's workgroups zeroed) or p update (BICG_P) fused with the zero-guess FP32 sweep of
// plan pl's level (2D, 16 x 16) on the updated vector: v0 (FP64) and its FP32 copy w32 are
// written as by launch_bicg, and the sweep's result goes to out.  pfold (BICG_S only, may be
// null): the nfold rhat.ap partials; the launch then also does the FIN_ALPHA finalize (alpha
// = rho / their fold, stored in scal), so that finalize is not launched.
bool bicg_sweep_supported(const te_dft_plan* pl);
int launch_bicg_sweep(int mode, const te_dft_plan* pl, double* scal, double* v0, const double* a0,
                      const double* a1, float* w32, float* out, double* pc, int nslots, hipStream_t s,
                      const double* pfold = nullptr, int nfold = 0);
}  // namespace te
