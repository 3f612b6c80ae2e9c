// Launch variants chosen per patch shape by the autotuner's table.  Every variant of a
// family computes each patch with the same arithmetic in the same order (only the mapping of
// patches onto workgroups and waves changes), so results do not depend on the table.
// thunderegg_amd/autotune.py times the variants on the GPU and writes the table; the library
// loads it on first use (tune_table.txt beside the shared library, or the file named by
// TE_TUNE_TABLE; TE_TUNE_TABLE=none keeps the hand defaults).
#pragma once

namespace te {

enum TuneFamily {
  TUNE_DFT_GENERIC = 0,  // DFT patch solve, VALU kernel (any n): patches per workgroup x threads per patch
  TUNE_DFT_MFMA_2D = 1,  // DFT patch solve, 2D matrix-core kernel (n 16, 32): workgroups of the launch
  TUNE_C3_SWEEP = 2,     // 3D 16^3 fused sweeps (modes 1, 2): slices per load batch
  TUNE_FAMILIES
};

// the variant for (family, dim, n, dtype): the one forced by a timing call on this thread,
// else the table's, else 0 (the hand default)
int tune_variant(int family, int dim, int n, int dtype);
int tune_num_variants(int family);
// force `variant` for (family, dim, n, dtype) on this thread until tune_unforce (timing calls)
void tune_force(int family, int dim, int n, int dtype, int variant);
void tune_unforce();

struct DftGenericVariant {
  int ppw;  // patches per workgroup
  int tpp;  // threads per patch
};
DftGenericVariant dft_generic_variant(int v);

// 2D matrix-core DFT launches: the workgroup count policy of variant v (dft_mfma.hip)
enum DftMfmaGrid {
  MFMA_GRID_DEFAULT = 0,    // n 16: one round of the plain kernel when every workgroup is resident at
                            // once, else the prefetching kernel on the resident count; n 32: <= 4096
  MFMA_GRID_RESIDENT = 1,   // the prefetching kernel (n 16) on the resident workgroup count / <= 2048 (n 32)
  MFMA_GRID_HALF = 2,       // half the resident count (n 16) / <= 8192 (n 32)
  MFMA_GRID_ALL = 3,        // one workgroup per 4 patches, plain kernel
  MFMA_GRID_VARIANTS
};

// 3D fused sweeps, variant v: slices per load batch (v 0: 2 in FP32, 1 in FP64; v 1, 2, 3:
// 1, 2, 4 slices; v 4, 5: 1, 2 slices; v 6, 7: 2, 1 slices; v 8, 9: 1, 2 slices), the minimum
// waves per SIMD the compiler is asked to fit the registers for (v 4, 5: 5, which the FP32
// sweep from zero and one-slice batch reach without spilling; v 8, 9: 6, which the FP32
// one-slice interface sweep with f ahead reaches without spilling; else, and in FP64, where
// the patch's LDS allows 3 workgroups per CU, no request), and whether the f loads of all four
// slices of a wave are issued at its start (v 6 - 9, FP32 interface sweeps; else per batch)
constexpr int kC3Variants = 10;
inline int c3_sweep_batch(int v, int dtype)
{
  if (v == 1 || v == 4 || v == 7 || v == 8) return 1;
  if (v == 2 || v == 5 || v == 6 || v == 9) return 2;
  if (v == 3) return 4;
  return dtype == 1 ? 2 : 1;  // TE_F32 = 1
}
inline int c3_sweep_waves(int v) { return v == 4 || v == 5 ? 5 : (v == 8 || v == 9 ? 6 : 1); }
inline bool c3_sweep_f_ahead(int v) { return v >= 6 && v <= 9; }

}  // namespace te
