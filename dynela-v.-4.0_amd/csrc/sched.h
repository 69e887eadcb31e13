// sched.h -- launch schedules of the element kernels and the corrector,
// and the candidates the autotuner (autotune.hip) chooses among.
//
// A schedule changes how the work is laid out on the GPU, never what is
// computed: every element is computed by the same lanes' arithmetic in any
// block size or occupancy, and the corrector sums a node's incidence slots in
// slot order whatever its batch; so results are the same bits under every
// schedule (tests/test_gpu_autotune.py).  That needs the library built with
// -ffp-contract=on (build.py): fusing multiply-adds in the back end follows
// the IR the vectorizers leave, which the waves-per-SIMD request changes,
// and made two occupancy variants of the mixed kernel round one Newton
// iterate 1 ulp apart.
#pragma once

#include "../../include/dnlg.h"

namespace dnlg {

// Element-kernel variants (precision paths) of a job.
// kVarMixedX: the mixed kernel for the tabular and network laws and the
// Direct integration (an FP64 return throughout).
enum ElemVariant { kVarFP64 = 0, kVarMixed = 1, kVarFP64Ann = 2, kVarMixedRR32 = 3, kVarMixedX = 4 };
constexpr int kElemVariants = 5;

// Threads per element-kernel block: 8 / 16 / 32 hexes, 32 / 64 / 128 quads
// (one lane per node and Gauss point).
constexpr int kSchedBlocks = 3;
constexpr int kHexBlocks[kSchedBlocks] = {64, 128, 256};
constexpr int kQuadBlocks[kSchedBlocks] = {128, 256, 512};
// Waves per SIMD requested from the compiler (amdgpu_waves_per_eu): each
// variant has its default (0: none requested, the registers the kernel
// needs set the occupancy -- 4 for both mixed variants) and one alternative
// that trades register spills for occupancy.
constexpr int kFP64Waves = 3, kFP64WavesAlt = 2;
constexpr int kMixedWavesAlt = 5, kRR32WavesAlt = 5;
// Incidence slots the corrector loads together (FP32 forces: 2, 4, 8; FP64
// forces: 2, 4).
// Cycles per captured graph (rounded up to a multiple of the time-step
// frequency, and to an even count, when the jobs share their phase).

inline int default_elem_block(int type) { return type == DNLG_EL_HEX8N_3D ? 128 : 256; }
inline int default_waves(int variant) { return (variant == kVarFP64 || variant == kVarFP64Ann) ? kFP64Waves : 0; }
inline int alt_waves(int variant) {
  return variant == kVarMixed ? kMixedWavesAlt : (variant == kVarMixedRR32 ? kRR32WavesAlt : kFP64WavesAlt);
}
inline int default_corr_batch(bool f32) { return f32 ? 8 : 4; }
constexpr int kDefaultGraphCycles = 16;

// Whether (block, waves) is a schedule the element kernels of this type and
// variant are built for.  The plane-strain quad, the tetrahedron, the
// network-law variant and the mixed table-law variant take their defaults
// only.
inline bool elem_sched_valid(int type, int variant, int block, int waves) {
  if (type == DNLG_EL_QUA4N_2D || type == DNLG_EL_TET4N_3D || variant == kVarFP64Ann || variant == kVarMixedX)
    return block == default_elem_block(type) && waves == default_waves(variant);
  const int *bl = type == DNLG_EL_HEX8N_3D ? kHexBlocks : kQuadBlocks;
  bool ok = false;
  for (int k = 0; k < kSchedBlocks; k++) ok = ok || bl[k] == block;
  return ok && (waves == default_waves(variant) || waves == alt_waves(variant));
}
inline bool corr_batch_valid(bool f32, int b) { return b == 2 || b == 4 || (f32 && b == 8); }

}  // namespace dnlg
