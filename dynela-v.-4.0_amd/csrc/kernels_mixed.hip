// kernels_mixed.hip -- mixed-precision element kernel (placeholder until the
// FP32 strain/stress path lands; jobs cannot select it yet).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace dnlg {

int elem_block_threads_mixed(int type) { return type == DNLG_EL_HEX8N_3D ? 128 : 256; }

void launch_elem_mixed(int, JobDev *const *, const Task *, int, hipStream_t) {}

}  // namespace dnlg
