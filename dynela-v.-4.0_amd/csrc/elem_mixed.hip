// elem_mixed.hip -- launchers of the mixed-precision element kernel
// (elem_mixed.h), one instantiation per schedule (sched.h).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "control.h"
#include "kernels.h"
#include "material.h"
#include "polar_fp64.h"
#include "sched.h"
#include "elem_mixed.h"

namespace dnlg {

// One schedule of the mixed kernel of one element type and return precision.
template <int NN, int DIM, bool AXI, bool RR64, int BLOCK>
void launch_mixed_block(int waves, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int derive,
                        hipStream_t s, const hipEvent_t *ev) {
  constexpr int ALT = RR64 ? kMixedWavesAlt : kRR32WavesAlt;
  if (waves == ALT)
    launch_elem(k_elem_mixed<NN, DIM, BLOCK, AXI, RR64, ALT>, BLOCK, nblk, s, ev, jobs, tasks, cs, derive);
  else
    launch_elem(k_elem_mixed<NN, DIM, BLOCK, AXI, RR64, 0>, BLOCK, nblk, s, ev, jobs, tasks, cs, derive);
}

template <int NN, int DIM, bool AXI, bool RR64>
void launch_mixed_type(int block, int waves, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int derive,
                       hipStream_t s, const hipEvent_t *ev) {
  constexpr int B0 = NN == 8 ? 64 : 128;
  if (block == B0)
    launch_mixed_block<NN, DIM, AXI, RR64, B0>(waves, jobs, tasks, nblk, cs, derive, s, ev);
  else if (block == 4 * B0)
    launch_mixed_block<NN, DIM, AXI, RR64, 4 * B0>(waves, jobs, tasks, nblk, cs, derive, s, ev);
  else
    launch_mixed_block<NN, DIM, AXI, RR64, 2 * B0>(waves, jobs, tasks, nblk, cs, derive, s, ev);
}

template <bool RR64>
void launch_mixed_rr(int type, int block, int waves, JobDev *const *jobs, const Task *tasks, int nblk, int cs,
                     int derive, hipStream_t s, const hipEvent_t *ev) {
  if (type == DNLG_EL_HEX8N_3D)
    launch_mixed_type<8, 3, false, RR64>(block, waves, jobs, tasks, nblk, cs, derive, s, ev);
  else if (type == DNLG_EL_QUA4N_AX)
    launch_mixed_type<4, 2, true, RR64>(block, waves, jobs, tasks, nblk, cs, derive, s, ev);
  else  // plane strain: default schedule only
    launch_elem(k_elem_mixed<4, 2, 256, false, RR64, 0>, 256, nblk, s, ev, jobs, tasks, cs, derive);
}

void launch_elem_mixed(int type, int block, int waves, JobDev *const *jobs, const Task *tasks, int nblk, int cs,
                       int derive, hipStream_t s, const hipEvent_t *ev, bool rr64, bool xl) {
  if (nblk <= 0) return;
  if (xl) {
    // the table laws and the Direct integration: default schedules only
    if (type == DNLG_EL_HEX8N_3D)
      launch_elem(k_elem_mixed<8, 3, 128, false, true, 0, true>, 128, nblk, s, ev, jobs, tasks, cs, derive);
    else if (type == DNLG_EL_QUA4N_AX)
      launch_elem(k_elem_mixed<4, 2, 256, true, true, 0, true>, 256, nblk, s, ev, jobs, tasks, cs, derive);
    else
      launch_elem(k_elem_mixed<4, 2, 256, false, true, 0, true>, 256, nblk, s, ev, jobs, tasks, cs, derive);
    return;
  }
  if (rr64)
    launch_mixed_rr<true>(type, block, waves, jobs, tasks, nblk, cs, derive, s, ev);
  else
    launch_mixed_rr<false>(type, block, waves, jobs, tasks, nblk, cs, derive, s, ev);
}

}  // namespace dnlg
