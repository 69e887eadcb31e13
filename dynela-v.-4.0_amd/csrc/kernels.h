// kernels.h -- host-side launchers of the device kernels.
#pragma once

#include "engine.h"

namespace dnlg {

void upload_tables(const double *hexN, const double *hexD, const double *quaN, const double *quaD);
int elem_block_threads(int type);
int elem_block_threads_mixed(int type);

// With ev != nullptr the launch records the kernel's start/stop in ev[0..1]
// (hipExtLaunchKernel), so its duration is measured on the launch stream.
void launch_elem_fp64(int type, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s,
                      const hipEvent_t *ev = nullptr);
void launch_elem_mixed(int type, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s,
                       const hipEvent_t *ev = nullptr);
void launch_correct(int dim, bool f32, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s);
void launch_courant(int dim, JobDev *const *jobs, const Task *tasks, int nblk, int force, hipStream_t s);
// close the last cycle (end of step, Courant time step) and, with open != 0,
// open the next one; tasks: max(1, ceil(n_bcn / 256)) blocks per job over
// the job's BC node list
void launch_cycle_step(JobDev *const *jobs, const Task *tasks, int nblk, int open, hipStream_t s);
void launch_force_dt(JobDev *const *jobs, int n, hipStream_t s);
void launch_set_up_to(JobDev *const *jobs, int n, const double *up_to, hipStream_t s);
void launch_prep(int type, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s);
void launch_mass_gather(int dim, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s);
// per job: {t, dt, up_to, status, inc} snapshot for the host loop
void launch_snapshot(JobDev *const *jobs, int n, double *out, hipStream_t s);

}  // namespace dnlg
