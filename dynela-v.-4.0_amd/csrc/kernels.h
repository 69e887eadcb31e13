// kernels.h -- host-side launchers of the device kernels.
#pragma once

#include "engine.h"

namespace dnlg {

// Shape tables of every unit that holds a copy (kernels.hip, elem_fp64.hip).
void upload_tables(const double *hexN, const double *hexD, const double *quaN, const double *quaD);
void upload_tables_fp64(const double *hexN, const double *hexD, const double *quaN, const double *quaD);

// cs: the control slot that holds the cycle's record (derive != 0: the
// previous cycle's, from which the element kernel derives the new record
// and stores it in slot 1 - cs).  With ev != nullptr the launch records the
// kernel's start/stop in ev[0..1] (hipExtLaunchKernel), so its duration is
// measured on the launch stream.  block, waves: the schedule (sched.h).
// ann: the variant for jobs with a network law (DNLG_LAW_ANN2SIG).
void launch_elem_fp64(int type, int block, int waves, JobDev *const *jobs, const Task *tasks, int nblk, int cs,
                      int derive, hipStream_t s, const hipEvent_t *ev = nullptr, bool ann = false);
// rr64: the radial return in FP64 (DNLG_PREC_MIXED), else in FP32
// (DNLG_PREC_MIXED_RR32)
void launch_elem_mixed(int type, int block, int waves, JobDev *const *jobs, const Task *tasks, int nblk, int cs,
                       int derive, hipStream_t s, const hipEvent_t *ev, bool rr64, bool xl);
// fold != 0: BC nodes also get the next cycle's field0 BC values (the next
// cycle's element launch then derives its record, derive != 0)
// batch: the incidence slots loaded together (2, 4 or 8 for FP32 forces; 2
// or 4 for FP64)
void launch_correct(int dim, bool f32, int batch, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int fold,
                    hipStream_t s);
void launch_courant(int dim, JobDev *const *jobs, const Task *tasks, int nblk, int cs, int force, hipStream_t s);
// close the cycle in slot cs (end of step, Courant time step) and, with
// open != 0, open the next one; the record goes to slot `out`.  Open:
// out = 1 - cs, tasks max(1, ceil(n_bcn / 256)) blocks per job over the
// job's BC node list; close only: out = 0, one block per job
void launch_cycle_step(JobDev *const *jobs, const Task *tasks, int nblk, int cs, int open, int out, hipStream_t s);
void launch_force_dt(JobDev *const *jobs, int n, hipStream_t s);
void launch_set_up_to(JobDev *const *jobs, int n, const double *up_to, hipStream_t s);
void launch_prep(int type, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s);
void launch_mass_gather(int dim, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s);
// per job: {t, dt, up_to, status, inc, power-iteration stall, result-file
// hold} snapshot for the host loop
constexpr int kSnapDoubles = 7;
void launch_snapshot(JobDev *const *jobs, int n, double *out, hipStream_t s);
// power-iteration time step: begin (force: the pre-loop's forced step,
// otherwise the jobs whose cycle in slot cs is due), one iteration = elem
// (per element type, over the Courant task lists) + node (per dim and
// precision) + step (one block per job), finish (host-driven loops: the
// converged step into slot 0), pending flags (1 = still iterating)
void launch_pi_begin(JobDev *const *jobs, int n, int cs, int force, hipStream_t s);
void launch_pi_elem(int type, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s);
void launch_pi_node(int dim, JobDev *const *jobs, const Task *tasks, int nblk, hipStream_t s);
void launch_pi_step(JobDev *const *jobs, const Task *tasks, int nblk, int cs, hipStream_t s);
void launch_pi_finish(JobDev *const *jobs, int n, int preloop, hipStream_t s);
void launch_pi_pending(JobDev *const *jobs, int n, int *out, hipStream_t s);
// history files: the IP-to-node extrapolation tables ([node][ip]) and their
// upload; the launches of one row-taking moment (preloop != 0: the pre-loop
// pass, slot 0; otherwise the end of the cycle in slot cs): evaluation (item
// ranges, blk >= 0, and kinetic-energy node ranges, blk = -1 - partial), then
// commit (one block per job with history or result files)
void extrapolation_tables(double hexE[64], double quaE[16]);
void upload_extrapolation();
void launch_history(JobDev *const *jobs, const Task *eval, int n_eval, const Task *per_job, int n_jobs, int cs,
                    int preloop, hipStream_t s);
// Node::fieldScalar of an IP field at every node (src, comp as HistItem;
// comp < 0 with a tensor source: all six components, out [6][N]); needs the
// job's nf_ptr / nf_pairs lists
void launch_node_field(JobDev *job, int64_t N, int src, int comp, double *out, hipStream_t s);

// Global energies of one job (energy.hip): kinetic energy of the field0
// speeds, stored internal energy and plastic work, by a fixed-order
// Neumaier-compensated reduction.  part: 6 doubles per block,
// energy_blocks(N, G) blocks; out: {KE, IE, W_p, KE + IE} (device).  The job
// must be between launch sequences (its record in control slot 0).
constexpr int64_t kEnergyItemsPerBlock = 4096;
int energy_blocks(int64_t N, int64_t G);
// The file-scope device scratch of the allocation-free variant: partials of
// at most kEnergyScratchBlocks blocks (jobs of up to 16.7 M nodes and
// integration points) and the results, 192 KiB per GPU.  energy_scratch_init
// takes its address on the current device (once per device, at context
// creation); energy_scratch returns it (nullptr before).
constexpr int kEnergyScratchBlocks = 4096;
constexpr int kMaxDevices = 64;
int energy_scratch_init(int device);
double *energy_scratch(int device);
void launch_energies(JobDev *job, int64_t N, int64_t G, int dim, double vol_scale, double *part, double *out,
                     hipStream_t s);

}  // namespace dnlg
