// control.h -- device-side readers of a job's control record and status,
// shared by the cycle kernels and the history kernels.
#pragma once

#include <hip/hip_ext.h>

#include "engine.h"

namespace dnlg {

// The job's status between launches of one cycle's predecessor and its own
// (no launch that could raise an error runs beside the reader).
__device__ __forceinline__ int job_status(CJobDev *C) {
  const int a = C->sync.st[0];
  return a != 0 ? a : C->sync.st[1];
}

// Read through the constant space (scalar loads); only for a record that no
// block of the reading launch writes.
template <typename R = const DNLG_CONST Control>
__device__ __forceinline__ Control load_control(CJobDev *C, int slot) {
  R &s = ((R *)C->c)[slot];
  Control c;
  c.t = s.t;
  c.dt = s.dt;
  c.up_to = s.up_to;
  c.inc = s.inc;
  c.f0 = s.f0;
  c.running = s.running;
  c.courant_due = s.courant_due;
  c.density_due = s.density_due;
  c.preloop = s.preloop;
  c.pending = s.pending;
  return c;
}

__device__ __forceinline__ int sym_idx(int i, int j) { return (i == j) ? (i == 0 ? 0 : (i == 1 ? 3 : 5)) : (i + j == 1 ? 1 : (i + j == 2 ? 2 : 4)); }

// Sticky job status of the cycle with increment parity `par`.
__device__ __forceinline__ void set_status(JobDev *J, int st, int par) { atomicCAS(&J->sync.st[par], 0, st); }

// ---------------------------------------------------------------------------
// Cycle control.  Between two cycles a job's control record goes through
//   close  Explicit::endStep (Explicit.C:539-552): t += dt, field1 becomes
//          field0; then Solver::computeTimeStep (Solver.C:196-236, Courant
//          branch: omega_max = 2 / dt_crit) when the time step was due;
//   open   Explicit::updateTimes (Explicit.C:414-427): last-step clip,
//          increment, and the engine's per-cycle flags.
// k_cycle_step does it where a launch sequence starts, after time-step
// cycles, and (close only) where a sequence ends; between the other cycles of
// a sequence the next cycle's element kernel derives the new record itself
// (cycle_control) and the corrector has already written the new cycle's
// field0 BC values.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void store_control(JobDev *J, int slot, const Control &c) {
  DNLG_GLOBAL Control &d = *gbl(&J->c[slot]);
  d.t = c.t;
  d.dt = c.dt;
  d.up_to = c.up_to;
  d.inc = c.inc;
  d.f0 = c.f0;
  d.running = c.running;
  d.courant_due = c.courant_due;
  d.density_due = c.density_due;
  d.preloop = c.preloop;
  d.pending = c.pending;
}

// Close the cycle that ran last (if any) and, with `open`, open the next.
// status: the job's status after that cycle; freq > 0: the maximum
// frequency of the time-step update (Courant: 2 / min L/c; power iteration:
// sqrt(fmax)), dt = safety Omega_s / freq.  Returns whether the field0 BC
// values of the new cycle differ from what the corrector left in field0
// (first cycle of an advance, a ramped speed BC, or a speed BC under a
// changed time step).
__device__ __forceinline__ bool close_open(CJobDev *C, Control &c, int status, double freq, bool open) {
  const double dt_last = c.dt;
  const bool resumed = !c.pending;
  if (c.pending && status == 0) {
    c.t += c.dt;
    c.f0 = 1 - c.f0;
    if (freq > 0.0) c.dt = C->safety * C->omegaS / freq;
  }
  c.pending = 0;
  bool bc_work = false;
  if (open) {
    c.preloop = 0;
    // a job whose power iteration stalled waits for the host-driven
    // iteration (k_pi_finish clears the flag), a held one for its result
    // file
    c.running = (status == 0 && c.t < c.up_to && !C->pi.stall && !C->res.hold) ? 1 : 0;
    if (c.running) {
      if (c.t + c.dt > c.up_to) c.dt = c.up_to - c.t;
      c.inc++;
      c.courant_due = (c.inc % C->dt_freq == 0) ? 1 : 0;
      // densities are consumed by the Courant step, read back at the end
      // and, when a history item reads them, after every cycle
      c.density_due = (C->density_always || c.courant_due || !(c.t + c.dt < c.up_to)) ? 1 : 0;
      c.pending = 1;
      bc_work = resumed || (C->bc_flags & 2) || ((C->bc_flags & 1) && c.dt != dt_last);
    }
  } else {
    c.running = 0;
  }
  return bc_work;
}

// Whether the Courant kernel computed partial minima in the cycle that ran
// last (Explicit.C:243-251: every dt_freq increments, while t_{n+1} < t_up_to).
__device__ __forceinline__ bool courant_ran(const Control &c, int status) {
  return c.pending && status == 0 && c.courant_due && c.t + c.dt < c.up_to;
}

// The record of the cycle that a launch works on.  derive == 0: slot cs
// holds it.  derive != 0 (the first launch of a cycle that no k_cycle_step
// opened, after a cycle without a time-step update): slot cs holds the
// previous cycle; every block derives the new record from it and the job's
// block 0 stores it in slot 1 - cs for the cycle's later launches.  The
// previous cycle's failures are in its own parity word, which no launch of
// the new cycle writes; a job that did not run in it stays stopped (it
// finished or failed), so reading both words is then race-free too.
// The record is uniform over the block; it is returned in scalar registers
// (the FP64 arithmetic of the derivation runs on the vector unit).
__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uniform_l(int64_t v) {
  const unsigned lo = (unsigned)uniform_i((int)(v & 0xffffffffLL));
  const unsigned hi = (unsigned)uniform_i((int)(v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double uniform_d(double v) { return __longlong_as_double(uniform_l(__double_as_longlong(v))); }

__device__ __forceinline__ Control cycle_control(JobDev *J, CJobDev *C, int cs, int derive, int blk) {
  Control c = load_control(C, cs);
  if (derive) {
    const int st0 = C->sync.st[0], st1 = C->sync.st[1];
    const int status = c.pending ? ((c.inc & 1) ? st1 : st0) : (st0 != 0 ? st0 : st1);
    close_open(C, c, status, 0.0, true);
    if (blk == 0 && threadIdx.x == 0) store_control(J, 1 - cs, c);
    c.t = uniform_d(c.t);
    c.dt = uniform_d(c.dt);
    c.inc = uniform_l(c.inc);
    c.f0 = uniform_i(c.f0);
    c.running = uniform_i(c.running);
    c.courant_due = uniform_i(c.courant_due);
    c.density_due = uniform_i(c.density_due);
    c.preloop = uniform_i(c.preloop);
    c.pending = uniform_i(c.pending);
  }
  return c;
}


// One launch of an element kernel template: with events (hipExtLaunchKernelGGL
// records start / stop on the stream) or without.
template <typename K>
void launch_elem(K kern, int block, int nblk, hipStream_t s, const hipEvent_t *ev, JobDev *const *jobs,
                 const Task *tasks, int cs, int derive) {
  if (ev)
    hipExtLaunchKernelGGL(kern, dim3(nblk), dim3(block), 0, s, ev[0], ev[1], 0, jobs, tasks, cs, derive);
  else
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(block), 0, s, jobs, tasks, cs, derive);
}

}  // namespace dnlg
