// control.h -- device-side readers of a job's control record and status,
// shared by the cycle kernels and the history kernels.
#pragma once

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

}  // namespace dnlg
