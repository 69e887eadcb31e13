// batch.h -- host-side job / context records and the batched cycle loop.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "engine.h"
#include "sched.h"

struct dnlg_ctx;

// Host side of one history file: where its rows go and how far the device
// ring has been drained.
struct dnlg_history {
  std::string path;   // empty: rows kept in memory only
  FILE *fp = nullptr;
  int32_t n_items = 0, cap = 0;
  double *ring = nullptr;  // device ring (HistFile::ring)
  int64_t drained = 0;
  std::vector<double> rows;  // every row taken, 1 + n_items doubles each
};

struct dnlg_job {
  dnlg_ctx *ctx = nullptr;
  // launch schedule (sched.h): set explicitly (dnlg_job_set_schedule) or
  // resolved from the context's rules when a batch is built
  dnlg_schedule sched{};
  bool sched_set = false;
  int type = 0, dim = 0, nn = 0, prec = 0, track = 0;
  int64_t N = 0, NE = 0, G = 0;
  double t_end = 0.0;
  dnlg::JobDev host{};          // host mirror holding the device pointers
  dnlg::JobDev *dev = nullptr;  // device record (scalars live here)
  std::vector<void *> allocs;
  bool started = false;
  int32_t n_cmin = 0;
  int32_t n_bcn = 0;  // nodes with constant BCs
  double t_now = 0.0;  // time, target and increment at the last progress snapshot
  double up_to = 0.0;
  int64_t inc_now = 0;
  int status = 0;
  // history files: host records, the device tables' host copies, and the
  // device tables themselves (replaced whenever a file is added)
  std::vector<dnlg_history> hist;
  std::vector<dnlg::HistFile> hist_files;
  std::vector<dnlg::HistItem> hist_items;
  std::vector<int2> hist_inc;
  std::vector<int32_t> hist_ke_cols;
  std::vector<void *> hist_tables;
  // result (VTK) files: base name, next file index, fields written
  std::string vtk_base;
  int vtk_index = 0;
  std::vector<int32_t> vtk_fields;
  ~dnlg_job() {
    for (auto &h : hist)
      if (h.fp) std::fclose(h.fp);
  }
};

namespace dnlg {
struct Batch;
struct BatchDeleter {
  void operator()(Batch *b) const;
};
using BatchPtr = std::unique_ptr<Batch, BatchDeleter>;
}  // namespace dnlg

struct dnlg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // a second stream and two events: a batch with jobs of both dimensions
  // runs the 3D jobs' and the 2D jobs' cycle kernels as two branches
  // (branches = false: all on `stream`)
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  bool branches = true;
  std::string err;
  bool timing = false;
  std::map<std::string, std::pair<double, int64_t>> ktime;
  std::vector<hipEvent_t> events;  // pool for timed launches
  std::vector<dnlg::BatchPtr> cache;
  int graph_cycles = 16;  // cycles per captured HIP graph (0 = plain launches)
  int pi_slots = 2;       // power iterations per time-step cycle inside a sequence
  // schedule rules (dnlg_ctx_add_schedule): the first rule matching a job's
  // element type, precision and size gives its schedule
  struct SchedRule {
    int type, prec;
    int64_t min_elems, max_elems;
    dnlg_schedule sched;
  };
  std::vector<SchedRule> sched_rules;
  // device workspace of the energy reduction (grown on demand)
  double *energy_ws = nullptr;
  int64_t energy_ws_doubles = 0;
};

namespace dnlg {

// The schedule a job runs with: its own if set, else the first matching
// rule of the context, else the defaults (sched.h); fields left 0 in a rule
// take the defaults.
dnlg_schedule resolve_schedule(const dnlg_ctx *ctx, const dnlg_job *job);
dnlg_schedule resolve_schedule(const dnlg_ctx *ctx, const dnlg_job *job, int64_t ne);
int job_variant(const dnlg_job *job);
bool schedule_valid(int type, int precision, int law, const dnlg_schedule &s, int stress_method = DNLG_STRESS_NR);

// Advance the jobs together toward up_to[i] (clipped to each job's end
// time); with max_cycles >= 0, return after about that many cycles (whole
// graphs) even if some job has not arrived.  Returns the first failing job's
// status, or a launch error.
int batch_run(dnlg_ctx *ctx, const std::vector<dnlg_job *> &jobs, const std::vector<double> &up_to,
              int64_t max_cycles = -1);
void batch_cache_clear(dnlg_ctx *ctx);
// Copies the rows that the job's history files took since the last drain
// from the device rings to the host records and files (api.hip).
int history_drain(dnlg_job *job);
// The set-up half of the pre-loop for a job that has not been solved: the
// reference Jacobian (detJ0) and the lumped masses (Model.C:754-775,
// 537-590), without the forced time step and the element pass; marks the
// job started.  Blocking.
int job_setup(dnlg_ctx *ctx, dnlg_job *job);
// Writes the result file of a job held at the end of a cycle (the written
// cycle's field0 is the other ping-pong slot) and releases it (api.hip).
int result_file_release(dnlg_job *job);
void batch_cache_forget(dnlg_ctx *ctx, const dnlg_job *job);

}  // namespace dnlg
