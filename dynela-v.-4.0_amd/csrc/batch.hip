// batch.hip -- block task lists, the pre-loop, and the cycle loop over a
// batch of jobs (one launch per phase covers every job of the batch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <map>

#include "batch.h"
#include "kernels.h"

namespace dnlg {

constexpr int kSnap = kSnapDoubles;  // snapshot doubles per job (launch_snapshot)

struct TaskList {
  Task *d = nullptr;
  int n = 0;
};

// Element types, indexed as the per-type task lists below.
constexpr int kNumTypes = 4;
constexpr int kTypes[kNumTypes] = {DNLG_EL_QUA4N_AX, DNLG_EL_HEX8N_3D, DNLG_EL_QUA4N_2D, DNLG_EL_TET4N_3D};

inline bool variant_mixed(int mx) { return mx == kVarMixed || mx == kVarMixedRR32 || mx == kVarMixedX; }

// One element launch of variant mx with schedule (block, waves) over a task
// list.
void launch_variant(int type, int mx, int block, int waves, JobDev *const *jobs, const Task *tasks, int nblk, int cs,
                    int derive, hipStream_t s, const hipEvent_t *ev) {
  if (variant_mixed(mx))
    launch_elem_mixed(type, block, waves, jobs, tasks, nblk, cs, derive, s, ev, mx != kVarMixedRR32, mx == kVarMixedX);
  else
    launch_elem_fp64(type, block, waves, jobs, tasks, nblk, cs, derive, s, ev, mx == kVarFP64Ann);
}

// Element launches of a batch: one per (element type, variant, schedule).
struct ElemGroup {
  int type, variant, block, waves;
  TaskList tl;
};
// Corrector launches: one per (dimension, force precision, slot batch).
struct CorrGroup {
  int dim;
  bool f32;
  int batch;
  TaskList tl;
};

struct Batch {
  std::vector<dnlg_job *> jobs;
  JobDev **d_jobs = nullptr;
  std::vector<ElemGroup> elem;
  std::vector<CorrGroup> corr;
  int graph_cycles = kDefaultGraphCycles;  // cycles per captured graph (the jobs' schedules)
  TaskList prep[kNumTypes];     // [type]
  TaskList cour[kNumTypes];     // [type]: Courant lengths and power-iteration products
  TaskList node[2][2];  // [dim==3?][mixed?]: mass assembly, power iteration
  TaskList step;        // cycle control + field0 BCs, over the BC node lists
  TaskList close;       // end of a launch sequence, one block per job
  // history files: evaluation (item ranges and kinetic-energy node ranges,
  // hist_blocks() of each per job), one block per job with history or
  // result files (commit); hist_cap: the smallest ring (rows) of the
  // batch's files
  TaskList hist_eval, hist_jobs;
  int hist_cap = 0;
  bool has_courant = false, has_pi = false;  // time-step methods present
  int *d_pend = nullptr;                     // power-iteration pending flags
  std::vector<int> h_pend;
  double *d_snap = nullptr;
  std::vector<double> h_snap;
  // captured graphs by (kind, increment phase), see graph_for
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;
  std::vector<void *> allocs;
  ~Batch() {
    for (auto &g : graphs) hipGraphExecDestroy(g.second);
    for (void *p : allocs) hipFree(p);
  }
};

namespace {

int upload_tasks(Batch &B, std::vector<Task> h, TaskList &out) {
  out.n = (int)h.size();
  out.d = nullptr;
  if (h.empty()) return 0;
  for (Task &t : h) t.jd = B.jobs[t.job]->dev;
  void *p = nullptr;
  if (hipMalloc(&p, h.size() * sizeof(Task)) != hipSuccess) return DNLG_E_OOM;
  B.allocs.push_back(p);
  if (hipMemcpy(p, h.data(), h.size() * sizeof(Task), hipMemcpyHostToDevice) != hipSuccess) return DNLG_E_HIP;
  out.d = (Task *)p;
  return 0;
}

// Contiguous item ranges per job, `per` items per block.
std::vector<Task> make_tasks(const std::vector<int> &idx, const std::vector<int64_t> &count, int64_t per) {
  std::vector<Task> t;
  for (size_t k = 0; k < idx.size(); k++) {
    const int64_t n = count[k];
    int32_t blk = 0;
    for (int64_t f = 0; f < n; f += per, blk++) t.push_back(Task{idx[k], blk, f, std::min(f + per, n)});
  }
  return t;
}

// Workgroups are dealt round-robin over the 8 XCDs, each with its own L2.
// Place logical block L at launch slot (L % per) * 8 + L / per so that each
// XCD walks one contiguous range of elements and their shared nodes stay in
// that XCD's L2.  Padding slots carry empty ranges.
std::vector<Task> xcd_order(const std::vector<Task> &t) {
  const int n = (int)t.size();
  if (n < 16) return t;
  const int per = (n + 7) / 8;
  std::vector<Task> o((size_t)per * 8, Task{t[0].job, 0, 0, 0});
  for (int L = 0; L < n; L++) o[(size_t)(L % per) * 8 + L / per] = t[L];
  return o;
}

int build_batch(dnlg_ctx *ctx, const std::vector<dnlg_job *> &jobs, Batch &B) {
  B.jobs = jobs;
  const int n = (int)jobs.size();
  void *p = nullptr;
  if (hipMalloc(&p, n * sizeof(JobDev *)) != hipSuccess) return DNLG_E_OOM;
  B.allocs.push_back(p);
  B.d_jobs = (JobDev **)p;
  std::vector<JobDev *> h(n);
  for (int i = 0; i < n; i++) h[i] = jobs[i]->dev;
  if (hipMemcpy(B.d_jobs, h.data(), n * sizeof(JobDev *), hipMemcpyHostToDevice) != hipSuccess) return DNLG_E_HIP;
  if (hipMalloc(&p, kSnap * n * sizeof(double)) != hipSuccess) return DNLG_E_OOM;
  B.allocs.push_back(p);
  B.d_snap = (double *)p;
  B.h_snap.assign(kSnap * n, 0.0);
  for (auto *j : jobs) {
    B.has_pi = B.has_pi || is_pi(j->host.dt_method);
    B.has_courant = B.has_courant || j->host.dt_method == DNLG_DT_COURANT;
  }
  if (hipMalloc(&p, n * sizeof(int)) != hipSuccess) return DNLG_E_OOM;
  B.allocs.push_back(p);
  B.d_pend = (int *)p;
  B.h_pend.assign(n, 0);
  int st;
  // schedules, then element groups in (type, variant, block, waves) order.
  // Jobs without a schedule of their own take the rule for the element
  // count of all such jobs of their type and precision together: they are
  // launched together, so that count is the launch's size, and one schedule
  // per kind keeps them in one element launch and one corrector launch.
  {
    std::map<std::array<int, 3>, int64_t> total;
    for (auto *j : jobs)
      if (!j->sched_set) total[{j->type, j->prec, job_variant(j)}] += j->NE;
    for (auto *j : jobs)
      if (!j->sched_set) j->sched = resolve_schedule(ctx, j, total[{j->type, j->prec, job_variant(j)}]);
  }
  B.graph_cycles = 0;
  for (auto *j : jobs) B.graph_cycles = std::max(B.graph_cycles, (int)j->sched.graph_cycles);
  std::map<std::array<int, 4>, std::vector<int>> groups;
  for (int i = 0; i < n; i++)
    groups[{jobs[i]->type, job_variant(jobs[i]), jobs[i]->sched.elem_block, jobs[i]->sched.waves}].push_back(i);
  for (auto &g : groups) {
    std::vector<int64_t> cnt;
    for (int i : g.second) cnt.push_back(jobs[i]->NE);
    ElemGroup eg{g.first[0], g.first[1], g.first[2], g.first[3], TaskList{}};
    const int nn = type_nodes(eg.type);
    if ((st = upload_tasks(B, xcd_order(make_tasks(g.second, cnt, eg.block / nn)), eg.tl))) return st;
    B.elem.push_back(eg);
  }
  for (int hx = 0; hx < kNumTypes; hx++) {
    const int type = kTypes[hx];
    std::vector<int> idx;
    std::vector<int64_t> cnt;
    for (int i = 0; i < n; i++)
      if (jobs[i]->type == type) {
        idx.push_back(i);
        cnt.push_back(jobs[i]->NE);
      }
    if ((st = upload_tasks(B, make_tasks(idx, cnt, 256), B.prep[hx]))) return st;
    if ((st = upload_tasks(B, make_tasks(idx, cnt, 256), B.cour[hx]))) return st;
  }
  {
    // at least one block per job (k_cycle_step counts max(1, ceil(n_bcn / 256)))
    std::vector<Task> t;
    for (int i = 0; i < n; i++) {
      const int64_t nb = jobs[i]->n_bcn;
      if (nb == 0) t.push_back(Task{i, 0, 0, 0});
      for (int64_t f = 0, blk = 0; f < nb; f += 256, blk++) t.push_back(Task{i, (int32_t)blk, f, std::min(f + 256, nb)});
    }
    if ((st = upload_tasks(B, t, B.step))) return st;
    t.clear();
    for (int i = 0; i < n; i++) t.push_back(Task{i, 0, 0, 0});
    if ((st = upload_tasks(B, t, B.close))) return st;
  }
  for (int d3 = 0; d3 < 2; d3++)
    for (int mx = 0; mx < 2; mx++) {
      std::vector<int> idx;
      std::vector<int64_t> cnt;
      for (int i = 0; i < n; i++)
        if ((jobs[i]->dim == 3) == (d3 == 1) && is_mixed(jobs[i]->prec) == (mx == 1)) {
          idx.push_back(i);
          cnt.push_back(jobs[i]->N);
        }
      if ((st = upload_tasks(B, make_tasks(idx, cnt, kNodeBlock), B.node[d3][mx]))) return st;
    }
  {
    std::map<std::array<int, 3>, std::vector<int>> cg;
    for (int i = 0; i < n; i++) cg[{jobs[i]->dim, is_mixed(jobs[i]->prec) ? 1 : 0, jobs[i]->sched.corr_batch}].push_back(i);
    for (auto &g : cg) {
      std::vector<int64_t> cnt;
      for (int i : g.second) cnt.push_back(jobs[i]->N);
      CorrGroup c{g.first[0], g.first[1] == 1, g.first[2], TaskList{}};
      if ((st = upload_tasks(B, make_tasks(g.second, cnt, kNodeBlock), c.tl))) return st;
      B.corr.push_back(c);
    }
  }
  {
    // ranges of about equal size, hist_blocks(count) of them
    auto ranges = [](std::vector<Task> &t, int job, int64_t count, bool ke) {
      const int nb = hist_blocks(count);
      const int64_t per = (count + nb - 1) / nb;
      for (int b = 0; b < nb; b++) {
        const int64_t f = b * per, e = std::min(count, f + per);
        if (f < e || (ke && f == 0)) t.push_back(Task{job, ke ? -1 - b : b, f, std::max(f, e)});
      }
    };
    std::vector<Task> eval, per_job;
    for (int i = 0; i < n; i++) {
      if (jobs[i]->hist.empty() && !jobs[i]->host.res.on) continue;
      per_job.push_back(Task{i, 0, 0, 0});
      if (jobs[i]->hist.empty()) continue;
      ranges(eval, i, (int64_t)jobs[i]->hist_items.size(), false);
      if (!jobs[i]->hist_ke_cols.empty()) ranges(eval, i, jobs[i]->N, true);
      for (const auto &h : jobs[i]->hist) B.hist_cap = B.hist_cap ? std::min(B.hist_cap, h.cap) : h.cap;
    }
    if ((st = upload_tasks(B, eval, B.hist_eval)) || (st = upload_tasks(B, per_job, B.hist_jobs))) return st;
  }
  (void)ctx;
  return 0;
}

// Phases of one cycle.  pre = end of the previous cycle + updateTimes +
// field0 BCs (k_cycle_step), element = the element kernels, post =
// predictor/corrector + Courant.  cs: the control slot of the cycle (see
// Control); the last cycle of a launch sequence stays open until
// launch_close (end of step and time step only), which leaves the record in
// slot 0.
void launch_pre(dnlg_ctx *ctx, Batch &B, int cs) {
  launch_cycle_step(B.d_jobs, B.step.d, B.step.n, cs, 1, 1 - cs, ctx->stream);
}

void launch_close(dnlg_ctx *ctx, Batch &B, int cs) {
  launch_cycle_step(B.d_jobs, B.close.d, B.close.n, cs, 0, 0, ctx->stream);
}

// ev: nullptr, or two events per element launch (up to kElemVariants per element type); each
// launch is bracketed by event records on the engine's stream.
int launch_element(dnlg_ctx *ctx, Batch &B, const hipEvent_t *ev, int cs, int derive) {
  hipStream_t s = ctx->stream;
  int k = 0;
  for (const ElemGroup &g : B.elem) {
    if (g.tl.n <= 0) continue;
    const hipEvent_t *e = ev ? &ev[2 * k] : nullptr;
    launch_variant(g.type, g.variant, g.block, g.waves, B.d_jobs, g.tl.d, g.tl.n, cs, derive, s, e);
    k++;
  }
  return k;
}

// courant = false leaves out the Courant kernel: only for a cycle known not
// to be a time-step cycle of any running job (see launch_cycles).  fold: the
// corrector also writes the next cycle's field0 BCs, and the next cycle's
// element launch derives its control record.
// One power iteration for every job of the batch that is iterating.
void launch_pi_iteration(dnlg_ctx *ctx, Batch &B, int cs) {
  hipStream_t s = ctx->stream;
  for (int hx = 0; hx < kNumTypes; hx++) launch_pi_elem(kTypes[hx], B.d_jobs, B.cour[hx].d, B.cour[hx].n, s);
  for (int d3 = 0; d3 < 2; d3++)
    for (int mx = 0; mx < 2; mx++) launch_pi_node(d3 ? 3 : 2, B.d_jobs, B.node[d3][mx].d, B.node[d3][mx].n, s);
  launch_pi_step(B.d_jobs, B.close.d, B.close.n, cs, s);
}

// Power iterations driven from the host until every job that iterates has
// converged (or failed after 1000 iterations): 8 iterations between checks.
int run_pi_host(dnlg_ctx *ctx, Batch &B, int cs) {
  hipStream_t s = ctx->stream;
  const int n = (int)B.jobs.size();
  for (int round = 0; round < 200; round++) {
    for (int k = 0; k < 8; k++) launch_pi_iteration(ctx, B, cs);
    launch_pi_pending(B.d_jobs, n, B.d_pend, s);
    if (hipMemcpyAsync(B.h_pend.data(), B.d_pend, n * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return DNLG_E_HIP;
    bool pending = false;
    for (int v : B.h_pend) pending = pending || v != 0;
    if (!pending) return 0;
  }
  return DNLG_E_HIP;  // unreachable: k_pi_step stops a job after 1000 iterations
}

void launch_history(dnlg_ctx *ctx, Batch &B, int cs, int preloop) {
  launch_history(B.d_jobs, B.hist_eval.d, B.hist_eval.n, B.hist_jobs.d, B.hist_jobs.n, cs, preloop, ctx->stream);
}

void launch_post(dnlg_ctx *ctx, Batch &B, int cs, bool courant, bool fold = false) {
  hipStream_t s = ctx->stream;
  for (const CorrGroup &g : B.corr) launch_correct(g.dim, g.f32, g.batch, B.d_jobs, g.tl.d, g.tl.n, cs, fold ? 1 : 0, s);
  // Explicit::endStep's history rows, before the time-step computation
  launch_history(ctx, B, cs, 0);
  if (courant && B.has_courant)
    for (int hx = 0; hx < kNumTypes; hx++)
      launch_courant(type_dim(kTypes[hx]), B.d_jobs, B.cour[hx].d, B.cour[hx].n, cs, 0, s);
  if (courant && B.has_pi) {
    // a few warm-started iterations in the sequence; a job that has not
    // converged stalls at the next k_cycle_step and the host finishes it
    launch_pi_begin(B.d_jobs, (int)B.jobs.size(), cs, 0, s);
    for (int k = 0; k < ctx->pi_slots; k++) launch_pi_iteration(ctx, B, cs);
  }
}

// Cycles j0 .. j1-1 of a launch sequence whose running jobs share the
// increment phase `phase` = increment mod freq (or phase < 0: no common
// phase), starting from control slot cs; returns the slot of the last
// cycle.  Cycle j has increment phase + j + 1 and updates the time step iff
// that is a multiple of freq.  With a common phase, a cycle without a
// time-step update that is not the sequence's last hands the opening of the
// next cycle to the corrector (field0 BCs) and the next element launch
// (control record), so k_cycle_step runs where the sequence starts and after
// time-step cycles only.  Every cycle moves the record to the other slot.
// One cycle's element, corrector and Courant kernels with the jobs of each
// dimension in a branch of their own: the 3D branch on the engine's stream,
// the 2D branch on the side stream, forked after the cycle opens and joined
// before the end-of-step launches.  The branches touch disjoint jobs, so
// either order gives the same results; running them together lets one
// branch's kernels fill the GPU while the other's launches drain (the
// element kernels are bound by their arithmetic, the correctors by HBM).
// Returns false (nothing launched) when the batch holds one dimension only.
bool launch_branches(dnlg_ctx *ctx, Batch &B, int el_cs, int derive, int cs, bool courant, bool fold) {
  bool dims[2] = {false, false};  // [dim == 3]
  for (const ElemGroup &g : B.elem)
    if (g.tl.n > 0) dims[type_dim(g.type) == 3] = true;
  if (!ctx->branches || !dims[0] || !dims[1]) return false;
  hipStream_t main = ctx->stream, side = ctx->side;
  hipEventRecord(ctx->fork_ev, main);
  hipStreamWaitEvent(side, ctx->fork_ev, 0);
  for (int d3 = 1; d3 >= 0; d3--) {
    hipStream_t s = d3 ? main : side;
    for (const ElemGroup &g : B.elem)
      if (g.tl.n > 0 && (type_dim(g.type) == 3) == (d3 == 1))
        launch_variant(g.type, g.variant, g.block, g.waves, B.d_jobs, g.tl.d, g.tl.n, el_cs, derive, s, nullptr);
    for (const CorrGroup &g : B.corr)
      if ((g.dim == 3) == (d3 == 1)) launch_correct(g.dim, g.f32, g.batch, B.d_jobs, g.tl.d, g.tl.n, cs, fold ? 1 : 0, s);
    if (courant && B.has_courant)
      for (int hx = 0; hx < kNumTypes; hx++)
        if ((type_dim(kTypes[hx]) == 3) == (d3 == 1))
          launch_courant(type_dim(kTypes[hx]), B.d_jobs, B.cour[hx].d, B.cour[hx].n, cs, 0, s);
  }
  hipEventRecord(ctx->join_ev, side);
  hipStreamWaitEvent(main, ctx->join_ev, 0);
  // Explicit::endStep's history rows (the Courant kernels above only write
  // their partial minima, which the history reads nowhere)
  launch_history(ctx, B, cs, 0);
  if (courant && B.has_pi) {
    launch_pi_begin(B.d_jobs, (int)B.jobs.size(), cs, 0, main);
    for (int k = 0; k < ctx->pi_slots; k++) launch_pi_iteration(ctx, B, cs);
  }
  return true;
}

int launch_cycles(dnlg_ctx *ctx, Batch &B, int j0, int j1, int phase, int freq, int cs) {
  auto due = [&](int j) { return phase < 0 || (phase + j + 1) % freq == 0; };
  bool folded = false;
  for (int j = j0; j < j1; j++) {
    // the element launch reads the record in el_cs (derive: and writes the
    // new one into the other slot); cs: the cycle's record from then on
    int el_cs, derive;
    if (folded) {
      el_cs = cs;
      derive = 1;
    } else {
      launch_pre(ctx, B, cs);
      el_cs = 1 - cs;
      derive = 0;
    }
    cs = 1 - cs;
    folded = phase >= 0 && j + 1 < j1 && !due(j);
    if (!launch_branches(ctx, B, el_cs, derive, cs, due(j), folded)) {
      launch_element(ctx, B, nullptr, el_cs, derive);
      launch_post(ctx, B, cs, due(j), folded);
    }
  }
  return cs;
}

// Explicit.C:143-178: reference Jacobian (detJ0), lumped mass, forced time
// step, then one element pass on the zero field0 increment.
int run_preloop(dnlg_ctx *ctx, Batch &B) {
  hipStream_t s = ctx->stream;
  const int n = (int)B.jobs.size();
  for (int hx = 0; hx < kNumTypes; hx++) launch_prep(kTypes[hx], B.d_jobs, B.prep[hx].d, B.prep[hx].n, s);
  for (int d3 = 0; d3 < 2; d3++)
    for (int mx = 0; mx < 2; mx++) launch_mass_gather(d3 ? 3 : 2, B.d_jobs, B.node[d3][mx].d, B.node[d3][mx].n, s);
  for (int hx = 0; hx < kNumTypes; hx++)
    launch_courant(type_dim(kTypes[hx]), B.d_jobs, B.cour[hx].d, B.cour[hx].n, 0, 1, s);
  launch_force_dt(B.d_jobs, n, s);
  if (B.has_pi) {
    // cold start of the power iteration (Model.C:966-970) on the reference
    // Jacobian and the lumped mass
    launch_pi_begin(B.d_jobs, n, 0, 1, s);
    int st = run_pi_host(ctx, B, 0);
    if (st) return st;
    launch_pi_finish(B.d_jobs, n, 1, s);
  }
  for (const ElemGroup &g : B.elem)
    launch_variant(g.type, g.variant, g.block, g.waves, B.d_jobs, g.tl.d, g.tl.n, 0, 0, s, nullptr);
  // the pre-loop's history rows (Explicit.C:182)
  launch_history(ctx, B, 0, 1);
  if (hipGetLastError() != hipSuccess) return DNLG_E_HIP;
  return 0;
}

}  // namespace

static int variant_of(int precision, int law, int stress_method) {
  if (precision == DNLG_PREC_MIXED)
    return (law == DNLG_LAW_TABULAR || law == DNLG_LAW_ANN2SIG || stress_method == DNLG_STRESS_DIRECT) ? kVarMixedX
                                                                                                    : kVarMixed;
  if (precision == DNLG_PREC_MIXED_RR32) return kVarMixedRR32;
  return law == DNLG_LAW_ANN2SIG ? kVarFP64Ann : kVarFP64;
}

int job_variant(const dnlg_job *j) { return variant_of(j->prec, j->host.law, j->host.stress_method); }

bool schedule_valid(int type, int precision, int law, const dnlg_schedule &s, int stress_method) {
  const int v = variant_of(precision, law, stress_method);
  return elem_sched_valid(type, v, s.elem_block, s.waves) && corr_batch_valid(precision != DNLG_PREC_FP64, s.corr_batch) &&
         s.graph_cycles >= 2 && s.graph_cycles <= 4096;
}

dnlg_schedule resolve_schedule(const dnlg_ctx *ctx, const dnlg_job *job) { return resolve_schedule(ctx, job, job->NE); }

// The schedule rule for a launch of `ne` elements of job's type and
// precision (a job alone: its own count; in a batch: the count of the
// launch it shares with the other jobs of its kind).
dnlg_schedule resolve_schedule(const dnlg_ctx *ctx, const dnlg_job *job, int64_t ne) {
  const int v = job_variant(job);
  dnlg_schedule s{default_elem_block(job->type), default_waves(v), default_corr_batch(is_mixed(job->prec)),
                  ctx->graph_cycles > 0 ? ctx->graph_cycles : kDefaultGraphCycles};
  for (const auto &r : ctx->sched_rules)
    if (r.type == job->type && r.prec == job->prec && ne >= r.min_elems && (r.max_elems <= 0 || ne < r.max_elems)) {
      dnlg_schedule t = s;
      if (r.sched.elem_block) t.elem_block = r.sched.elem_block;
      if (r.sched.waves) t.waves = r.sched.waves;
      if (r.sched.corr_batch) t.corr_batch = r.sched.corr_batch;
      if (r.sched.graph_cycles) t.graph_cycles = r.sched.graph_cycles;
      if (schedule_valid(job->type, job->prec, job->host.law, t, job->host.stress_method)) s = t;
      break;
    }
  return s;
}

int job_setup(dnlg_ctx *ctx, dnlg_job *job) {
  if (job->started) return 0;
  Batch P;
  int st = build_batch(ctx, {job}, P);
  if (st) return st;
  hipStream_t s = ctx->stream;
  for (int hx = 0; hx < kNumTypes; hx++) launch_prep(kTypes[hx], P.d_jobs, P.prep[hx].d, P.prep[hx].n, s);
  for (int d3 = 0; d3 < 2; d3++)
    for (int mx = 0; mx < 2; mx++) launch_mass_gather(d3 ? 3 : 2, P.d_jobs, P.node[d3][mx].d, P.node[d3][mx].n, s);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return DNLG_E_HIP;
  Sync sy;
  if (hipMemcpy(&sy, &job->dev->sync, sizeof(Sync), hipMemcpyDeviceToHost) != hipSuccess) return DNLG_E_HIP;
  job->status = sy.st[0] != 0 ? sy.st[0] : sy.st[1];  // a negative reference Jacobian
  job->started = true;
  return 0;
}

namespace {

Batch *find_or_build(dnlg_ctx *ctx, const std::vector<dnlg_job *> &jobs, int &st) {
  for (auto &b : ctx->cache)
    if (b->jobs == jobs) return b.get();
  BatchPtr B(new Batch());
  st = build_batch(ctx, jobs, *B);
  if (st) return nullptr;
  if (ctx->cache.size() >= 8) ctx->cache.erase(ctx->cache.begin());
  ctx->cache.push_back(std::move(B));
  return ctx->cache.back().get();
}

template <typename F>
int capture(dnlg_ctx *ctx, F &&f, hipGraphExec_t *out) {
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return DNLG_E_HIP;
  f();
  if (hipStreamEndCapture(ctx->stream, &g) != hipSuccess) return DNLG_E_HIP;
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return DNLG_E_HIP;
  *out = ex;
  return 0;
}

// Graph kinds: whole cycles, and for the timed mode the pre and post phases
// of one sampled cycle (the element kernels run between them, bracketed by
// events) plus the remaining cycles of a unit.
enum GraphKind { kWhole = 0, kPre = 1, kPost = 2, kRest = 3 };

// Cycles per graph: with phase < 0 (running jobs at different increments
// modulo the time-step frequency) `cycles` cycles with the Courant kernel in
// every cycle (it returns at once on the other cycles); with a common phase
// p = increment mod freq, a multiple of freq cycles with the Courant kernel
// only in the cycles whose increment is a multiple of freq, so the phase is
// the same at the end of the graph.
int graph_len(int cycles, int phase, int freq) {
  int len = phase < 0 ? cycles : freq * std::max(1, (cycles + freq - 1) / freq);
  // an even count: each cycle moves the control record to the other slot,
  // and every graph starts and ends with it in slot 0
  if (len % 2) len += phase < 0 ? 1 : freq;
  return len;
}

hipGraphExec_t graph_for(dnlg_ctx *ctx, Batch &B, int kind, int phase, int freq, int &st) {
  const auto key = std::make_pair(kind * 65536 + (phase < 0 ? 0 : freq), phase);
  auto it = B.graphs.find(key);
  if (it != B.graphs.end()) return it->second;
  const int len = graph_len(B.graph_cycles, phase, freq);
  hipGraphExec_t g = nullptr;
  switch (kind) {
    case kWhole:
      st = capture(ctx, [&] { launch_cycles(ctx, B, 0, len, phase, freq, 0); }, &g);
      break;
    case kPre:  // slot 0 -> 1
      st = capture(ctx, [&] { launch_pre(ctx, B, 0); }, &g);
      break;
    case kPost:
      st = capture(ctx, [&] { launch_post(ctx, B, 1, true); }, &g);
      break;
    default:  // the len - 1 cycles after a sampled one, slot 1 -> 0
      st = capture(ctx, [&] { launch_cycles(ctx, B, 1, len, phase, freq, 1); }, &g);
      break;
  }
  if (st) return nullptr;
  B.graphs[key] = g;
  return g;
}

}  // namespace

int batch_run(dnlg_ctx *ctx, const std::vector<dnlg_job *> &jobs, const std::vector<double> &up_to,
              int64_t max_cycles) {
  const int n = (int)jobs.size();
  hipStream_t s = ctx->stream;
  int st = 0;
  std::vector<dnlg_job *> fresh;
  for (auto *j : jobs)
    if (!j->started && j->status == 0) fresh.push_back(j);
  if (!fresh.empty()) {
    Batch P;
    if ((st = build_batch(ctx, fresh, P)) || (st = run_preloop(ctx, P))) {
      ctx->err = "pre-loop launch failed";
      return st;
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
      ctx->err = "pre-loop failed";
      return DNLG_E_HIP;
    }
    for (auto *j : fresh) j->started = true;
  }
  Batch *B = find_or_build(ctx, jobs, st);
  if (!B) {
    ctx->err = "building the batch failed";
    return st;
  }
  // target times, clipped to each job's end time
  for (int i = 0; i < n; i++) B->h_snap[i] = std::min(up_to[i], jobs[i]->t_end);
  if (hipMemcpyAsync(B->d_snap, B->h_snap.data(), n * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) {
    ctx->err = "setting the target times failed";
    return DNLG_E_HIP;
  }
  launch_set_up_to(B->d_jobs, n, B->d_snap, s);
  const bool use_graph = ctx->graph_cycles > 0;
  const bool timed = ctx->timing;
  constexpr int kTimedChunk = 64;  // sampled cycles between event read-backs
  std::vector<int> ev_used;        // events used per timed cycle
  auto collect = [&]() -> int {
    auto &acc = ctx->ktime["element"];
    size_t base = 0;
    for (int used : ev_used) {
      for (int k = 0; k < used; k++) {
        float ms = 0.f;
        const hipError_t e = hipEventElapsedTime(&ms, ctx->events[base + 2 * k], ctx->events[base + 2 * k + 1]);
        if (e != hipSuccess) {
          ctx->err = std::string("kernel timing failed: ") + hipGetErrorString(e);
          return DNLG_E_HIP;
        }
        acc.first += ms;
        acc.second += 1;
      }
      base += 2 * used;
    }
    ev_used.clear();
    return 0;
  };
  // DNLG_PROGRESS=<seconds>: a progress line on stderr at that period
  const char *prog_env = std::getenv("DNLG_PROGRESS");
  const double prog_period = prog_env ? std::atof(prog_env) : 0.0;
  auto prog_last = std::chrono::steady_clock::now();
  int64_t launched = 0;
  while (true) {
    if (prog_period > 0.0) {
      const auto now = std::chrono::steady_clock::now();
      if (std::chrono::duration<double>(now - prog_last).count() >= prog_period) {
        prog_last = now;
        int active = 0;
        for (int i = 0; i < n; i++)
          if ((int)B->h_snap[kSnap * i + 3] == 0 && B->h_snap[kSnap * i] < B->h_snap[kSnap * i + 2]) active++;
        std::fprintf(stderr, "[dnlg] batch of %d jobs: %lld cycles launched, %d active\n", n, (long long)launched,
                     active);
        std::fflush(stderr);
      }
    }
    launch_snapshot(B->d_jobs, n, B->d_snap, s);
    if (hipMemcpyAsync(B->h_snap.data(), B->d_snap, kSnap * n * sizeof(double), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      ctx->err = "device error during the cycle loop";
      return DNLG_E_HIP;
    }
    if ((st = collect())) return st;
    for (auto *j : jobs)
      if ((st = history_drain(j))) return st;
    bool held = false;
    for (int i = 0; i < n; i++)
      if (B->h_snap[kSnap * i + 6] != 0.0) {
        held = true;
        if ((st = result_file_release(jobs[i]))) return st;
      }
    if (held) continue;
    bool stalled = false;
    for (int i = 0; i < n; i++) stalled = stalled || B->h_snap[kSnap * i + 5] != 0.0;
    if (stalled) {
      // power iterations that the launch sequence's slots did not finish
      if ((st = run_pi_host(ctx, *B, 0))) {
        ctx->err = "power iteration failed";
        return st;
      }
      launch_pi_finish(B->d_jobs, n, 0, s);
      continue;
    }
    double need = 0.0;
    int phase = -2, freq = 0;  // common increment phase of the running jobs, -1 if none
    for (int i = 0; i < n; i++) {
      const double t = B->h_snap[kSnap * i], dt = B->h_snap[kSnap * i + 1], up = B->h_snap[kSnap * i + 2];
      const int js = (int)B->h_snap[kSnap * i + 3];
      jobs[i]->status = js;
      jobs[i]->t_now = t;
      jobs[i]->up_to = up;
      jobs[i]->inc_now = (int64_t)B->h_snap[kSnap * i + 4];
      if (js != 0 || !(t < up)) continue;
      need = std::max(need, dt > 0 ? std::ceil((up - t) / dt) : 1.0);
      const int f = jobs[i]->host.dt_freq;
      const int p = (int)((int64_t)B->h_snap[kSnap * i + 4] % f);
      if (phase == -2) {
        phase = p;
        freq = f;
      } else if (phase != p || freq != f) {
        phase = -1;
      }
    }
    if (need <= 0.0) break;
    if (phase < 0) {
      phase = -1;
      freq = 1;
    }
    if (max_cycles >= 0) {
      if (launched >= max_cycles) break;
      need = std::min(need, (double)(max_cycles - launched));
    }
    if (use_graph && phase > 0) {
      // A captured sequence is specific to the increment phase it starts
      // at, and capturing one costs as much as a few hundred cycles: bring
      // the jobs to phase 0 with at most freq - 1 directly launched cycles,
      // so that each batch captures the phase-0 sequence once and reuses it
      // across solve calls.
      const int k = (int)std::min(need, (double)(freq - phase));
      const int cs = launch_cycles(ctx, *B, 0, k, phase, freq, 0);
      launch_close(ctx, *B, cs);
      launched += k;
      const hipError_t le = hipGetLastError();
      if (le != hipSuccess) {
        ctx->err = std::string("kernel launch failed: ") + hipGetErrorString(le);
        return DNLG_E_HIP;
      }
      continue;
    }
    int chunk = (int)std::min(need, 4096.0);
    if (B->hist_cap > 0) {
      // at most one history row per cycle and file between two drains
      const int per = use_graph ? graph_len(B->graph_cycles, phase, freq) : 1;
      chunk = std::min(chunk, std::max(1, B->hist_cap / per) * per);
    }
    launched += chunk;
    hipGraphExec_t g_whole = nullptr, g_pre = nullptr, g_post = nullptr, g_rest = nullptr;
    if (use_graph) {
      if (timed) {
        g_pre = graph_for(ctx, *B, kPre, -1, 1, st);
        if (!st) g_post = graph_for(ctx, *B, kPost, -1, 1, st);
        if (!st && graph_len(B->graph_cycles, phase, freq) > 1) g_rest = graph_for(ctx, *B, kRest, phase, freq, st);
      } else {
        g_whole = graph_for(ctx, *B, kWhole, phase, freq, st);
      }
      if (st) {
        ctx->err = "graph capture failed";
        return st;
      }
    }
    int cs = 0;  // control slot of the cycle in flight (slot 0 between sequences)
    if (timed) {
      // units of one sampled cycle + the rest of a graph's cycles
      const int per = use_graph ? graph_len(B->graph_cycles, phase, freq) : 1;
      const int units = std::min((chunk + per - 1) / per, kTimedChunk);
      launched += (int64_t)units * per - chunk;
      while (ctx->events.size() < (size_t)2 * B->elem.size() * units) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return DNLG_E_HIP;
        ctx->events.push_back(e);
      }
      size_t base = 0;
      for (int k = 0; k < units; k++) {
        if (use_graph) {
          if (hipGraphLaunch(g_pre, s) != hipSuccess) return DNLG_E_HIP;
        } else {
          launch_pre(ctx, *B, cs);
        }
        cs = 1 - cs;
        const int used = launch_element(ctx, *B, &ctx->events[base], cs, 0);
        ev_used.push_back(used);
        base += 2 * used;
        if (use_graph) {
          // the sampled cycle runs in slot 1; the rest graph ends in slot 0
          if (hipGraphLaunch(g_post, s) != hipSuccess) return DNLG_E_HIP;
          if (g_rest && hipGraphLaunch(g_rest, s) != hipSuccess) return DNLG_E_HIP;
          cs = 0;
        } else {
          launch_post(ctx, *B, cs, true);
        }
      }
    } else if (use_graph) {
      const int len = graph_len(B->graph_cycles, phase, freq);
      const int reps = (chunk + len - 1) / len;
      launched += (int64_t)reps * len - chunk;
      for (int r = 0; r < reps; r++)
        if (hipGraphLaunch(g_whole, s) != hipSuccess) {
          ctx->err = "graph launch failed";
          return DNLG_E_HIP;
        }
    } else {
      cs = launch_cycles(ctx, *B, 0, chunk, phase, freq, 0);
    }
    launch_close(ctx, *B, cs);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
      ctx->err = std::string("kernel launch failed: ") + hipGetErrorString(le);
      return DNLG_E_HIP;
    }
  }
  for (auto *j : jobs)
    if (j->status != 0) {
      ctx->err = "job failed with status " + std::to_string(j->status);
      return j->status;
    }
  return 0;
}

void BatchDeleter::operator()(Batch *b) const { delete b; }

void batch_cache_clear(dnlg_ctx *ctx) { ctx->cache.clear(); }

void batch_cache_forget(dnlg_ctx *ctx, const dnlg_job *job) {
  auto &c = ctx->cache;
  c.erase(std::remove_if(c.begin(), c.end(),
                         [&](const BatchPtr &b) {
                           return std::find(b->jobs.begin(), b->jobs.end(), job) != b->jobs.end();
                         }),
          c.end());
}

}  // namespace dnlg
