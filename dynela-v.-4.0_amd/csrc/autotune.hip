// autotune.hip -- launch schedules of jobs and the autotuner that measures
// them on the device.
//
// The reference splits each parallel element loop into fixed OpenMP chunks
// (Parallel.C:78-113).  Here the layout of the element kernels (threads per
// block, waves per SIMD), the corrector's slot batching and the length of
// the captured graphs are chosen per element type, precision and job size
// from a table of rules that this tuner measured (dnlPython.autotune writes
// it); a schedule never changes what is computed (sched.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "../../include/dnlg.h"
#include "batch.h"

using namespace dnlg;

namespace {

// Normalised copy: 0 fields take the job's defaults.
dnlg_schedule normalised(const dnlg_job *job, const dnlg_schedule &in) {
  const int v = job_variant(job);
  dnlg_schedule s = in;
  if (!s.elem_block) s.elem_block = default_elem_block(job->type);
  if (!s.waves) s.waves = default_waves(v);
  if (!s.corr_batch) s.corr_batch = default_corr_batch(is_mixed(job->prec));
  if (!s.graph_cycles) s.graph_cycles = job->ctx->graph_cycles > 0 ? job->ctx->graph_cycles : kDefaultGraphCycles;
  return s;
}

int set_schedule(dnlg_job *job, const dnlg_schedule &s) {
  job->sched = s;
  job->sched_set = true;
  batch_cache_forget(job->ctx, job);  // the next advance builds the job's batch anew
  return DNLG_OK;
}

// ms per cycle of the job under its current schedule: one untimed graph's
// worth of cycles (graph capture, first launches), then cycles_per_trial
// cycles timed.  DNLG_E_STATE when the job reaches its end time first.
int trial(dnlg_job *job, int64_t cycles, double *ms) {
  dnlg_ctx *ctx = job->ctx;
  const std::vector<dnlg_job *> jobs{job};
  const std::vector<double> up{job->t_end};
  int st = batch_run(ctx, jobs, up, std::max<int64_t>(2, job->sched.graph_cycles));
  if (st) return st;
  const int64_t inc0 = job->inc_now;
  const auto t0 = std::chrono::steady_clock::now();
  if ((st = batch_run(ctx, jobs, up, cycles))) return st;
  const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  const int64_t done = job->inc_now - inc0;
  if (done < cycles / 2 || !(job->t_now < job->t_end)) return DNLG_E_STATE;
  *ms = wall / (double)done;
  return DNLG_OK;
}

// The candidate with the smallest median time, each measured `rounds`
// times round-robin.
bool same(const dnlg_schedule &a, const dnlg_schedule &b) {
  return a.elem_block == b.elem_block && a.waves == b.waves && a.corr_batch == b.corr_batch &&
         a.graph_cycles == b.graph_cycles;
}

int pick(dnlg_job *job, const std::vector<dnlg_schedule> &cand, int64_t cycles, int rounds, dnlg_schedule *best,
         double *best_ms, const dnlg_schedule *ref = nullptr, double *ref_ms = nullptr) {
  std::vector<std::vector<double>> t(cand.size());
  for (int r = 0; r < rounds; r++)
    for (size_t c = 0; c < cand.size(); c++) {
      set_schedule(job, cand[c]);
      double ms = 0.0;
      const int st = trial(job, cycles, &ms);
      if (st) return st;
      t[c].push_back(ms);
    }
  size_t bi = 0;
  double bm = 0.0;
  for (size_t c = 0; c < cand.size(); c++) {
    std::vector<double> v = t[c];
    std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
    const double med = v[v.size() / 2];
    if (c == 0 || med < bm) {
      bm = med;
      bi = c;
    }
    if (ref && ref_ms && same(cand[c], *ref)) *ref_ms = med;
  }
  *best = cand[bi];
  *best_ms = bm;
  return DNLG_OK;
}

}  // namespace

extern "C" {

int dnlg_schedule_valid(int elem_type, int precision, int law, const dnlg_schedule *sched) {
  if (!sched) return 0;
  if (!type_known(elem_type)) return 0;
  const int v = precision == DNLG_PREC_MIXED ? (law == DNLG_LAW_TABULAR || law == DNLG_LAW_ANN2SIG ? kVarMixedX : kVarMixed)
                : precision == DNLG_PREC_MIXED_RR32 ? kVarMixedRR32
                : (law == DNLG_LAW_ANN2SIG ? kVarFP64Ann : kVarFP64);
  dnlg_schedule s = *sched;
  if (!s.elem_block) s.elem_block = default_elem_block(elem_type);
  if (!s.waves) s.waves = default_waves(v);
  if (!s.corr_batch) s.corr_batch = default_corr_batch(precision != DNLG_PREC_FP64);
  if (!s.graph_cycles) s.graph_cycles = kDefaultGraphCycles;
  return schedule_valid(elem_type, precision, law, s) ? 1 : 0;
}

int dnlg_ctx_add_schedule(dnlg_ctx *ctx, int elem_type, int precision, int64_t min_elems, int64_t max_elems,
                          const dnlg_schedule *sched) {
  if (!ctx || !sched) return DNLG_E_INVALID;
  // the law is not part of a rule: a rule that names a non-default schedule
  // applies to the closed-form laws; network-law jobs keep their defaults
  if (!dnlg_schedule_valid(elem_type, precision, DNLG_LAW_JOHNSON_COOK, sched)) {
    ctx->err = "dnlg_ctx_add_schedule: the kernels are not built for this schedule";
    return DNLG_E_INVALID;
  }
  ctx->sched_rules.push_back(dnlg_ctx::SchedRule{elem_type, precision, min_elems, max_elems, *sched});
  batch_cache_clear(ctx);
  return DNLG_OK;
}

int dnlg_ctx_clear_schedules(dnlg_ctx *ctx) {
  if (!ctx) return DNLG_E_INVALID;
  ctx->sched_rules.clear();
  batch_cache_clear(ctx);
  return DNLG_OK;
}

int dnlg_job_set_schedule(dnlg_job *job, const dnlg_schedule *sched) {
  if (!job || !sched) return DNLG_E_INVALID;
  const dnlg_schedule s = normalised(job, *sched);
  if (!schedule_valid(job->type, job->prec, job->host.law, s, job->host.stress_method)) {
    job->ctx->err = "dnlg_job_set_schedule: the kernels are not built for this schedule";
    return DNLG_E_INVALID;
  }
  return set_schedule(job, s);
}

int dnlg_job_get_schedule(dnlg_job *job, dnlg_schedule *sched) {
  if (!job || !sched) return DNLG_E_INVALID;
  *sched = job->sched_set ? job->sched : resolve_schedule(job->ctx, job);
  return DNLG_OK;
}

int dnlg_job_autotune(dnlg_job *job, int64_t cycles_per_trial, int rounds, dnlg_schedule *best, double *ms_per_cycle) {
  // ms_per_cycle[0]: the best schedule, [1]: the starting one
  if (!job || cycles_per_trial < 2 || rounds < 1) return DNLG_E_INVALID;
  dnlg_ctx *ctx = job->ctx;
  if (job->status != DNLG_OK) return DNLG_E_STATE;
  if (ctx->graph_cycles <= 0) {
    ctx->err = "dnlg_job_autotune: graphs are switched off (DNLG_GRAPH_CYCLES=0)";
    return DNLG_E_STATE;
  }
  dnlg_schedule cur;
  dnlg_job_get_schedule(job, &cur);
  const dnlg_schedule start = cur;
  const int v = job_variant(job);
  double ms = 0.0, start_ms = -1.0;
  int st;
  // 1. element kernel: threads per block x waves
  {
    std::vector<dnlg_schedule> cand;
    const int *bl = job->type == DNLG_EL_HEX8N_3D ? kHexBlocks : kQuadBlocks;
    for (int b = 0; b < kSchedBlocks; b++)
      for (int w : {default_waves(v), alt_waves(v)}) {
        dnlg_schedule s = cur;
        s.elem_block = bl[b];
        s.waves = w;
        if (schedule_valid(job->type, job->prec, job->host.law, s, job->host.stress_method)) cand.push_back(s);
      }
    if (cand.size() > 1 && (st = pick(job, cand, cycles_per_trial, rounds, &cur, &ms, &start, &start_ms))) return st;
  }
  // 2. corrector slot batch
  {
    std::vector<dnlg_schedule> cand;
    for (int b : {2, 4, 8}) {
      dnlg_schedule s = cur;
      s.corr_batch = b;
      if (schedule_valid(job->type, job->prec, job->host.law, s, job->host.stress_method)) cand.push_back(s);
    }
    if ((st = pick(job, cand, cycles_per_trial, rounds, &cur, &ms, start_ms < 0 ? &start : nullptr,
                   start_ms < 0 ? &start_ms : nullptr)))
      return st;
  }
  // 3. cycles per graph
  {
    std::vector<dnlg_schedule> cand;
    for (int g : {20, 40, 80}) {
      dnlg_schedule s = cur;
      s.graph_cycles = g;
      cand.push_back(s);
    }
    if ((st = pick(job, cand, cycles_per_trial, rounds, &cur, &ms))) return st;
  }
  set_schedule(job, cur);
  if (best) *best = cur;
  if (ms_per_cycle) {
    ms_per_cycle[0] = ms;
    ms_per_cycle[1] = start_ms;
  }
  return DNLG_OK;
}

}  // extern "C"
