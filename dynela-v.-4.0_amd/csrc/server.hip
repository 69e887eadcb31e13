// server.hip -- continuous-batching job server over one context.
//
// The reference runs one model per process (DynELA.C:21,47-52); a serving
// deployment instead keeps a queue of submitted models and advances every
// admitted job together, a few cycles at a time, so that jobs join and leave
// the running batch between steps.  Admission is by device-memory footprint
// (the same figure dnlg_footprint_bytes reports): a queued job is created on
// the device only when its footprint fits in the budget that running and
// finished-but-not-taken jobs leave free.  Queued jobs are admitted in
// submission order; a job that does not fit does not block smaller ones
// behind it.  Each job's arithmetic does not depend on what else runs
// beside it (per-job task lists, per-job reductions), so a job's results are
// the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dnlg.h"
#include "batch.h"

using namespace dnlg;

namespace {

enum { kQueued = 0, kRunning = 1, kDone = 2 };

struct ServerJob {
  int64_t ticket = 0;
  int state = kQueued;
  int status = DNLG_OK;
  int64_t footprint = 0;
  dnlg_job *job = nullptr;
  // descriptor copies, kept until admission
  dnlg_mesh mesh{};
  std::vector<double> coords;
  std::vector<int32_t> conn;
  dnlg_material mat{};
  std::vector<double> table;  // tabular law points
  std::vector<dnlg_boundary> bnd;
  std::vector<int64_t> iptr, cptr;
  std::vector<int32_t> iidx, cidx;
  dnlg_solver_params prm{};
  std::string err;

  void release_descriptors() {
    std::vector<double>().swap(coords);
    std::vector<int32_t>().swap(conn);
    std::vector<dnlg_boundary>().swap(bnd);
    std::vector<double>().swap(table);
    std::vector<int64_t>().swap(iptr);
    std::vector<int64_t>().swap(cptr);
    std::vector<int32_t>().swap(iidx);
    std::vector<int32_t>().swap(cidx);
  }
};

}  // namespace

struct dnlg_server {
  dnlg_ctx *ctx = nullptr;
  int64_t budget = 0;
  int max_jobs = 0;
  int64_t in_use = 0;  // footprints of running and finished-but-not-taken jobs
  int64_t cycle_budget = 0;  // > 0: a job still running after this many increments is stopped
  int64_t next_ticket = 1;
  std::deque<ServerJob *> queue;
  std::vector<ServerJob *> running;
  std::map<int64_t, std::unique_ptr<ServerJob>> all;
};

namespace {

int srv_fail(dnlg_server *s, int st, const char *msg) {
  if (s && s->ctx) s->ctx->err = msg;
  return st;
}

// Create the device job of an admitted entry from its descriptor copies.
int admit(dnlg_server *s, ServerJob *j) {
  dnlg_mesh mesh = j->mesh;
  mesh.coords = j->coords.data();
  mesh.conn = j->conn.data();
  dnlg_bcs bcs{};
  bcs.n_boundaries = (int32_t)j->bnd.size();
  bcs.boundaries = j->bnd.empty() ? nullptr : j->bnd.data();
  bcs.initial_ptr = j->iptr.empty() ? nullptr : j->iptr.data();
  bcs.initial_idx = j->iidx.empty() ? nullptr : j->iidx.data();
  bcs.constant_ptr = j->cptr.empty() ? nullptr : j->cptr.data();
  bcs.constant_idx = j->cidx.empty() ? nullptr : j->cidx.data();
  dnlg_job *job = nullptr;
  const int st = dnlg_job_create(s->ctx, &mesh, &j->mat, &bcs, &j->prm, &job);
  j->release_descriptors();
  if (st != DNLG_OK) {
    j->state = kDone;
    j->status = st;
    j->err = s->ctx->err;
    return st;
  }
  j->job = job;
  j->state = kRunning;
  s->in_use += j->footprint;
  s->running.push_back(j);
  return DNLG_OK;
}

}  // namespace

extern "C" {

int dnlg_server_create(dnlg_ctx *ctx, int64_t hbm_budget, int max_jobs, dnlg_server **out) {
  if (!ctx || !out || max_jobs < 0) return DNLG_E_INVALID;
  *out = nullptr;
  if (hbm_budget <= 0) {
    if (hipSetDevice(ctx->device) != hipSuccess) return DNLG_E_HIP;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return DNLG_E_HIP;
    hbm_budget = (int64_t)(0.85 * (double)free_b);
  }
  dnlg_server *s = new dnlg_server();
  s->ctx = ctx;
  s->budget = hbm_budget;
  s->max_jobs = max_jobs > 0 ? max_jobs : 1024;
  *out = s;
  return DNLG_OK;
}

int dnlg_server_destroy(dnlg_server *s) {
  if (!s) return DNLG_E_INVALID;
  for (auto &kv : s->all)
    if (kv.second->job) dnlg_job_destroy(kv.second->job);
  delete s;
  return DNLG_OK;
}

int dnlg_server_submit(dnlg_server *s, const dnlg_mesh *mesh, const dnlg_material *mat, const dnlg_bcs *bcs,
                       const dnlg_solver_params *prm, int64_t *ticket) {
  if (!s || !mesh || !mat || !prm || !ticket) return srv_fail(s, DNLG_E_INVALID, "null argument");
  if (mesh->n_nodes <= 0 || mesh->n_elems <= 0 || !mesh->coords || !mesh->conn)
    return srv_fail(s, DNLG_E_INVALID, "empty mesh");
  const int64_t fp = dnlg_model_footprint_bytes(mesh, mat, bcs, prm);
  if (fp < 0) return srv_fail(s, DNLG_E_INVALID, "unsupported element type");
  if (fp > s->budget) return srv_fail(s, DNLG_E_CAPACITY, "job footprint exceeds the server's memory budget");
  std::unique_ptr<ServerJob> j(new ServerJob());
  j->ticket = s->next_ticket++;
  j->footprint = fp;
  j->mesh = *mesh;
  const int dim = type_dim(mesh->elem_type);
  const int nn = type_nodes(mesh->elem_type);
  j->coords.assign(mesh->coords, mesh->coords + (size_t)dim * mesh->n_nodes);
  j->conn.assign(mesh->conn, mesh->conn + (size_t)nn * mesh->n_elems);
  j->mat = *mat;
  if (mat->law == DNLG_LAW_TABULAR || mat->law == DNLG_LAW_ANN2SIG) {
    // the job keeps its own copy of the table (point pairs or the packed
    // network); an oversized one is not kept, and job creation then
    // rejects the law
    const int64_t n = mat->law == DNLG_LAW_TABULAR ? 2 * mat->n_table : mat->n_table;
    if (mat->table && n > 0 && n <= (int64_t)2 << 20) {
      j->table.assign(mat->table, mat->table + n);
      j->mat.table = j->table.data();
    } else {
      j->mat.table = nullptr;
    }
  } else {
    j->mat.table = nullptr;
  }
  j->prm = *prm;
  if (bcs && bcs->n_boundaries > 0) {
    if (!bcs->boundaries) return srv_fail(s, DNLG_E_INVALID, "boundaries missing");
    j->bnd.assign(bcs->boundaries, bcs->boundaries + bcs->n_boundaries);
    const int64_t N = mesh->n_nodes;
    if (bcs->initial_ptr) {
      j->iptr.assign(bcs->initial_ptr, bcs->initial_ptr + N + 1);
      if (j->iptr[N] > 0) {
        if (!bcs->initial_idx) return srv_fail(s, DNLG_E_INVALID, "malformed boundary lists");
        j->iidx.assign(bcs->initial_idx, bcs->initial_idx + j->iptr[N]);
      }
    }
    if (bcs->constant_ptr) {
      j->cptr.assign(bcs->constant_ptr, bcs->constant_ptr + N + 1);
      if (j->cptr[N] > 0) {
        if (!bcs->constant_idx) return srv_fail(s, DNLG_E_INVALID, "malformed boundary lists");
        j->cidx.assign(bcs->constant_idx, bcs->constant_idx + j->cptr[N]);
      }
    }
  }
  *ticket = j->ticket;
  s->queue.push_back(j.get());
  s->all[j->ticket] = std::move(j);
  return DNLG_OK;
}

int dnlg_server_step(dnlg_server *s, int max_cycles, int *n_running, int *n_queued) {
  if (!s || max_cycles <= 0) return DNLG_E_INVALID;
  // admission, in submission order, by footprint
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    ServerJob *j = *it;
    if ((int)s->running.size() >= s->max_jobs) break;
    if (j->footprint <= s->budget - s->in_use) {
      admit(s, j);  // a failed creation ends the entry with its status
      it = s->queue.erase(it);
    } else {
      ++it;
    }
  }
  int st = DNLG_OK;
  if (!s->running.empty()) {
    std::vector<dnlg_job *> jobs;
    std::vector<double> up;
    for (ServerJob *j : s->running) {
      jobs.push_back(j->job);
      up.push_back(j->prm.t_end);
    }
    if (hipSetDevice(s->ctx->device) != hipSuccess) return srv_fail(s, DNLG_E_HIP, "hipSetDevice");
    st = batch_run(s->ctx, jobs, up, max_cycles);
    bool job_failed = false;
    for (dnlg_job *jb : jobs) job_failed = job_failed || jb->status != 0;
    if (st != DNLG_OK && !job_failed) return st;  // failure of the launch sequence itself
    st = DNLG_OK;
    // retire finished and failed jobs
    std::vector<ServerJob *> keep;
    for (ServerJob *j : s->running) {
      if (j->job->status != 0 || !(j->job->t_now < j->job->up_to)) {
        j->state = kDone;
        j->status = j->job->status;
      } else if (s->cycle_budget > 0 && j->job->inc_now >= s->cycle_budget) {
        j->state = kDone;
        j->status = DNLG_E_CAPACITY;
        j->err = "cycle budget exhausted before t_end";
      } else {
        keep.push_back(j);
      }
    }
    s->running.swap(keep);
  }
  if (n_running) *n_running = (int)s->running.size();
  if (n_queued) *n_queued = (int)s->queue.size();
  return st;
}

int dnlg_server_set_cycle_budget(dnlg_server *s, int64_t cycles) {
  if (!s || cycles < 0) return DNLG_E_INVALID;
  s->cycle_budget = cycles;
  return DNLG_OK;
}

int dnlg_server_poll(dnlg_server *s, int64_t ticket, int *state, int *status) {
  if (!s) return DNLG_E_INVALID;
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return srv_fail(s, DNLG_E_INVALID, "unknown ticket");
  if (state) *state = it->second->state;
  if (status) *status = it->second->status;
  return DNLG_OK;
}

int dnlg_server_take(dnlg_server *s, int64_t ticket, dnlg_job **job) {
  if (!s || !job) return DNLG_E_INVALID;
  *job = nullptr;
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return srv_fail(s, DNLG_E_INVALID, "unknown ticket");
  ServerJob *j = it->second.get();
  if (j->state != kDone) return srv_fail(s, DNLG_E_STATE, "job not finished");
  const int status = j->status;
  if (j->job) {
    s->in_use -= j->footprint;
    *job = j->job;
    j->job = nullptr;
  }
  if (!j->err.empty()) s->ctx->err = j->err;
  s->all.erase(it);
  return status;
}

}  // extern "C"
