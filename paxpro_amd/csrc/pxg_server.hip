// Serving loop of libpaxgpu: a persistent worker thread admits propagation jobs
// of mixed kinds and grid sizes by HBM budget into slots (one plan and one HIP
// stream each, default priority), advances every running job in chunks of
// z-steps so that their kernels overlap on the device, and records latency
// from submission to result-on-host.  Plans are recycled (pxg_plan_reset) so
// steady-state serving does not allocate device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/paxgpu.h"
#include "pxg_internal.hpp"

namespace {

using clk = std::chrono::steady_clock;

struct Job {
  uint64_t id = 0;
  pxg_job spec{};
  std::vector<pxg_shape> shapes;
  std::vector<pxg_part> parts;
  std::vector<pxg_module> modules;
  int state = PXG_JOB_QUEUED;
  int err = PXG_OK;
  std::string msg;
  clk::time_point t_submit, t_done;
  pxg_plan plan = nullptr;
  size_t bytes = 0;
  unsigned done = 0, in_flight = 0;
  int chunks = 0;
  bool cancel = false;
  hipEvent_t ev = nullptr;
};

struct Pooled {
  pxg_plan plan;
  pxg_problem prob;
  size_t bytes;
};

bool same_shape(const pxg_problem &a, const pxg_problem &b)
{
  return a.kind == b.kind && a.precision == b.precision && a.nx == b.nx && a.batch == b.batch &&
         a.store_dsx == b.store_dsx && a.store_dsz == b.store_dsz && (a.store_dsz == 0 || a.nz == b.nz);
}

}  // namespace

struct pxg_server_s {
  pxg_ctx ctx;
  int slots;
  size_t budget;
  unsigned chunk;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Job>> queue;
  std::vector<std::shared_ptr<Job>> running;
  std::map<uint64_t, std::shared_ptr<Job>> all;
  std::vector<Pooled> pool;
  size_t reserved = 0;  // running + pooled plan bytes
  uint64_t next_id = 1;
  bool stop = false;
  std::thread worker;
  // statistics
  std::vector<double> lat_ms;
  double cell_steps = 0.0;
  uint64_t n_sub = 0, n_done = 0, n_fail = 0, n_cancel = 0, n_rej = 0;
  double busy_s = 0.0;
  clk::time_point busy_since;
  std::vector<hipEvent_t> free_events;

  void finish(const std::shared_ptr<Job> &j, int state, int err, const std::string &msg)
  {
    j->state = state;
    j->err = err;
    j->msg = msg;
    j->t_done = clk::now();
    if (state == PXG_JOB_DONE) {
      lat_ms.push_back(std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count());
      const double cells = j->spec.prob.kind == PXG_FFT3D || j->spec.prob.kind == PXG_ADI
                               ? static_cast<double>(j->spec.prob.nx) * j->spec.prob.nx
                               : static_cast<double>(j->spec.prob.nx);
      cell_steps += cells * j->spec.nsteps;
      ++n_done;
    } else if (state == PXG_JOB_CANCELLED) {
      ++n_cancel;
    } else if (err == PXG_ERR_BUDGET) {
      ++n_rej;
    } else {
      ++n_fail;
    }
    if (j->plan) release_plan(j);
    if (j->ev) {
      free_events.push_back(j->ev);
      j->ev = nullptr;
    }
    cv_done.notify_all();
  }

  void release_plan(const std::shared_ptr<Job> &j)
  {
    (void)pxg_sync(j->plan);
    if (static_cast<int>(pool.size()) < slots) {
      pool.push_back({j->plan, j->spec.prob, j->bytes});
    } else {
      pxg_plan_destroy(j->plan);
      reserved -= j->bytes;
    }
    j->plan = nullptr;
  }

  // A plan for the job: recycle a pooled plan of the same shape, else create
  // one (evicting idle pooled plans when the budget needs room).
  bool acquire_plan(const std::shared_ptr<Job> &j)
  {
    const pxg_problem &p = j->spec.prob;
    for (size_t i = 0; i < pool.size(); ++i) {
      if (same_shape(pool[i].prob, p)) {
        j->plan = pool[i].plan;
        j->bytes = pool[i].bytes;
        pool.erase(pool.begin() + i);
        if (pxg_plan_reset(j->plan, &p) != PXG_OK) return false;
        return true;
      }
    }
    size_t need = 0;
    if (pxg_plan_bytes(&p, &need) != PXG_OK) return false;
    while (reserved + need > budget && !pool.empty()) {
      pxg_plan_destroy(pool.front().plan);
      reserved -= pool.front().bytes;
      pool.erase(pool.begin());
    }
    if (reserved + need > budget) return false;
    if (pxg_plan_create(ctx, &p, &j->plan) != PXG_OK) return false;
    j->bytes = need;
    reserved += need;
    return true;
  }

  int setup(const std::shared_ptr<Job> &j)
  {
    const pxg_job &s = j->spec;
    int rc = PXG_OK;
    if (s.material == PXG_MAT_CSG)
      rc = pxg_set_material_csg(j->plan, j->shapes.data(), static_cast<int>(j->shapes.size()), j->parts.data(),
                                static_cast<int>(j->parts.size()), j->modules.data(),
                                static_cast<int>(j->modules.size()), 0);
    else if (s.material == PXG_MAT_CURVED_WG)
      rc = pxg_set_material_curved_wg(j->plan, &s.wg);
    else
      rc = pxg_set_material_none(j->plan);
    // work-group shapes: timed on the first plan of each shape, cached in the
    // context for the next ones (zeroes the field, set just below)
    if (rc == PXG_OK && s.prob.kind == PXG_FFT3D) rc = pxg_plan_autotune(j->plan, 3);
    if (rc == PXG_OK) rc = pxg_set_field(j->plan, 0, s.field);
    return rc;
  }

  void loop()
  {
    (void)hipSetDevice(pxg::ctx_device(ctx));
    std::unique_lock<std::mutex> lk(mu);
    while (!stop) {
      bool progressed = false;
      // admission: FIFO with backfill of jobs that fit
      for (auto it = queue.begin(); it != queue.end() && static_cast<int>(running.size()) < slots;) {
        std::shared_ptr<Job> j = *it;
        if (j->cancel) {
          it = queue.erase(it);
          finish(j, PXG_JOB_CANCELLED, PXG_ERR_CANCELLED, "cancelled while queued");
          continue;
        }
        size_t need = 0;
        const int brc = pxg_plan_bytes(&j->spec.prob, &need);
        if (brc != PXG_OK || need > budget) {
          it = queue.erase(it);
          finish(j, PXG_JOB_FAILED, brc != PXG_OK ? brc : PXG_ERR_BUDGET,
                 brc != PXG_OK ? pxg_last_error() : "job larger than the HBM budget");
          continue;
        }
        if (!acquire_plan(j)) {
          ++it;  // does not fit now; later jobs may
          continue;
        }
        const int rc = setup(j);
        it = queue.erase(it);
        if (rc != PXG_OK) {
          finish(j, PXG_JOB_FAILED, rc, pxg_last_error());
          continue;
        }
        if (running.empty()) busy_since = clk::now();
        j->state = PXG_JOB_RUNNING;
        if (free_events.empty()) {
          hipEvent_t e;
          (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
          free_events.push_back(e);
        }
        j->ev = free_events.back();
        free_events.pop_back();
        running.push_back(j);
        progressed = true;
      }
      // progress of running jobs
      for (size_t i = 0; i < running.size();) {
        std::shared_ptr<Job> j = running[i];
        if (j->in_flight) {
          if (hipEventQuery(j->ev) != hipSuccess) {
            ++i;
            continue;
          }
          j->done += j->in_flight;
          j->in_flight = 0;
          ++j->chunks;
          progressed = true;
        }
        if (j->cancel) {
          running.erase(running.begin() + i);
          finish(j, PXG_JOB_CANCELLED, PXG_ERR_CANCELLED, "cancelled");
          continue;
        }
        if (j->spec.inject_failure_after_chunks >= 0 && j->chunks >= j->spec.inject_failure_after_chunks) {
          running.erase(running.begin() + i);
          finish(j, PXG_JOB_FAILED, PXG_ERR_INJECTED, "failure injected by the host");
          continue;
        }
        if (j->done >= j->spec.nsteps) {
          const int rc = pxg_get_field(j->plan, 0, j->spec.result);
          running.erase(running.begin() + i);
          finish(j, rc == PXG_OK ? PXG_JOB_DONE : PXG_JOB_FAILED, rc, rc == PXG_OK ? "" : pxg_last_error());
          continue;
        }
        const unsigned c = std::min(chunk, j->spec.nsteps - j->done);
        int rc = pxg_step(j->plan, j->spec.first_step + j->done, c);
        if (rc == PXG_OK) rc = hipEventRecord(j->ev, pxg::plan_stream(j->plan)) == hipSuccess ? PXG_OK : PXG_ERR_DEVICE;
        if (rc != PXG_OK) {
          running.erase(running.begin() + i);
          finish(j, PXG_JOB_FAILED, rc, pxg_last_error());
          continue;
        }
        j->in_flight = c;
        progressed = true;
        ++i;
      }
      if (running.empty() && busy_since != clk::time_point()) {
        busy_s += std::chrono::duration<double>(clk::now() - busy_since).count();
        busy_since = clk::time_point();
      }
      if (!progressed) {
        if (running.empty() && queue.empty())
          cv_work.wait(lk, [&] { return stop || !queue.empty(); });
        else {
          lk.unlock();
          std::this_thread::sleep_for(std::chrono::microseconds(20));
          lk.lock();
        }
      }
    }
    for (auto &j : running) finish(j, PXG_JOB_CANCELLED, PXG_ERR_CANCELLED, "server stopped");
    running.clear();
    for (auto &j : queue) finish(j, PXG_JOB_CANCELLED, PXG_ERR_CANCELLED, "server stopped");
    queue.clear();
    for (auto &p : pool) pxg_plan_destroy(p.plan);
    pool.clear();
    for (auto e : free_events) (void)hipEventDestroy(e);
    free_events.clear();
  }
};

extern "C" {

int pxg_server_create(pxg_ctx ctx, int slots, size_t hbm_budget, unsigned chunk_steps, pxg_server *out)
{
  if (!ctx || !out || slots < 1 || chunk_steps < 1) return PXG_ERR_INVALID;
  auto *s = new pxg_server_s;
  s->ctx = ctx;
  s->slots = slots;
  s->budget = hbm_budget ? hbm_budget : static_cast<size_t>(64) << 30;
  s->chunk = chunk_steps;
  s->worker = std::thread([s] { s->loop(); });
  *out = s;
  return PXG_OK;
}

int pxg_server_submit(pxg_server s, const pxg_job *job, uint64_t *id)
{
  if (!s || !job || !id || !job->field || !job->result || job->nsteps < 1 || job->first_step < 1 ||
      job->prob.batch != 1)
    return PXG_ERR_INVALID;
  auto j = std::make_shared<Job>();
  j->spec = *job;
  if (job->material == PXG_MAT_CSG) {
    if (job->n_shapes < 0 || job->n_parts < 0 || job->n_modules < 0) return PXG_ERR_INVALID;
    j->shapes.assign(job->shapes, job->shapes + job->n_shapes);
    j->parts.assign(job->parts, job->parts + job->n_parts);
    j->modules.assign(job->modules, job->modules + job->n_modules);
  }
  j->t_submit = clk::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    j->id = s->next_id++;
    s->queue.push_back(j);
    s->all[j->id] = j;
    ++s->n_sub;
  }
  s->cv_work.notify_one();
  *id = j->id;
  return PXG_OK;
}

int pxg_server_status(pxg_server s, uint64_t id, int *state)
{
  if (!s || !state) return PXG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return PXG_ERR_INVALID;
  *state = it->second->state;
  return PXG_OK;
}

int pxg_server_wait(pxg_server s, uint64_t id, double *latency_ms, double timeout_s)
{
  if (!s) return PXG_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return PXG_ERR_INVALID;
  std::shared_ptr<Job> j = it->second;
  auto pred = [&] { return j->state >= PXG_JOB_DONE; };
  if (timeout_s > 0) {
    if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return PXG_ERR_STATE;
  } else {
    s->cv_done.wait(lk, pred);
  }
  if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
  s->all.erase(id);
  return j->state == PXG_JOB_DONE ? PXG_OK : j->err;
}

int pxg_server_cancel(pxg_server s, uint64_t id)
{
  if (!s) return PXG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return PXG_ERR_INVALID;
  it->second->cancel = true;
  s->cv_work.notify_one();
  return PXG_OK;
}

int pxg_server_stats_get(pxg_server s, pxg_server_stats *st)
{
  if (!s || !st) return PXG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(st, 0, sizeof(*st));
  st->submitted = s->n_sub;
  st->completed = s->n_done;
  st->failed = s->n_fail;
  st->cancelled = s->n_cancel;
  st->rejected = s->n_rej;
  st->running = s->running.size();
  st->queued = s->queue.size();
  std::vector<double> v = s->lat_ms;
  if (!v.empty()) {
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) { return v[std::min(v.size() - 1, static_cast<size_t>(q * (v.size() - 1) + 0.5))]; };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    double sum = 0;
    for (double x : v) sum += x;
    st->mean_ms = sum / v.size();
    st->max_ms = v.back();
  }
  st->cell_steps = s->cell_steps;
  st->busy_s = s->busy_s + (s->busy_since != clk::time_point()
                                ? std::chrono::duration<double>(clk::now() - s->busy_since).count()
                                : 0.0);
  st->hbm_reserved = s->reserved;
  st->hbm_budget = s->budget;
  return PXG_OK;
}

int pxg_server_destroy(pxg_server s)
{
  if (!s) return PXG_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->worker.join();
  delete s;
  return PXG_OK;
}

}  // extern "C"
