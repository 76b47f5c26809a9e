// Serving loop of libpaxgpu: a persistent worker thread admits propagation jobs
// of mixed kinds and grid sizes by HBM budget into slots (one plan and one HIP
// stream each, default priority), advances every running job in chunks of
// z-steps so that their kernels overlap on the device, and records latency
// from submission to result-on-host.  Plans are recycled (pxg_plan_reset) so
// steady-state serving does not allocate device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
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

// Jobs whose chunks can share launches (pxg_step_group): FFT3D, batch 1, no
// solution storage, a power-of-two line length.
bool groupable(const pxg_problem &p)
{
  return p.kind == PXG_FFT3D && p.batch == 1 && p.store_dsz == 0 && p.nx >= 16 && p.nx <= 4096 &&
         (p.nx & (p.nx - 1)) == 0;
}

bool same_group(const pxg_problem &a, const pxg_problem &b) { return a.nx == b.nx && a.precision == b.precision; }

}  // namespace

struct pxg_server_s {
  pxg_ctx ctx;
  int slots;
  size_t budget;
  unsigned chunk;
  // mu guards queue, all, the job fields that other threads read (state, err,
  // msg, t_done, cancel) and the statistics.  running, pool and the plans are
  // the worker's alone; the worker takes mu only to move jobs and record results.
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Job>> queue;
  std::map<uint64_t, std::shared_ptr<Job>> all;
  std::vector<std::shared_ptr<Job>> running;
  std::vector<Pooled> pool;
  std::atomic<size_t> reserved{0};  // running + pooled plan bytes
  std::atomic<uint64_t> n_running{0};
  std::atomic<int> max_group{pxg::MAX_GROUP};  // 1: every job's chunks run alone
  // largest nx that groups (0: any).  Measured on MI355X, 64-sphere phantoms,
  // mixed precision, chunks of 64 steps: shared launches serve 1.45x the
  // Mcell*steps/s of chunks alone at 256^2 and 1.38x at 512^2, but 0.95x at 1024^2,
  // whose single-request passes already fill the device.
  std::atomic<int> group_max_nx{512};
  std::atomic<uint64_t> n_chunks{0}, n_grouped{0};  // chunks issued, of them in groups of >= 2
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

  // Records the end of a job (called with mu held).
  void finish_locked(const std::shared_ptr<Job> &j, int state, int err, const std::string &msg)
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
    cv_done.notify_all();
  }

  // Worker only, without mu: returns the job's plan to the pool (or frees it)
  // and its event, then records the end of the job.
  void finish(const std::shared_ptr<Job> &j, int state, int err, const std::string &msg)
  {
    if (j->plan) release_plan(j);
    if (j->ev) {
      free_events.push_back(j->ev);
      j->ev = nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    finish_locked(j, state, err, msg);
  }

  void drop_plan(pxg_plan plan, size_t bytes)
  {
    pxg_plan_destroy(plan);
    reserved -= bytes;
  }

  void release_plan(const std::shared_ptr<Job> &j)
  {
    if (pxg_sync(j->plan) == PXG_OK && static_cast<int>(pool.size()) < slots)
      pool.push_back({j->plan, j->spec.prob, j->bytes});
    else
      drop_plan(j->plan, j->bytes);
    j->plan = nullptr;
  }

  size_t pooled_bytes() const
  {
    size_t b = 0;
    for (const Pooled &p : pool) b += p.bytes;
    return b;
  }

  bool pooled_shape(const pxg_problem &p) const
  {
    for (const Pooled &q : pool)
      if (same_shape(q.prob, p)) return true;
    return false;
  }

  enum Acquire { ACQ_OK, ACQ_NO_ROOM, ACQ_ERROR };

  // A plan for the job: recycle a pooled plan of the same shape, else create
  // one (evicting idle pooled plans when the budget needs room).  ACQ_NO_ROOM:
  // the job waits for memory; ACQ_ERROR: *rc and *msg say why it cannot run
  // (a failed recycle has destroyed the pooled plan and released its bytes).
  Acquire acquire_plan(const std::shared_ptr<Job> &j, int *rc, std::string *msg)
  {
    const pxg_problem &p = j->spec.prob;
    for (size_t i = 0; i < pool.size(); ++i) {
      if (!same_shape(pool[i].prob, p)) continue;
      const Pooled got = pool[i];
      pool.erase(pool.begin() + static_cast<long>(i));
      const int r = j->spec.inject_plan_failure == PXG_INJECT_RESET ? PXG_ERR_INJECTED
                                                                    : pxg_plan_reset(got.plan, &p);
      if (r != PXG_OK) {
        *msg = r == PXG_ERR_INJECTED ? "plan recycling failure injected by the host" : pxg_last_error();
        *rc = r;
        drop_plan(got.plan, got.bytes);
        return ACQ_ERROR;
      }
      j->plan = got.plan;
      j->bytes = got.bytes;
      return ACQ_OK;
    }
    size_t need = 0;
    if ((*rc = pxg_plan_bytes(&p, &need)) != PXG_OK) {
      *msg = pxg_last_error();
      return ACQ_ERROR;
    }
    while (reserved + need > budget && !pool.empty()) {
      drop_plan(pool.front().plan, pool.front().bytes);
      pool.erase(pool.begin());
    }
    if (reserved + need > budget) return ACQ_NO_ROOM;
    if (j->spec.inject_plan_failure == PXG_INJECT_CREATE) {
      *rc = PXG_ERR_INJECTED;
      *msg = "plan creation failure injected by the host";
      return ACQ_ERROR;
    }
    const int r = pxg_plan_create(ctx, &p, &j->plan);
    if (r != PXG_OK) {
      j->plan = nullptr;
      // the context's own budget is shared with plans outside this server:
      // wait while jobs of ours still hold memory, else the job cannot run
      if (r == PXG_ERR_BUDGET && !running.empty()) return ACQ_NO_ROOM;
      *rc = r;
      *msg = pxg_last_error();
      return ACQ_ERROR;
    }
    j->bytes = need;
    reserved += need;
    return ACQ_OK;
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
    if (rc == PXG_OK && (s.prob.kind == PXG_FFT3D || s.prob.kind == PXG_ADI)) rc = pxg_plan_autotune(j->plan, 3);
    if (rc == PXG_OK) rc = pxg_set_field(j->plan, 0, s.field);
    return rc;
  }

  // Next queued job to admit (FIFO with backfill): the first one that is
  // cancelled, cannot ever fit (both finished here), or fits now, i.e. a
  // pooled plan of its shape exists or its bytes fit once idle pooled plans are
  // evicted.  Called with mu held; the job is taken out of the queue.
  std::shared_ptr<Job> next_admission_locked()
  {
    const size_t evictable = pooled_bytes();
    for (auto it = queue.begin(); it != queue.end();) {
      std::shared_ptr<Job> j = *it;
      if (j->cancel) {
        it = queue.erase(it);
        finish_locked(j, PXG_JOB_CANCELLED, PXG_ERR_CANCELLED, "cancelled while queued");
        continue;
      }
      size_t need = 0;
      const int brc = pxg_plan_bytes(&j->spec.prob, &need);
      if (brc != PXG_OK || need > budget) {
        it = queue.erase(it);
        finish_locked(j, PXG_JOB_FAILED, brc != PXG_OK ? brc : PXG_ERR_BUDGET,
                      brc != PXG_OK ? pxg_last_error() : "job larger than the HBM budget");
        continue;
      }
      if (pooled_shape(j->spec.prob) || reserved - evictable + need <= budget) {
        queue.erase(it);
        return j;
      }
      ++it;
    }
    return nullptr;
  }

  // Admits queued jobs into free slots; plan work runs without mu.
  bool admit()
  {
    bool progressed = false;
    while (static_cast<int>(running.size()) < slots) {
      std::shared_ptr<Job> j;
      {
        std::lock_guard<std::mutex> lk(mu);
        j = next_admission_locked();
      }
      if (!j) break;
      int rc = PXG_OK;
      std::string msg;
      const Acquire a = acquire_plan(j, &rc, &msg);
      if (a == ACQ_NO_ROOM) {
        std::lock_guard<std::mutex> lk(mu);
        queue.push_front(j);  // keeps its turn; waits for running jobs to free memory
        break;
      }
      if (a == ACQ_OK && (rc = setup(j)) != PXG_OK) msg = pxg_last_error();
      if (rc != PXG_OK) {
        finish(j, PXG_JOB_FAILED, rc, msg);
        progressed = true;
        continue;
      }
      if (free_events.empty()) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
          finish(j, PXG_JOB_FAILED, PXG_ERR_DEVICE, "hipEventCreate failed");
          continue;
        }
        free_events.push_back(e);
      }
      j->ev = free_events.back();
      free_events.pop_back();
      const bool first = running.empty();
      running.push_back(j);
      ++n_running;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (first) busy_since = clk::now();
        j->state = PXG_JOB_RUNNING;
      }
      progressed = true;
    }
    return progressed;
  }

  void retire(std::shared_ptr<Job> j, int state, int err, const std::string &msg)
  {
    running.erase(std::find(running.begin(), running.end(), j));
    --n_running;
    finish(j, state, err, msg);
  }

  unsigned next_chunk(const Job &j) const { return std::min(chunk, j.spec.nsteps - j.done); }

  // A running job of j's group whose chunk is in flight and whose next chunk
  // will be as long as j's: j waits for it rather than starting a chunk alone.
  bool partner_in_flight(const Job &j) const
  {
    const unsigned c = next_chunk(j);
    for (const auto &r : running) {
      if (r.get() == &j || !r->in_flight || !groupable(r->spec.prob) || !same_group(r->spec.prob, j.spec.prob))
        continue;
      const unsigned left = r->spec.nsteps - r->done - r->in_flight;
      if (std::min(chunk, left) == c) return true;
    }
    return false;
  }

  // Enqueues the next chunk of the jobs in g (one job, or a group of jobs with
  // equal chunk lengths) and records each job's event on its own stream.
  void issue(const std::vector<std::shared_ptr<Job>> &g)
  {
    const unsigned c = next_chunk(*g[0]);
    int rc;
    if (g.size() == 1) {
      rc = pxg_step(g[0]->plan, g[0]->spec.first_step + g[0]->done, c);
    } else {
      pxg_plan plans[pxg::MAX_GROUP];
      unsigned first[pxg::MAX_GROUP];
      for (size_t i = 0; i < g.size(); ++i) {
        plans[i] = g[i]->plan;
        first[i] = g[i]->spec.first_step + g[i]->done;
      }
      rc = pxg_step_group(plans, static_cast<int>(g.size()), first, c);
      n_grouped += g.size();
    }
    n_chunks += g.size();
    const std::string msg = rc == PXG_OK ? std::string() : std::string(pxg_last_error());
    for (const auto &j : g) {
      int r = rc;
      if (r == PXG_OK)
        r = hipEventRecord(j->ev, pxg::plan_stream(j->plan)) == hipSuccess ? PXG_OK : PXG_ERR_DEVICE;
      if (r != PXG_OK) {
        retire(j, PXG_JOB_FAILED, r, rc != PXG_OK ? msg : "hipEventRecord failed");
        continue;
      }
      j->in_flight = c;
    }
  }

  // Advances running jobs by one chunk each; finished jobs copy their result
  // out and leave.  Jobs that are ready together, share a group shape and have
  // chunks of equal length run in shared launches (at most max_group per
  // launch); a ready job of such a shape whose partner is in flight waits for
  // it.  Runs without mu except to read cancellation flags.
  bool advance()
  {
    bool progressed = false;
    std::vector<std::shared_ptr<Job>> ready;
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
      bool cancel;
      {
        std::lock_guard<std::mutex> lk(mu);
        cancel = j->cancel;
      }
      if (cancel) {
        retire(j, PXG_JOB_CANCELLED, PXG_ERR_CANCELLED, "cancelled");
        continue;
      }
      if (j->spec.inject_failure_after_chunks >= 0 && j->chunks >= j->spec.inject_failure_after_chunks) {
        retire(j, PXG_JOB_FAILED, PXG_ERR_INJECTED, "failure injected by the host");
        continue;
      }
      if (j->done >= j->spec.nsteps) {
        const int rc = pxg_get_field(j->plan, 0, j->spec.result);
        retire(j, rc == PXG_OK ? PXG_JOB_DONE : PXG_JOB_FAILED, rc, rc == PXG_OK ? "" : pxg_last_error());
        continue;
      }
      ready.push_back(j);
      ++i;
    }
    const size_t gmax = static_cast<size_t>(std::max(1, std::min<int>(max_group, pxg::MAX_GROUP)));
    std::vector<bool> taken(ready.size(), false);
    for (size_t a = 0; a < ready.size(); ++a) {
      if (taken[a]) continue;
      const std::shared_ptr<Job> &j = ready[a];
      std::vector<std::shared_ptr<Job>> g{j};
      taken[a] = true;
      const int gnx = group_max_nx;
      if (gmax > 1 && groupable(j->spec.prob) && (gnx == 0 || j->spec.prob.nx <= gnx)) {
        const unsigned c = next_chunk(*j);
        for (size_t b = a + 1; b < ready.size() && g.size() < gmax; ++b)
          if (!taken[b] && groupable(ready[b]->spec.prob) && same_group(j->spec.prob, ready[b]->spec.prob) &&
              next_chunk(*ready[b]) == c) {
            g.push_back(ready[b]);
            taken[b] = true;
          }
        if (g.size() == 1 && partner_in_flight(*j)) continue;
      }
      issue(g);
      progressed = true;
    }
    if (running.empty() && busy_since != clk::time_point()) {
      std::lock_guard<std::mutex> lk(mu);
      busy_s += std::chrono::duration<double>(clk::now() - busy_since).count();
      busy_since = clk::time_point();
    }
    return progressed;
  }

  void loop()
  {
    (void)hipSetDevice(pxg::ctx_device(ctx));
    for (;;) {
      {
        std::lock_guard<std::mutex> lk(mu);
        if (stop) break;
      }
      bool progressed = admit();
      progressed = advance() || progressed;
      if (!progressed) {
        std::unique_lock<std::mutex> lk(mu);
        if (running.empty())
          cv_work.wait(lk, [&] { return stop || !queue.empty(); });
        else
          cv_work.wait_for(lk, std::chrono::microseconds(20), [&] { return stop; });
      }
    }
    while (!running.empty()) retire(running.front(), PXG_JOB_CANCELLED, PXG_ERR_CANCELLED, "server stopped");
    std::deque<std::shared_ptr<Job>> left;
    {
      std::lock_guard<std::mutex> lk(mu);
      left.swap(queue);
      for (auto &j : left) finish_locked(j, PXG_JOB_CANCELLED, PXG_ERR_CANCELLED, "server stopped");
    }
    for (auto &p : pool) drop_plan(p.plan, p.bytes);
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
  st->running = s->n_running.load();
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
  st->chunks = s->n_chunks;
  st->grouped_chunks = s->n_grouped;
  return PXG_OK;
}

int pxg_server_set_grouping(pxg_server s, int max_group, int max_nx)
{
  if (!s || max_group < 1 || max_group > pxg::MAX_GROUP || max_nx < 0) return PXG_ERR_INVALID;
  s->max_group = max_group;
  s->group_max_nx = max_nx;
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
