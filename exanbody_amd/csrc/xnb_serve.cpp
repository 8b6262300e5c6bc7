// xnb_serve.cpp -- continuous-batching request loop over the C ABI (include/xnb_amd.h).
//
// Jobs are visible from their arrival time on.  The loop thread admits waiting jobs in
// arrival order while the budgeted device bytes and the active-job limit allow, and hands
// them to a set-up thread (xnb_create, xnb_nbh_dist, xnb_set_particles, xnb_init on the
// job's own stream), which works while the loop thread keeps advancing the batch.  Jobs whose
// set-up is done join the batch (xnb_batch_attach) between batch steps; finished jobs leave
// it (xnb_batch_detach, output copy, xnb_destroy), freeing their budget.
// Latency = completion - arrival.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/xnb_amd.h"

namespace {

using clk = std::chrono::steady_clock;

double seconds_since(clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); }

struct Active {
  int job = -1;
  int slot = -1;
  xnb_ctx* ctx = nullptr;
  size_t bytes = 0;
};

struct Prepared {
  int job = -1;
  xnb_ctx* ctx = nullptr;
  int rc = XNB_OK;
};

// Set-up thread: creates and initialises the contexts of admitted jobs, in admission order,
// and hands them back ready to attach.  Each context is touched by one thread at a time: by
// this thread until it is in `ready`, by the loop thread afterwards.
class Preparer {
 public:
  Preparer(const xnb_job* jobs, const xnb_serve_config* cfg) : jobs_(jobs), cfg_(cfg), th_([this] { work(); }) {}
  ~Preparer() { stop(); }
  void push(int job) {
    {
      std::lock_guard<std::mutex> g(m_);
      todo_.push_back(job);
    }
    cv_.notify_all();
  }
  // Takes the prepared jobs; waits for one until `deadline` when none is ready.
  std::vector<Prepared> take(bool wait, clk::time_point deadline) {
    std::unique_lock<std::mutex> g(m_);
    if (wait) cv_done_.wait_until(g, deadline, [this] { return !ready_.empty(); });
    std::vector<Prepared> out(ready_.begin(), ready_.end());
    ready_.clear();
    return out;
  }
  // Stops after the job in hand; returns the jobs never started.
  std::vector<int> stop() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    if (th_.joinable()) th_.join();
    std::vector<int> left(todo_.begin(), todo_.end());
    todo_.clear();
    return left;
  }

 private:
  void work() {
    for (;;) {
      int q;
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [this] { return stop_ || !todo_.empty(); });
        if (stop_) return;
        q = todo_.front();
        todo_.pop_front();
      }
      const xnb_job& J = jobs_[q];
      Prepared p;
      p.job = q;
      int rc = xnb_create(&p.ctx, &J.domain, cfg_->device);
      if (!rc) rc = xnb_nbh_dist(p.ctx, J.run.rcut, J.run.rcut_inc, nullptr, nullptr, nullptr);
      if (!rc) rc = xnb_set_particles(p.ctx, J.n, J.r, J.v, J.id, nullptr);
      if (!rc) rc = xnb_init(p.ctx, &J.run.lj, J.run.rcut, J.run.rcut_inc, J.run.precision);
      if (!rc && cfg_->autotune) rc = xnb_autotune(p.ctx, &J.run.lj, J.run.rcut, J.run.precision, nullptr);
      if (!rc) rc = xnb_synchronize(p.ctx);
      p.rc = rc;
      {
        std::lock_guard<std::mutex> g(m_);
        ready_.push_back(p);
      }
      cv_done_.notify_all();
    }
  }
  const xnb_job* jobs_;
  const xnb_serve_config* cfg_;
  std::mutex m_;
  std::condition_variable cv_, cv_done_;
  std::deque<int> todo_;
  std::deque<Prepared> ready_;
  bool stop_ = false;
  std::thread th_;  // last: started once the members above exist
};

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t lo = (size_t)std::floor(pos), hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
}

}  // namespace

extern "C" {

// Slots (atoms and their periodic ghosts, times the batch's slot head room of 1.3 and the
// capacity growth step): two copies of the particle state, backup, scratch and the ghost
// scheme, ~500 B per slot.  Brick lists: 4 B per neighbour within the list radius (u16
// entries in both directions, lane and wave padding, the brick head room).  Cells: the
// per-cell tables, ~64 B per cell with the ghost layers.  Plus a fixed 16 MiB for the scalar
// buffers, reduction partials and launch tables.  Checked against counted allocations from
// 300 to 200k atoms, densities 0.6 to 1.6 and list radii 2.8 to 4.4 (measured peaks 0.65-0.8
// of this from 30k atoms on).
static size_t job_bytes(double slots, double atoms, double neighbours, double cells) {
  const double b = slots * 500.0 + atoms * 4.0 * std::max(neighbours, 0.0) + cells * 64.0;
  return (size_t)b + ((size_t)16 << 20);
}

size_t xnb_job_bytes(size_t n) { return job_bytes(1.4 * (double)n, (double)n, 110.0, (double)n); }

size_t xnb_job_bytes_for(const xnb_job* J) {
  if (!J) return 0;
  const xnb_domain& d = J->domain;
  const double rl = J->run.rcut + J->run.rcut_inc;
  const double gl = d.cell_size > 0 ? std::max(1.0, std::ceil(rl / d.cell_size)) : 1.0;
  double vol = 1.0, ghost = 1.0, cells = 1.0;
  for (int a = 0; a < 3; a++) {
    const double D = std::max<double>(1.0, (double)d.grid_dims[a]);
    vol *= d.bounds_max[a] - d.bounds_min[a];
    cells *= D + 2.0 * gl;
    if (d.periodic[a]) ghost *= (D + 2.0 * gl) / D;
  }
  const double n = (double)J->n;
  const double nb = vol > 0 ? n / vol * (4.0 / 3.0) * M_PI * rl * rl * rl : 0.0;
  return job_bytes(n * ghost, n, nb, cells);
}

int xnb_serve_run(const xnb_job* jobs, int njobs, const xnb_serve_config* cfg, xnb_job_result* results,
                  xnb_serve_stats* stats) {
  if (njobs < 0 || (njobs > 0 && (!jobs || !results)) || !cfg || cfg->steps_per_iter < 1) return XNB_E_ARG;
  for (int q = 1; q < njobs; q++)
    if (jobs[q].arrival_s < jobs[q - 1].arrival_s) return XNB_E_ARG;  // arrival order
  // every result is written, also when the loop stops on an error: a job that never
  // completed keeps status XNB_E_STATE and is counted as failed
  for (int q = 0; q < njobs; q++) {
    results[q] = xnb_job_result{};
    results[q].status = XNB_E_STATE;
  }
  xnb_batch* B = nullptr;
  int e = xnb_batch_create(&B, cfg->device);
  if (e) return e;
  const size_t budget = cfg->hbm_budget_bytes ? cfg->hbm_budget_bytes : (size_t)-1;
  const int max_active = cfg->max_active > 0 ? cfg->max_active : 1 << 30;
  std::vector<Active> act;
  std::vector<size_t> need(njobs > 0 ? njobs : 0);
  int next = 0, done_jobs = 0, failed = 0, peak_active = 0, in_setup = 0;
  size_t bytes = 0, peak_bytes = 0, peak_dev = 0;
  double atom_steps = 0.0, active_sum = 0.0;
  int64_t iters = 0;
  const clk::time_point t0 = clk::now();
  auto finish = [&](int job, int status, double t) {
    results[job].finish_s = t;
    results[job].latency_s = t - jobs[job].arrival_s;
    results[job].status = status;
    if (status) failed++;
    done_jobs++;
  };
  auto ctx_peak = [](xnb_ctx* c) {
    size_t pk = 0;
    if (c) xnb_device_bytes(c, nullptr, &pk);
    return pk;
  };
  Preparer prep(jobs, cfg);
  while (done_jobs < njobs) {
    double now = seconds_since(t0);
    // admit arrived jobs in order; their set-up runs on the set-up thread
    while (next < njobs && jobs[next].arrival_s <= now && (int)act.size() + in_setup < max_active) {
      need[next] = xnb_job_bytes_for(&jobs[next]);
      // wait for memory (a lone job always runs)
      if ((!act.empty() || in_setup) && bytes + need[next] > budget) break;
      results[next] = xnb_job_result{};
      results[next].status = XNB_E_STATE;
      results[next].start_s = now;
      bytes += need[next];
      peak_bytes = std::max(peak_bytes, bytes);
      in_setup++;
      prep.push(next++);
    }
    // jobs whose set-up is done join the batch; with nothing to step, wait for one (or for
    // the next arrival)
    const clk::time_point tnow = clk::now();
    clk::time_point until = tnow + std::chrono::seconds(1);
    if (next < njobs) {  // an arrival still to come ends the wait; one already waiting for memory does not
      const clk::time_point arr =
          t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(jobs[next].arrival_s));
      if (arr > tnow) until = std::min(until, arr);
    }
    for (const Prepared& p : prep.take(act.empty() && in_setup > 0, until)) {
      in_setup--;
      int rc = p.rc;
      Active a;
      a.job = p.job;
      a.ctx = p.ctx;
      a.bytes = need[p.job];
      if (!rc) rc = xnb_batch_attach(B, a.ctx, &jobs[p.job].run, &a.slot);
      if (rc) {
        results[p.job].peak_device_bytes = ctx_peak(a.ctx);
        if (a.ctx) xnb_destroy(a.ctx);
        bytes -= a.bytes;
        finish(p.job, rc, seconds_since(t0));
      } else {
        act.push_back(a);
      }
    }
    peak_active = std::max(peak_active, (int)act.size());
    if (act.empty()) {
      if (in_setup > 0) continue;
      if (next < njobs) {  // idle until the next arrival
        const double wait = jobs[next].arrival_s - seconds_since(t0);
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        continue;
      }
      break;
    }
    int n_active = 0;
    e = xnb_batch_step(B, cfg->steps_per_iter, &n_active);
    if (e) break;
    iters++;
    active_sum += (double)act.size();
    size_t held = 0;
    for (const Active& a : act) {
      size_t cur = 0;
      xnb_device_bytes(a.ctx, &cur, nullptr);
      held += cur;
    }
    peak_dev = std::max(peak_dev, held);
    // retire finished (or failed) jobs
    for (size_t q = 0; q < act.size();) {
      int64_t steps_done = 0;
      int state = 0;
      xnb_batch_box_status(B, act[q].slot, &steps_done, &state);
      if (state == 0) {
        q++;
        continue;
      }
      const Active a = act[q];
      const xnb_job& J = jobs[a.job];
      xnb_run_params p = J.run;
      int rc = xnb_batch_detach(B, a.slot, &p);
      if (!rc && (J.r_out || J.v_out || J.id_out)) {
        std::vector<double> rr(J.r_out ? 0 : 3 * J.n), vv(J.v_out ? 0 : 3 * J.n);
        std::vector<uint64_t> ii(J.id_out ? 0 : J.n);
        std::vector<uint32_t> cells(J.n);
        rc = xnb_get_particles(a.ctx, 0, cells.data(), J.r_out ? J.r_out : rr.data(), J.v_out ? J.v_out : vv.data(), nullptr,
                               nullptr, J.id_out ? J.id_out : ii.data(), nullptr);
      }
      results[a.job].n_rebuilds = p.n_rebuilds;
      results[a.job].n_violations = p.n_violations;
      results[a.job].peak_device_bytes = ctx_peak(a.ctx);
      if (!rc) atom_steps += (double)J.n * (double)steps_done;
      xnb_destroy(a.ctx);
      bytes -= a.bytes;
      finish(a.job, rc, seconds_since(t0));
      act.erase(act.begin() + (long)q);
    }
  }
  // left over after an error: jobs in the batch, prepared or never started
  const std::vector<int> unstarted = prep.stop();
  const double wall = seconds_since(t0);
  for (const Active& a : act) {
    xnb_batch_detach(B, a.slot, nullptr);
    xnb_destroy(a.ctx);
    finish(a.job, e ? e : XNB_E_STATE, wall);
  }
  for (const Prepared& p : prep.take(false, clk::now())) {
    if (p.ctx) xnb_destroy(p.ctx);
    finish(p.job, e ? e : (p.rc ? p.rc : XNB_E_STATE), wall);
  }
  for (int q : unstarted) finish(q, e ? e : XNB_E_STATE, wall);
  for (; next < njobs; next++) finish(next, e ? e : XNB_E_STATE, wall);  // never admitted
  xnb_batch_destroy(B);
  if (stats) {
    std::vector<double> lat;
    for (int q = 0; q < njobs; q++)
      if (results[q].status == XNB_OK) lat.push_back(results[q].latency_s);
    stats->wall_s = wall;
    stats->atom_steps = atom_steps;
    stats->atom_steps_per_s = wall > 0 ? atom_steps / wall : 0.0;
    stats->p50_latency_s = percentile(lat, 0.50);
    stats->p99_latency_s = percentile(lat, 0.99);
    stats->max_latency_s = lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end());
    stats->mean_active = iters ? active_sum / (double)iters : 0.0;
    stats->completed = done_jobs - failed;
    stats->failed = failed;
    stats->peak_active = peak_active;
    stats->peak_bytes = peak_bytes;
    stats->peak_device_bytes = peak_dev;
  }
  return e;
}

}  // extern "C"
