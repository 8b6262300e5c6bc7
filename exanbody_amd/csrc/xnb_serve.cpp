// xnb_serve.cpp -- continuous-batching request loop over the C ABI (include/xnb_amd.h).
//
// Jobs are visible from their arrival time on.  Each loop iteration admits waiting jobs in
// arrival order while the budgeted device bytes and the active-job limit allow (xnb_create,
// xnb_set_particles, xnb_nbh_dist, xnb_init on the job's own stream, then
// xnb_batch_attach), advances the batch, and retires finished jobs (xnb_batch_detach, output
// copy, xnb_destroy), freeing their budget.  Latency = completion - arrival.
#include <algorithm>
#include <chrono>
#include <cmath>
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

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t lo = (size_t)std::floor(pos), hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
}

}  // namespace

extern "C" {

size_t xnb_job_bytes(size_t n) {
  // two copies of the particle state (~140 B/slot), backup and scratch (~40 B/slot), brick
  // lists (~260 B/atom), 1.3x slots for ghosts, plus a fixed part for cells and tables
  return n * 900 + ((size_t)16 << 20);
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
  int next = 0, done_jobs = 0, failed = 0, peak_active = 0;
  size_t bytes = 0, peak_bytes = 0;
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
  while (done_jobs < njobs) {
    double now = seconds_since(t0);
    // admit arrived jobs in order
    while (next < njobs && jobs[next].arrival_s <= now && (int)act.size() < max_active) {
      const xnb_job& J = jobs[next];
      const size_t need = xnb_job_bytes(J.n);
      if (!act.empty() && bytes + need > budget) break;  // wait for memory (a lone job always runs)
      xnb_job_result& R = results[next];
      R = xnb_job_result{};
      R.start_s = now;
      Active a;
      a.job = next;
      a.bytes = need;
      int rc = xnb_create(&a.ctx, &J.domain, cfg->device);
      if (!rc) rc = xnb_nbh_dist(a.ctx, J.run.rcut, J.run.rcut_inc, nullptr, nullptr, nullptr);
      if (!rc) rc = xnb_set_particles(a.ctx, J.n, J.r, J.v, J.id, nullptr);
      if (!rc) rc = xnb_init(a.ctx, &J.run.lj, J.run.rcut, J.run.rcut_inc, J.run.precision);
      if (!rc && cfg->autotune) rc = xnb_autotune(a.ctx, &J.run.lj, J.run.rcut, J.run.precision, nullptr);
      if (!rc) rc = xnb_batch_attach(B, a.ctx, &J.run, &a.slot);
      if (rc) {
        if (a.ctx) xnb_destroy(a.ctx);
        finish(next, rc, seconds_since(t0));
      } else {
        act.push_back(a);
        bytes += need;
        peak_bytes = std::max(peak_bytes, bytes);
      }
      next++;
      now = seconds_since(t0);
    }
    peak_active = std::max(peak_active, (int)act.size());
    if (act.empty()) {
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
      if (!rc) atom_steps += (double)J.n * (double)steps_done;
      xnb_destroy(a.ctx);
      bytes -= a.bytes;
      finish(a.job, rc, seconds_since(t0));
      act.erase(act.begin() + (long)q);
    }
  }
  const double wall = seconds_since(t0);
  for (const Active& a : act) {  // left over after an error
    xnb_batch_detach(B, a.slot, nullptr);
    xnb_destroy(a.ctx);
    finish(a.job, e ? e : XNB_E_STATE, wall);
  }
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
  }
  return e;
}

}  // extern "C"
