// Serving loop: inversion requests (a problem description plus the outer-SMC settings) are admitted against
// a device-memory budget as they arrive, and run concurrently on one GPU.  No counterpart in the reference,
// whose main.f90 runs one inversion per MPI job.
//
// Each admitted request gets its own context (its own HIP streams), its problem uploaded once and kept
// resident for the whole outer SMC, and a host thread that drives smcf_fault_smc_exec.  Requests that arrive
// while others run are admitted as soon as the memory the finished ones held is free again (first fit in
// arrival order), and their kernels overlap on the device with those of the requests already running.
// Every random number is keyed by the request's seed, so a request's result does not depend on what else
// runs beside it.
//
// Admission needs both the device bytes (smcf_problem_bytes plus a fixed per-request overhead for the context,
// its streams and the runtime's per-stream state) and a free running place: at most max_running requests run at
// once.  HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default, at most 32 accepted
// here) and each request uses two streams, so the default is GPU_MAX_HW_QUEUES / 2, never more than 16.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/smcf.h"
#include "smcf_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kRequestOverheadBytes = (size_t)64 << 20;  // context, two streams, runtime state per request
constexpr int kMaxRunningCap = 16;

int default_max_running() {
  int q = 4;  // HIP's default number of hardware queues per process
  if (const char* e = std::getenv("GPU_MAX_HW_QUEUES")) {
    int v = std::atoi(e);
    if (v > 0) q = std::min(v, 32);
  }
  return std::max(1, std::min(kMaxRunningCap, q / 2));
}

// A request owns copies of everything it was submitted with (the caller may free its arrays at once).
struct Request {
  int64_t id = 0;
  // problem description with its arrays
  smcf_problem_desc desc{};
  std::vector<int> nxi, neta, nsar_index;
  std::vector<double> xbend, obs_points, obs_unitvec, obs_sigma, dvec, sar_sinv;
  int max_batch = 0;
  // outer SMC settings
  std::vector<double> range;
  int N = 0, max_stages = 0;
  uint64_t seed = 0;
  size_t bytes = 0;
  // timing (seconds since the server started)
  double t_submit = 0, t_start = 0, t_done = 0;
  // results
  std::vector<double> particles, likelihood, gamma_trace;
  int nstage = 0;
  double evidence = 0;
  int rc = SMCF_OK;
  std::string err;
  bool done = false;
};

}  // namespace

struct smcf_server {
  int device = 0;
  size_t budget = 0;
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<Request>> queue;  // submitted, not yet admitted (arrival order)
  std::map<int64_t, std::shared_ptr<Request>> all;
  size_t in_use = 0;
  int running = 0;
  int max_running = 1;
  int64_t next_id = 1;
  bool stopping = false;
  std::thread admitter;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  void run_request(std::shared_ptr<Request> r) {
    smcf_ctx* ctx = nullptr;
    smcf_problem* pb = nullptr;
    int rc = smcf_ctx_create(device, &ctx);
    if (rc == SMCF_OK) rc = smcf_problem_create(ctx, &r->desc, r->max_batch, &pb);
    if (rc == SMCF_OK) {
      int nf = (int)r->range.size() / 2;
      r->particles.assign((size_t)nf * r->N, 0.0);
      r->likelihood.assign(r->N, 0.0);
      r->gamma_trace.assign(r->max_stages + 1, 0.0);
      rc = smcf_fault_smc_exec(pb, r->range.data(), r->N, r->seed, r->max_stages, r->max_batch,
                               r->particles.data(), r->likelihood.data(), &r->nstage, &r->evidence,
                               r->gamma_trace.data());
    }
    std::string err = rc == SMCF_OK ? std::string() : std::string(smcf_last_error());
    if (pb) smcf_problem_destroy(pb);
    if (ctx) smcf_ctx_destroy(ctx);
    std::lock_guard<std::mutex> lk(mu);
    r->rc = rc;
    r->err = err;
    r->t_done = now();
    r->done = true;
    in_use -= r->bytes;
    --running;
    cv.notify_all();
  }

  // admits queued requests (first fit in arrival order) whenever memory is released or a request arrives
  void admit_loop() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      bool admitted = false;
      for (auto it = queue.begin(); it != queue.end();) {
        auto r = *it;
        if (running < max_running && in_use + r->bytes <= budget) {
          in_use += r->bytes;
          ++running;
          r->t_start = now();
          it = queue.erase(it);
          // detached: the worker's last act is to decrement `running` under the lock, and destroy waits for 0
          std::thread(&smcf_server::run_request, this, r).detach();
          admitted = true;
        } else {
          ++it;
        }
      }
      if (stopping && queue.empty()) return;
      if (!admitted) cv.wait(lk);
    }
  }
};

extern "C" {

int smcf_server_create(int device, size_t hbm_budget_bytes, smcf_server** out) {
  if (!out || hbm_budget_bytes == 0) return SMCF_EINVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    smcf::set_error("no such device");
    return SMCF_EINVAL;
  }
  auto* s = new smcf_server();
  s->device = device;
  s->budget = hbm_budget_bytes;
  s->max_running = default_max_running();
  s->t0 = Clock::now();
  s->admitter = std::thread(&smcf_server::admit_loop, s);
  *out = s;
  return SMCF_OK;
}

int smcf_server_submit(smcf_server* s, const smcf_problem_desc* d, int max_batch, const double* range, int N,
                       uint64_t seed, int max_stages, int64_t* id) {
  if (!s || !d || !range || N < 1 || max_stages < 1 || max_batch < 1 || !id) return SMCF_EINVAL;
  auto r = std::make_shared<Request>();
  int np = d->nplane;
  if (np < 1 || !d->nxi_ls || !d->neta_ls || d->nobs < 1) return SMCF_EINVAL;
  r->nxi.assign(d->nxi_ls, d->nxi_ls + np);
  r->neta.assign(d->neta_ls, d->neta_ls + np);
  if (np > 1 && d->xbend) r->xbend.assign(d->xbend, d->xbend + (np - 1));
  size_t nobs = (size_t)d->nobs;
  r->obs_points.assign(d->obs_points, d->obs_points + 2 * nobs);
  r->obs_unitvec.assign(d->obs_unitvec, d->obs_unitvec + 3 * nobs);
  r->obs_sigma.assign(d->obs_sigma, d->obs_sigma + nobs);
  r->dvec.assign(d->dvec, d->dvec + nobs);
  if (d->npath > 0 && d->nsar_index) r->nsar_index.assign(d->nsar_index, d->nsar_index + d->npath + 1);
  if (d->nsar_total > 0 && d->sar_sinv && !r->nsar_index.empty()) {
    size_t n2 = 0;
    for (int p = 0; p < d->npath; ++p) {
      size_t np_ = (size_t)(r->nsar_index[p + 1] - r->nsar_index[p]);
      n2 += np_ * np_;
    }
    r->sar_sinv.assign(d->sar_sinv, d->sar_sinv + n2);
  }
  r->desc = *d;
  r->desc.nxi_ls = r->nxi.data();
  r->desc.neta_ls = r->neta.data();
  r->desc.xbend = r->xbend.empty() ? nullptr : r->xbend.data();
  r->desc.nsar_index = r->nsar_index.empty() ? nullptr : r->nsar_index.data();
  r->desc.obs_points = r->obs_points.data();
  r->desc.obs_unitvec = r->obs_unitvec.data();
  r->desc.obs_sigma = r->obs_sigma.data();
  r->desc.dvec = r->dvec.data();
  r->desc.sar_sinv = r->sar_sinv.empty() ? nullptr : r->sar_sinv.data();
  int nf = d->fix_xbend ? 2 * np + 4 : 3 * np + 3;
  r->range.assign(range, range + 2 * (size_t)nf);
  r->N = N;
  r->seed = seed;
  r->max_stages = max_stages;
  r->max_batch = max_batch;
  int rc = smcf_problem_bytes(&r->desc, max_batch, &r->bytes);
  if (rc) return rc;
  r->bytes += kRequestOverheadBytes;
  std::lock_guard<std::mutex> lk(s->mu);
  if (r->bytes > s->budget) {
    smcf::set_error("request needs " + std::to_string(r->bytes) + " device bytes, more than the budget " +
                    std::to_string(s->budget));
    return SMCF_ENOMEM;
  }
  r->id = s->next_id++;
  r->t_submit = s->now();
  s->queue.push_back(r);
  s->all[r->id] = r;
  *id = r->id;
  s->cv.notify_all();
  return SMCF_OK;
}

int smcf_server_wait(smcf_server* s, int64_t id, double* particles, double* likelihood, int* nstage,
                     double* evidence, double* gamma_trace, double* times) {
  if (!s) return SMCF_EINVAL;
  std::shared_ptr<Request> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) {
      smcf::set_error("unknown request id");
      return SMCF_EINVAL;
    }
    r = it->second;
    s->cv.wait(lk, [&] { return r->done; });
    s->all.erase(it);
  }
  if (times) {
    times[0] = r->t_submit;
    times[1] = r->t_start;
    times[2] = r->t_done;
  }
  if (r->rc != SMCF_OK) {
    smcf::set_error(r->err);
    return r->rc;
  }
  if (particles) std::memcpy(particles, r->particles.data(), sizeof(double) * r->particles.size());
  if (likelihood) std::memcpy(likelihood, r->likelihood.data(), sizeof(double) * r->likelihood.size());
  if (nstage) *nstage = r->nstage;
  if (evidence) *evidence = r->evidence;
  if (gamma_trace) std::memcpy(gamma_trace, r->gamma_trace.data(), sizeof(double) * r->gamma_trace.size());
  return SMCF_OK;
}

int smcf_server_set_max_running(smcf_server* s, int max_running) {
  if (!s || max_running < 1 || max_running > kMaxRunningCap) {
    smcf::set_error("max_running must be 1 .. 16");
    return SMCF_EINVAL;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->max_running = max_running;
  s->cv.notify_all();
  return SMCF_OK;
}

int smcf_server_status(smcf_server* s, int* queued, int* running, size_t* bytes_in_use) {
  if (!s) return SMCF_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  if (queued) *queued = (int)s->queue.size();
  if (running) *running = s->running;
  if (bytes_in_use) *bytes_in_use = s->in_use;
  return SMCF_OK;
}

int smcf_server_destroy(smcf_server* s) {
  if (!s) return SMCF_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
    s->cv.notify_all();
  }
  s->admitter.join();
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [&] { return s->running == 0; });
  }
  delete s;
  return SMCF_OK;
}

}  // extern "C"
