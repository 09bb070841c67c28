// Serving loop: inversion requests (a problem description plus the outer-SMC settings) are admitted against
// a device-memory budget as they arrive and batched continuously on one GPU.  No counterpart in the reference,
// whose main.f90 runs one inversion per MPI job.
//
// The server owns one context and one persistent stage engine.  An admitted request's problem is uploaded once
// and joins the engine as a member for the whole outer SMC; a host thread per request drives smcf_fault_smc_exec,
// whose waves are queued to the engine.  Every stage kernel therefore runs once over the instances of all running
// requests (continuous batching: a request that arrives while others are on the GPU joins the next stage launch
// once admitted, and its geometries fill its slots as they free up).  When its outer SMC has finished the request
// leaves the engine, its problem is freed and the next queued requests are admitted in arrival order (first fit
// under the budget).  Every random number is keyed by the request's seed and an instance's arithmetic does not
// depend on its neighbours, so a request's result does not depend on what else runs beside it.
//
// Admission needs the device bytes (smcf_problem_bytes plus a fixed per-request allowance for the engine's tables
// and staging) and a free running place: at most max_running requests are members at once (each holds a host
// thread).
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

constexpr size_t kRequestOverheadBytes = (size_t)16 << 20;  // engine tables and admission staging per request
constexpr int kMaxRunningCap = 256;
constexpr int kDefaultMaxRunning = 64;

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
  // while admitted
  smcf_problem* pb = nullptr;
  std::thread worker;
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
  std::map<smcf::Problem*, std::shared_ptr<Request>> live;  // admitted, by problem
  size_t in_use = 0;
  int running = 0;
  int max_running = kDefaultMaxRunning;
  int64_t next_id = 1;
  bool stopping = false;
  bool engine_failed = false;
  smcf_ctx* ctx = nullptr;
  smcf::Engine* eng = nullptr;
  std::thread admitter, engine_thread;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  // engine thread: the request's problem is a member now; its outer SMC starts
  void added(smcf::Problem* P) {
    std::shared_ptr<Request> r;
    {
      std::lock_guard<std::mutex> lk(mu);
      r = live[P];
    }
    r->worker = std::thread([this, r] {
      (void)hipSetDevice(device);
      int nf = (int)r->range.size() / 2;
      r->particles.assign((size_t)nf * r->N, 0.0);
      r->likelihood.assign(r->N, 0.0);
      r->gamma_trace.assign(r->max_stages + 1, 0.0);
      int rc = smcf_fault_smc_exec(r->pb, r->range.data(), r->N, r->seed, r->max_stages, r->max_batch,
                                   r->particles.data(), r->likelihood.data(), &r->nstage, &r->evidence,
                                   r->gamma_trace.data());
      r->rc = rc;
      if (rc) r->err = smcf_last_error();
      smcf::serve_engine_leave(eng, &r->pb->p);
    });
  }

  // engine thread: the request's problem has left the engine; free it and admit the next ones
  void removed(smcf::Problem* P) {
    std::shared_ptr<Request> r;
    {
      std::lock_guard<std::mutex> lk(mu);
      r = live[P];
      live.erase(P);
    }
    finish(r);
  }

  void finish(std::shared_ptr<Request> r) {
    if (r->worker.joinable()) r->worker.join();
    if (r->pb) smcf_problem_destroy(r->pb);
    r->pb = nullptr;
    std::lock_guard<std::mutex> lk(mu);
    r->t_done = now();
    r->done = true;
    in_use -= r->bytes;
    --running;
    cv.notify_all();
  }

  void engine_main() {
    (void)hipSetDevice(device);
    int rc = smcf::serve_engine_run(eng);
    if (rc == SMCF_OK) return;
    // the engine failed: its members' outer SMCs have returned with its error; end them and refuse the rest
    std::string msg;
    smcf::serve_engine_error(eng, &msg);
    std::vector<std::shared_ptr<Request>> left;
    {
      std::lock_guard<std::mutex> lk(mu);
      engine_failed = true;
      for (auto& kv : live) left.push_back(kv.second);
      live.clear();
      cv.notify_all();
    }
    for (auto& r : left) {
      if (r->rc == SMCF_OK) {
        r->rc = rc;
        r->err = msg;
      }
      finish(r);
    }
  }

  // admits queued requests (first fit in arrival order) whenever memory is released or a request arrives
  void admit_loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      bool admitted = false;
      for (auto it = queue.begin(); it != queue.end(); ++it) {
        auto r = *it;
        if (engine_failed) {
          queue.erase(it);
          r->rc = SMCF_ESTATE;
          r->err = "the server's stage engine failed";
          r->t_done = now();
          r->done = true;
          cv.notify_all();
          admitted = true;
          break;
        }
        if (running < max_running && in_use + r->bytes <= budget) {
          in_use += r->bytes;
          ++running;
          r->t_start = now();
          queue.erase(it);
          lk.unlock();
          smcf_problem* pb = nullptr;
          int rc;
          {
            std::lock_guard<std::mutex> cl(smcf::serve_engine_capture_lock(eng));
            rc = smcf_problem_create(ctx, &r->desc, r->max_batch, &pb);
          }
          std::string err = rc ? std::string(smcf_last_error()) : std::string();
          lk.lock();
          if (rc) {
            r->rc = rc;
            r->err = err;
            r->t_done = now();
            r->done = true;
            in_use -= r->bytes;
            --running;
            cv.notify_all();
          } else {
            r->pb = pb;
            live[&pb->p] = r;
            smcf::serve_engine_add(eng, &pb->p);
          }
          admitted = true;
          break;  // the queue may have changed while the problem was created: scan again
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
  smcf_ctx* ctx = nullptr;
  int rc = smcf_ctx_create(device, &ctx);
  if (rc) return rc;
  auto* s = new smcf_server();
  s->device = device;
  s->budget = hbm_budget_bytes;
  s->t0 = Clock::now();
  s->ctx = ctx;
  s->eng = smcf::serve_engine_create(ctx, [s](smcf::Problem* P) { s->added(P); },
                                     [s](smcf::Problem* P) { s->removed(P); });
  if (!s->eng) {
    smcf_ctx_destroy(ctx);
    delete s;
    return SMCF_EHIP;
  }
  s->engine_thread = std::thread(&smcf_server::engine_main, s);
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
  if (d->mover != SMCF_MOVER_GIBBS) {
    smcf::set_error("the server batches Gibbs-mover requests (HMC inversions run through smcf_fault_smc_exec)");
    return SMCF_EUNSUPPORTED;
  }
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
    smcf::set_error("max_running must be 1 .. " + std::to_string(kMaxRunningCap));
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
  smcf::serve_engine_stop(s->eng);
  s->engine_thread.join();
  smcf::serve_engine_destroy(s->eng);
  smcf_ctx_destroy(s->ctx);
  delete s;
  return SMCF_OK;
}

}  // extern "C"
