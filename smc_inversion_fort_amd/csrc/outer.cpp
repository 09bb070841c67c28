// Outer fault SMC (smc_fault.f90:785-1075) on top of the batched device likelihood.
//
// Per outer stage the host does the O(N) work in the reference's own arithmetic order:
// gamma bisection with CV 0.2 (smc_fault.f90:392-485), plain FP64 normalisation and
// moments, 0.04 * covariance with an unblocked lower Cholesky (487-570), systematic
// residual resampling with a real(16) running u (572-590, here __float128), and the
// descending reorder through qsort_ (592-637, sort.cpp).  The expensive part, 2N inner
// SMC runs (work_mcmc_sampling, 639-783), is issued as device batches in waves: wave 0
// re-evaluates every parent that has children, wave k carries the k-th child of every
// parent with more than k children.  After the descending sort each wave is a prefix of
// the parents.  A parent with no child has no effect on any output under keyed draws, so
// its re-evaluation is skipped.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/smcf.h"

namespace smcf {
void set_error(const std::string& msg);  // api.cpp: the message smcf_last_error returns
}

namespace {

struct Philox {
  // Philox4x32-10, counter (c0, c1, c2, site << 24 | stage), two 64-bit words out
  static void draw(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2, uint32_t c1, uint32_t c0,
                   uint64_t out[2]) {
    uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = (site << 24) | (stage & 0xFFFFFFu);
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
    for (int r = 0; r < 10; ++r) {
      uint64_t p0 = (uint64_t)0xD2511F53u * x0, p1 = (uint64_t)0xCD9E8D57u * x2;
      uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0, y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
      x0 = y0; x1 = (uint32_t)p1; x2 = y2; x3 = (uint32_t)p0;
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    out[0] = ((uint64_t)x1 << 32) | x0;
    out[1] = ((uint64_t)x3 << 32) | x2;
  }
};

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

double u01(uint64_t w) { return (double)(w >> 11) * 0x1.0p-53; }
double clampu(double u) { return u < 1e-5 ? 1e-5 : u; }

enum : uint32_t { kOuterInit = 16, kOuterResample = 17, kOuterProp = 18, kOuterMH = 19 };

// fault_find_next_gamma + calc_mean_std_vector (sequential sums)
double find_next_gamma(double gprev, const std::vector<double>& L, std::vector<double>& w, double* ev_out) {
  int N = (int)L.size();
  double lmin = L[0];
  for (int i = 0; i < N; ++i) lmin = std::min(lmin, L[i]);
  double lower = gprev, upper = 1.0, err = 1.0, gamma = 1.0;
  int cnt = 0;
  while (err > 1e-8) {
    gamma = (lower + upper) / 2.0;
    double dg = gamma - gprev;
    for (int i = 0; i < N; ++i) w[i] = std::exp(-dg * (L[i] - lmin));
    double mean = 0.0;
    for (int i = 0; i < N; ++i) mean = mean + w[i];
    mean = mean / N;
    double sd = 0.0;
    for (int i = 0; i < N; ++i) sd = sd + (w[i] - mean) * (w[i] - mean);
    sd = std::sqrt(sd / N);
    double cv = sd / mean;
    if (cv > 0.2) upper = gamma; else lower = gamma;
    err = std::fabs(cv - 0.2);
    if (std::fabs(gamma - 1.0) < 1e-8) break;
    if (cnt > 1000) break;
    cnt++;
  }
  double ev = 0.0, dg = gamma - gprev;
  for (int i = 0; i < N; ++i) ev = ev + std::exp(-dg * (L[i] - lmin));
  ev = ev / N;
  *ev_out = -std::log(ev) + dg * lmin;
  return gamma;
}

// x := L x with L lower non-unit, dtrmv('l','n','n') column order
void trmv_lower(int n, const std::vector<double>& A, double* x) {
  for (int j = n - 1; j >= 0; --j)
    if (x[j] != 0.0) {
      double t = x[j];
      for (int i = n - 1; i > j; --i) x[i] += t * A[j * n + i];
      x[j] = x[j] * A[j * n + j];
    }
}

double box_muller(uint64_t w0, uint64_t w1) {
  const double pi = 3.141592653589793;
  return std::sqrt(-2.0 * std::log(clampu(u01(w0)))) * std::sin(2.0 * pi * clampu(u01(w1)));
}

}  // namespace

extern "C" {

// key of the inner SMC run of one outer evaluation (outer stage, role 0 = parent / 1 = child, slot)
uint64_t smcf_inner_key(uint64_t seed, uint32_t outer_stage, uint32_t role, uint32_t slot) {
  uint64_t a = splitmix64(seed);
  uint64_t b = ((uint64_t)outer_stage << 33) ^ ((uint64_t)role << 32) ^ (uint64_t)slot;
  return splitmix64(a ^ splitmix64(b));
}

}  // extern "C"

namespace {

struct Gsmc {
  const double* qmean;      // nd
  const double* qcov_chol;  // nd x nd column-major, lower Cholesky factor of qcov
  const double* qcov_inv;   // nd x nd column-major
  double log_det;
};

// fault_calc_neglog_q (smc_fault.f90:9-27), same summation order
double neglog_q(const double* x, const Gsmc& g, int nd) {
  const double pi = 3.141592653589793;
  double delta[64], tmp = 0.0;
  for (int i = 0; i < nd; ++i) delta[i] = x[i] - g.qmean[i];
  for (int j = 0; j < nd; ++j)
    for (int i = 0; i < nd; ++i) tmp = tmp + delta[i] * g.qcov_inv[j * nd + i] * delta[j];
  return tmp / 2.0 + std::log(2.0 * pi) * nd / 2.0 + g.log_det / 2.0;
}

// per-stage dumps of fault_smc_exec (smc_fault.f90:923-936, 1054-1067): <output_dir><iter>.csv, one line per
// particle, its ndim coordinates then its likelihood, each in Fortran F20.5 (C "%20.5f" for values that fit)
std::mutex g_out_mu;
std::map<const smcf_problem*, std::string> g_out_dir;

std::string output_dir_of(const smcf_problem* pb) {
  std::lock_guard<std::mutex> lk(g_out_mu);
  auto it = g_out_dir.find(pb);
  return it == g_out_dir.end() ? std::string() : it->second;
}

// where each stage's reweight and resample runs (smcf_fault_smc_set_outer_stage): 0 host, 1 / 2 device
std::map<const smcf_problem*, int> g_outer_mode;

int outer_mode_of(const smcf_problem* pb) {
  std::lock_guard<std::mutex> lk(g_out_mu);
  auto it = g_outer_mode.find(pb);
  return it == g_outer_mode.end() ? 0 : it->second;
}

int write_stage(const std::string& dir, int iter, int N, int nd, const std::vector<double>& X,
                const std::vector<double>& L) {
  if (dir.empty()) return SMCF_OK;
  std::string name = dir + std::to_string(iter) + ".csv";
  FILE* f = std::fopen(name.c_str(), "w");
  if (!f) return SMCF_EINVAL;
  for (int i = 0; i < N; ++i) {
    for (int d = 0; d < nd; ++d) std::fprintf(f, "%20.5f", X[(size_t)i * nd + d]);
    std::fprintf(f, "%20.5f\n", L[i]);
  }
  return std::fclose(f) == 0 ? SMCF_OK : SMCF_EINVAL;
}

// batched evaluation helper: calls of at most max_batch geometries, each streamed through the slots
int eval_batch(smcf_problem* pb, int nd, int max_batch, const std::vector<double>& th, const std::vector<uint64_t>& keys,
               std::vector<double>& out) {
  int B = (int)keys.size();
  out.resize(B);
  for (int s = 0; s < B; s += max_batch) {
    int e = std::min(B, s + max_batch);
    int r = smcf_fault_calc_likelihood_batch(pb, e - s, th.data() + (size_t)s * nd, keys.data() + s, out.data() + s,
                                             nullptr);
    if (r) return r;
  }
  return SMCF_OK;
}

int ndim_fault_of(smcf_problem* pb, int* nd) {
  int nnode, ndof, nds, nobs;
  return smcf_problem_dims(pb, &nnode, &ndof, &nds, nd, &nobs);
}

int fault_smc(smcf_problem* pb, const double* range, int N, uint64_t seed, int max_stages, int max_batch,
              double* particles, double* likelihood, int* nstage, double* evidence, double* gamma_trace,
              const Gsmc* g) {
  int nd = 0;
  int rc = ndim_fault_of(pb, &nd);
  if (rc) return rc;
  if (!range || N < 1 || max_batch < 1 || !particles || !likelihood || nd > 64) return SMCF_EINVAL;
  std::vector<double> X((size_t)nd * N), L(N), w(N), Xn((size_t)nd * N), Ln(N), cov(nd * nd), mean(nd);
  std::vector<double> Q(N, 0.0), Qn(N, 0.0);
  std::vector<int> assigned(N), sidx(N);
  if ((rc = smcf_fault_sample_init_particles(nd, N, seed, range, g ? g->qmean : nullptr, g ? g->qcov_chol : nullptr,
                                             X.data())))
    return rc;
  {
    std::vector<uint64_t> keys(N);
    for (int i = 0; i < N; ++i) keys[i] = smcf_inner_key(seed, 0, 0, i);
    if ((rc = eval_batch(pb, nd, max_batch, X, keys, L))) return rc;
  }
  if (g)
    for (int i = 0; i < N; ++i) Q[i] = neglog_q(&X[(size_t)i * nd], *g, nd);
  const std::string out_dir = output_dir_of(pb);
  const int omode = outer_mode_of(pb);
  if ((rc = write_stage(out_dir, 0, N, nd, X, L))) return rc;
  double gamma = 0.0, ev_total = 0.0;
  int iter = 1;
  while (1.0 - gamma > 1e-8 && iter <= max_stages) {
    double gprev = gamma, ev;
    int info = 0;  // ignored, as in the reference
    if (omode) {
      if ((rc = smcf_outer_reweight_resample(smcf_problem_context(pb), nd, N, gprev, seed, iter, omode == 2, X.data(),
                                             L.data(), g ? Q.data() : nullptr, &gamma, &ev, w.data(), cov.data(),
                                             &info, assigned.data(), sidx.data())))
        return rc;
    } else {
      if ((rc = smcf_fault_find_next_gamma(N, gprev, L.data(), g ? Q.data() : nullptr, w.data(), &gamma, &ev)))
        return rc;
      smcf_fault_normalize_weights(N, w.data());
      smcf_fault_calc_mean_particles(nd, N, X.data(), w.data(), mean.data());
      smcf_fault_calc_cov_particles(nd, N, X.data(), w.data(), mean.data(), cov.data(), &info);
      smcf_fault_resample_particles(N, w.data(), seed, iter, assigned.data());
      smcf_fault_reorder_to_send(nd, N, assigned.data(), X.data(), L.data(), Q.data(), sidx.data());
    }
    ev_total = ev_total + ev;
    if (gamma_trace) gamma_trace[iter - 1] = gamma;
    int acc = 0;
    if ((rc = smcf_work_mcmc_sampling(pb, seed, iter, gamma, range, cov.data(), N, assigned.data(), X.data(),
                                      g ? Q.data() : nullptr, g ? g->qmean : nullptr, g ? g->qcov_inv : nullptr,
                                      g ? g->log_det : 0.0, max_batch, Xn.data(), Ln.data(), Qn.data(), &acc)))
      return rc;
    X = Xn;
    L = Ln;
    Q = Qn;
    if ((rc = write_stage(out_dir, iter, N, nd, X, L))) return rc;
    iter++;
  }
  std::memcpy(particles, X.data(), sizeof(double) * X.size());
  std::memcpy(likelihood, L.data(), sizeof(double) * N);
  if (nstage) *nstage = iter - 1;
  if (evidence) *evidence = ev_total;
  return SMCF_OK;
}

}  // namespace

extern "C" {

// fault_sample_init_particles (smc_fault.f90:29-64): uniform in the range per dimension, or qmean + chol z
int smcf_fault_sample_init_particles(int nd, int N, uint64_t seed, const double* range, const double* qmean,
                                     const double* qcov_chol, double* particles) {
  if (nd < 1 || nd > 64 || N < 1 || !particles || (!qmean && !range) || (qmean && !qcov_chol)) return SMCF_EINVAL;
  uint64_t okey = splitmix64(seed ^ 0x5EEDF00Dull);
  std::vector<double> chol;
  if (qmean) chol.assign(qcov_chol, qcov_chol + (size_t)nd * nd);
  for (int i = 0; i < N; ++i) {
    if (qmean) {
      double z[64];
      for (int d = 0; d < nd; ++d) {
        uint64_t ww[2];
        Philox::draw(okey, kOuterInit, 0, i, d, 0, ww);
        z[d] = box_muller(ww[0], ww[1]);
      }
      trmv_lower(nd, chol, z);
      for (int d = 0; d < nd; ++d) particles[(size_t)i * nd + d] = z[d] + qmean[d];
    } else {
      for (int d = 0; d < nd; ++d) {
        uint64_t ww[2];
        Philox::draw(okey, kOuterInit, 0, i, d, 0, ww);
        double r = clampu(u01(ww[0]));
        particles[(size_t)i * nd + d] = range[2 * d] + r * (range[2 * d + 1] - range[2 * d]);
      }
    }
  }
  return SMCF_OK;
}

// calc_mean_std_vector (smc_fault.f90:372-390): sequential sums, population standard deviation
int smcf_fault_mean_std(int n, const double* vec, double* mean, double* std_out) {
  if (n < 1 || !vec) return SMCF_EINVAL;
  double m = 0.0;
  for (int i = 0; i < n; ++i) m = m + vec[i];
  m = m / n;
  double sd = 0.0;
  for (int i = 0; i < n; ++i) sd = sd + (vec[i] - m) * (vec[i] - m);
  sd = std::sqrt(sd / n);
  if (mean) *mean = m;
  if (std_out) *std_out = sd;
  return SMCF_OK;
}

int smcf_fault_find_next_gamma(int N, double gamma_prev, const double* likelihood, const double* q, double* weights,
                               double* gamma, double* neglog_evidence) {
  if (N < 1 || !likelihood || !weights) return SMCF_EINVAL;
  std::vector<double> R(likelihood, likelihood + N), w(N);
  if (q)  // tempering of L - q (smc_fault.f90:411-416, 433-438, 479-481)
    for (int i = 0; i < N; ++i) R[i] = likelihood[i] - q[i];
  double ev = 0.0;
  double g = find_next_gamma(gamma_prev, R, w, &ev);
  std::copy(w.begin(), w.end(), weights);
  if (gamma) *gamma = g;
  if (neglog_evidence) *neglog_evidence = ev;
  return SMCF_OK;
}

// fault_normalize_weights (smc_fault.f90:487-501): plain sequential FP64 sum
int smcf_fault_normalize_weights(int N, double* weights) {
  if (N < 1 || !weights) return SMCF_EINVAL;
  double s = 0.0;
  for (int i = 0; i < N; ++i) s = s + weights[i];
  for (int i = 0; i < N; ++i) weights[i] = weights[i] / s;
  return SMCF_OK;
}

// fault_calc_mean_particles (smc_fault.f90:503-520)
int smcf_fault_calc_mean_particles(int nd, int N, const double* particles, const double* weights, double* mean) {
  if (nd < 1 || N < 1 || !particles || !weights || !mean) return SMCF_EINVAL;
  for (int d = 0; d < nd; ++d) mean[d] = 0.0;
  for (int i = 0; i < N; ++i)
    for (int d = 0; d < nd; ++d) mean[d] = mean[d] + weights[i] * particles[(size_t)i * nd + d];
  return SMCF_OK;
}

// fault_calc_cov_particles (smc_fault.f90:522-570): lower weighted covariance x 0.04, then dpotrf('L') unblocked;
// cov receives the lower factor (upper triangle zero), info the first non-positive pivot (1-based; 0 if none)
int smcf_fault_calc_cov_particles(int nd, int N, const double* particles, const double* weights, const double* mean,
                                  double* cov, int* info) {
  if (nd < 1 || N < 1 || !particles || !weights || !mean || !cov) return SMCF_EINVAL;
  for (int k = 0; k < nd * nd; ++k) cov[k] = 0.0;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < nd; ++j)
      for (int r = j; r < nd; ++r)
        cov[j * nd + r] = cov[j * nd + r] + weights[i] * (particles[(size_t)i * nd + r] - mean[r]) *
                                                (particles[(size_t)i * nd + j] - mean[j]);
  for (int k = 0; k < nd * nd; ++k) cov[k] = cov[k] * 0.04;
  if (info) *info = 0;
  for (int j = 0; j < nd; ++j) {
    double ajj = cov[j * nd + j];
    for (int k = 0; k < j; ++k) ajj -= cov[k * nd + j] * cov[k * nd + j];
    if (ajj <= 0.0 || std::isnan(ajj)) {
      cov[j * nd + j] = ajj;
      if (info) *info = j + 1;
      break;
    }
    ajj = std::sqrt(ajj);
    cov[j * nd + j] = ajj;
    for (int r = j + 1; r < nd; ++r) {
      double t = cov[j * nd + r];
      for (int k = 0; k < j; ++k) t -= cov[k * nd + r] * cov[k * nd + j];
      cov[j * nd + r] = t * (1.0 / ajj);
    }
  }
  return SMCF_OK;
}

// fault_resample_particles (smc_fault.f90:572-590): systematic residual resampling with a real(16) running u,
// u drawn in the keyed stream of (seed, stage)
int smcf_fault_resample_particles(int N, const double* weights, uint64_t seed, int stage, int* assigned_num) {
  if (N < 1 || !weights || !assigned_num) return SMCF_EINVAL;
  uint64_t okey = splitmix64(seed ^ 0x5EEDF00Dull);
  uint64_t wq[2];
  Philox::draw(okey, kOuterResample, (uint32_t)stage, 0, 0, 0, wq);
  __float128 u = (__float128)((double)(wq[0] >> 11) * 0x1.0p-53) + (__float128)((double)(wq[1] >> 11) * 0x1.0p-106);
  if (u < (__float128)1e-5) u = (__float128)1e-5;
  __float128 dn = (__float128)N;
  u = u / dn;
  for (int i = 0; i < N; ++i) {
    __float128 t = ((__float128)weights[i] - u) * dn;
    long long f = (long long)t;
    if ((__float128)f > t) f -= 1;  // floor
    int a = (int)f + 1;
    assigned_num[i] = a;
    u = u + (__float128)a / dn - (__float128)weights[i];
  }
  return SMCF_OK;
}

// fault_reorder_to_send (smc_fault.f90:592-637) with numprocs = 1: particles, likelihood, q and counts reordered by
// descending assigned_num through qsort_ (sort.cpp); sorted_idx receives the 1-based permutation
int smcf_fault_reorder_to_send(int nd, int N, int* assigned_num, double* particles, double* likelihood, double* q,
                               int* sorted_idx) {
  if (nd < 1 || N < 1 || !assigned_num || !particles || !likelihood || !sorted_idx) return SMCF_EINVAL;
  std::vector<int> ta(assigned_num, assigned_num + N);
  std::vector<double> tx(particles, particles + (size_t)N * nd), tl(likelihood, likelihood + N);
  std::vector<double> tq;
  if (q) tq.assign(q, q + N);
  for (int i = 0; i < N; ++i) sorted_idx[i] = i + 1;
  int n = N;
  qsort_(assigned_num, sorted_idx, &n);
  for (int i = 0; i < N; ++i) {
    int o = sorted_idx[i] - 1;
    assigned_num[i] = ta[o];
    std::memcpy(particles + (size_t)i * nd, &tx[(size_t)o * nd], sizeof(double) * nd);
    likelihood[i] = tl[o];
    if (q) q[i] = tq[o];
  }
  return SMCF_OK;
}

// work_mcmc_sampling (smc_fault.f90:639-783) with numprocs = 1.  Wave 0 re-evaluates every parent that has
// children, wave k carries the k-th child of every parent with more than k children, each wave one batched call.
int smcf_work_mcmc_sampling(smcf_problem* pb, uint64_t seed, int stage, double gamma, const double* range,
                            const double* cov_chol, int N, const int* assigned_num, const double* particles,
                            const double* q, const double* qmean, const double* qcov_inv, double log_qcov_det,
                            int max_batch, double* particles_new, double* likelihood_new, double* q_new,
                            int* acc_count) {
  int nd = 0;
  int rc = ndim_fault_of(pb, &nd);
  if (rc) return rc;
  if (!range || !cov_chol || N < 1 || !assigned_num || !particles || !particles_new || !likelihood_new ||
      max_batch < 1 || nd > 64 || (q && (!qmean || !qcov_inv)))
    return SMCF_EINVAL;
  const bool gs = q != nullptr;
  Gsmc g{qmean, nullptr, qcov_inv, log_qcov_det};
  uint64_t okey = splitmix64(seed ^ 0x5EEDF00Dull);
  std::vector<int> id_start(N);
  id_start[0] = 0;
  for (int i = 0; i < N - 1; ++i) id_start[i + 1] = id_start[i] + assigned_num[i];
  if (id_start[N - 1] + assigned_num[N - 1] != N) {
    smcf::set_error("work_mcmc_sampling: the assigned counts do not sum to N");
    return SMCF_ESTATE;
  }
  std::vector<double> cov(cov_chol, cov_chol + (size_t)nd * nd), Lcur(N, 0.0), qcur(N, 0.0);
  std::vector<double> cur(particles, particles + (size_t)N * nd);
  if (gs) std::copy(q, q + N, qcur.begin());
  // wave 0: parents with children
  std::vector<double> th, out;
  std::vector<uint64_t> keys;
  std::vector<int> who;
  for (int i = 0; i < N; ++i)
    if (assigned_num[i] > 0) {
      th.insert(th.end(), particles + (size_t)i * nd, particles + (size_t)(i + 1) * nd);
      keys.push_back(smcf_inner_key(seed, stage, 0, i));
      who.push_back(i);
    }
  if ((rc = eval_batch(pb, nd, max_batch, th, keys, out))) return rc;
  for (size_t k = 0; k < who.size(); ++k) Lcur[who[k]] = out[k];
  int amax = *std::max_element(assigned_num, assigned_num + N);
  int acc_n = 0;
  for (int k = 0; k < amax; ++k) {
    th.clear(); keys.clear(); who.clear();
    std::vector<int> slot;
    for (int i = 0; i < N; ++i) {
      if (assigned_num[i] <= k) continue;
      int j = id_start[i] + k;
      double z[64];
      for (int d = 0; d < nd; ++d) {
        uint64_t wz[2];
        Philox::draw(okey, kOuterProp, (uint32_t)stage, j, d, 0, wz);
        z[d] = box_muller(wz[0], wz[1]);
      }
      trmv_lower(nd, cov, z);
      double cand[64];
      for (int d = 0; d < nd; ++d) cand[d] = cur[(size_t)i * nd + d] + z[d];
      for (int d = 0; d < nd; ++d) {
        double lo = range[2 * d], hi = range[2 * d + 1];
        if (hi - lo > 1e-3) {
          if (cand[d] < lo) cand[d] = std::min(2.0 * lo - cand[d], hi);
          if (cand[d] > hi) cand[d] = std::max(2.0 * hi - cand[d], lo);
        }
      }
      th.insert(th.end(), cand, cand + nd);
      keys.push_back(smcf_inner_key(seed, stage, 1, j));
      who.push_back(i);
      slot.push_back(j);
    }
    if ((rc = eval_batch(pb, nd, max_batch, th, keys, out))) return rc;
    for (size_t m = 0; m < who.size(); ++m) {
      int i = who[m], j = slot[m];
      uint64_t wm[2];
      Philox::draw(okey, kOuterMH, (uint32_t)stage, j, 0, 0, wm);
      double mh = clampu(u01(wm[0]));
      bool acc;
      double qc = 0.0;
      if (gs) {
        qc = neglog_q(&th[m * nd], g, nd);
        acc = gamma * (Lcur[i] - out[m]) + (1.0 - gamma) * (qcur[i] - qc) > std::log(mh);
      } else {
        acc = gamma * (Lcur[i] - out[m]) > std::log(mh);
      }
      if (acc) {
        std::memcpy(&cur[(size_t)i * nd], &th[m * nd], sizeof(double) * nd);
        Lcur[i] = out[m];
        qcur[i] = qc;
        acc_n++;
      }
      likelihood_new[j] = Lcur[i];
      if (q_new) q_new[j] = qcur[i];
      std::memcpy(particles_new + (size_t)j * nd, &cur[(size_t)i * nd], sizeof(double) * nd);
    }
  }
  if (acc_count) *acc_count = acc_n;
  return SMCF_OK;
}

int smcf_fault_smc_set_outer_stage(smcf_problem* pb, int mode) {
  if (!pb || mode < 0 || mode > 2) return SMCF_EINVAL;
  std::lock_guard<std::mutex> lk(g_out_mu);
  if (mode)
    g_outer_mode[pb] = mode;
  else
    g_outer_mode.erase(pb);
  return SMCF_OK;
}

int smcf_fault_smc_set_output_dir(smcf_problem* pb, const char* output_dir) {
  if (!pb) return SMCF_EINVAL;
  std::lock_guard<std::mutex> lk(g_out_mu);
  if (output_dir && output_dir[0])
    g_out_dir[pb] = output_dir;
  else
    g_out_dir.erase(pb);
  return SMCF_OK;
}

int smcf_fault_smc_exec(smcf_problem* pb, const double* range, int N, uint64_t seed, int max_stages,
                        int max_batch, double* particles, double* likelihood, int* nstage, double* evidence,
                        double* gamma_trace) {
  return fault_smc(pb, range, N, seed, max_stages, max_batch, particles, likelihood, nstage, evidence,
                   gamma_trace, nullptr);
}

int smcf_fault_smc_exec_gsmc(smcf_problem* pb, const double* range, int N, uint64_t seed, int max_stages,
                             int max_batch, const double* qmean, const double* qcov_chol, const double* qcov_inv,
                             double log_qcov_det, double* particles, double* likelihood, int* nstage,
                             double* evidence, double* gamma_trace) {
  if (!qmean || !qcov_chol || !qcov_inv) return SMCF_EINVAL;
  Gsmc g{qmean, qcov_chol, qcov_inv, log_qcov_det};
  return fault_smc(pb, range, N, seed, max_stages, max_batch, particles, likelihood, nstage, evidence,
                   gamma_trace, &g);
}

}  // extern "C"

namespace smcf {
// the outer-stage mode of a problem (0 host, 1 / 2 device), for callers that run outer SMCs beside a stage engine
int outer_stage_mode(const smcf_problem* pb) { return outer_mode_of(pb); }
}  // namespace smcf
