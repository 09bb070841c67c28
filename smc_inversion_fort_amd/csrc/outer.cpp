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

int fault_smc(smcf_problem* pb, const double* range, int N, uint64_t seed, int max_stages, int max_batch,
              double* particles, double* likelihood, int* nstage, double* evidence, double* gamma_trace,
              const Gsmc* g) {
  int nnode, ndof, nds, nd, nobs;
  int rc = smcf_problem_dims(pb, &nnode, &ndof, &nds, &nd, &nobs);
  if (rc) return rc;
  if (!range || N < 1 || max_batch < 1 || !particles || !likelihood || nd > 64) return SMCF_EINVAL;
  uint64_t okey = splitmix64(seed ^ 0x5EEDF00Dull);
  std::vector<double> X((size_t)nd * N), L(N), w(N), Xn((size_t)nd * N), Ln(N), Lcur(N), cov(nd * nd), mean(nd);
  std::vector<double> Q(N, 0.0), Qn(N, 0.0), R(N);
  std::vector<int> assigned(N), sidx(N), id_start(N);
  std::vector<double> chol;
  if (g) chol.assign(g->qcov_chol, g->qcov_chol + (size_t)nd * nd);
  for (int i = 0; i < N; ++i) {
    if (g) {  // fault_sample_init_particles, gsmc branch: qmean + chol(qcov) z
      double z[64];
      for (int d = 0; d < nd; ++d) {
        uint64_t ww[2];
        Philox::draw(okey, kOuterInit, 0, i, d, 0, ww);
        z[d] = box_muller(ww[0], ww[1]);
      }
      trmv_lower(nd, chol, z);
      for (int d = 0; d < nd; ++d) X[(size_t)i * nd + d] = z[d] + g->qmean[d];
    } else {
      for (int d = 0; d < nd; ++d) {
        uint64_t ww[2];
        Philox::draw(okey, kOuterInit, 0, i, d, 0, ww);
        double r = clampu(u01(ww[0]));
        X[(size_t)i * nd + d] = range[2 * d] + r * (range[2 * d + 1] - range[2 * d]);
      }
    }
  }
  // batched evaluation helper: calls of at most max_batch geometries, each streamed through the slots
  auto eval = [&](const std::vector<double>& th, const std::vector<uint64_t>& keys, std::vector<double>& out) -> int {
    int B = (int)keys.size();
    out.resize(B);
    for (int s = 0; s < B; s += max_batch) {
      int e = std::min(B, s + max_batch);
      int r = smcf_fault_calc_likelihood_batch(pb, e - s, th.data() + (size_t)s * nd, keys.data() + s,
                                               out.data() + s, nullptr);
      if (r) return r;
    }
    return SMCF_OK;
  };
  {
    std::vector<uint64_t> keys(N);
    for (int i = 0; i < N; ++i) keys[i] = smcf_inner_key(seed, 0, 0, i);
    std::vector<double> out;
    if ((rc = eval(X, keys, out))) return rc;
    L = out;
  }
  if (g)
    for (int i = 0; i < N; ++i) Q[i] = neglog_q(&X[(size_t)i * nd], *g, nd);
  const std::string out_dir = output_dir_of(pb);
  if ((rc = write_stage(out_dir, 0, N, nd, X, L))) return rc;
  double gamma = 0.0, ev_total = 0.0;
  int iter = 1;
  while (1.0 - gamma > 1e-8 && iter <= max_stages) {
    double gprev = gamma, ev;
    if (g) {  // tempering of L - q
      for (int i = 0; i < N; ++i) R[i] = L[i] - Q[i];
      gamma = find_next_gamma(gprev, R, w, &ev);
    } else {
      gamma = find_next_gamma(gprev, L, w, &ev);
    }
    ev_total = ev_total + ev;
    if (gamma_trace) gamma_trace[iter - 1] = gamma;
    // fault_normalize_weights / fault_calc_mean_particles / fault_calc_cov_particles
    double s = 0.0;
    for (int i = 0; i < N; ++i) s = s + w[i];
    for (int i = 0; i < N; ++i) w[i] = w[i] / s;
    for (int d = 0; d < nd; ++d) mean[d] = 0.0;
    for (int i = 0; i < N; ++i)
      for (int d = 0; d < nd; ++d) mean[d] = mean[d] + w[i] * X[(size_t)i * nd + d];
    std::fill(cov.begin(), cov.end(), 0.0);
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < nd; ++j)
        for (int r = j; r < nd; ++r)
          cov[j * nd + r] = cov[j * nd + r] + w[i] * (X[(size_t)i * nd + r] - mean[r]) * (X[(size_t)i * nd + j] - mean[j]);
    for (auto& c : cov) c = c * 0.04;
    for (int j = 0; j < nd; ++j) {  // dpotrf('L') unblocked; stops at a non-positive pivot (info ignored)
      double ajj = cov[j * nd + j];
      for (int k = 0; k < j; ++k) ajj -= cov[k * nd + j] * cov[k * nd + j];
      if (ajj <= 0.0 || std::isnan(ajj)) { cov[j * nd + j] = ajj; break; }
      ajj = std::sqrt(ajj);
      cov[j * nd + j] = ajj;
      for (int r = j + 1; r < nd; ++r) {
        double t = cov[j * nd + r];
        for (int k = 0; k < j; ++k) t -= cov[k * nd + r] * cov[k * nd + j];
        cov[j * nd + r] = t * (1.0 / ajj);
      }
    }
    // fault_resample_particles with a real(16) running u
    {
      uint64_t wq[2];
      Philox::draw(okey, kOuterResample, iter, 0, 0, 0, wq);
      __float128 u = (__float128)((double)(wq[0] >> 11) * 0x1.0p-53) + (__float128)((double)(wq[1] >> 11) * 0x1.0p-106);
      if (u < (__float128)1e-5) u = (__float128)1e-5;
      __float128 dn = (__float128)N;
      u = u / dn;
      for (int i = 0; i < N; ++i) {
        __float128 t = ((__float128)w[i] - u) * dn;
        long long f = (long long)t;
        if ((__float128)f > t) f -= 1;  // floor
        int a = (int)f + 1;
        assigned[i] = a;
        u = u + (__float128)a / dn - (__float128)w[i];
      }
    }
    // fault_reorder_to_send (numprocs = 1): descending by assigned_num through qsort_
    {
      std::vector<int> ta = assigned;
      std::vector<double> tx = X, tl = L, tq = Q;
      for (int i = 0; i < N; ++i) sidx[i] = i + 1;
      int n = N;
      qsort_(assigned.data(), sidx.data(), &n);
      for (int i = 0; i < N; ++i) {
        int o = sidx[i] - 1;
        assigned[i] = ta[o];
        std::memcpy(&X[(size_t)i * nd], &tx[(size_t)o * nd], sizeof(double) * nd);
        L[i] = tl[o];
        Q[i] = tq[o];
      }
    }
    id_start[0] = 0;
    for (int i = 0; i < N - 1; ++i) id_start[i + 1] = id_start[i] + assigned[i];
    int total = id_start[N - 1] + assigned[N - 1];
    if (total != N) return SMCF_ESTATE;
    // wave 0: parents with children
    std::vector<double> th;
    std::vector<uint64_t> keys;
    std::vector<int> who;
    for (int i = 0; i < N; ++i)
      if (assigned[i] > 0) {
        th.insert(th.end(), &X[(size_t)i * nd], &X[(size_t)i * nd] + nd);
        keys.push_back(smcf_inner_key(seed, iter, 0, i));
        who.push_back(i);
      }
    std::vector<double> out;
    if ((rc = eval(th, keys, out))) return rc;
    for (size_t q = 0; q < who.size(); ++q) Lcur[who[q]] = out[q];
    std::vector<double> cur = X, qcur = Q;
    int amax = *std::max_element(assigned.begin(), assigned.end());
    for (int k = 0; k < amax; ++k) {
      th.clear(); keys.clear(); who.clear();
      std::vector<int> slot;
      for (int i = 0; i < N; ++i) {
        if (assigned[i] <= k) continue;
        int j = id_start[i] + k;
        double z[64];
        for (int d = 0; d < nd; ++d) {
          uint64_t wz[2];
          Philox::draw(okey, kOuterProp, iter, j, d, 0, wz);
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
        keys.push_back(smcf_inner_key(seed, iter, 1, j));
        who.push_back(i);
        slot.push_back(j);
      }
      if ((rc = eval(th, keys, out))) return rc;
      for (size_t q = 0; q < who.size(); ++q) {
        int i = who[q], j = slot[q];
        uint64_t wm[2];
        Philox::draw(okey, kOuterMH, iter, j, 0, 0, wm);
        double mh = clampu(u01(wm[0]));
        bool acc;
        double qc = 0.0;
        if (g) {
          qc = neglog_q(&th[q * nd], *g, nd);
          acc = gamma * (Lcur[i] - out[q]) + (1.0 - gamma) * (qcur[i] - qc) > std::log(mh);
        } else {
          acc = gamma * (Lcur[i] - out[q]) > std::log(mh);
        }
        if (acc) {
          std::memcpy(&cur[(size_t)i * nd], &th[q * nd], sizeof(double) * nd);
          Lcur[i] = out[q];
          qcur[i] = qc;
        }
        Ln[j] = Lcur[i];
        Qn[j] = qcur[i];
        std::memcpy(&Xn[(size_t)j * nd], &cur[(size_t)i * nd], sizeof(double) * nd);
      }
    }
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
