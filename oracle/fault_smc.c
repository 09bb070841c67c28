/*
 * Oracle part 3: outer fault SMC (numprocs = 1).  Test infrastructure only (see oracle.h).
 *
 * Reference files followed (smc_fault.f90):
 *   29-64     fault_sample_init_particles (uniform branch)
 *   392-485   fault_find_next_gamma (CV 0.2, tol 1e-8, cap 1000)
 *   487-570   normalize / mean / covariance x 0.04 + dpotrf('L')
 *   572-590   fault_resample_particles (real(16) u)
 *   592-637   fault_reorder_to_send + sort.cpp qsort_
 *   639-783   work_mcmc_sampling       785-1075 fault_smc_exec
 * With keyed draws a parent with zero children has no effect on any output, so its
 * re-evaluation (smc_fault.f90:695) is skipped.
 */
#include "oracle.h"
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>

typedef __float128 q128;
enum { SITE_OUTER_INIT = 16, SITE_OUTER_RESAMPLE = 17, SITE_OUTER_PROP = 18, SITE_OUTER_MH = 19 };

void orc_fault_normalize_weights(double* w, int N) {
    double s = 0.0;
    for (int i = 0; i < N; ++i) s = s + w[i];
    for (int i = 0; i < N; ++i) w[i] = w[i] / s;
}

void orc_fault_calc_mean_particles(const double* x, const double* w, double* mean, int N, int nd) {
    for (int d = 0; d < nd; ++d) mean[d] = 0.0;
    for (int i = 0; i < N; ++i)
        for (int d = 0; d < nd; ++d) mean[d] = mean[d] + w[i] * x[(size_t)i * nd + d];
}

/* covariance (lower), x 0.04, then unblocked lower Cholesky; info as LAPACK */
void orc_fault_calc_cov_particles(const double* x, const double* w, const double* mean,
                                  double* cov, int N, int nd, int* info) {
    for (int k = 0; k < nd * nd; ++k) cov[k] = 0.0;
    for (int i = 0; i < N; ++i) {
        double wt = w[i];
        for (int j = 0; j < nd; ++j)
            for (int r = j; r < nd; ++r)
                cov[j * nd + r] = cov[j * nd + r] + wt * (x[(size_t)i * nd + r] - mean[r]) *
                                                    (x[(size_t)i * nd + j] - mean[j]);
    }
    for (int k = 0; k < nd * nd; ++k) cov[k] = cov[k] * 0.04;
    *info = 0;
    for (int j = 0; j < nd; ++j) {
        double ajj = cov[j * nd + j];
        for (int k = 0; k < j; ++k) ajj -= cov[k * nd + j] * cov[k * nd + j];
        if (ajj <= 0.0 || isnan(ajj)) { cov[j * nd + j] = ajj; *info = j + 1; return; }
        ajj = sqrt(ajj);
        cov[j * nd + j] = ajj;
        for (int r = j + 1; r < nd; ++r) {
            double t = cov[j * nd + r];
            for (int k = 0; k < j; ++k) t -= cov[k * nd + r] * cov[k * nd + j];
            cov[j * nd + r] = t * (1.0 / ajj);
        }
    }
}

void orc_fault_reorder(int* assigned, double* particles, double* likelihood, int N, int nd,
                       int* sorted_idx) {
    int* ta = (int*)malloc(sizeof(int) * N);
    double* tp = (double*)malloc(sizeof(double) * (size_t)N * nd);
    double* tl = (double*)malloc(sizeof(double) * N);
    memcpy(ta, assigned, sizeof(int) * N);
    memcpy(tp, particles, sizeof(double) * (size_t)N * nd);
    memcpy(tl, likelihood, sizeof(double) * N);
    for (int i = 0; i < N; ++i) sorted_idx[i] = i + 1;
    orc_qsort(assigned, sorted_idx, N);
    for (int i = 0; i < N; ++i) {
        int o = sorted_idx[i] - 1;
        assigned[i] = ta[o];
        memcpy(particles + (size_t)i * nd, tp + (size_t)o * nd, sizeof(double) * nd);
        likelihood[i] = tl[o];
    }
    free(ta); free(tp); free(tl);
}

static double draw_u(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2, uint32_t c1, int which) {
    uint64_t w[2];
    orc_draw_u64x2(key, site, stage, c2, c1, 0, w);
    return orc_uniform_raw(w[which]);
}

/* x := L x with L lower, non-unit (dtrmv 'L','N','N' column order) */
static void trmv_lower(int n, const double* A, double* x) {
    for (int j = n - 1; j >= 0; --j) {
        if (x[j] != 0.0) {
            double t = x[j];
            for (int i = n - 1; i > j; --i) x[i] += t * A[j * n + i];
            x[j] = x[j] * A[j * n + j];
        }
    }
}

/* Outer proposal of work_mcmc_sampling (smc_fault.f90:711-747): cand = cur + L z (dtrmv of the lower
 * Cholesky factor of 0.04 cov), then each dimension whose range is wider than 1e-3 is reflected once into
 * [lo, hi] (and clipped there); z is overwritten with L z. */
void orc_fault_propose(int nd, const double* chol, const double* cur, double* z, const double* range, double* cand) {
    trmv_lower(nd, chol, z);
    for (int d = 0; d < nd; ++d) cand[d] = cur[d] + z[d];
    for (int d = 0; d < nd; ++d) {
        double lo = range[2 * d], hi = range[2 * d + 1];
        if (hi - lo > 1e-3) {
            if (cand[d] < lo) cand[d] = fmin(2.0 * lo - cand[d], hi);
            if (cand[d] > hi) cand[d] = fmax(2.0 * hi - cand[d], lo);
        }
    }
}

/* fault_calc_neglog_q, smc_fault.f90:9-27 */
double orc_fault_calc_neglog_q(const double* x, const double* qmean, const double* qinv, double log_det, int nd) {
    double pi = 2.0 * asin(1.0), tmp = 0.0, delta[64];
    for (int i = 0; i < nd; ++i) delta[i] = x[i] - qmean[i];
    for (int j = 0; j < nd; ++j)
        for (int i = 0; i < nd; ++i) tmp = tmp + delta[i] * qinv[j * nd + i] * delta[j];
    return tmp / 2.0 + log(2.0 * pi) * nd / 2.0 + log_det / 2.0;
}

void orc_fault_smc_exec(const orc_problem* pb, const double* range, int N, uint64_t seed,
                        double* X, double* L, int max_stages, orc_fault_result* res) {
    orc_fault_smc_exec_gsmc(pb, range, N, seed, X, L, max_stages, res, NULL);
}

void orc_fault_smc_exec_gsmc(const orc_problem* pb, const double* range, int N, uint64_t seed,
                             double* X, double* L, int max_stages, orc_fault_result* res, const orc_gsmc* g) {
    int nd = pb->fault.fix_xbend ? 2 * pb->fault.nplane + 4 : 3 * pb->fault.nplane + 3;
    uint64_t okey = orc_splitmix64(seed ^ 0x5EEDF00Dull);
    double* Q = (double*)malloc(sizeof(double) * N);
    double* Qn = (double*)malloc(sizeof(double) * N);
    for (int i = 0; i < N; ++i) {
        if (g) {  /* qmean + chol(qcov) z, smc_fault.f90:41-52 */
            double z[64];
            for (int d = 0; d < nd; ++d) {
                uint64_t w[2];
                orc_draw_u64x2(okey, SITE_OUTER_INIT, 0, i, d, 0, w);
                z[d] = orc_box_muller_sin(orc_clamp_uniform(orc_uniform_raw(w[0])),
                                          orc_clamp_uniform(orc_uniform_raw(w[1])));
            }
            trmv_lower(nd, g->qcov_chol, z);
            for (int d = 0; d < nd; ++d) X[(size_t)i * nd + d] = z[d] + g->qmean[d];
        } else {
            for (int d = 0; d < nd; ++d) {
                double r = orc_clamp_uniform(draw_u(okey, SITE_OUTER_INIT, 0, i, d, 0));
                X[(size_t)i * nd + d] = range[2 * d] + r * (range[2 * d + 1] - range[2 * d]);
            }
        }
    }
    #pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < N; ++i)
        L[i] = orc_fault_calc_likelihood(pb, X + (size_t)i * nd, orc_inner_key(seed, 0, 0, i), NULL);
    for (int i = 0; i < N; ++i)
        Q[i] = g ? orc_fault_calc_neglog_q(X + (size_t)i * nd, g->qmean, g->qcov_inv, g->log_det, nd) : 0.0;
    double* w = (double*)malloc(sizeof(double) * N);
    double* mean = (double*)malloc(sizeof(double) * nd);
    double* cov = (double*)malloc(sizeof(double) * nd * nd);
    int* assigned = (int*)malloc(sizeof(int) * N);
    int* sidx = (int*)malloc(sizeof(int) * N);
    int* id_start = (int*)malloc(sizeof(int) * N);
    double* Xn = (double*)malloc(sizeof(double) * (size_t)N * nd);
    double* Ln = (double*)malloc(sizeof(double) * N);
    double* Lcur = (double*)malloc(sizeof(double) * N);
    double gamma = 0.0, evidence = 0.0;
    int iter = 1;
    while (1.0 - gamma > 1e-8 && iter <= max_stages) {
        double gprev = gamma, ev;
        if (g) {  /* tempering of L - q (smc_fault.f90:411-416, 433-438, 479-481) */
            for (int i = 0; i < N; ++i) Ln[i] = L[i] - Q[i];
            gamma = orc_find_next_gamma(gprev, Ln, N, w, &ev, 0.2, 1e-8, 1000);
        } else {
            gamma = orc_find_next_gamma(gprev, L, N, w, &ev, 0.2, 1e-8, 1000);
        }
        evidence = evidence + ev;
        if (res && res->gamma_trace) res->gamma_trace[iter - 1] = gamma;
        orc_fault_normalize_weights(w, N);
        orc_fault_calc_mean_particles(X, w, mean, N, nd);
        int info;
        orc_fault_calc_cov_particles(X, w, mean, cov, N, nd, &info);
        uint64_t wq[2];
        orc_draw_u64x2(okey, SITE_OUTER_RESAMPLE, iter, 0, 0, 0, wq);
        orc_resample_quad(N, w, (double)(wq[0] >> 11) * 0x1.0p-53, (double)(wq[1] >> 11) * 0x1.0p-106,
                          assigned);
        orc_fault_reorder(assigned, X, L, N, nd, sidx);
        for (int i = 0; i < N; ++i) Qn[i] = Q[i];
        for (int i = 0; i < N; ++i) Q[i] = Qn[sidx[i] - 1];
        id_start[0] = 0;
        for (int i = 0; i < N - 1; ++i) id_start[i + 1] = id_start[i] + assigned[i];
        /* wave 0: parent re-evaluations (only parents with children matter) */
        #pragma omp parallel for schedule(dynamic)
        for (int i = 0; i < N; ++i)
            if (assigned[i] > 0)
                Lcur[i] = orc_fault_calc_likelihood(pb, X + (size_t)i * nd, orc_inner_key(seed, iter, 0, i), NULL);
        #pragma omp parallel for schedule(dynamic)
        for (int i = 0; i < N; ++i) {
            double cur[64], cand[64], z[64];
            memcpy(cur, X + (size_t)i * nd, sizeof(double) * nd);
            double lc = Lcur[i], qc = Q[i];
            for (int j = id_start[i]; j < id_start[i] + assigned[i]; ++j) {
                for (int d = 0; d < nd; ++d) {
                    uint64_t wz[2];
                    orc_draw_u64x2(okey, SITE_OUTER_PROP, iter, j, d, 0, wz);
                    z[d] = orc_box_muller_sin(orc_clamp_uniform(orc_uniform_raw(wz[0])),
                                              orc_clamp_uniform(orc_uniform_raw(wz[1])));
                }
                orc_fault_propose(nd, cov, cur, z, range, cand);
                double lcand = orc_fault_calc_likelihood(pb, cand, orc_inner_key(seed, iter, 1, j), NULL);
                double qcand = g ? orc_fault_calc_neglog_q(cand, g->qmean, g->qcov_inv, g->log_det, nd) : 0.0;
                double mh = orc_clamp_uniform(draw_u(okey, SITE_OUTER_MH, iter, j, 0, 0));
                int acc = g ? (gamma * (lc - lcand) + (1.0 - gamma) * (qc - qcand) > log(mh))
                            : (gamma * (lc - lcand) > log(mh));
                if (acc) {
                    memcpy(cur, cand, sizeof(double) * nd);
                    lc = lcand;
                    qc = qcand;
                }
                Ln[j] = lc;
                Qn[j] = qc;
                memcpy(Xn + (size_t)j * nd, cur, sizeof(double) * nd);
            }
        }
        memcpy(X, Xn, sizeof(double) * (size_t)N * nd);
        memcpy(L, Ln, sizeof(double) * N);
        memcpy(Q, Qn, sizeof(double) * N);
        iter++;
    }
    free(Q); free(Qn);
    if (res) { res->nstage = iter - 1; res->evidence = evidence; }
    free(w); free(mean); free(cov); free(assigned); free(sidx); free(id_start); free(Xn); free(Ln); free(Lcur);
}
