/*
 * Oracle part 2: inner slip SMC.  Test infrastructure only (see oracle.h).
 *
 * Reference files followed (smc_slip.f90 unless noted):
 *   8-54      RNG correction, Box-Muller, cdf_norm / pdf_norm
 *   56-104    slip_calc_likelihood (Gram form)
 *   106-137   slip_calc_prior       139-182  slip_calc_grad
 *   184-287   slip_gen_init_particles (sequential Gibbs chain + Newton)
 *   289-387   slip_calc_mean_std_vector / slip_find_next_gamma
 *   465-589   normalize (real(16)), mean/cov-diag (real(16)), resampling (real(16) u)
 *   591-917   slip_hmc_sampling / slip_leapfrog
 *   919-1214  slip_hmc_tuning
 *   1347-1443 slip_calc_gsd_gsg      1445-1663 slip_smc_exec
 *   smc_fault.f90:66-292 fault_calc_likelihood
 * The Gibbs-sweep mover (mover=1) has no reference counterpart: it is the
 * truncated-MVN Gibbs mover of the benchmark configuration, specified in
 * orc_gibbs_move below and restated identically on the device.  Mover 2 (HMC with
 * hmc_ntau leapfrog steps of the reference's slip_leapfrog and a step size adapted
 * toward a target acceptance, ORC_HMC_ADAPT_GAIN in oracle.h) has no reference
 * counterpart either: it replaces the ESJD tuning pass of slip_hmc_tuning.
 */
#include "oracle.h"
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

typedef __float128 q128;

enum { SITE_GIBBS_INIT = 1, SITE_RESAMPLE = 2, SITE_TUNE_DTAU = 3, SITE_TUNE_MOM = 4,
       SITE_TUNE_MH = 5, SITE_TUNE_RESAMPLE = 6, SITE_HMC_MOM = 7, SITE_HMC_MH = 8,
       SITE_GIBBS_MOVE = 9 };

static double draw_uniform(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2, uint32_t c1,
                           int which) {
    uint64_t w[2];
    orc_draw_u64x2(key, site, stage, c2, c1, 0, w);
    return orc_uniform_raw(w[which]);
}
static double draw_bm(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2, uint32_t c1) {
    uint64_t w[2];
    orc_draw_u64x2(key, site, stage, c2, c1, 0, w);
    return orc_box_muller_sin(orc_clamp_uniform(orc_uniform_raw(w[0])),
                              orc_clamp_uniform(orc_uniform_raw(w[1])));
}
double orc_draw_normal(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2, uint32_t c1) {
    return draw_bm(key, site, stage, c2, c1);
}
double orc_draw_clamped_uniform(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2, uint32_t c1,
                                uint32_t word) {
    return orc_clamp_uniform(draw_uniform(key, site, stage, c2, c1, (int)word));
}
static void draw_quad(uint64_t key, uint32_t site, uint32_t stage, double* hi, double* lo) {
    uint64_t w[2];
    orc_draw_u64x2(key, site, stage, 0, 0, 0, w);
    *hi = (double)(w[0] >> 11) * 0x1.0p-53;
    *lo = (double)(w[1] >> 11) * 0x1.0p-106;
}

/* ---------------- InSAR covariance, init.f90:82-119 ----------------
 * Per path p (rows a .. a+n-1, a = nsar_index(p) - 1): C(i, j) = s_i s_j exp(-dist_ij / 10) with dist the
 * distance of the two points, and s_i = obs_sigma(nsar_index(p) + i) for 1-based i -- one row past the
 * observation's own (init.f90:107-108), so the last row of a path takes the next row's sigma.  The
 * inverse, column-major, is written to sinv_out block after block; it is formed by LU with partial
 * pivoting and one forward/back substitution per unit column (the factors of dgetrf; dgetri applies the
 * same factors).  Returns 0, or -1 when a path's last sigma index lies past nobs (the reference reads
 * past the array there) or a pivot is zero. */
int orc_calc_sigma_sar_mat(int npath, const int* nsar_index, int nobs, const double* obs_points,
                           const double* obs_sigma, double* sinv_out) {
    size_t off = 0;
    for (int ip = 0; ip < npath; ++ip) {
        int a = nsar_index[ip] - 1, n = nsar_index[ip + 1] - nsar_index[ip];
        if (n <= 0) continue;
        if (a + n >= nobs) return -1;
        double* A = (double*)malloc(sizeof(double) * n * n);
        int* piv = (int*)malloc(sizeof(int) * n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double dx = obs_points[2 * (a + i)] - obs_points[2 * (a + j)];
                double dy = obs_points[2 * (a + i) + 1] - obs_points[2 * (a + j) + 1];
                double dist = sqrt(dx * dx + dy * dy);
                A[(size_t)j * n + i] = obs_sigma[a + 1 + i] * obs_sigma[a + 1 + j] * exp(-dist / 1e1);
            }
        /* LU, row pivoting: P A = L U, unit L below the diagonal */
        for (int k = 0; k < n; ++k) {
            int pr = k;
            double best = fabs(A[(size_t)k * n + k]);
            for (int i = k + 1; i < n; ++i)
                if (fabs(A[(size_t)k * n + i]) > best) { best = fabs(A[(size_t)k * n + i]); pr = i; }
            piv[k] = pr;
            if (best == 0.0) { free(A); free(piv); return -1; }
            if (pr != k)
                for (int j = 0; j < n; ++j) {
                    double t = A[(size_t)j * n + k]; A[(size_t)j * n + k] = A[(size_t)j * n + pr]; A[(size_t)j * n + pr] = t;
                }
            double rp = 1.0 / A[(size_t)k * n + k];
            for (int i = k + 1; i < n; ++i) A[(size_t)k * n + i] *= rp;
            for (int j = k + 1; j < n; ++j) {
                double u = A[(size_t)j * n + k];
                if (u != 0.0)
                    for (int i = k + 1; i < n; ++i) A[(size_t)j * n + i] -= A[(size_t)k * n + i] * u;
            }
        }
        double* X = sinv_out + off;
        for (int c = 0; c < n; ++c) {
            double* x = X + (size_t)c * n;
            for (int i = 0; i < n; ++i) x[i] = (i == c) ? 1.0 : 0.0;
            for (int k = 0; k < n; ++k)
                if (piv[k] != k) { double t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
            for (int j = 0; j < n; ++j)           /* L y = P e_c */
                for (int i = j + 1; i < n; ++i) x[i] -= A[(size_t)j * n + i] * x[j];
            for (int j = n - 1; j >= 0; --j) {    /* U x = y */
                x[j] /= A[(size_t)j * n + j];
                for (int i = 0; i < j; ++i) x[i] -= A[(size_t)j * n + i] * x[j];
            }
        }
        free(A); free(piv);
        off += (size_t)n * n;
    }
    return 0;
}

/* ---------------- Gram quantities, smc_slip.f90:1347-1443 ---------------- */
void orc_slip_calc_gsd_gsg(const double* gmat, const double* dvec, int nobs, int ndim, int npath,
                           const int* nsar_index, int nsar_total, const double* sar_sinv,
                           double sigma_sar2, double sigma_gnss2, const double* obs_sigma,
                           double* gsdvec, double* gsgmat, double* dsd) {
    double* sd = (double*)calloc(nobs, sizeof(double));
    double* gts = (double*)calloc((size_t)ndim * nobs, sizeof(double));
    size_t off = 0;
    for (int ip = 0; ip < npath; ++ip) {
        int n = nsar_index[ip + 1] - nsar_index[ip];
        int r0 = nsar_index[ip] - 1;
        const double* A = sar_sinv + off;
        /* dgemv('n'): y = alpha*A*x, column-oriented accumulation */
        for (int j = 0; j < n; ++j) {
            double t = (1.0 / sigma_sar2) * dvec[r0 + j];
            for (int i = 0; i < n; ++i) sd[r0 + i] += t * A[(size_t)j * n + i];
        }
        /* gts(:, r0+c) = (1/s2) * G_path^T * Sinv(:, c) */
        for (int c = 0; c < n; ++c)
            for (int d = 0; d < ndim; ++d) {
                double s = 0.0;
                for (int k = 0; k < n; ++k) s += gmat[(size_t)d * nobs + r0 + k] * A[(size_t)c * n + k];
                gts[(size_t)(r0 + c) * ndim + d] = (1.0 / sigma_sar2) * s;
            }
        off += (size_t)n * n;
    }
    for (int io = nsar_total; io < nobs; ++io)
        sd[io] = dvec[io] / (sigma_gnss2 * (obs_sigma[io] * obs_sigma[io]));
    for (int d = 0; d < ndim; ++d) {
        double s = 0.0;
        for (int io = 0; io < nobs; ++io) s = s + gmat[(size_t)d * nobs + io] * sd[io];
        gsdvec[d] = s;
    }
    double t = 0.0;
    for (int io = 0; io < nobs; ++io) t = t + dvec[io] * sd[io];
    *dsd = t;
    for (int io = nsar_total; io < nobs; ++io)
        for (int d = 0; d < ndim; ++d)
            gts[(size_t)io * ndim + d] = gmat[(size_t)d * nobs + io] / ((obs_sigma[io] * obs_sigma[io]) * sigma_gnss2);
    for (int j = 0; j < ndim; ++j)
        for (int i = 0; i < ndim; ++i) {
            double s = 0.0;
            for (int k = 0; k < nobs; ++k) s += gts[(size_t)k * ndim + i] * gmat[(size_t)j * nobs + k];
            gsgmat[(size_t)j * ndim + i] = s;
        }
    free(gts); free(sd);
}

/* ---------------- likelihood / prior / gradient ---------------- */
double orc_slip_calc_likelihood(const double* s, int ndim, const double* H, const double* b,
                                double dsd, int nsar_total, int ngnss, double lss2, double lsg2) {
    double v = 0.0;
    for (int i = 0; i < ndim; ++i)
        for (int j = 0; j < i; ++j) v = v + s[i] * H[(size_t)i * ndim + j] * s[j];
    v = v * 2.0;
    for (int i = 0; i < ndim; ++i) v = v + s[i] * H[(size_t)i * ndim + i] * s[i];
    double bs = 0.0;
    for (int i = 0; i < ndim; ++i) bs = bs + b[i] * s[i];
    return (v - 2.0 * bs + dsd) / 2.0 + nsar_total * lss2 / 2.0 + 3.0 * ngnss * lsg2 / 2.0;
}

double orc_slip_calc_prior(const double* s, const double* alpha2_full, const int* li,
                           const double* lv, int nnode, double* lsvec) {
    int n = 2 * nnode;
    for (int i = 0; i < n; ++i) {
        double t = 0.0;
        for (int j = 0; j < 9; ++j) t = t + lv[9 * i + j] * s[li[9 * i + j] - 1];
        lsvec[i] = t;
    }
    double p = 0.0;
    for (int i = 0; i < n; ++i) p = p + (lsvec[i] * lsvec[i]) / (2.0 * alpha2_full[i]);
    return p;
}

void orc_slip_calc_grad(double* grad, const double* s, int ndim, int nnode, int ndof,
                        const int* li, const double* lv, const int* lti, const double* ltv,
                        double gamma, const double* alpha2_full, const double* b,
                        const double* H, double* lsvec) {
    double* g1 = (double*)calloc(ndim, sizeof(double));
    for (int j = 0; j < ndim; ++j) {
        double t = s[j];
        for (int i = 0; i < ndim; ++i) g1[i] += t * H[(size_t)j * ndim + i];
    }
    for (int i = 0; i < ndim; ++i) { g1[i] = g1[i] - b[i]; g1[i] = g1[i] * gamma; }
    for (int n = 0; n < 2 * nnode; ++n) {
        double t = 0.0;
        for (int j = 0; j < 9; ++j) t = t + lv[9 * n + j] * s[li[9 * n + j] - 1];
        lsvec[n] = t / alpha2_full[n];
    }
    for (int d = 0; d < 2 * ndof; ++d) {
        double t = 0.0;
        for (int j = 0; j < 9; ++j) t = t + ltv[9 * d + j] * lsvec[lti[9 * d + j] - 1];
        grad[d] = g1[d] + t;
    }
    free(g1);
}

double orc_cdf_norm(double x, double mu, double sigma2) {
    return (1.0 + erf((x - mu) / sqrt(2.0 * sigma2))) / 2.0;
}
double orc_pdf_norm(double x, double mu, double sigma2) {
    double pi = 2.0 * asin(1.0);
    return exp(-((x - mu) * (x - mu)) / (2.0 * sigma2)) / sqrt(2.0 * pi * sigma2);
}

/* ---------------- prior draw: one sequential Gibbs chain, smc_slip.f90:209-264 ---------------- */
void orc_slip_gibbs_init_chain(double* particles, int P, int ndim, const double* ll,
                               const double* alpha2_full, double max_slip, const double* bm) {
    double* cur = (double*)calloc(ndim, sizeof(double));
    for (int k = 0; k < P; ++k) {
        for (int i = 0; i < ndim; ++i) {
            double s2 = 1.0 / (ll[(size_t)i * ndim + i] / alpha2_full[i]);
            double mu = 0.0;
            for (int j = 0; j < ndim; ++j) {
                if (j == i) continue;
                double lij = ll[(size_t)j * ndim + i];
                if (lij == 0.0) continue; /* exact: subtracting a signed zero leaves mu unchanged */
                mu = mu - lij / alpha2_full[i] * cur[j] * s2;
            }
            double v1 = bm[(size_t)k * ndim + i];
            double x = mu + v1 * sqrt(s2);
            double fx = orc_cdf_norm(x, mu, s2);
            double f1 = orc_cdf_norm(max_slip, mu, s2);
            double f0 = orc_cdf_norm(0.0, mu, s2);
            double y = mu, err = 1.0, yp = y + err;
            int cnt = 0;
            double target = (f1 - f0) * fx + f0;
            while (err > 1e-8 && cnt < 100) {
                y = y - (orc_cdf_norm(y, mu, s2) - target) / orc_pdf_norm(y, mu, s2);
                err = fabs(y - yp);
                yp = y;
                cnt++;
            }
            cur[i] = y;
        }
        memcpy(particles + (size_t)k * ndim, cur, sizeof(double) * ndim);
    }
    free(cur);
}

/* ---------------- tempering: smc_slip.f90:289-387 / smc_fault.f90:372-485 ---------------- */
double orc_find_next_gamma(double gamma_prev, const double* L, int P, double* w,
                           double* neglog_evidence, double cv_threshold, double tol, int cap) {
    double lmin = L[0];
    for (int i = 0; i < P; ++i) if (L[i] < lmin) lmin = L[i];
    double lower = gamma_prev, upper = 1.0, err = 1.0, gamma = 1.0;
    int cnt = 0;
    while (err > tol) {
        gamma = (lower + upper) / 2.0;
        double dg = gamma - gamma_prev;
        for (int i = 0; i < P; ++i) w[i] = exp(-dg * (L[i] - lmin));
        double mean = 0.0;
        for (int i = 0; i < P; ++i) mean = mean + w[i];
        mean = mean / P;
        double sd = 0.0;
        for (int i = 0; i < P; ++i) sd = sd + (w[i] - mean) * (w[i] - mean);
        sd = sqrt(sd / P);
        double cv = sd / mean;
        if (cv > cv_threshold) upper = gamma; else lower = gamma;
        err = fabs(cv - cv_threshold);
        if (fabs(gamma - 1.0) < tol) break;
        if (cnt > cap) break;
        cnt++;
    }
    double ev = 0.0, dg = gamma - gamma_prev;
    for (int i = 0; i < P; ++i) ev = ev + exp(-dg * (L[i] - lmin));
    ev = ev / P;
    *neglog_evidence = -log(ev) + dg * lmin;
    return gamma;
}

void orc_slip_normalize_weights(double* w, int P) {
    q128 s = 0;
    for (int i = 0; i < P; ++i) s = s + (q128)w[i];
    for (int i = 0; i < P; ++i) w[i] = (double)((q128)w[i] / s);
}

void orc_slip_calc_mean_particles(const double* x, const double* w, double* mean, int P, int nd) {
    q128* m = (q128*)calloc(nd, sizeof(q128));
    for (int i = 0; i < P; ++i)
        for (int d = 0; d < nd; ++d) m[d] = m[d] + (q128)(w[i] * x[(size_t)i * nd + d]);
    for (int d = 0; d < nd; ++d) mean[d] = (double)m[d];
    free(m);
}

void orc_slip_calc_cov_particles(const double* x, const double* w, const double* mean,
                                 double* cov_diag, int P, int nd) {
    q128* c = (q128*)calloc(nd, sizeof(q128));
    for (int i = 0; i < P; ++i)
        for (int d = 0; d < nd; ++d) {
            double di = x[(size_t)i * nd + d] - mean[d];
            c[d] = c[d] + (q128)(w[i] * di * di);
        }
    for (int d = 0; d < nd; ++d) {
        cov_diag[d] = (double)c[d];
        if (fabs(cov_diag[d]) < 1e-3) cov_diag[d] = 1e-3;
    }
    free(c);
}

/* systematic residual resampling with a real(16) running u, smc_slip.f90:572-589 */
void orc_resample_quad(int P, const double* w, double u_hi, double u_lo, int* assigned) {
    q128 u = (q128)u_hi + (q128)u_lo;
    if (u < (q128)1e-5) u = (q128)1e-5;
    q128 dn = (q128)P;
    u = u / dn;
    for (int i = 0; i < P; ++i) {
        q128 t = ((q128)w[i] - u) * dn;
        int a = (int)floorq(t) + 1;
        assigned[i] = a;
        u = u + (q128)a / dn - (q128)w[i];
    }
}

/* ---------------- standard normal quantile: Wichura (1988) AS 241 PPND16 ---------------- */
double orc_ndtri(double p) {
    double q = p - 0.5, r, z;
    if (fabs(q) <= 0.425) {
        r = 0.180625 - q * q;
        z = q * (((((((2509.0809287301226727 * r + 33430.575583588128105) * r + 67265.770927008700853) * r +
                     45921.953931549871457) * r + 13731.693765509461125) * r + 1971.5909503065514427) * r +
                  133.14166789178437745) * r + 3.387132872796366608) /
            (((((((5226.495278852545925 * r + 28729.085735721942674) * r + 39307.89580009271061) * r +
                 21213.794301586595867) * r + 5394.1960214247511077) * r + 687.1870074920579083) * r +
              42.313330701600911252) * r + 1.0);
        return z;
    }
    r = q < 0 ? p : 1.0 - p;
    r = sqrt(-log(r));
    if (r <= 5.0) {
        r -= 1.6;
        z = (((((((7.7454501427834140764e-4 * r + 0.0227238449892691845833) * r + 0.24178072517745061177) * r +
                 1.27045825245236838258) * r + 3.64784832476320460504) * r + 5.7694972214606914055) * r +
              4.6303378461565452959) * r + 1.42343711074968357734) /
            (((((((1.05075007164441684324e-9 * r + 5.475938084995344946e-4) * r + 0.0151986665636164571966) * r +
                 0.14810397642748007459) * r + 0.68976733498510000455) * r + 1.6763848301838038494) * r +
              2.05319162663775882187) * r + 1.0);
    } else {
        r -= 5.0;
        z = (((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r + 0.0012426609473880784386) * r +
                 0.026532189526576123093) * r + 0.29656057182850489123) * r + 1.7848265399172913358) * r +
              5.4637849111641143699) * r + 6.6579046435011037772) /
            (((((((2.04426310338993978564e-15 * r + 1.4215117583164458887e-7) * r + 1.8463183175100546818e-5) * r +
                 7.868691311456132591e-4) * r + 0.0148753612908506148525) * r + 0.13692988092273580531) * r +
              0.59983220655588793769) * r + 1.0);
    }
    return q < 0 ? -z : z;
}

/* log Q(x) (standard normal survival) and Mills ratio Q/phi for x >= 35 from the
 * asymptotic series of the Mills ratio, R(x) = (1/x)(1 - x^-2 + 3x^-4 - 15x^-6 + ...) */
static double mills_series(double x) {
    double r = 1.0 / (x * x);
    return 1.0 + r * (-1.0 + r * (3.0 + r * (-15.0 + r * (105.0 + r * (-945.0 + r * (10395.0 + r * -135135.0))))));
}
static double logq_far(double x) {
    return -0.5 * x * x - log(x) - 0.91893853320467274178 + log(mills_series(x));
}

/* upper-tail truncation beyond 35 sigma: solve log Q(z) = log Q(a) + log(1 - u (1 - Q(b)/Q(a)))
 * by Newton (d/dz log Q = -1/R) from the exponential-tail start */
static double tn_far(double a, double b, double u) {
    double la = logq_far(a), lb = logq_far(b);
    double T = la + log1p(-u * (-expm1(lb - la)));
    double z = a - log1p(-u * (-expm1(-a * (b - a)))) / a;
    for (int it = 0; it < 4; ++it) {
        double g = logq_far(z) - T;
        z = z + g * (mills_series(z) / z);
    }
    return fmin(fmax(z, a), b);
}

/* standard normal truncated to [a, b], by inversion with tail-aware probabilities.
 * One erfc per bound gives t = Q(|x|), from which both Phi(x) and Q(x) follow without
 * cancellation; a bound is skipped when its tail is below 1e-20 of the interval's mass
 * (x <= -10 with x^2 > y^2 + 100 for the lower bound, mirrored for the upper).  The target
 * p = Phi(a) + u m and its complement q = Q(b) + (1 - u) m (m the mass) are both formed
 * directly, and the smaller one is inverted: z = ndtri(p) or z = -ndtri(q).  Beyond 35 sigma
 * (Q underflows near 38) the Mills-ratio Newton solve takes over. */
static double tn_std(double a, double b, double u) {
    if (a > 35.0) return tn_far(a, b, u);
    if (b < -35.0) return -tn_far(-b, -a, 1.0 - u);  /* mirror: quantile u of z is 1-u of -z */
    double ta = (a <= -10.0 && a * a > b * b + 100.0) ? 0.0 : 0.5 * erfc(fabs(a) * M_SQRT1_2);
    double tb = (b >= 10.0 && b * b > a * a + 100.0) ? 0.0 : 0.5 * erfc(fabs(b) * M_SQRT1_2);
    double pa = a < 0.0 ? ta : 1.0 - ta, qa = a < 0.0 ? 1.0 - ta : ta;
    double pb = b < 0.0 ? tb : 1.0 - tb, qb = b < 0.0 ? 1.0 - tb : tb;
    double m = a >= 0.0 ? qa - qb : pb - pa;
    double p = pa + u * m, q = qb + (1.0 - u) * m;
    double z = p <= q ? orc_ndtri(p) : -orc_ndtri(q);
    return fmin(fmax(z, a), b);
}

double orc_trunc_normal(double mu, double sigma, double lo, double hi, double u) {
    double a = (lo - mu) / sigma, b = (hi - mu) / sigma;
    double x = mu + sigma * tn_std(a, b, u);
    return fmin(fmax(x, lo), hi);
}

/* the same draw parameterised by the square root of the precision, as the Gibbs mover uses it:
 * a = (lo - mu) s, b = (hi - mu) s, x = mu + z (1 / s) */
double orc_trunc_normal_prec(double mu, double s, double lo, double hi, double u) {
    double a = (lo - mu) * s, b = (hi - mu) * s;
    double x = mu + tn_std(a, b, u) * (1.0 / s);
    return fmin(fmax(x, lo), hi);
}

/* Single-precision draw of the mixed-precision Gibbs mover (no reference counterpart; the device
 * restates it in trunc_normal_prec_f32).  Inverse CDF, p in (0, 1/2], returning z <= 0 with Phi(z) = p:
 * for p > 0.0017 (w = -ln(4 p (1 - p)) < 5), z = sqrt(2) erfinv(2p - 1) from the central polynomial of
 * Giles' single-precision erfinv (M. Giles, "Approximating the erfinv function", GPU Computing Gems Jade
 * Edition, 2011); below that the tail of AS 241 PPND7 (Wichura 1988, Appl. Statist. 37, 477-484), the
 * 7-digit version of orc_ndtri. */
float orc_ndtri7_lower(float p) {
    float w = -logf(4.0f * p * (1.0f - p));
    if (w < 5.0f) {
        float t = w - 2.5f;
        float c = 2.81022636e-08f;
        c = fmaf(c, t, 3.43273939e-07f);
        c = fmaf(c, t, -3.5233877e-06f);
        c = fmaf(c, t, -4.39150654e-06f);
        c = fmaf(c, t, 0.00021858087f);
        c = fmaf(c, t, -0.00125372503f);
        c = fmaf(c, t, -0.00417768164f);
        c = fmaf(c, t, 0.246640727f);
        c = fmaf(c, t, 1.50140941f);
        return 1.41421356237f * c * (2.0f * p - 1.0f);
    }
    float r = sqrtf(-logf(p)), z;
    if (r <= 5.0f) {
        r -= 1.6f;
        z = (((0.17023821103f * r + 1.3067284816f) * r + 2.7568153900f) * r + 1.4234372777f) /
            ((0.12021132975f * r + 0.73700164250f) * r + 1.0f);
    } else {
        r -= 5.0f;
        z = (((0.017337203997f * r + 0.42868294337f) * r + 3.0812263860f) * r + 6.6579051150f) /
            ((0.012258202635f * r + 0.24197894225f) * r + 1.0f);
    }
    return -z;
}

/* erfc(x), x >= 0, in single precision: Numerical Recipes in C (2nd ed., section 6.2, erfcc),
 * t exp(-x^2 + c(t)), t = 1 / (1 + x/2), c a degree-9 Chebyshev-fitted polynomial; fractional error
 * below 1.2e-7 for every x >= 0.  The device forms t with its hardware reciprocal and the exponential
 * with its hardware exp2 (FMA-split exponent), and takes sqrt(-ln p) in orc_ndtri7_lower from its
 * hardware log2 and square root; each differs from the libm function used here by about an ulp. */
float orc_erfc_cheb32(float x) {
    float t = 1.0f / (1.0f + 0.5f * x);
    float c = -1.26551223f +
              t * (1.00002368f +
                   t * (0.37409196f +
                        t * (0.09678418f +
                             t * (-0.18628806f +
                                  t * (0.27886807f +
                                       t * (-1.13520398f +
                                            t * (1.48851587f + t * (-0.82215223f + t * 0.17087277f))))))));
    return t * expf(c - x * x);
}

/* The tn_std construction in single precision: a = (lo - mu) s and b = (hi - mu) s formed in double
 * and rounded; one erfc (orc_erfc_cheb32) per bound within 10 sigma; p = Phi(a) + u m, q = Q(b) + (1 - u) m, the
 * smaller inverted.
 * Returns NaN (the mover then takes orc_trunc_normal_prec) when a bound lies beyond 8 sigma on the
 * far side of the mean or the interval holds less than 1e-30 of the mass. */
double orc_trunc_normal_prec_f32(double mu, double s, double lo, double hi, float u) {
    double a64 = (lo - mu) * s, b64 = (hi - mu) * s;
    if (a64 > 8.0 || b64 < -8.0) return NAN;
    float a = (float)fmax(a64, -40.0), b = (float)fmin(b64, 40.0);
    const float r2 = 0.70710678118654752f;
    /* a bound 10 or more standard deviations out: its tail (< 7.7e-24) is taken as zero, below 2^-24 of
     * the interval's mass since the other bound lies within 8 sigma */
    float ta = fabsf(a) < 10.0f ? 0.5f * orc_erfc_cheb32(fabsf(a) * r2) : 0.0f;
    float tb = fabsf(b) < 10.0f ? 0.5f * orc_erfc_cheb32(fabsf(b) * r2) : 0.0f;
    float pa = a < 0.0f ? ta : 1.0f - ta, qa = a < 0.0f ? 1.0f - ta : ta;
    float pb = b < 0.0f ? tb : 1.0f - tb, qb = b < 0.0f ? 1.0f - tb : tb;
    float m = a >= 0.0f ? qa - qb : pb - pa;
    if (!(m >= 1e-30f)) return NAN;
    float p = pa + u * m, q = qb + (1.0f - u) * m;
    float z = orc_ndtri7_lower(fminf(p, q));
    z = p <= q ? z : -z;
    z = fminf(fmaxf(z, a), b);
    double x = mu + (double)z * (1.0 / s);
    return fmin(fmax(x, lo), hi);
}

static double open_uniform(uint64_t w) { return ((double)(w >> 11) + 0.5) * 0x1.0p-53; }

/* ---------------- per-instance state ---------------- */
typedef struct {
    const orc_problem* pb;
    int ndim, nnode, ndof, P;
    double *H, *b, dsd, *ll, *alpha2, lss2, lsg2;
    int *li, *lti;
    double *lv, *ltv;
} inst_t;

static void inst_build(inst_t* I, const orc_problem* pb, const double* theta, const double* gmat);
static void inst_free(inst_t* I);

static double inst_likelihood(const inst_t* I, const double* s) {
    return orc_slip_calc_likelihood(s, I->ndim, I->H, I->b, I->dsd, I->pb->nsar_total,
                                    I->pb->ngnss, I->lss2, I->lsg2);
}
static double inst_prior(const inst_t* I, const double* s, double* lsvec) {
    return orc_slip_calc_prior(s, I->alpha2, I->li, I->lv, I->nnode, lsvec);
}

/* ---------------- HMC: smc_slip.f90:732-917 ---------------- */
static void leapfrog(const inst_t* I, double gamma, double post_cur, double* p, const double* cov,
                     double* cand, const double* cur, double dtau, int ntau, double* ham_cur,
                     double* like_c, double* prior_c, double* post_c, double* ham_c, double* grad,
                     double* lsvec) {
    int nd = I->ndim;
    double smax = I->pb->max_slip;
    double h = post_cur;
    for (int i = 0; i < nd; ++i) h = h + p[i] * cov[i] * p[i] / 2.0;
    *ham_cur = h;
    orc_slip_calc_grad(grad, cand, nd, I->nnode, I->ndof, I->li, I->lv, I->lti, I->ltv, gamma,
                       I->alpha2, I->b, I->H, lsvec);
    for (int i = 0; i < nd; ++i) p[i] = p[i] - grad[i] * 5e-1 * dtau;
    for (int it = 0; it <= ntau; ++it) {
        for (int i = 0; i < nd; ++i) {
            cand[i] = (it < ntau) ? cand[i] + cov[i] * p[i] * dtau : cand[i] + cov[i] * p[i] * 5e-1 * dtau;
            if (fabs(cand[i]) > 1e3) {
                memcpy(cand, cur, sizeof(double) * nd);
                *ham_c = 1e20;
                return;
            }
            while (cand[i] < 0.0 || cand[i] > smax) {
                if (cand[i] > smax) { cand[i] = 2.0 * smax - cand[i]; p[i] = -p[i]; }
                if (cand[i] < 0.0) { cand[i] = -cand[i]; p[i] = -p[i]; }
            }
        }
        if (it == ntau) break;
        orc_slip_calc_grad(grad, cand, nd, I->nnode, I->ndof, I->li, I->lv, I->lti, I->ltv, gamma,
                           I->alpha2, I->b, I->H, lsvec);
        for (int i = 0; i < nd; ++i) p[i] = p[i] - grad[i] * dtau;
    }
    *like_c = inst_likelihood(I, cand);
    *prior_c = inst_prior(I, cand, lsvec);
    *post_c = gamma * *like_c + *prior_c;
    h = *post_c;
    for (int i = 0; i < nd; ++i) h = h + p[i] * cov[i] * p[i] / 2.0;
    *ham_c = h;
}

typedef struct {
    double log_dtau_upper, log_dtau_lower;
    int ntau_upper, ntau_lower;
    double* dtau_ls;
    int* ntau_ls;
    double* score;  /* optional copy of the children's scores (-1 where not scored) */
} hmc_tune_t;

static int hmc_tuning(const inst_t* I, uint64_t key, int stage, double gamma, const double* X,
                      const double* L, const double* PR, const double* cov, const int* assigned,
                      const int* id_start, hmc_tune_t* T) {
    int P = I->P, nd = I->ndim, tf = 5, cnt = 0;
    for (int i = 0; i < P; i += tf)
        for (int j = id_start[i]; j < id_start[i] + assigned[i]; ++j) {
            cnt++;
            double u1 = orc_clamp_uniform(draw_uniform(key, SITE_TUNE_DTAU, stage, j, 0, 0));
            double u2 = orc_clamp_uniform(draw_uniform(key, SITE_TUNE_DTAU, stage, j, 0, 1));
            T->ntau_ls[j] = (int)(T->ntau_lower + u1 * (T->ntau_upper - T->ntau_lower));
            double ld = T->log_dtau_lower + u2 * (T->log_dtau_upper - T->log_dtau_lower);
            T->dtau_ls[j] = exp(ld);
        }
    if (cnt == 0) return 0;
    double* score = (double*)malloc(sizeof(double) * P);
    for (int j = 0; j < P; ++j) score[j] = -1.0;
    #pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < P; i += tf) {
        double* cur = (double*)malloc(sizeof(double) * nd);
        double* cand = (double*)malloc(sizeof(double) * nd);
        double* p = (double*)malloc(sizeof(double) * nd);
        double* grad = (double*)malloc(sizeof(double) * nd);
        double* lsv = (double*)malloc(sizeof(double) * 2 * I->nnode);
        memcpy(cur, X + (size_t)i * nd, sizeof(double) * nd);
        double lc = L[i], pc = PR[i], postc = gamma * lc + pc;
        for (int j = id_start[i]; j < id_start[i] + assigned[i]; ++j) {
            double dtau = T->dtau_ls[j];
            memcpy(cand, cur, sizeof(double) * nd);
            for (int d = 0; d < nd; ++d) p[d] = draw_bm(key, SITE_TUNE_MOM, stage, j, d) / sqrt(cov[d]);
            double hc, lcand = 0, pcand = 0, postcand = 0, hcand;
            leapfrog(I, gamma, postc, p, cov, cand, cur, dtau, 50, &hc, &lcand, &pcand, &postcand,
                     &hcand, grad, lsv);
            double dham = hc - hcand, sc = 0.0;
            for (int d = 0; d < nd; ++d) sc = sc + ((cand[d] - cur[d]) * (cand[d] - cur[d])) / cov[d];
            sc = sc * exp(fmin(0.0, dham));
            score[j] = sc;
            double mh = orc_clamp_uniform(draw_uniform(key, SITE_TUNE_MH, stage, j, 0, 0));
            if (hc - hcand > log(mh)) {
                memcpy(cur, cand, sizeof(double) * nd);
                lc = lcand; pc = pcand; postc = postcand;
            }
        }
        free(cur); free(cand); free(p); free(grad); free(lsv);
    }
    if (T->score) memcpy(T->score, score, sizeof(double) * P);
    double ss = 0.0;
    for (int i = 0; i < P; i += tf)
        for (int j = id_start[i]; j < id_start[i] + assigned[i]; ++j) ss = ss + score[j];
    if (ss < 1e-5) {
        ss = 0.0;
        for (int i = 0; i < P; i += tf)
            for (int j = id_start[i]; j < id_start[i] + assigned[i]; ++j) { score[j] = 1.0; ss = ss + 1.0; }
    }
    double uh, ul;
    draw_quad(key, SITE_TUNE_RESAMPLE, stage, &uh, &ul);
    q128 u = (q128)uh + (q128)ul;
    if (u < (q128)1e-5) u = (q128)1e-5;
    q128 dn = (q128)P;
    u = u / dn;
    double* dnew = (double*)malloc(sizeof(double) * P);
    int* nnew = (int*)malloc(sizeof(int) * P);
    int c2 = 0;
    for (int i = 0; i < P; i += tf)
        for (int j = id_start[i]; j < id_start[i] + assigned[i]; ++j) {
            double wj = score[j] / ss;
            int a = (int)floorq(((q128)wj - u) * dn) + 1;
            u = u + (q128)a / dn - (q128)wj;
            for (int k = 0; k < a && c2 < P; ++k) { dnew[c2] = T->dtau_ls[j]; nnew[c2] = T->ntau_ls[j]; c2++; }
        }
    int ok = (c2 == P);
    for (int j = 0; j < c2; ++j) { T->dtau_ls[j] = dnew[j]; T->ntau_ls[j] = nnew[j]; }
    q128 m = 0;
    for (int j = 0; j < P; ++j) m = m + (q128)T->dtau_ls[j];
    m = m / dn;
    T->log_dtau_upper = (double)(logq(m) + (q128)log(5.0));
    T->log_dtau_lower = (double)(logq(m) - (q128)log(5.0));
    free(dnew); free(nnew); free(score);
    return ok ? 0 : -1;
}

int orc_hmc_tuning_step(const orc_problem* pb, const double* theta, const double* gmat, uint64_t key, int stage,
                        double gamma, const double* X, const double* L, const double* PR, const double* cov,
                        const int* assigned, double* dtau_io, double* log_lo, double* log_hi, double* score_out) {
    inst_t I;
    inst_build(&I, pb, theta, gmat);
    int P = I.P;
    int* id_start = (int*)malloc(sizeof(int) * P);
    id_start[0] = 0;
    for (int k = 0; k < P - 1; ++k) id_start[k + 1] = id_start[k] + assigned[k];
    hmc_tune_t T;
    T.log_dtau_upper = *log_hi; T.log_dtau_lower = *log_lo; T.ntau_lower = 1; T.ntau_upper = 100;
    T.dtau_ls = dtau_io;
    T.ntau_ls = (int*)malloc(sizeof(int) * P);
    for (int k = 0; k < P; ++k) T.ntau_ls[k] = 50;
    T.score = score_out;
    int cnt = 0;
    for (int i = 0; i < P; i += 5) cnt += assigned[i];
    int rc = hmc_tuning(&I, key, stage, gamma, X, L, PR, cov, assigned, id_start, &T);
    *log_hi = T.log_dtau_upper; *log_lo = T.log_dtau_lower;
    free(T.ntau_ls); free(id_start);
    inst_free(&I);
    return cnt == 0 ? 1 : rc;
}

/* ---------------- HMC sampling over all parents, smc_slip.f90:591-730 ---------------- */
static void hmc_sampling(const inst_t* I, uint64_t key, int stage, double gamma, const double* X,
                         const double* L, const double* PR, double* Xn, double* Ln, double* PRn,
                         const double* cov, const int* assigned, const int* id_start,
                         const double* dtau_ls, int ntau, double* acc) {
    int P = I->P, nd = I->ndim;
    #pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < P; ++i) {
        if (assigned[i] == 0) continue;
        double* cur = (double*)malloc(sizeof(double) * nd);
        double* cand = (double*)malloc(sizeof(double) * nd);
        double* p = (double*)malloc(sizeof(double) * nd);
        double* grad = (double*)malloc(sizeof(double) * nd);
        double* lsv = (double*)malloc(sizeof(double) * 2 * I->nnode);
        memcpy(cur, X + (size_t)i * nd, sizeof(double) * nd);
        double lc = L[i], pc = PR[i], postc = gamma * lc + pc;
        for (int j = id_start[i]; j < id_start[i] + assigned[i]; ++j) {
            memcpy(cand, cur, sizeof(double) * nd);
            for (int d = 0; d < nd; ++d) p[d] = draw_bm(key, SITE_HMC_MOM, stage, j, d) / sqrt(cov[d]);
            double hc, lcand = 0, pcand = 0, postcand = 0, hcand;
            leapfrog(I, gamma, postc, p, cov, cand, cur, dtau_ls[j], ntau, &hc, &lcand, &pcand, &postcand,
                     &hcand, grad, lsv);
            if (acc) acc[j] = exp(fmin(0.0, hc - hcand));
            double mh = orc_clamp_uniform(draw_uniform(key, SITE_HMC_MH, stage, j, 0, 0));
            if ((hc - hcand) > log(mh)) {
                memcpy(cur, cand, sizeof(double) * nd);
                lc = lcand; pc = pcand; postc = postcand;
            }
            memcpy(Xn + (size_t)j * nd, cur, sizeof(double) * nd);
            Ln[j] = lc; PRn[j] = pc;
        }
        free(cur); free(cand); free(p); free(grad); free(lsv);
    }
}

/* ---------------- Gibbs-sweep mover (benchmark configuration; no reference counterpart) --------
 * Target at stage gamma: exp(-gamma*likelihood(s) - prior_quadratic(s)) on [0, max_slip]^ndim,
 * with precision Q = gamma*H + LL/alpha^2 (formed entrywise as gamma*H_ki + LL_ki/alpha^2) and
 * linear term gamma*b.  Each child j starts from its parent and runs `sweeps` full
 * systematic-scan sweeps (i = 1..ndim); coordinate i is redrawn from N(mu_i, 1/Q_ii)
 * truncated to [0, max_slip] with
 *   mu_i = s_i - (sum_k Q_ki s_k - gamma*b_i) * (1 / Q_ii)       (column i of Q)
 * using the open uniform u = ((w>>11) + 1/2) 2^-53 of word (i & 1) of draw
 * (site 9, stage, c2=j, c1=i>>1, c0=sweep): coordinates 2m and 2m+1 share one Philox block.
 * The truncated normal is parameterised by s = sqrt(Q_ii): a = -mu s, b = (max_slip - mu) s.
 */
static void gibbs_move(const inst_t* I, uint64_t key, int stage, double gamma, int sweeps,
                       const double* X, double* Xn, double* Ln, double* PRn, const int* assigned,
                       const int* id_start) {
    int P = I->P, nd = I->ndim;
    double smax = I->pb->max_slip;
    double a2 = I->alpha2[0];
    double* Q = (double*)malloc(sizeof(double) * (size_t)nd * nd);
    double* hb = (double*)malloc(sizeof(double) * nd);
    for (size_t k = 0; k < (size_t)nd * nd; ++k) Q[k] = gamma * I->H[k] + I->ll[k] / a2;
    for (int i = 0; i < nd; ++i) hb[i] = gamma * I->b[i];
    /* per-coordinate constants of the stage: 1 / Q_ii and sqrt(Q_ii) */
    double* ip = (double*)malloc(sizeof(double) * nd);
    double* sq = (double*)malloc(sizeof(double) * nd);
    for (int i = 0; i < nd; ++i) {
        ip[i] = 1.0 / Q[(size_t)i * nd + i];
        sq[i] = sqrt(Q[(size_t)i * nd + i]);
    }
    #pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < P; ++i) {
        double* lsv = (double*)malloc(sizeof(double) * 2 * I->nnode);
        for (int j = id_start[i]; j < id_start[i] + assigned[i]; ++j) {
            double* s = Xn + (size_t)j * nd;
            memcpy(s, X + (size_t)i * nd, sizeof(double) * nd);
            for (int sw = 0; sw < sweeps; ++sw)
                for (int c = 0; c < nd; ++c) {
                    /* sum_k Q(k, c) s_k with 8 partial sums (the mover fixes no order) */
                    const double* qc = Q + (size_t)c * nd;
                    double ps[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    int k = 0;
                    for (; k + 8 <= nd; k += 8)
                        for (int r = 0; r < 8; ++r) ps[r] += qc[k + r] * s[k + r];
                    for (; k < nd; ++k) ps[k & 7] += qc[k] * s[k];
                    double qs = ((ps[0] + ps[1]) + (ps[2] + ps[3])) + ((ps[4] + ps[5]) + (ps[6] + ps[7]));
                    double mu = s[c] - (qs - hb[c]) * ip[c];
                    uint64_t w[2];
                    orc_draw_u64x2(key, SITE_GIBBS_MOVE, stage, j, c >> 1, sw, w);
                    s[c] = orc_trunc_normal_prec(mu, sq[c], 0.0, smax, open_uniform(w[c & 1]));
                }
            Ln[j] = inst_likelihood(I, s);
            PRn[j] = inst_prior(I, s, lsv);
        }
        free(lsv);
    }
    free(Q); free(hb); free(ip); free(sq);
}

/* ---------------- instance setup shared by slip_smc_exec and fault_calc_likelihood ------------- */
static void inst_build(inst_t* I, const orc_problem* pb, const double* theta, const double* gmat) {
    memset(I, 0, sizeof(*I));
    I->pb = pb;
    int npatch;
    orc_mesh_sizes(&pb->fault, &I->nnode, &I->ndof, &npatch, NULL);
    I->ndim = 2 * I->ndof;
    I->P = pb->nparticle_slip;
    int nnode = I->nnode, ndof = I->ndof, nd = I->ndim;
    int* cny = (int*)malloc(sizeof(int) * 4 * npatch);
    double* coor = (double*)malloc(sizeof(double) * 2 * nnode);
    int* n2e = (int*)malloc(sizeof(int) * 4 * nnode);
    int* n2es = (int*)malloc(sizeof(int) * nnode);
    int* id_dof = (int*)malloc(sizeof(int) * ndof);
    orc_discretize_fault(&pb->fault, theta, cny, coor, n2e, n2es, id_dof);
    double* luni = (double*)malloc(sizeof(double) * (size_t)nnode * nnode);
    double* lmat = (double*)malloc(sizeof(double) * (size_t)2 * nnode * 2 * ndof);
    orc_gen_laplacian(&pb->fault, theta, nnode, id_dof, ndof, luni, lmat);
    I->li = (int*)malloc(sizeof(int) * 9 * 2 * nnode);
    I->lv = (double*)malloc(sizeof(double) * 9 * 2 * nnode);
    I->lti = (int*)malloc(sizeof(int) * 9 * 2 * ndof);
    I->ltv = (double*)malloc(sizeof(double) * 9 * 2 * ndof);
    orc_gen_sparse_lmat(lmat, nnode, ndof, I->li, I->lv, I->lti, I->ltv);
    I->ll = (double*)malloc(sizeof(double) * (size_t)nd * nd);
    orc_calc_ll(lmat, nnode, ndof, I->ll);
    double la2;
    orc_get_sigma(theta, pb->fault.fix_xbend, pb->fault.nplane, &I->lss2, &I->lsg2, &la2);
    I->alpha2 = (double*)malloc(sizeof(double) * 2 * nnode);
    for (int k = 0; k < 2 * nnode; ++k) I->alpha2[k] = exp(la2);
    I->H = (double*)malloc(sizeof(double) * (size_t)nd * nd);
    I->b = (double*)malloc(sizeof(double) * nd);
    orc_slip_calc_gsd_gsg(gmat, pb->dvec, pb->nobs, nd, pb->npath, pb->nsar_index, pb->nsar_total,
                          pb->sar_sinv, exp(I->lss2), exp(I->lsg2), pb->obs_sigma, I->b, I->H, &I->dsd);
    free(luni); free(lmat); free(id_dof); free(n2es); free(n2e); free(coor); free(cny);
}

static void inst_free(inst_t* I) {
    free(I->li); free(I->lv); free(I->lti); free(I->ltv); free(I->ll); free(I->alpha2);
    free(I->H); free(I->b);
}

/* ---------------- slip_smc_exec, smc_slip.f90:1445-1663 ---------------- */
static void smc_run(const inst_t* I, uint64_t key, double* X, double* L, double* PR, double* mean,
                    orc_slip_result* res) {
    int P = I->P, nd = I->ndim;
    double* bm = (double*)malloc(sizeof(double) * (size_t)P * nd);
    #pragma omp parallel for
    for (int k = 0; k < P; ++k)
        for (int i = 0; i < nd; ++i) bm[(size_t)k * nd + i] = draw_bm(key, SITE_GIBBS_INIT, 0, k, i);
    orc_slip_gibbs_init_chain(X, P, nd, I->ll, I->alpha2, I->pb->max_slip, bm);
    free(bm);
    #pragma omp parallel for
    for (int k = 0; k < P; ++k) {
        double* lsv = (double*)malloc(sizeof(double) * 2 * I->nnode);
        L[k] = inst_likelihood(I, X + (size_t)k * nd);
        PR[k] = inst_prior(I, X + (size_t)k * nd, lsv);
        free(lsv);
    }
    double* w = (double*)malloc(sizeof(double) * P);
    double* cov = (double*)malloc(sizeof(double) * nd);
    int* assigned = (int*)malloc(sizeof(int) * P);
    int* id_start = (int*)malloc(sizeof(int) * P);
    double* Xn = (double*)malloc(sizeof(double) * (size_t)P * nd);
    double* Ln = (double*)malloc(sizeof(double) * P);
    double* PRn = (double*)malloc(sizeof(double) * P);
    hmc_tune_t T;
    T.log_dtau_upper = 0.0; T.log_dtau_lower = -1e1; T.ntau_lower = 1; T.ntau_upper = 100;
    T.dtau_ls = (double*)malloc(sizeof(double) * P);
    T.ntau_ls = (int*)malloc(sizeof(int) * P);
    T.score = NULL;
    for (int k = 0; k < P; ++k) { T.dtau_ls[k] = exp(-5.0); T.ntau_ls[k] = 50; }
    double* acc = (double*)malloc(sizeof(double) * P);
    double log_dtau = ORC_HMC_LOG_DTAU0;
    int iter = 1;
    double gamma = 0.0, neglog = 0.0;
    while (1.0 - gamma > 10e-8) {
        if (I->pb->stage_cap > 0 && iter > I->pb->stage_cap) break;
        double ev;
        gamma = orc_find_next_gamma(gamma, L, P, w, &ev, 2.0, 1e-10, 100000);
        neglog = neglog + ev;
        if (res && res->gamma_trace && iter <= 201) res->gamma_trace[iter - 1] = gamma;
        if (res && res->evid_trace && iter <= 201) res->evid_trace[iter - 1] = ev;
        if (iter > 200) { neglog = 1e10; break; }
        orc_slip_normalize_weights(w, P);
        orc_slip_calc_mean_particles(X, w, mean, P, nd);
        orc_slip_calc_cov_particles(X, w, mean, cov, P, nd);
        double uh, ul;
        draw_quad(key, SITE_RESAMPLE, iter, &uh, &ul);
        orc_resample_quad(P, w, uh, ul, assigned);
        id_start[0] = 0;
        for (int k = 0; k < P - 1; ++k) id_start[k + 1] = id_start[k] + assigned[k];
        if (I->pb->mover == 1) {
            gibbs_move(I, key, iter, gamma, I->pb->gibbs_sweeps, X, Xn, Ln, PRn, assigned, id_start);
        } else if (I->pb->mover == 2) {
            double dt = exp(log_dtau);
            for (int k = 0; k < P; ++k) T.dtau_ls[k] = dt;
            hmc_sampling(I, key, iter, gamma, X, L, PR, Xn, Ln, PRn, cov, assigned, id_start, T.dtau_ls,
                         I->pb->hmc_ntau, acc);
            double sa = 0.0;
            for (int k = 0; k < P; ++k) sa = sa + acc[k];
            double abar = sa / (double)P;
            if (res && res->accept_trace && iter <= 201) res->accept_trace[iter - 1] = abar;
            if (res && res->dtau_trace && iter <= 201) res->dtau_trace[iter - 1] = dt;
            log_dtau = log_dtau + ORC_HMC_ADAPT_GAIN * (abar - I->pb->hmc_accept);
        } else {
            hmc_tuning(I, key, iter, gamma, X, L, PR, cov, assigned, id_start, &T);
            if (res && res->dtau_trace && iter <= 201) {
                double m = 0.0;
                for (int k = 0; k < P; ++k) m = m + T.dtau_ls[k];
                res->dtau_trace[iter - 1] = m / (double)P;
            }
            hmc_sampling(I, key, iter, gamma, X, L, PR, Xn, Ln, PRn, cov, assigned, id_start, T.dtau_ls, 50,
                         NULL);
        }
        memcpy(X, Xn, sizeof(double) * (size_t)P * nd);
        memcpy(L, Ln, sizeof(double) * P);
        memcpy(PR, PRn, sizeof(double) * P);
        iter++;
    }
    if (res) { res->nstage = iter - 1; res->neglog = neglog; }
    free(w); free(cov); free(assigned); free(id_start); free(Xn); free(Ln); free(PRn);
    free(T.dtau_ls); free(T.ntau_ls); free(acc);
}

void orc_slip_smc_exec(const orc_problem* pb, const double* theta, const double* gmat,
                       uint64_t key, double* particles, double* likelihood, double* prior,
                       double* mean, orc_slip_result* res) {
    inst_t I;
    inst_build(&I, pb, theta, gmat);
    smc_run(&I, key, particles, likelihood, prior, mean, res);
    inst_free(&I);
}

/* Test hooks on one instance (theta, gmat): the tempered gradient and one leapfrog trajectory
 * (slip_leapfrog, smc_slip.f90:732-917) from (cur, p0) with mass diag(cov). */
void orc_inst_grad(const orc_problem* pb, const double* theta, const double* gmat, double gamma,
                   const double* s, double* grad) {
    inst_t I;
    inst_build(&I, pb, theta, gmat);
    double* lsv = (double*)malloc(sizeof(double) * 2 * I.nnode);
    orc_slip_calc_grad(grad, s, I.ndim, I.nnode, I.ndof, I.li, I.lv, I.lti, I.ltv, gamma, I.alpha2, I.b, I.H, lsv);
    free(lsv);
    inst_free(&I);
}

void orc_hmc_leapfrog(const orc_problem* pb, const double* theta, const double* gmat, double gamma,
                      const double* cur, const double* p0, const double* cov, double dtau, int ntau,
                      double* cand, double* p, double* ham_cur, double* ham_cand) {
    inst_t I;
    inst_build(&I, pb, theta, gmat);
    int nd = I.ndim;
    double* grad = (double*)malloc(sizeof(double) * nd);
    double* lsv = (double*)malloc(sizeof(double) * 2 * I.nnode);
    double post = gamma * inst_likelihood(&I, cur) + inst_prior(&I, cur, lsv);
    double lc, pc, postc;
    memcpy(cand, cur, sizeof(double) * nd);
    memcpy(p, p0, sizeof(double) * nd);
    leapfrog(&I, gamma, post, p, cov, cand, cur, dtau, ntau, ham_cur, &lc, &pc, &postc, ham_cand, grad, lsv);
    free(grad); free(lsv);
    inst_free(&I);
}

/* fault_calc_likelihood, smc_fault.f90:66-292 */
double orc_fault_calc_likelihood(const orc_problem* pb, const double* theta, uint64_t key,
                                 orc_slip_result* res) {
    if (pb->test_like) {
        int nf = pb->fault.fix_xbend ? 2 * pb->fault.nplane + 4 : 3 * pb->fault.nplane + 3;
        double v = 0.0;
        for (int d = 0; d < nf; ++d) {
            double z = (theta[d] - pb->test_like[d]) / pb->test_like[nf + d];
            v += 0.5 * z * z;
        }
        if (res) res->nstage = 0, res->neglog = v;
        return v;
    }
    int nnode, ndof, npatch;
    orc_mesh_sizes(&pb->fault, &nnode, &ndof, &npatch, NULL);
    int nd = 2 * ndof, P = pb->nparticle_slip;
    double* g = (double*)malloc(sizeof(double) * (size_t)pb->nobs * nd);
    orc_calc_greens_func(&pb->fault, theta, pb->nobs, pb->nsar_total, pb->ngnss, pb->obs_points,
                         pb->obs_unitvec, g);
    double* X = (double*)malloc(sizeof(double) * (size_t)P * nd);
    double* L = (double*)malloc(sizeof(double) * P);
    double* PR = (double*)malloc(sizeof(double) * P);
    double* mean = (double*)malloc(sizeof(double) * nd);
    orc_slip_result r = {0};
    if (res) r = *res;
    orc_slip_smc_exec(pb, theta, g, key, X, L, PR, mean, &r);
    if (res) *res = r;
    free(g); free(X); free(L); free(PR); free(mean);
    return r.neglog;
}

void orc_fault_calc_likelihood_batch(const orc_problem* pb, const double* theta, int ndim_fault,
                                     int B, const uint64_t* keys, double* out) {
    for (int b = 0; b < B; ++b)
        out[b] = orc_fault_calc_likelihood(pb, theta + (size_t)b * ndim_fault, keys[b], NULL);
}
