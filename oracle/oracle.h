/*
 * CPU oracle for the nested SMC fault inversion hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference Fortran (KaiNakao/smc_inversion_fort: the Fortran sources under src/ and src/sort.cpp) used
 * as the checker for the HIP library and as the timed host-CPU baseline.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product never links or calls it.
 *
 * Conventions follow the reference: Fortran column-major arrays, 1-based index
 * VALUES inside integer arrays (id_dof, cny, node_to_elem_val, ELL indices),
 * FP64 everywhere, __float128 where the reference uses real(16).
 *
 * Random numbers: the reference draws from the Fortran intrinsic stream.  Here
 * every draw is keyed (Philox4x32-10, see orc_philox4x32) by its logical
 * position, identical to the HIP path.  Core routines take already-drawn
 * numbers as arguments so that golden vectors made with the reference's own
 * stream can drive them too.
 *
 * Pinning status per function is listed in oracle/README.md.
 */
#ifndef SMCF_ORACLE_H
#define SMCF_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- RNG (shared keyed-draw specification) ---------------- */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* two raw 64-bit words for draw (key, site, stage, c2, c1, c0) */
void orc_draw_u64x2(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2,
                    uint32_t c1, uint32_t c0, uint64_t out[2]);
double orc_uniform_raw(uint64_t word);       /* (w >> 11) * 2^-53 in [0,1) */
double orc_clamp_uniform(double u);          /* smc_slip.f90:8-15 */
double orc_box_muller_sin(double r, double s); /* smc_slip.f90:26-38 (p only) */
uint64_t orc_splitmix64(uint64_t x);
uint64_t orc_inner_key(uint64_t seed, uint32_t outer_stage, uint32_t role, uint32_t slot);

/* ---------------- Okada DC3D0 (DC3Dfortran.f90:2-607,1236-1286,1400-1435) ---------------- */
void orc_dc3d0(double alpha, double x, double y, double z, double depth, double dip,
               double pot1, double pot2, double pot3, double pot4, double u[12], int* iret);

/* ---------------- geometry / mesh / Laplacian (init.f90:121-541) ---------------- */
typedef struct {
    int nplane;
    const int* nxi_ls;   /* nplane */
    const int* neta_ls;  /* nplane */
    double xmin, xmax, zmin;
    int fix_xbend;
    const double* xbend; /* nplane-1 (may be NULL if nplane==1) */
} orc_fault_desc;

void orc_get_geometry(int iplane, const orc_fault_desc* f, const double* theta,
                      int* nxi, int* neta, double* xl, double* yl, double* xr, double* yr,
                      double* dip, double* dip_rad, double* lxi, double* leta);
void orc_mesh_sizes(const orc_fault_desc* f, int* nnode, int* ndof, int* npatch, int* ndof_index);
void orc_discretize_fault(const orc_fault_desc* f, const double* theta, int* cny, double* coor,
                          int* node_to_elem_val, int* node_to_elem_size, int* id_dof);
void orc_gen_laplacian(const orc_fault_desc* f, const double* theta, int nnode, const int* id_dof,
                       int ndof, double* luni, double* lmat);
void orc_gen_sparse_lmat(const double* lmat, int nnode, int ndof, int* lmat_index,
                         double* lmat_val, int* ltmat_index, double* ltmat_val);
void orc_calc_ll(const double* lmat, int nnode, int ndof, double* llmat);
void orc_get_sigma(const double* theta, int fix_xbend, int nplane, double* log_sigma_sar2,
                   double* log_sigma_gnss2, double* log_alpha2);

/* ---------------- Green's function (gfunc.f90:6-348) ---------------- */
void orc_call_dc3d0(double xs, double ys, double zs, double xo, double yo, double uxi,
                    double ueta, double dip, double area, double strike, double uret[3]);
void orc_calc_greens_func(const orc_fault_desc* f, const double* theta, int nobs, int nsar_total,
                          int ngnss, const double* obs_points, const double* obs_unitvec,
                          double* gmat /* nobs x 2ndof */);

/* ---------------- inner slip SMC (smc_slip.f90) ---------------- */
void orc_slip_calc_gsd_gsg(const double* gmat, const double* dvec, int nobs, int ndim, int npath,
                           const int* nsar_index, int nsar_total, const double* sar_sinv,
                           double sigma_sar2, double sigma_gnss2, const double* obs_sigma,
                           double* gsdvec, double* gsgmat, double* dsd);
double orc_slip_calc_likelihood(const double* svec, int ndim, const double* gsgmat,
                                const double* gsdvec, double dsd, int nsar_total, int ngnss,
                                double log_sigma_sar2, double log_sigma_gnss2);
double orc_slip_calc_prior(const double* svec, const double* alpha2_full, const int* lmat_index,
                           const double* lmat_val, int nnode, double* lsvec);
void orc_slip_calc_grad(double* grad, const double* svec, int ndim, int nnode, int ndof,
                        const int* lmat_index, const double* lmat_val, const int* ltmat_index,
                        const double* ltmat_val, double gamma, const double* alpha2_full,
                        const double* gsdvec, const double* gsgmat, double* lsvec);
double orc_cdf_norm(double x, double mu, double sigma2);
double orc_pdf_norm(double x, double mu, double sigma2);
/* sequential Gibbs prior chain; bm = Box-Muller normals (ndim x P, in draw order) */
void orc_slip_gibbs_init_chain(double* particles, int P, int ndim, const double* llmat,
                               const double* alpha2_full, double max_slip, const double* bm);
double orc_find_next_gamma(double gamma_prev, const double* L, int P, double* weights,
                           double* neglog_evidence, double cv_threshold, double tol, int cap);
void orc_slip_normalize_weights(double* w, int P);
void orc_slip_calc_mean_particles(const double* particles, const double* w, double* mean, int P, int ndim);
void orc_slip_calc_cov_particles(const double* particles, const double* w, const double* mean,
                                 double* cov_diag, int P, int ndim);
/* u_hi + u_lo is the (unclamped) quad uniform; clamping is done inside */
void orc_resample_quad(int P, const double* w, double u_hi, double u_lo, int* assigned);
double orc_ndtri(double p);
double orc_trunc_normal_prec(double mu, double s, double lo, double hi, double u);
/* InSAR inverse covariance blocks of calc_sigma_sar_mat (init.f90:82-119), column-major, concatenated */
int orc_calc_sigma_sar_mat(int npath, const int* nsar_index, int nobs, const double* obs_points,
                           const double* obs_sigma, double* sinv_out);
/* single-precision draw of the mixed-precision Gibbs mover (NaN where it declines) and its PPND7 */
double orc_trunc_normal_prec_f32(double mu, double s, double lo, double hi, float u);
float orc_ndtri7_lower(float p);
float orc_erfc_cheb32(float x);
double orc_trunc_normal(double mu, double sigma, double lo, double hi, double u);

typedef struct {
    /* observations */
    int nobs, nsar_total, ngnss, npath;
    const int* nsar_index;     /* npath+1, 1-based */
    const double* obs_points;  /* 2 x nobs */
    const double* obs_unitvec; /* 3 x nobs */
    const double* obs_sigma;   /* nobs */
    const double* dvec;        /* nobs */
    const double* sar_sinv;    /* concatenated inverse SAR covariance blocks (may be NULL) */
    orc_fault_desc fault;
    double max_slip;
    int nparticle_slip;
    int mover;                 /* 0 = reference HMC (ntau 50, ESJD tuning), 1 = Gibbs sweeps,
                                  2 = HMC with hmc_ntau steps and a step size adapted toward hmc_accept */
    int gibbs_sweeps;
    int hmc_ntau;              /* leapfrog steps of mover 2 */
    double hmc_accept;         /* target mean acceptance probability of mover 2 */
    int stage_cap;             /* test hook: stop after this many stages (0: run to gamma = 1) */
    const double* test_like;   /* test hook: when set, the fault likelihood is the Gaussian
                                  0.5 sum_d ((theta_d - m_d) / s_d)^2 with m = test_like[0..nd) and
                                  s = test_like[nd..2 nd), and no inner SMC runs */
} orc_problem;

/* Step-size adaptation of mover 2 (no reference counterpart): the stage's children are moved with
 * dtau = exp(log_dtau); afterwards log_dtau += ORC_HMC_ADAPT_GAIN * (abar - hmc_accept), abar = the mean over
 * the P children of min(1, exp(H_cur - H_cand)) summed in child order.  First stage: log_dtau = log(0.1). */
#define ORC_HMC_ADAPT_GAIN 2.0
#define ORC_HMC_LOG_DTAU0 (-2.302585092994045684)

typedef struct {
    int nstage;            /* stages executed */
    double neglog;         /* -log Z estimate (or 1e10 sentinel) */
    double* gamma_trace;   /* optional, >= 202 entries */
    double* evid_trace;    /* optional, per-stage neglog evidence increment */
    double* accept_trace;  /* optional (mover 2), per-stage mean acceptance probability */
    double* dtau_trace;    /* optional (movers 0 and 2), per-stage step size (mover 0: mean over children) */
} orc_slip_result;

/* slip_smc_exec for one theta with prepared G etc.  particles (ndim x P) out. */
void orc_slip_smc_exec(const orc_problem* pb, const double* theta, const double* gmat,
                       uint64_t key, double* particles, double* likelihood, double* prior,
                       double* mean, orc_slip_result* res);
/* fault_calc_likelihood: returns -log Zhat for theta */
double orc_fault_calc_likelihood(const orc_problem* pb, const double* theta, uint64_t key,
                                 orc_slip_result* res);
/* batched evaluation with OpenMP over thetas (CPU-baseline helper) */
void orc_fault_calc_likelihood_batch(const orc_problem* pb, const double* theta, int ndim_fault,
                                     int B, const uint64_t* keys, double* out);

/* ---------------- outer fault SMC (smc_fault.f90) ---------------- */
void orc_fault_normalize_weights(double* w, int N);
void orc_fault_calc_mean_particles(const double* particles, const double* w, double* mean, int N, int nd);
void orc_fault_calc_cov_particles(const double* particles, const double* w, const double* mean,
                                  double* cov, int N, int nd, int* info);
void orc_qsort(int* arr, int* index, int n);   /* sort.cpp:3-7 */
void orc_fault_propose(int nd, const double* chol, const double* cur, double* z, const double* range,
                       double* cand);
void orc_fault_reorder(int* assigned, double* particles, double* likelihood, int N, int nd,
                       int* sorted_idx);
typedef struct {
    int nstage;
    double evidence;       /* fault_evidence (negative log) */
    double* gamma_trace;   /* optional */
} orc_fault_result;
void orc_inst_grad(const orc_problem* pb, const double* theta, const double* gmat, double gamma,
                   const double* s, double* grad);
/* One slip_hmc_tuning call (smc_slip.f90:919-1214) on one instance with given parents: draws the tuning
 * children's dtau from [log_lo, log_hi], runs their trajectories (ntau 50) in the keyed streams of `stage`,
 * scores them, resamples the scores into dtau_io (P values in, P out), updates the log-range and returns 0,
 * 1 when there were no tuning children (dtau_io and the range unchanged), -1 when the resampled count was
 * not P.  score_out (P, may be NULL) receives each child's score (-1 for children that were not scored). */
int orc_hmc_tuning_step(const orc_problem* pb, const double* theta, const double* gmat, uint64_t key, int stage,
                        double gamma, const double* X, const double* L, const double* PR, const double* cov,
                        const int* assigned, double* dtau_io, double* log_lo, double* log_hi, double* score_out);
/* Box-Muller normal of the keyed stream (sites of the reference's draws) */
double orc_draw_normal(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2, uint32_t c1);
double orc_draw_clamped_uniform(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2, uint32_t c1,
                                uint32_t word);
void orc_hmc_leapfrog(const orc_problem* pb, const double* theta, const double* gmat, double gamma,
                      const double* cur, const double* p0, const double* cov, double dtau, int ntau,
                      double* cand, double* p, double* ham_cur, double* ham_cand);
typedef struct {
    const double* qmean;      /* nd */
    const double* qcov_chol;  /* nd x nd, lower Cholesky factor of qcov (main.f90:332) */
    const double* qcov_inv;   /* nd x nd */
    double log_det;
} orc_gsmc;
double orc_fault_calc_neglog_q(const double* x, const double* qmean, const double* qinv, double log_det, int nd);
void orc_fault_smc_exec_gsmc(const orc_problem* pb, const double* range, int N, uint64_t seed, double* particles,
                             double* likelihood, int max_stages, orc_fault_result* res, const orc_gsmc* g);
void orc_fault_smc_exec(const orc_problem* pb, const double* range /* 2 x nd */, int N, uint64_t seed,
                        double* particles /* nd x N out */, double* likelihood /* N out */,
                        int max_stages, orc_fault_result* res);

#ifdef __cplusplus
}
#endif
#endif
