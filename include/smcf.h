/*
 * libsmcf: MI355X implementation of the nested SMC fault-inversion hot path.
 *
 * C ABI (plain pointers and sizes).  Arrays follow the reference's Fortran
 * conventions: column-major, 1-based index VALUES inside integer arrays, FP64.
 * Every entry point returns SMCF_OK (0) or an error code; smcf_last_error()
 * gives the message.  The reference has no error codes; its numeric sentinels
 * are reproduced in the outputs (-log Z = 1e10 after more than 200 inner stages,
 * smc_slip.f90:1557-1560).  All host buffers are caller-allocated, as in the
 * reference (main.f90:349-390); device state lives in the context/problem.
 *
 * A context owns one HIP stream on one device and is not thread-safe; use one
 * context per host thread.
 */
#ifndef SMCF_H
#define SMCF_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SMCF_OK 0
#define SMCF_EINVAL 1
#define SMCF_EHIP 2
#define SMCF_ENOMEM 3
#define SMCF_EUNSUPPORTED 4
#define SMCF_ESTATE 5

/* inner movers */
#define SMCF_MOVER_HMC 0   /* reference HMC: ntau 50 + ESJD dtau tuning (smc_slip.f90:591-1214) */
#define SMCF_MOVER_GIBBS 1 /* truncated-MVN Gibbs sweeps (benchmark configuration) */
#define SMCF_MOVER_HMC_ACCEPT 2 /* HMC with hmc_ntau leapfrog steps (the reference's slip_leapfrog) and one step size
                                   per inversion, adapted after every stage toward the mean acceptance probability
                                   hmc_accept (log dtau += 2.0 * (mean acceptance - hmc_accept), first dtau 0.1);
                                   no reference counterpart: it replaces the ESJD tuning pass */

/* arithmetic of the inner movers (no counterpart in the reference, which is FP64 throughout).  Gibbs: all three.
 * HMC movers: FP64 or MIXED_BF16, where the leapfrog gradients of the waves with many columns come from split-BF16
 * products of Q_g = gamma H + L^T L / alpha2 and the trajectory's final likelihood, the Hamiltonian and the
 * Metropolis test stay FP64 (the sampler stays exact; trajectories follow a force within ~2^-16 of the gradient). */
#define SMCF_PRECISION_FP64 0  /* every operation in FP64 (default; parity with the reference's arithmetic) */
#define SMCF_PRECISION_MIXED 1 /* Q.S products in FP32 on the matrix cores and single-precision inverse-CDF
                                  draws; conditional means, corrections, likelihood, weights and evidence FP64 */
#define SMCF_PRECISION_MIXED_BF16 2 /* as MIXED, with the Q.S products formed from split-BF16 operands (x = hi + lo,
                                       three BF16 matrix-core products hi.hi + hi.lo + lo.hi, FP32 accumulation):
                                       about 2^-16 relative error per term at 5x the FP32 matrix rate */

typedef struct smcf_ctx smcf_ctx;
typedef struct smcf_problem smcf_problem;

const char* smcf_last_error(void);
int smcf_ctx_create(int device, smcf_ctx** out);
int smcf_ctx_destroy(smcf_ctx* ctx);
/* kernel timing with HIP events on the context's stream (off by default) */
int smcf_set_profiling(smcf_ctx* ctx, int on);
/* cumulative device time of kernels named `name` since profiling was enabled */
int smcf_kernel_stats(smcf_ctx* ctx, const char* name, long* count, double* total_ms);

/*
 * Observation + fault-mesh + sampler settings.  Mirrors what main.f90 reads and
 * broadcasts (main.f90:107-307) and the fixed workspace sizes of main.f90:349-390.
 * Rows 1..nsar_total are InSAR, the rest GNSS triplets (init.f90:24-80).
 */
typedef struct {
  int nplane;
  const int* nxi_ls;          /* nplane */
  const int* neta_ls;         /* nplane */
  double xmin, xmax, zmin;
  int fix_xbend;
  const double* xbend;        /* nplane-1, may be NULL when nplane == 1 */
  int nobs, nsar_total, ngnss, npath;
  const int* nsar_index;      /* npath+1, 1-based */
  const double* obs_points;   /* 2 x nobs */
  const double* obs_unitvec;  /* 3 x nobs */
  const double* obs_sigma;    /* nobs */
  const double* dvec;         /* nobs */
  const double* sar_sinv;     /* inverse SAR covariance per path (calc_sigma_sar_mat, init.f90:82-119), each block
                                 column-major n_p x n_p, concatenated in path order; may be NULL when nsar_total == 0 */
  double max_slip;
  int nparticle_slip;
  int mover;                  /* SMCF_MOVER_* */
  int gibbs_sweeps;
  int precision;              /* SMCF_PRECISION_* (MIXED needs the Gibbs mover; HMC movers take FP64 or MIXED_BF16) */
  int hmc_ntau;               /* SMCF_MOVER_HMC_ACCEPT: leapfrog steps per trajectory (0: 10) */
  double hmc_accept;          /* SMCF_MOVER_HMC_ACCEPT: target mean acceptance probability (0: 0.65) */
} smcf_problem_desc;

/* Upload a problem; device buffers are sized for up to max_batch geometries in flight. */
int smcf_problem_create(smcf_ctx* ctx, const smcf_problem_desc* desc, int max_batch, smcf_problem** out);
/* Device bytes smcf_problem_create would allocate for this description (no device work; no counterpart in the
 * reference, used for admission by memory budget), and the bytes a created problem holds. */
int smcf_problem_bytes(const smcf_problem_desc* desc, int max_batch, size_t* bytes);
int smcf_problem_device_bytes(const smcf_problem* pb, size_t* bytes);
/* The context a problem was created on. */
smcf_ctx* smcf_problem_context(const smcf_problem* pb);
/* Rows per coordinate block of the Gibbs sweep kernels (no reference counterpart): the value a problem of this form
 * and ndim created now would use, from the library's tuning table (measured on the MI355X by the bench's
 * --autotune mode).  The rows set the order in which a conditional mean's terms are added, so results are
 * reproducible for a given value; smcf_set_gibbs_block_rows overrides the table for ndim_lo..ndim_hi (FP64: 16 or
 * 32; split BF16: 16, 32 or 48) for problems created afterwards. */
int smcf_gibbs_block_rows(int precision, int ndim);
int smcf_set_gibbs_block_rows(int precision, int ndim_lo, int ndim_hi, int rows);
int smcf_problem_destroy(smcf_problem* pb);
int smcf_problem_dims(const smcf_problem* pb, int* nnode, int* ndof, int* ndim_slip, int* ndim_fault,
                      int* nobs);

/* DC3D0 for n points (replaces DC3Dfortran.f90:2-118).  in: 10 x n
 * (alpha, x, y, z, depth, dip, pot1..pot4); u: 12 x n; iret: n (left at its input
 * value when depth < 0, as the reference does). */
int smcf_dc3d0(smcf_ctx* ctx, int n, const double* in, double* u, int* iret);

/* call_dc3d0 (gfunc.f90:65-100) for n point sources: in 10 x n = (xsource, ysource, zsource, xobs, yobs, uxi, ueta,
 * dip, area, strike), uret 3 x n (cm, geographic frame). */
int smcf_call_dc3d0(smcf_ctx* ctx, int n, const double* in, double* uret);
/* calc_responce (gfunc.f90:102-194) at npts observation points (pts 2 x npts): displacement uobs (3 x npts) of the
 * slip distribution slip_dist (2 x nnode, (u_xi, u_eta) per node) over the target patches target_id (1-based,
 * ntarget of them) of the plane whose frame is (xf, yf, zf, strike, dip); cny 4 x npatch and coor 2 x nnode as
 * discretize_fault builds them.  calc_responce_dist (gfunc.f90:196-240) is this over the SAR points and one
 * point per GNSS station. */
int smcf_calc_responce(smcf_ctx* ctx, int npts, const double* pts, int npatch, const int* cny, int nnode,
                       const double* coor, const double* slip_dist, double xf, double yf, double zf, double strike,
                       double dip, int ntarget, const int* target_id, double* uobs);

/* Green's matrices for B geometries (replaces calc_greens_func, gfunc.f90:242-348).
 * theta: ndim_fault x B; gmat: nobs x ndim_slip x B. */
int smcf_calc_greens_func(smcf_problem* pb, int B, const double* theta, double* gmat);

/* gen_laplacian + gen_sparse_lmat + calc_ll for one theta (init.f90:295-541).
 * Any output pointer may be NULL.  lmat_index/val: 9 x 2nnode; ltmat: 9 x 2ndof;
 * llmat: ndim x ndim. */
int smcf_laplacian(smcf_problem* pb, const double* theta, int* lmat_index, double* lmat_val,
                   int* ltmat_index, double* ltmat_val, double* llmat);

/* discretize_fault (init.f90:198-293) and the dense operators of gen_laplacian (init.f90:295-474) for one theta:
 * cny 4 x npatch, coor 2 x nnode, node_to_elem_val 4 x nnode (-1 padded), node_to_elem_size nnode, id_dof ndof
 * (1-based ids), luni nnode x nnode, lmat 2 nnode x 2 ndof.  Any output may be NULL. */
int smcf_discretize_fault(smcf_problem* pb, const double* theta, int* cny, double* coor, int* node_to_elem_val,
                          int* node_to_elem_size, int* id_dof, double* luni, double* lmat);

/* Gram quantities for B geometries (replaces slip_calc_gsd_gsg, smc_slip.f90:1347-1443).
 * gsgmat: ndim x ndim x B; gsdvec: ndim x B; dsd: B.  NULL outputs are skipped. */
int smcf_slip_calc_gsd_gsg(smcf_problem* pb, int B, const double* theta, double* gsgmat, double* gsdvec,
                           double* dsd);

/* Full inner SMC for one theta (replaces slip_smc_exec, smc_slip.f90:1445-1663, called
 * from fault_calc_likelihood, smc_fault.f90:66-292).  Outputs may be NULL.
 * particles: ndim x P (column-major); gamma_trace: >= 202 entries. */
int smcf_slip_smc_exec(smcf_problem* pb, const double* theta, uint64_t key, double* particles,
                       double* likelihood, double* prior, double* mean, int* nstage, double* neglog,
                       double* gamma_trace);

/* Batched fault_calc_likelihood (smc_fault.f90:66-292): -log Z for B thetas, each with its own
 * RNG key.  This is what work_eval_init_particles / work_mcmc_sampling waves call.  Any B >= 1:
 * the geometries stream through the problem's max_batch device slots, a finished inner SMC's slot
 * taking the next geometry; results do not depend on B, the slot count or the order. */
int smcf_fault_calc_likelihood_batch(smcf_problem* pb, int B, const double* theta, const uint64_t* keys,
                                     double* neglog, int* nstage);
/* Per-geometry latency of the last batch call on this problem (no counterpart in the reference): seconds from
 * the call's start to the stage boundary at which geometry j's result was read back, for j < B. */
int smcf_batch_times(const smcf_problem* pb, int B, double* seconds);
/* Seconds from the start of the last batch call to the launch of geometry j's set-up in a free slot (0 for the
 * geometries of the first fill); smcf_batch_times minus this is the geometry's own latency in the stream. */
int smcf_batch_admit_times(const smcf_problem* pb, int B, double* seconds);

/* Plain mean of each geometry's final slip particles (ndim_slip x B, column b for theta column b) and its -log Z:
 * the per-particle loop of calc_slip_map (main.f90:600-640, fault_calc_likelihood then the average over the
 * nparticle_slip particles), batched.  B <= max_batch. */
int smcf_slip_mean_batch(smcf_problem* pb, int B, const double* theta, const uint64_t* keys, double* mean,
                         double* neglog);

/* Component procedures of the inner SMC for one geometry theta, on the device (the geometry is set up in the
 * problem's first slot; particles in the reference layout, ndim x n column per particle, n <= nparticle_slip).
 *   smcf_slip_calc_likelihood   slip_calc_likelihood   (smc_slip.f90:56-104), Gram form
 *   smcf_slip_calc_prior        slip_calc_prior        (smc_slip.f90:106-137)
 *   smcf_slip_calc_grad         slip_calc_grad         (smc_slip.f90:139-182), tempered by gamma; grad ndim x n
 *   smcf_slip_leapfrog          slip_leapfrog          (smc_slip.f90:732-917) for n independent trajectories from
 *                               (cur, p0) with mass diagonal cov_diag; cand / p_out ndim x n, ham_cur / ham_cand n
 *                               (1e20 where the trajectory escaped |x| > 1e3; the candidate is then cur), and the
 *                               candidates' likelihood and prior (n each; any output may be NULL)
 *   smcf_slip_gen_init_particles slip_gen_init_particles (smc_slip.f90:184-287): the sequential prior Gibbs chain
 *                               of all nparticle_slip particles in the keyed stream of `key`, with likelihood
 *                               and prior
 * The gradient and leapfrog need a problem created with an HMC mover (SMCF_EUNSUPPORTED otherwise). */
int smcf_slip_calc_likelihood(smcf_problem* pb, const double* theta, int n, const double* particles, double* out);
int smcf_slip_calc_prior(smcf_problem* pb, const double* theta, int n, const double* particles, double* out);
int smcf_slip_calc_grad(smcf_problem* pb, const double* theta, double gamma, int n, const double* particles,
                        double* grad);
int smcf_slip_leapfrog(smcf_problem* pb, const double* theta, double gamma, int n, const double* cur,
                       const double* p0, const double* cov_diag, double dtau, int ntau, double* cand, double* p_out,
                       double* ham_cur, double* ham_cand, double* likelihood_cand, double* prior_cand);
int smcf_slip_gen_init_particles(smcf_problem* pb, const double* theta, uint64_t key, double* particles,
                                 double* likelihood, double* prior);
/* slip_calc_gsd_gsg (smc_slip.f90:1347-1443) from a given Green's matrix (nobs x ndim) and sigma^2 values, with the
 * problem's observation set (data, sigmas, InSAR path covariance) */
int smcf_slip_calc_gsd_gsg_gmat(smcf_problem* pb, const double* gmat, double sigma_sar2, double sigma_gnss2,
                                double* gsgmat, double* gsdvec, double* dsd);
/* The same procedures from the arrays of the reference's argument lists (no problem needed): gsgmat ndim x ndim
 * and gsdvec ndim of slip_calc_gsd_gsg, alpha2_full 2 nnode, the ELL of L (lmat_index / lmat_val, 9 x 2 nnode,
 * 1-based columns) and of L^T (ltmat_index / ltmat_val, 9 x ndim, 1-based rows) of gen_sparse_lmat.
 *   smcf_slip_likelihood_gram  slip_calc_likelihood (smc_slip.f90:56-104)
 *   smcf_slip_prior_ell        slip_calc_prior (106-137)
 *   smcf_slip_grad_ell         slip_calc_grad (139-182), grad ndim x n */
int smcf_slip_likelihood_gram(smcf_ctx* ctx, int ndim, int n, const double* particles, const double* gsgmat,
                              const double* gsdvec, double dsd, int nsar_total, int ngnss, double log_sigma_sar2,
                              double log_sigma_gnss2, double* out);
int smcf_slip_prior_ell(smcf_ctx* ctx, int nnode, int ndim, int n, const double* particles, const double* alpha2_full,
                        const int* lmat_index, const double* lmat_val, double* out);
int smcf_slip_grad_ell(smcf_ctx* ctx, int nnode, int ndim, int n, const double* particles, double gamma,
                       const double* gsgmat, const double* gsdvec, const double* alpha2_full, const int* lmat_index,
                       const double* lmat_val, const int* ltmat_index, const double* ltmat_val, double* grad);
/* SMCF_MOVER_HMC_ACCEPT (no reference counterpart): after smcf_slip_smc_exec, the step size the next stage of
 * that inversion would use and the mean acceptance probability of its last stage */
int smcf_slip_hmc_accept_state(const smcf_problem* pb, double* next_dtau, double* mean_accept);
/* slip_calc_mean_particles + slip_calc_cov_particles (smc_slip.f90:485-570): weighted mean and variance diagonal
 * (real(16) sums as double-double, variances below 1e-3 floored at 1e-3) of P particles (ndim x P) with
 * normalised weights */
int smcf_slip_calc_moments(smcf_ctx* ctx, int ndim, int P, const double* particles, const double* weights,
                           double* mean, double* cov_diag);

/* Inner-SMC components on the device (for parity checks and reuse). */
/* slip_find_next_gamma / fault_find_next_gamma (smc_slip.f90:314-387, smc_fault.f90:392-485) */
int smcf_find_next_gamma(smcf_ctx* ctx, int P, double gamma_prev, const double* likelihood, double cv,
                         double tol, int cap, double* gamma, double* weights, double* neglog_evidence);
/* Allocation-free smcf_find_next_gamma.  No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops and harnesses that forbid hipMalloc).  Same arguments
 * and results, bit for bit, and no hipMalloc/hipFree: the device copies of the likelihoods and weights live in a
 * fixed scratch array that the library holds on each device for the life of the process (its address is resolved
 * by smcf_ctx_create).  P <= 3584, else SMCF_EINVAL.  Calls on one GPU share that array: calls that overlap on one
 * GPU (from several contexts or threads) race and must be serialised by the caller. */
int smcf_find_next_gamma_noalloc(smcf_ctx* ctx, int P, double gamma_prev, const double* likelihood, double cv,
                                 double tol, int cap, double* gamma, double* weights, double* neglog_evidence);
/* slip_normalize_weights + slip_resample_particles (smc_slip.f90:465-483, 572-589);
 * u_hi + u_lo is the unclamped real(16) uniform. */
int smcf_normalize_resample(smcf_ctx* ctx, int P, const double* weights, double u_hi, double u_lo,
                            double* weights_norm, int* assigned_num);
/* truncated-normal inversion of the Gibbs mover (no reference counterpart) */
int smcf_trunc_normal(smcf_ctx* ctx, int n, const double* mu, const double* sigma, double lo, double hi,
                      const double* u, double* out);
/* single-precision draw of the mixed-precision Gibbs mover (no reference counterpart): sq = sqrt of the
 * conditional precision, u single-precision uniforms; NaN where the mover falls back to the FP64 draw */
int smcf_trunc_normal_f32(smcf_ctx* ctx, int n, const double* mu, const double* sq, double lo, double hi,
                          const float* u, double* out);
/* raw Philox4x32-10 blocks: ctr 4 x n, key 2 x n, out 4 x n */
int smcf_philox(smcf_ctx* ctx, int n, const uint32_t* ctr, const uint32_t* key, uint32_t* out);

/* Outer SMC (replaces fault_smc_exec, smc_fault.f90:785-1075, with numprocs = 1 semantics).
 * range: 2 x ndim_fault (range.dat layout); particles: ndim_fault x N out; likelihood: N out;
 * gamma_trace: >= max_stages entries (may be NULL).  Inner evaluations of one wave are issued as
 * smcf_fault_calc_likelihood_batch calls of at most max_batch geometries (any value >= 1; each call
 * streams through the problem's slots, so a whole wave per call keeps the device fullest). */
/* Per-stage dumps of the outer SMC (smc_fault.f90:923-936, 1054-1067): with a directory set, smcf_fault_smc_exec
 * and smcf_fault_smc_exec_gsmc write <output_dir><iter>.csv for iter = 0 (initial particles) and every stage, one
 * line per particle: its coordinates and its -log Z, each as Fortran F20.5.  output_dir is used as given (the
 * reference appends the stage number directly, so it normally ends in '/'); NULL or "" stops the dumps. */
int smcf_fault_smc_set_output_dir(smcf_problem* pb, const char* output_dir);
int smcf_fault_smc_exec(smcf_problem* pb, const double* range, int N, uint64_t seed, int max_stages, int max_batch,
                        double* particles, double* likelihood, int* nstage, double* evidence, double* gamma_trace);
/* Outer SMC with the Gaussian proposal q = N(qmean, qcov) (the gsmc branch of fault_smc_exec,
 * smc_fault.f90:41-52, 411-481, 708-775; q from fault_calc_neglog_q, smc_fault.f90:9-27).
 * qcov_chol is the lower Cholesky factor of qcov (main.f90:332 factors it before the call), qcov_inv its
 * inverse, both ndim x ndim column-major; log_qcov_det = log det qcov.  Initial particles are
 * qmean + chol z; weights temper L - q; the MH test is gamma (L_cur - L_cand) + (1 - gamma)(q_cur - q_cand). */
int smcf_fault_smc_exec_gsmc(smcf_problem* pb, const double* range, int N, uint64_t seed, int max_stages,
                             int max_batch, const double* qmean, const double* qcov_chol, const double* qcov_inv,
                             double log_qcov_det, double* particles, double* likelihood, int* nstage,
                             double* evidence, double* gamma_trace);
/* Stage components of the outer SMC (smc_fault.f90, one rank), host code in the reference's arithmetic order;
 * smcf_fault_smc_exec is their composition.  Particles ndim x N column-major.  Random numbers are keyed by the
 * outer seed and stage, as in smcf_fault_smc_exec.
 *   smcf_fault_sample_init_particles  fault_sample_init_particles (29-64): uniform in range (2 x ndim), or, with
 *                                     qmean != NULL, qmean + qcov_chol z (the gsmc branch)
 *   smcf_fault_mean_std               calc_mean_std_vector (372-390): sequential sums, population std
 *   smcf_fault_find_next_gamma        fault_find_next_gamma (392-485): bisection to CV 0.2, tol 1e-8, at most 1000
 *                                     steps; with q != NULL the weights temper L - q (gsmc); neglog_evidence is the
 *                                     stage's increment (the reference adds it to fault_evidence)
 *   smcf_fault_normalize_weights      fault_normalize_weights (487-501)
 *   smcf_fault_calc_mean_particles    fault_calc_mean_particles (503-520)
 *   smcf_fault_calc_cov_particles     fault_calc_cov_particles (522-570): lower factor of 0.04 x the weighted
 *                                     covariance (dpotrf 'L'); info = first non-positive pivot, 0 if none
 *   smcf_fault_resample_particles     fault_resample_particles (572-590): real(16) running u
 *   smcf_fault_reorder_to_send        fault_reorder_to_send (592-637) with numprocs = 1 (q may be NULL)
 *   smcf_work_mcmc_sampling           work_mcmc_sampling (639-783): re-evaluates the parents, proposes
 *                                     cur + cov_chol z (reflected into range) for every child, Metropolis test
 *                                     gamma (L_cur - L_cand) [+ (1 - gamma)(q_cur - q_cand) when q != NULL];
 *                                     children at 0-based id_start(i) + k; q, qmean, qcov_inv, q_new may be NULL */
int smcf_fault_sample_init_particles(int ndim, int N, uint64_t seed, const double* range, const double* qmean,
                                     const double* qcov_chol, double* particles);
int smcf_fault_mean_std(int n, const double* vec, double* mean, double* std_out);
int smcf_fault_find_next_gamma(int N, double gamma_prev, const double* likelihood, const double* q, double* weights,
                               double* gamma, double* neglog_evidence);
int smcf_fault_normalize_weights(int N, double* weights);
int smcf_fault_calc_mean_particles(int ndim, int N, const double* particles, const double* weights, double* mean);
int smcf_fault_calc_cov_particles(int ndim, int N, const double* particles, const double* weights, const double* mean,
                                  double* cov, int* info);
int smcf_fault_resample_particles(int N, const double* weights, uint64_t seed, int stage, int* assigned_num);
int smcf_fault_reorder_to_send(int ndim, int N, int* assigned_num, double* particles, double* likelihood, double* q,
                               int* sorted_idx);
int smcf_work_mcmc_sampling(smcf_problem* pb, uint64_t seed, int stage, double gamma, const double* range,
                            const double* cov_chol, int N, const int* assigned_num, const double* particles,
                            const double* q, const double* qmean, const double* qcov_inv, double log_qcov_det,
                            int max_batch, double* particles_new, double* likelihood_new, double* q_new,
                            int* acc_count);
/* RNG key of the inner SMC of one outer evaluation: (seed, outer stage, role 0 parent / 1 child, slot) */
uint64_t smcf_inner_key(uint64_t seed, uint32_t outer_stage, uint32_t role, uint32_t slot);

/* The outer reweight and resample of one stage on the device, in one call: fault_find_next_gamma,
 * fault_normalize_weights, fault_calc_mean_particles, fault_calc_cov_particles, fault_resample_particles and
 * fault_reorder_to_send (smc_fault.f90:372-637) - the components above, for callers that keep the stage's O(N)
 * work on the GPU.  Host arrays: particles (ndim x N), likelihood and q (q may be NULL: plain tempering) are
 * reordered in place as fault_reorder_to_send does; weights receives the normalised weights in the order before
 * the reorder, cov_chol the lower factor of 0.04 x the weighted covariance, assigned_num the counts (reordered),
 * sorted_idx the 1-based permutation.  The stage's sums are block reductions in a fixed order: gamma, the evidence
 * increment, the weights and the factor agree with the host components to rounding; the counts are those of
 * smcf_fault_resample_particles for the same weights.  device_sort = 0 orders the parents with qsort_ on the host
 * (the reference's order); 1 with a stable radix sort on the device (descending counts, tied parents in ascending
 * index, where std::sort leaves them in its own order).  ndim <= 64. */
int smcf_outer_reweight_resample(smcf_ctx* ctx, int ndim, int N, double gamma_prev, uint64_t seed, int stage,
                                 int device_sort, double* particles, double* likelihood, double* q, double* gamma,
                                 double* neglog_evidence, double* weights, double* cov_chol, int* info,
                                 int* assigned_num, int* sorted_idx);
/* Where smcf_fault_smc_exec / _gsmc run each stage's reweight and resample: 0 the host components (default, the
 * reference's arithmetic order), 1 smcf_outer_reweight_resample with the host sort, 2 with the device sort. */
int smcf_fault_smc_set_outer_stage(smcf_problem* pb, int mode);

/*
 * Groups: many independent inversions in one stage loop (no counterpart in the reference, whose only batching is
 * the MPI scatter of one inversion's outer particles, smc_fault.f90:896-1069).  The members are problems of
 * different meshes, observation sets, particle counts and forms (Gibbs mover), created on the group's context;
 * every stage kernel runs once over the instances of all members (the Gibbs sweeps once per kernel class: form
 * and block rows of the tuning table), and with graphs on (the default) a stage is one HIP graph launch.  Each
 * member's results equal those of the same calls on the member alone.  A problem belongs to one group at a time
 * and is destroyed after its group (smcf_problem_destroy refuses a member: SMCF_ESTATE); group outer SMCs need the
 * host outer stage (smcf_fault_smc_set_outer_stage 0, the default).
 *   smcf_group_calc_likelihood_batch  smcf_fault_calc_likelihood_batch for every member at once: B[i] geometries
 *                                     (theta[i] ndim_fault x B[i], keys[i]) stream through member i's slots;
 *                                     nstage may be NULL, B[i] may be 0
 *   smcf_group_fault_smc_exec         smcf_fault_smc_exec for every member at once (range[i], N[i], seed[i];
 *                                     particles[i] ndim_fault x N[i], likelihood[i] N[i]; nstage and evidence
 *                                     one entry per member, may be NULL): each member's outer SMC
 *                                     runs on its own host thread, its waves served by the shared stage loop
 *   smcf_group_set_graph              1: capture each call's stage launches into a HIP graph (default); 0: launch
 *                                     them one by one (identical results) */
typedef struct smcf_group smcf_group;
int smcf_group_create(smcf_ctx* ctx, int n, smcf_problem* const* members, smcf_group** out);
int smcf_group_set_graph(smcf_group* g, int on);
int smcf_group_calc_likelihood_batch(smcf_group* g, const int* B, const double* const* theta,
                                     const uint64_t* const* keys, double* const* neglog, int* const* nstage);
int smcf_group_fault_smc_exec(smcf_group* g, const double* const* range, const int* N, const uint64_t* seed,
                              int max_stages, double* const* particles, double* const* likelihood, int* nstage,
                              double* evidence);
int smcf_group_destroy(smcf_group* g);

/*
 * Serving (no counterpart in the reference, whose main.f90 runs one inversion per job).  A server admits
 * inversion requests - a problem description (Gibbs mover, any form) plus the outer-SMC settings of
 * smcf_fault_smc_exec - against a device-memory budget as they arrive (first fit in arrival order, sized by
 * smcf_problem_bytes) and batches them continuously: the server holds one context and one stage engine, an admitted
 * request's problem joins it as a member (resident for the whole outer SMC, driven by a host thread of its own), and
 * every stage kernel runs once over the instances of all running requests.  A request's result does not depend on
 * what runs beside it (keyed RNG, per-instance arithmetic).  smcf_server_submit copies its inputs (requests with an
 * HMC mover are refused, SMCF_EUNSUPPORTED); smcf_server_wait blocks until the request is done, returns its results
 * as smcf_fault_smc_exec would (gamma_trace: max_stages + 1 entries) and times[3] = seconds from server start to
 * submission, admission and completion.  Each request id is waited for once.
 */
typedef struct smcf_server smcf_server;
int smcf_server_create(int device, size_t hbm_budget_bytes, smcf_server** out);
int smcf_server_submit(smcf_server* s, const smcf_problem_desc* desc, int max_batch, const double* range, int N,
                       uint64_t seed, int max_stages, int64_t* id);
int smcf_server_wait(smcf_server* s, int64_t id, double* particles, double* likelihood, int* nstage,
                     double* evidence, double* gamma_trace, double* times);
int smcf_server_status(smcf_server* s, int* queued, int* running, size_t* bytes_in_use);
/* At most max_running requests run at once (1 .. 256, default 64; each holds a host thread); the others wait in
 * arrival order.  Admission also counts a fixed 16 MiB per request (engine tables and admission staging) beside its
 * smcf_problem_bytes. */
int smcf_server_set_max_running(smcf_server* s, int max_running);
int smcf_server_destroy(smcf_server* s);

/* sort.cpp:3-7 - 1-based index sorted by descending arr(index) (std::sort, same tie order). */
void qsort_(int* arr, int* index, int* n);

#ifdef __cplusplus
}
#endif
#endif
