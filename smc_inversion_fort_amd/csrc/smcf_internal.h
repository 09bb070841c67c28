// Internal structures of libsmcf (host side + kernel launch interfaces).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/smcf.h"

namespace smcf {

// most lanes per slip component in the pipelined prior chain (a 64-lane wavefront holds 32 / chain_lanes instances
// of chain_lanes lanes per component)
constexpr int kChainLanes = 32;

void set_error(const std::string& msg);

#define SMCF_HIP(call)                                                                        \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      ::smcf::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                   \
      return SMCF_EHIP;                                                                       \
    }                                                                                         \
  } while (0)

// Mesh/observation description resident on the device (read-only, shared by a batch).
struct MeshDev {
  int nplane, nnode, ndof, npatch, ndim, ndim_fault;
  int npts, nobs, nsar, ngnss;
  int fix_xbend;
  int chain_lag;            // pipelined prior chain: rounds between consecutive particles (> LL band per component)
  int chain_w;              // per-lane LDS ring length (power of two >= chain_lag + 1)
  int chain_lanes;          // lanes per slip component (power of two, (chain_lanes - 1) * chain_lag >= ndof)
  double xmin, xmax, zmin;
  const double* xbend;      // nplane-1 (>=1 allocated)
  const int* plane_nxi;     // nplane
  const int* plane_neta;    // nplane
  const int* plane_node0;   // nplane: first node (0-based)
  const int* cny;           // 4 x npatch, 1-based node ids
  const int* patch_plane;   // npatch
  const int* node_plane;    // nnode
  const int* node_k;        // nnode: 1-based column index inside its plane
  const int* node_l;        // nnode: 1-based row index inside its plane
  const int* id_dof;        // ndof, 1-based node ids
  const int* dof_of_node;   // nnode, 0-based dof index or -1
  const int* n2e;           // 4 x nnode, 1-based patch ids (node_to_elem_val)
  const int* n2e_slot;      // 4 x nnode, position of the node inside that patch (0..3)
  const int* n2es;          // nnode
  const double* pts;        // 2 x npts (SAR points then GNSS stations)
  const int* pt_row0;       // npts: first observation row (0-based)
  const int* pt_nrow;       // npts: 1 (SAR) or 3 (GNSS)
  const double* unitvec;    // 3 x nobs
  const double* obs_sigma;  // nobs
  const double* dvec;       // nobs
  // InSAR paths (rows 0 .. nsar-1): inverse covariance blocks, column-major per path, concatenated
  int npath;
  const double* sinv;       // sum over paths of n_p^2
  const int* sar_r0;        // npath: first row (0-based)
  const int* sar_n;         // npath: rows
  const long* sar_off;      // npath: offset of the block in sinv
  const int* row_path;      // nsar: path of each SAR row
};

// Per-batch device state.  Index b is the instance (geometry particle) slot.
struct BatchDev {
  int cap;                  // instance capacity
  int P;                    // inner particles
  double* theta;            // ndim_fault x cap
  double* geo;              // 16 x nplane x cap
  double* scal;             // 8 x cap: alpha2, log_sigma_sar2, log_sigma_gnss2, like_const, sigma_gnss2
  double* G;                // nobs x ndim x cap (column-major per instance)
  double* gscr;             // 24 x npts x npatch x cap (Green's phase-1 partial responses)
  double* luni;             // 9 x nnode x cap (stencil slots)
  int* li;                  // 9 x 2nnode x cap (ELL of L, 1-based columns)
  double* lv;
  int* lti;                 // 9 x 2ndof x cap (ELL of L^T, 1-based rows)
  double* ltv;
  int* llc_idx;             // 25 x ndim x cap (sparse columns of L^T L, 0-based rows, -1 pad)
  double* llc_val;
  double* Gw;               // ndim x nsar x cap: SAR rows of Sigma^-1 G, (1/sigma_sar^2) Sinv_p^T G_p (nsar > 0)
  double* sdv;              // nobs x cap: Sigma^-1 d (the reference's sdvec)
  double* H;                // ndim x ndim x cap
  double* bvec;             // ndim x cap
  double* dsd;              // cap
  int kb;                   // columns per block of the blocked Gibbs precision (gibbs_kb)
  double* Q;                // ndim x ndim x cap (Gibbs precision)
  float* Q32;               // same layout in FP32 (mixed precision only, else null)
  double* qdiag;            // 4 x ndim x cap: gamma b_i, 1/Q_ii, sqrt(Q_ii), 1/sqrt(Q_ii) of the stage
  double* S;                // ndim x P x cap, dimension-major (S[i*P + p])
  double* S2;
  float* S32;               // FP32 copies of S and S2 read by the mixed-precision Gibbs products (else null)
  float* S32b;
  // split-BF16 copies (SMCF_PRECISION_MIXED_BF16 only, else null), bf16 bits:
  //   SX, SXb  particles of S and S2 as [ndp/8][P][hi 8 | lo 8], ndp = nd rounded up to 32 (pad rows zero)
  //   QX       the stage's precision in matrix-core fragment order (k_form_qx)
  uint16_t* SX;
  uint16_t* SXb;
  uint16_t* QX;
  //   HX       H in fragment order (64-row blocks), formed once per instance for the centred likelihood
  //   DX       D = S - mref of the stage's moved particles, [ndp/8][P][hi 8 | lo 8] (pad rows zero)
  uint16_t* HX;
  uint16_t* DX;
  double* mref;             // ndim x cap: reference particle of the stage (centred likelihood)
  double* gref;             // ndim x cap: H mref - b
  double* lref;             // cap: L(mref)
  double* Y;                // ndim x P x cap (H X of the HMC leapfrog; null for the Gibbs mover)
  double* like;             // P x cap
  double* prior;
  double* like2;
  double* prior2;
  double* w;                // P x cap
  int* parent;              // P x cap
  int* assigned;            // P x cap
  double* mean;             // ndim x cap
  double* cov;              // ndim x cap
  // stage state
  uint64_t* key;            // cap
  double* gamma;            // cap
  double* neglog;           // cap
  int* iter;                // cap
  int* active;              // cap: runs the mover this stage
  int* done;                // cap
  int* par;                 // cap: which particle buffer (S or S2) is current
  int* sel;                 // 2 x cap: slots being admitted (two halves, used alternately)
  double* gtrace;           // 202 x cap
};

// state of the HMC mover (per instance: P columns / parents / children)
struct HmcDev {
  double *Xc, *Lc, *Rc, *postc;  // chain state per parent
  double *Xt, *Xt2, *Pm, *LS, *hcur;  // trajectory per wave column (positions, ping-pong; momentum; L x; H_cur)
  double* Qg;                         // gamma H + L^T L / alpha2 of the stage (ndim x ndim per instance)
  int* accf;                          // Metropolis outcome per wave column
  int *wla, *wlw, *wln;               // work lists (instance, column tile) of the sampling and tuning passes
  double* part;                       // per-column partial sums of the wave kernels (32 x 6 x P per instance)
  void* QgX;                          // split-BF16 Q_g in fragment order (SMCF_PRECISION_MIXED_BF16)
  void *XXa, *XXb;                    // split-BF16 positions (ping-pong with Xt / Xt2), [d/8][column][hi 8 | lo 8]
  double *dtau, *dtmp, *score;   // per child
  double *ldlo, *ldhi;           // per instance log-dtau range
  int *wptr, *wcols, *tptr, *tcols, *wamax, *tamax;  // wave lists (sampling / tuning)
  int *ncol, *cpar, *cchild, *dead, *idstart, *flag;
};

// host side of one HMC pass (tuning or sampling) of a stage: the wave pointers read back once per stage and the
// work lists built from them (2 ints per item: instance, column tile; la: every instance's 64-column tiles, lw:
// those of the instances the wide kernels take, ln: the 16-column tiles of the others; off*[k] = wave k's first
// item)
struct HmcPassHost {
  std::vector<int> ptr;
  std::vector<int> la, lw, ln;
  std::vector<int> offa, offw, offn;
};

struct HmcHost {
  std::vector<int> amax;  // wave counts per instance: sampling (B), then tuning (B)
  HmcPassHost pass[2];    // sampling, tuning
};

// rows of the split-BF16 position copies (whole 32-row contraction steps; rows past ndim stay zero)
__host__ __device__ inline size_t hmc_xx_rows(int nd) { return (size_t)((nd + 31) / 32) * 32; }

// 16-byte fragments (hi and lo) of the split-BF16 Q_g of one instance
inline size_t hmc_qgx_frags(int nd) { return (size_t)((nd + 15) / 16) * ((nd + 31) / 32) * 64 * 2; }

// entries (ints) of one pass's work list: per instance at most P / 16 + P + 1 items of 2 ints
inline size_t hmc_work_entries(int cap, int P) { return (size_t)cap * 2 * ((size_t)P + P / 16 + 1); }

// One problem's part of a stage launch (a device table of these lets one launch cover instances of problems with
// different meshes, observation sets and particle counts): its per-slot buffers, dimensions and sampler settings.
struct MemberDev {
  BatchDev bt;
  int nd, nf, P, kb, sweeps;
  int chain_lag, chain_w, chain_lanes;  // pipelined prior chain (MeshDev)
  double max_slip;
};

// The instances one stage launch covers (device arrays built by the stage engine):
//   inst[g] = (member, slot), g < ninst, for the kernels with one workgroup (or grid row) per instance;
//   items = (member, slot, chunk, 0), one workgroup per 256 particles, for the chunked kernels, in XCD-aware
//   order: item w runs on XCD w % 8 and all chunks of one instance share w % 8, so an instance's matrices are
//   fetched into one L2; member -1 marks an empty item.
struct StageSet {
  const MemberDev* mem;
  const int2* inst;
  int ninst;
  int nd_max, P_max, kb_max;
};
struct ItemList {
  const int4* items;
  int n;
};

// Member table of one Gibbs-sweep launch, passed as a kernel argument (at most kGibbsArgMembers members per launch;
// the items' member field indexes it).  Kernel arguments are constant for the launch, so the compiler may reload
// these values instead of keeping them in registers: read from a device table they pushed the split-BF16 sweep
// kernel past 256 VGPRs into scratch memory.
constexpr int kGibbsArgMembers = 16;
struct GibbsMember {
  int nd, P, sweeps, pad;
  double max_slip;
  const double* Q;
  const void* QX;     // split BF16 (bf16x8 fragments)
  const float* Q32;   // FP32 form
  double* S;
  double* S2;
  void* SX;           // split BF16
  void* SXb;
  float* S32;         // FP32 form
  float* S32b;
  void* DX;           // split BF16
  const double* mref;
  const double* qdiag;
  const uint64_t* key;
  const int* iter;
  const int* active;
  const int* par;
};
struct GibbsArgs {
  GibbsMember m[kGibbsArgMembers];
};

struct Engine;

struct Problem {
  smcf_ctx* ctx;
  MeshDev m;
  BatchDev bt;
  HmcDev hm;
  HmcHost hh;
  MemberDev* mdev = nullptr;  // this problem as a one-entry member table (device)
  Engine* engine = nullptr;   // set while the problem is a member of a running stage engine (group, server): its
                              // batch calls are queued to that engine
  double max_slip;
  int P, mover, sweeps, precision;
  int hmc_ntau;       // leapfrog steps (50 for SMCF_MOVER_HMC, smc_slip.f90:669)
  double hmc_accept;  // target acceptance of SMCF_MOVER_HMC_ACCEPT
  std::vector<int> sar_r0, sar_n;  // host copies of the InSAR path layout (launch geometry)
  std::vector<long> sar_off;
  std::vector<void*> allocs;
  bool in_group = false;  // a member of a group (smcf_group_create)
  bool sizing = false;  // count the device bytes of a description without allocating (smcf_problem_bytes)
  size_t bytes = 0;     // device bytes allocated (or counted) for the problem
  int alloc_rc = 0;     // first failure of a problem allocation or upload (SMCF_* code), 0 while none
  std::vector<double> t_done;  // seconds from the start of the last batch call to each geometry's result
  std::vector<double> t_admit; // seconds from the start of the last batch call to each geometry's set-up launch
};

inline GibbsMember gibbs_member_of(const Problem& P) {
  GibbsMember g{};
  g.nd = P.m.ndim;
  g.P = P.P;
  g.sweeps = P.sweeps;
  g.max_slip = P.max_slip;
  g.Q = P.bt.Q;
  g.QX = P.bt.QX;
  g.Q32 = P.bt.Q32;
  g.S = P.bt.S;
  g.S2 = P.bt.S2;
  g.SX = P.bt.SX;
  g.SXb = P.bt.SXb;
  g.S32 = P.bt.S32;
  g.S32b = P.bt.S32b;
  g.DX = P.bt.DX;
  g.mref = P.bt.mref;
  g.qdiag = P.bt.qdiag;
  g.key = P.bt.key;
  g.iter = P.bt.iter;
  g.active = P.bt.active;
  g.par = P.bt.par;
  return g;
}

inline MemberDev member_of(const Problem& P) {
  MemberDev d{};
  d.bt = P.bt;
  d.nd = P.m.ndim;
  d.nf = P.m.ndim_fault;
  d.P = P.P;
  d.kb = P.bt.kb;
  d.sweeps = P.sweeps;
  d.chain_lag = P.m.chain_lag;
  d.chain_w = P.m.chain_w;
  d.chain_lanes = P.m.chain_lanes;
  d.max_slip = P.max_slip;
  return d;
}


// kernels (launchers return hipError_t of the launch)
// sel (optional): the B instance slots to work on, instead of slots 0 .. B-1
hipError_t launch_geometry(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel = nullptr);
hipError_t launch_greens(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel = nullptr);
hipError_t launch_laplacian(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel = nullptr);
// Gram quantities H, b, dsd (InSAR paths need the host copies of the path layout in P); with scal_given the
// sigmas already in bt.scal are used instead of those of theta
hipError_t launch_gram(const Problem& P, const MeshDev& m, BatchDev& bt, int B, hipStream_t s,
                       const int* sel = nullptr, bool scal_given = false);
// prior chains: entry k is (member, slot) = list[k] of the member table mem, or with list == nullptr member 0 and slot
// sel[k] (slot k when sel == nullptr).  Wavefront w runs the entries jobs[w] = (first, count), which belong to one
// member and number at most 32 / chain_lanes of it; with jobs == nullptr wavefront w runs entry w alone (nw entries).
// chain_w = the largest ring length of the members involved.
hipError_t launch_init_chain(const MemberDev* mem, const int2* list, const int* sel, const int2* jobs, int nw,
                             int chain_w, hipStream_t s);
// likelihood of the current particles of the given slots of one problem (matrix cores; Gram form; set-up)
hipError_t launch_like(const MeshDev& m, BatchDev& bt, int B, int use_active, hipStream_t s, const int* sel = nullptr);
// prior of the current particles
hipError_t launch_prior(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel = nullptr);
// mean and variance diagonal of P dimension-major particles S with normalised weights w
hipError_t launch_moments_raw(int nd, int P, const double* S, const double* w, double* mean, double* cov,
                              hipStream_t s);
// Stage kernels over a stage set (every instance of every member; instances not active this stage are skipped)
hipError_t launch_gamma_resample(const StageSet& ss, hipStream_t s);
hipError_t launch_moments(const StageSet& ss, hipStream_t s);
hipError_t launch_gather(const StageSet& ss, hipStream_t s);
// the stage's Gibbs precision: the FP64 blocked layout (any_full: FP64 and FP32 members), its FP32 copy (any_f32),
// the split-BF16 fragment layout with the FP64 diagonal blocks (any_bf), and the diagonal constants
hipError_t launch_form_q(const StageSet& ss, bool any_full, bool any_f32, bool any_bf, hipStream_t s);
// Gibbs sweeps of the items of one launch of a kernel class (precision, block rows); items index the launch's
// member table
hipError_t launch_gibbs(const GibbsArgs& a, const ItemList& it, int precision, int kb, hipStream_t s);
// likelihood of the moved particles: Gram form on the FP64 matrix cores (FP64 and FP32 members) or the centred
// split-BF16 form (split-BF16 members; launch_like_ref before the sweeps)
hipError_t launch_like_stage(const StageSet& ss, const ItemList& it, hipStream_t s);
hipError_t launch_like_ref(const StageSet& ss, hipStream_t s);
hipError_t launch_like_bf(const StageSet& ss, const ItemList& it, hipStream_t s);
// columns per block of the blocked Gibbs precision layout (SMCF_PRECISION_*)
int gibbs_kb(int precision, int nd);
bool gibbs_rows_valid(int precision, int rows);
void set_gibbs_rows(int precision, int lo, int hi, int rows);
// FP32 copies of the current particles of the given slots (mixed precision)
hipError_t launch_s32(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel = nullptr);
// split-BF16 copies of the current particles of the given slots (SMCF_PRECISION_MIXED_BF16)
hipError_t launch_sbf(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel = nullptr);
// centred split-BF16 likelihood (SMCF_PRECISION_MIXED_BF16): H in fragment order for the given slots (set-up)
hipError_t launch_h_bf(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel = nullptr);
// coordinates of the split-BF16 particle copies (nd rounded up to whole 32-row matrix-core steps)
__host__ __device__ inline int bf_rows(int nd) { return (nd + 31) & ~31; }
// bf16 values of one instance's fragment-ordered precision QX (KB-row blocks x 32-column steps)
__host__ __device__ inline size_t bf_q_elems(int nd, int kb) { return (size_t)((nd + kb - 1) / kb) * kb * bf_rows(nd) * 2; }

// HMC movers: SMCF_MOVER_HMC (ESJD tuning, ntau 50) or SMCF_MOVER_HMC_ACCEPT (ntau steps, step size adapted
// toward `accept`)
struct HmcCfg {
  int mover, ntau;
  double accept;
  int bf;  // leapfrog gradients from split-BF16 products (SMCF_PRECISION_MIXED_BF16)
};
hipError_t launch_hmc_init(BatchDev& bt, HmcDev& h, int B, hipStream_t s, const int* sel, int mover);
hipError_t launch_hmc_stage(const MeshDev& m, BatchDev& bt, HmcDev& h, HmcHost& hh, int B, double max_slip,
                            const HmcCfg& cfg, hipStream_t s);
// one instance in slot 0: the tempered gradient of columns 0 .. n-1 of h.Xt into G (ndim x n, column per
// particle), and n independent leapfrog trajectories from (h.Xt, h.Pm) with bt.cov as the mass diagonal
// (H(cur) into hc_out, H(cand) into hn_out, both device arrays)
hipError_t launch_prior_raw(int nn2, int n, int ld, const double* X, const double* a2, const int* li, const double* lv,
                            double* out, hipStream_t s);
hipError_t launch_grad_raw(int nd, int nn2, int n, int ld, double gamma, const double* H, const double* X,
                           const double* bv, const double* a2, const int* li, const double* lv, const int* lti,
                           const double* ltv, const int* ncol_active, double* Y, double* LS, double* G,
                           hipStream_t s);
// Gram-form likelihood of one instance's particles from raw arrays (k_like_mfma)
hipError_t launch_like_raw(int nd, int P, const double* H, const double* bv, const double* dsd, const double* scal,
                           const double* S, const int* zero, double* out, hipStream_t s);
hipError_t launch_grad_cols(const MeshDev& m, BatchDev& bt, HmcDev& h, int n, double gamma, double* G, hipStream_t s);
hipError_t launch_leapfrog_cols(const MeshDev& m, BatchDev& bt, HmcDev& h, int n, double gamma, double dtau, int ntau,
                                double max_slip, double* hc_out, double* hn_out, double* lk_out, double* pr_out,
                                hipStream_t s);

hipError_t launch_dc3d0_batch(int n, const double* in, double* u, int* iret, hipStream_t s);
hipError_t launch_call_dc3d0(int n, const double* in, double* uret, hipStream_t s);
hipError_t launch_calc_responce(int npts, const double* pts, const int* cny, const double* coor, const double* slip,
                                double xf, double yf, double zf, double strike, double dip, int ntarget,
                                const int* target, double* uobs, hipStream_t s);
hipError_t launch_find_next_gamma(int P, double gprev, const double* L, double cv, double tol, int cap,
                                  double* out3, double* w, hipStream_t s);
constexpr int kMaxOuterDim = 64;  // geometry parameters of the outer SMC (6 per plane at most, 3 nplane + 3)
hipError_t launch_outer_stage(int N, int nd, double gprev, const double* L, const double* q, const double* X,
                              uint64_t okey, int stage, double* w, double* scal, double* cov, int* assigned,
                              hipStream_t s);
hipError_t launch_outer_sort(int N, int nd, const int* assigned, int* idx, int* tmp, const double* X_in,
                             const double* L_in, const double* q_in, int* a_out, double* X_out, double* L_out,
                             double* q_out, int* sorted_idx, hipStream_t s);
hipError_t launch_normalize_resample(int P, const double* w, double uh, double ul, double* wn,
                                     int* assigned, hipStream_t s);
hipError_t launch_trunc_normal_f32(int n, const double* mu, const double* sq, double lo, double hi, const float* u,
                                   double* out, hipStream_t s);
hipError_t launch_trunc_normal(int n, const double* mu, const double* sig, double lo, double hi,
                               const double* u, double* out, hipStream_t s);
hipError_t launch_philox(int n, const uint32_t* ctr, const uint32_t* key, uint32_t* out, hipStream_t s);

}  // namespace smcf

struct smcf_problem {
  smcf::Problem p;
};

struct smcf_ctx {
  int device;
  hipStream_t stream;
  hipStream_t setup_stream;  // set-up of streamed-in instances, overlapping the stage loop
  double* scratch = nullptr;  // device address of the no-allocation entry point's scratch array on this device
  bool profiling;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  struct Stat {
    std::string name;
    long count;
    double total_ms;
  };
  std::vector<Stat> stats;
};

namespace smcf {

// Device time of a named section of the library's stream (smcf_set_profiling): events around the section,
// resolved into ctx->stats by resolve_profiles.
struct ProfScope {
  smcf_ctx* ctx;
  const char* name;
  hipStream_t st;
  hipEvent_t a{}, b{};
  ProfScope(smcf_ctx* c, const char* n, hipStream_t s = nullptr) : ctx(c), name(n), st(s ? s : c->stream) {
    if (ctx->profiling) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, st);
    }
  }
  ~ProfScope() {
    if (ctx->profiling) {
      (void)hipEventRecord(b, st);
      ctx->pending.push_back({name, {a, b}});
    }
  }
};
void resolve_profiles(smcf_ctx* ctx);

// the outer-stage mode set by smcf_fault_smc_set_outer_stage (0: host components)
int outer_stage_mode(const smcf_problem* pb);

// Set-up of n instances of one problem on stream s (slots sel[0 .. n-1], device array; thetas and keys already in
// their slots), left held (marked done) until the stage engine releases them: setup_pre forms everything up to
// the Gram quantities, the prior chain runs once over every instance admitted together (launch_init_chain), and
// setup_post forms the particle copies, the likelihood and the prior of the initial particles.
int setup_pre(Problem& P, int n, const int* sel, hipStream_t s);
int setup_post(Problem& P, int n, const int* sel, hipStream_t s);

// The inner SMC of N geometries of one problem streamed through its slots (the stage engine with one member):
// -log Z into neglog and the stage counts into nstage (either may be null); want_moments keeps the weighted mean
// and variance of every stage's particles.
int run_stream(Problem& P, int N, const double* theta, const uint64_t* keys, bool want_moments, double* neglog,
               int* nstage);
// batch call of a problem that is a member of a running engine: queued to it, returns when the engine has finished
// all N geometries
int engine_submit_wait(Problem& P, int N, const double* theta, const uint64_t* keys, double* neglog, int* nstage);

// Serving: one persistent engine on the server's context whose members (the admitted requests' problems) join
// (serve_engine_add, from any thread; `added` runs on the engine's thread once the problem is a member, so its
// client may start) and leave once their client has finished (serve_engine_leave; `removed` runs on the engine's
// thread after the problem left).  serve_engine_run serves until serve_engine_stop and the last member has left.
Engine* serve_engine_create(smcf_ctx* ctx, std::function<void(Problem*)> added, std::function<void(Problem*)> removed);
int serve_engine_run(Engine* E);
void serve_engine_add(Engine* E, Problem* P);
void serve_engine_leave(Engine* E, Problem* P);
void serve_engine_stop(Engine* E);
void serve_engine_destroy(Engine* E);
int serve_engine_error(Engine* E, std::string* msg);
// held by a thread that calls HIP beside the engine's thread (creating a problem): a stream capture of the stage
// graph must not see such calls
std::mutex& serve_engine_capture_lock(Engine* E);

}  // namespace smcf
