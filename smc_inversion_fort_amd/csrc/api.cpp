// C ABI of libsmcf: contexts, problems, the batched inner-SMC driver and component entry
// points.  The stage loop mirrors slip_smc_exec (smc_slip.f90:1445-1663) for many geometries
// at once: every kernel of a stage runs over all instances, instances that have reached
// gamma = 1 (or the 200-stage cap) are masked, and the host stops when all are done.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "smcf_internal.h"

namespace smcf {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace smcf

using namespace smcf;

static constexpr int kMaxInnerParticles = 3584;

// Scratch of smcf_find_next_gamma_noalloc: the device copies of the likelihoods and the weights (kNoAllocMaxP
// doubles each) and the two results, in one array that every device holds for the life of the process.  Host
// code never passes `scratch` itself to a kernel or a copy: its address on a device comes from
// hipGetSymbolAddress, resolved once per device by smcf_ctx_create and kept in g_scratch.
static constexpr int kNoAllocMaxP = kMaxInnerParticles;
static constexpr size_t kScratchDoubles = 2 * (size_t)kNoAllocMaxP + 4;
static __device__ __align__(256) double scratch[kScratchDoubles];
static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch;  // per device index; nullptr until resolved

// Device address of `scratch` on `device`, which must be the current device.
static int scratch_resolve(int device, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch.size() <= device) g_scratch.resize((size_t)device + 1, nullptr);
  if (!g_scratch[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return SMCF_EHIP;
    }
    g_scratch[device] = (double*)p;
  }
  *out = g_scratch[device];
  return SMCF_OK;
}


#define CHECK_LAUNCH(expr)                                                           \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                  \
      return SMCF_EHIP;                                                              \
    }                                                                                \
  } while (0)

void smcf::resolve_profiles(smcf_ctx* ctx) {
  for (auto& pe : ctx->pending) {
    float ms = 0.f;
    (void)hipEventSynchronize(pe.second.second);
    (void)hipEventElapsedTime(&ms, pe.second.first, pe.second.second);
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
    auto it = std::find_if(ctx->stats.begin(), ctx->stats.end(),
                           [&](const smcf_ctx::Stat& s) { return s.name == pe.first; });
    if (it == ctx->stats.end()) ctx->stats.push_back({pe.first, 1, (double)ms});
    else { it->count++; it->total_ms += ms; }
  }
  ctx->pending.clear();
}

namespace {

// Allocations of a problem.  After the first failure every later call returns that failure's code without
// allocating, so problem_build can issue its allocations in sequence and check P.alloc_rc once.
template <class T>
int dalloc(Problem& P, T** ptr, size_t n) {
  void* p = nullptr;
  *ptr = nullptr;
  if (P.alloc_rc) return P.alloc_rc;
  if (n == 0) n = 1;
  P.bytes += (n * sizeof(T) + 255) & ~(size_t)255;
  if (P.sizing) {
    *ptr = nullptr;
    return SMCF_OK;
  }
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc(") + std::to_string(n * sizeof(T)) + " bytes): " + hipGetErrorString(e));
    P.alloc_rc = SMCF_ENOMEM;
    return SMCF_ENOMEM;
  }
  P.allocs.push_back(p);
  *ptr = (T*)p;
  return SMCF_OK;
}

// Temporary device buffers of one component entry point's call, freed on every return path (hipFree waits for
// the device, so kernels still reading them have finished).
struct TmpBufs {
  std::vector<void*> p;
  TmpBufs() = default;
  TmpBufs(const TmpBufs&) = delete;
  TmpBufs& operator=(const TmpBufs&) = delete;
  ~TmpBufs() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  hipError_t alloc(T** out, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(T) * (n ? n : 1));
    if (e == hipSuccess) p.push_back(q);
    *out = (T*)q;
    return e;
  }
};

template <class T>
int upload(Problem& P, const T** dst, const std::vector<T>& v) {
  T* d = nullptr;
  int rc = dalloc(P, &d, v.size());
  if (rc) return rc;
  if (!v.empty() && !P.sizing) {
    hipError_t e = hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      set_error(std::string("hipMemcpy (problem upload): ") + hipGetErrorString(e));
      P.alloc_rc = SMCF_EHIP;
      return SMCF_EHIP;
    }
  }
  *dst = d;
  return SMCF_OK;
}

// hold = 1 keeps the slot marked done (skipped by the stage kernels) until k_activate releases it
__global__ void k_state_init(BatchDev bt, int B, const int* __restrict__ sel, int hold) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  if (sel) b = sel[b];
  bt.gamma[b] = 0.0;
  bt.neglog[b] = 0.0;
  bt.iter[b] = 1;
  bt.active[b] = 0;
  bt.done[b] = hold;
  bt.par[b] = 0;
}

}  // namespace

namespace smcf {

int setup_pre(Problem& P, int n, const int* sel, hipStream_t s) {
  smcf_ctx* ctx = P.ctx;
  MeshDev& m = P.m;
  BatchDev& bt = P.bt;
  hipLaunchKernelGGL(k_state_init, dim3((n + 63) / 64), dim3(64), 0, s, bt, n, sel, 1);
  if (P.mover != SMCF_MOVER_GIBBS) CHECK_LAUNCH(launch_hmc_init(bt, P.hm, n, s, sel, P.mover));
  {
    ProfScope ps(ctx, "geometry_greens", s);
    CHECK_LAUNCH(launch_geometry(m, bt, n, s, sel));
    CHECK_LAUNCH(launch_greens(m, bt, n, s, sel));
  }
  CHECK_LAUNCH(launch_laplacian(m, bt, n, s, sel));
  {
    ProfScope ps(ctx, "gram", s);
    CHECK_LAUNCH(launch_gram(P, m, bt, n, s, sel));
  }
  if (bt.HX) CHECK_LAUNCH(launch_h_bf(m, bt, n, s, sel));
  return SMCF_OK;
}

int setup_post(Problem& P, int n, const int* sel, hipStream_t s) {
  MeshDev& m = P.m;
  BatchDev& bt = P.bt;
  if (bt.S32) CHECK_LAUNCH(launch_s32(m, bt, n, s, sel));
  if (bt.SX) CHECK_LAUNCH(launch_sbf(m, bt, n, s, sel));
  CHECK_LAUNCH(launch_like(m, bt, n, 0, s, sel));
  CHECK_LAUNCH(launch_prior(m, bt, n, s, sel));
  return SMCF_OK;
}

}  // namespace smcf

extern "C" {

const char* smcf_last_error(void) { return g_err.c_str(); }

int smcf_ctx_create(int device, smcf_ctx** out) {
  if (!out) return SMCF_EINVAL;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    set_error("no HIP device available");
    return SMCF_EHIP;
  }
  if (device < 0 || device >= n) {
    set_error("device index out of range");
    return SMCF_EINVAL;
  }
  SMCF_HIP(hipSetDevice(device));
  double* scr = nullptr;
  int rs = scratch_resolve(device, &scr);
  if (rs) return rs;
  auto* c = new smcf_ctx();
  c->device = device;
  c->profiling = false;
  c->scratch = scr;
  hipError_t es = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (es == hipSuccess) {
    es = hipStreamCreateWithFlags(&c->setup_stream, hipStreamNonBlocking);
    if (es != hipSuccess) (void)hipStreamDestroy(c->stream);
  }
  if (es != hipSuccess) {
    delete c;
    set_error(std::string("hipStreamCreate: ") + hipGetErrorString(es));
    return SMCF_EHIP;
  }
  *out = c;
  return SMCF_OK;
}

int smcf_ctx_destroy(smcf_ctx* ctx) {
  if (!ctx) return SMCF_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamSynchronize(ctx->setup_stream);
  resolve_profiles(ctx);
  (void)hipStreamDestroy(ctx->setup_stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SMCF_OK;
}

int smcf_set_profiling(smcf_ctx* ctx, int on) {
  if (!ctx) return SMCF_EINVAL;
  ctx->profiling = on != 0;
  if (on) ctx->stats.clear();
  return SMCF_OK;
}

int smcf_kernel_stats(smcf_ctx* ctx, const char* name, long* count, double* total_ms) {
  if (!ctx || !name) return SMCF_EINVAL;
  (void)hipStreamSynchronize(ctx->stream);
  resolve_profiles(ctx);
  *count = 0;
  *total_ms = 0.0;
  for (auto& s : ctx->stats)
    if (s.name == name) { *count = s.count; *total_ms = s.total_ms; }
  return SMCF_OK;
}

// Builds a problem on ctx's device, or (sizing) only counts the device bytes it would allocate.
static int problem_build(smcf_ctx* ctx, const smcf_problem_desc* d, int max_batch, bool sizing, smcf_problem** out,
                         size_t* bytes) {
  if ((!ctx && !sizing) || !d || (!out && !sizing) || max_batch < 1 || d->nplane < 1 || d->nparticle_slip < 1) {
    set_error("invalid problem description");
    return SMCF_EINVAL;
  }
  if (d->nparticle_slip > kMaxInnerParticles) {
    set_error("nparticle_slip > " + std::to_string(kMaxInnerParticles) +
              ": the tempering workgroup keeps all weights in 64 KiB of LDS");
    return SMCF_EINVAL;
  }
  if (!sizing) SMCF_HIP(hipSetDevice(ctx->device));
  auto holder = std::make_unique<smcf_problem>();
  Problem& P = holder->p;
  P.ctx = ctx;
  P.sizing = sizing;
  P.max_slip = d->max_slip;
  P.P = d->nparticle_slip;
  P.mover = d->mover;
  P.sweeps = d->gibbs_sweeps > 0 ? d->gibbs_sweeps : 5;
  P.precision = d->precision;
  P.hmc_ntau = P.mover == SMCF_MOVER_HMC ? 50 : (d->hmc_ntau > 0 ? d->hmc_ntau : 10);
  P.hmc_accept = d->hmc_accept > 0.0 ? d->hmc_accept : 0.65;
  if (P.mover != SMCF_MOVER_HMC && P.mover != SMCF_MOVER_GIBBS && P.mover != SMCF_MOVER_HMC_ACCEPT) {
    set_error("unknown mover " + std::to_string(P.mover));
    return SMCF_EINVAL;
  }
  if (P.mover == SMCF_MOVER_HMC_ACCEPT && (P.hmc_ntau > 1000 || P.hmc_accept >= 1.0)) {
    set_error("hmc_ntau must be 1 .. 1000 and hmc_accept in (0, 1)");
    return SMCF_EINVAL;
  }
  if (P.precision != SMCF_PRECISION_FP64 && P.precision != SMCF_PRECISION_MIXED &&
      P.precision != SMCF_PRECISION_MIXED_BF16) {
    set_error("unknown precision " + std::to_string(P.precision));
    return SMCF_EINVAL;
  }
  if (P.precision == SMCF_PRECISION_MIXED && P.mover != SMCF_MOVER_GIBBS) {
    set_error("the FP32 form is implemented for the Gibbs mover only (HMC movers take SMCF_PRECISION_MIXED_BF16)");
    return SMCF_EUNSUPPORTED;
  }
  MeshDev& m = P.m;
  memset(&m, 0, sizeof(m));
  int np = d->nplane;
  m.nplane = np;
  m.fix_xbend = d->fix_xbend;
  m.xmin = d->xmin; m.xmax = d->xmax; m.zmin = d->zmin;
  m.ndim_fault = d->fix_xbend ? 2 * np + 4 : 3 * np + 3;
  // mesh topology (discretize_fault, init.f90:198-293; sizes main.f90:217-238)
  std::vector<int> pnxi(np), pneta(np), pnode0(np), cny, patch_plane, node_plane, node_k, node_l, id_dof;
  int nnode = 0, npatch = 0;
  for (int ip = 0; ip < np; ++ip) {
    pnxi[ip] = d->nxi_ls[ip];
    pneta[ip] = d->neta_ls[ip];
    if (pnxi[ip] < 1 || pneta[ip] < 1) { set_error("nxi/neta must be >= 1"); return SMCF_EINVAL; }
    pnode0[ip] = nnode;
    int W = pnxi[ip] + 1;
    for (int j = 1; j <= pneta[ip] + 1; ++j)
      for (int i = 1; i <= W; ++i) {
        node_plane.push_back(ip);
        node_k.push_back(i);
        node_l.push_back(j);
        bool dof = !(j == 1 || (ip == 0 && i == 1) || (ip == np - 1 && i == W));
        if (dof) id_dof.push_back(nnode + i + W * (j - 1));
      }
    for (int j = 1; j <= pneta[ip]; ++j)
      for (int i = 1; i <= pnxi[ip]; ++i) {
        int n1 = i + W * (j - 1) + nnode;
        cny.insert(cny.end(), {n1, n1 + 1, n1 + 1 + W, n1 + W});
        patch_plane.push_back(ip);
      }
    nnode += W * (pneta[ip] + 1);
    npatch += pnxi[ip] * pneta[ip];
  }
  int ndof = (int)id_dof.size();
  std::vector<int> dof_of_node(nnode, -1), n2e(4 * nnode, -1), n2e_slot(4 * nnode, 0), n2es(nnode, 0);
  for (int k = 0; k < ndof; ++k) dof_of_node[id_dof[k] - 1] = k;
  for (int e = 0; e < npatch; ++e)
    for (int a = 0; a < 4; ++a) {
      int nd0 = cny[4 * e + a] - 1;
      int c = n2es[nd0]++;
      n2e[4 * nd0 + c] = e + 1;
      n2e_slot[4 * nd0 + c] = a;
    }
  m.nnode = nnode; m.ndof = ndof; m.npatch = npatch; m.ndim = 2 * ndof;
  {
    // LL couples a DOF only with DOFs of the 5 x 5 node window on its own plane (one slip component);
    // the prior chain's pipeline lag must exceed the widest forward reach in DOF order
    int band = 0;
    for (int k = 0; k < ndof; ++k) {
      int ji = id_dof[k] - 1, ip = node_plane[ji], W = pnxi[ip] + 1;
      for (int wl = -2; wl <= 2; ++wl)
        for (int wk = -2; wk <= 2; ++wk) {
          int kk = node_k[ji] + wk, ll = node_l[ji] + wl;
          if (kk < 1 || kk > W || ll < 1 || ll > pneta[ip] + 1) continue;
          int dr = dof_of_node[pnode0[ip] + (kk - 1) + (ll - 1) * W];
          if (dr >= 0) band = std::max(band, dr - k);
        }
    }
    // L lanes per component need lag >= band + 1 and (L - 1) lag >= ndof; a chain runs (P - 1) lag + ndof rounds and
    // a wavefront holds 32 / L chains.  The fewest lanes whose chain is at most a quarter longer than with 32 are
    // taken (fewer chain wavefronts beside the stage kernels), with the ring within 128 values per lane.
    auto lag_for = [&](int L) { return std::max(band + 1, (ndof + L - 2) / (L - 1)); };
    auto rounds_for = [&](int L) { return (double)(P.P - 1) * lag_for(L) + ndof; };
    m.chain_lanes = kChainLanes;
    for (int L = 2; L < kChainLanes; L *= 2)
      if (lag_for(L) < 128 && rounds_for(L) <= 1.25 * rounds_for(kChainLanes)) {
        m.chain_lanes = L;
        break;
      }
    m.chain_lag = lag_for(m.chain_lanes);
    m.chain_w = 1;
    while (m.chain_w < m.chain_lag + 1) m.chain_w *= 2;
  }
  // observation points: SAR rows, then one point per GNSS triplet (gfunc.f90:212-239)
  m.nobs = d->nobs; m.nsar = d->nsar_total; m.ngnss = d->ngnss;
  if (m.nsar + 3 * m.ngnss != m.nobs) { set_error("nobs != nsar_total + 3*ngnss"); return SMCF_EINVAL; }
  std::vector<double> pts;
  std::vector<int> row0, nrow;
  for (int r = 0; r < m.nsar; ++r) {
    pts.push_back(d->obs_points[2 * r]); pts.push_back(d->obs_points[2 * r + 1]);
    row0.push_back(r); nrow.push_back(1);
  }
  for (int g = 0; g < m.ngnss; ++g) {
    int r = m.nsar + 3 * g;
    pts.push_back(d->obs_points[2 * r]); pts.push_back(d->obs_points[2 * r + 1]);
    row0.push_back(r); nrow.push_back(3);
  }
  m.npts = (int)row0.size();
  std::vector<double> xb(std::max(1, np - 1), 0.0);
  for (int k = 0; k < np - 1; ++k) xb[k] = d->xbend ? d->xbend[k] : 0.0;
  std::vector<double> uv(d->obs_unitvec, d->obs_unitvec + 3 * (size_t)m.nobs);
  std::vector<double> os(d->obs_sigma, d->obs_sigma + m.nobs), dv(d->dvec, d->dvec + m.nobs);
  // every device buffer allocated so far is freed on any return before the problem is handed out
  struct FreeOnFail {
    Problem* p;
    ~FreeOnFail() {
      if (p)
        for (void* q : p->allocs) (void)hipFree(q);
    }
  } guard{&P};
  upload(P, &m.xbend, xb);
  upload(P, &m.plane_nxi, pnxi);
  upload(P, &m.plane_neta, pneta);
  upload(P, &m.plane_node0, pnode0);
  upload(P, &m.cny, cny);
  upload(P, &m.patch_plane, patch_plane);
  upload(P, &m.node_plane, node_plane);
  upload(P, &m.node_k, node_k);
  upload(P, &m.node_l, node_l);
  upload(P, &m.id_dof, id_dof);
  upload(P, &m.dof_of_node, dof_of_node);
  upload(P, &m.n2e, n2e);
  upload(P, &m.n2e_slot, n2e_slot);
  upload(P, &m.n2es, n2es);
  upload(P, &m.pts, pts);
  upload(P, &m.pt_row0, row0);
  upload(P, &m.pt_nrow, nrow);
  upload(P, &m.unitvec, uv);
  upload(P, &m.obs_sigma, os);
  upload(P, &m.dvec, dv);
  // InSAR paths: rows nsar_index(p) .. nsar_index(p+1)-1 (1-based) with the inverse covariance blocks the
  // caller computed (calc_sigma_sar_mat, init.f90:82-119), column-major, concatenated in path order
  m.npath = 0;
  if (m.nsar > 0) {
    if (d->npath < 1 || !d->nsar_index || !d->sar_sinv) {
      set_error("InSAR rows need npath >= 1, nsar_index and sar_sinv");
      return SMCF_EINVAL;
    }
    std::vector<int> r0v, nv, rowp(m.nsar);
    std::vector<long> offv;
    long off = 0;
    for (int p = 0; p < d->npath; ++p) {
      int a = d->nsar_index[p] - 1, e = d->nsar_index[p + 1] - 1;
      if (a != (p == 0 ? 0 : r0v.back() + nv.back()) || e < a || e > m.nsar) {
        set_error("nsar_index must split the SAR rows 1..nsar_total into consecutive paths");
        return SMCF_EINVAL;
      }
      r0v.push_back(a); nv.push_back(e - a); offv.push_back(off);
      for (int r = a; r < e; ++r) rowp[r] = p;
      off += (long)(e - a) * (e - a);
    }
    if (r0v.back() + nv.back() != m.nsar) {
      set_error("nsar_index(npath+1) must be nsar_total + 1");
      return SMCF_EINVAL;
    }
    std::vector<double> sinv(d->sar_sinv, d->sar_sinv + off);
    m.npath = d->npath;
    upload(P, &m.sinv, sinv);
    upload(P, &m.sar_r0, r0v);
    upload(P, &m.sar_n, nv);
    upload(P, &m.sar_off, offv);
    upload(P, &m.row_path, rowp);
    P.sar_r0 = r0v; P.sar_n = nv; P.sar_off = offv;
  }
  if (P.alloc_rc) return P.alloc_rc;
  // batch buffers
  BatchDev& bt = P.bt;
  memset(&bt, 0, sizeof(bt));
  size_t C = (size_t)max_batch, nd = m.ndim, Pn = P.P;
  bt.cap = max_batch;
  bt.P = P.P;
  dalloc(P, &bt.theta, m.ndim_fault * C);
  dalloc(P, &bt.geo, 16 * np * C);
  dalloc(P, &bt.scal, 8 * C);
  dalloc(P, &bt.G, (size_t)m.nobs * nd * C);
  dalloc(P, &bt.gscr, (size_t)24 * m.npts * npatch * C);
  dalloc(P, &bt.luni, (size_t)9 * nnode * C);
  dalloc(P, &bt.li, (size_t)9 * 2 * nnode * C);
  dalloc(P, &bt.lv, (size_t)9 * 2 * nnode * C);
  dalloc(P, &bt.lti, (size_t)9 * 2 * ndof * C);
  dalloc(P, &bt.ltv, (size_t)9 * 2 * ndof * C);
  dalloc(P, &bt.llc_idx, (size_t)25 * nd * C);
  dalloc(P, &bt.llc_val, (size_t)25 * nd * C);
  bt.Gw = nullptr;
  if (m.nsar > 0) dalloc(P, &bt.Gw, (size_t)m.nsar * nd * C);
  dalloc(P, &bt.sdv, (size_t)m.nobs * C);
  dalloc(P, &bt.H, nd * nd * C);
  dalloc(P, &bt.bvec, nd * C);
  dalloc(P, &bt.dsd, C);
  bt.kb = gibbs_kb(P.precision, (int)nd);
  bt.Q = nullptr;
  if (P.mover == SMCF_MOVER_GIBBS) dalloc(P, &bt.Q, (nd + bt.kb) * nd * C);  // blocked layout, whole blocks
  bt.Q32 = nullptr;
  bt.S32 = bt.S32b = nullptr;
  bt.SX = bt.SXb = bt.QX = nullptr;
  if (P.precision == SMCF_PRECISION_MIXED) {
    dalloc(P, &bt.Q32, (nd + bt.kb) * nd * C);
    dalloc(P, &bt.S32, nd * Pn * C);
    dalloc(P, &bt.S32b, nd * Pn * C);
  }
  bt.HX = bt.DX = nullptr;
  bt.mref = bt.gref = bt.lref = nullptr;
  if (P.precision == SMCF_PRECISION_MIXED_BF16 && P.mover == SMCF_MOVER_GIBBS) {
    dalloc(P, &bt.QX, bf_q_elems((int)nd, bt.kb) * C);
    dalloc(P, &bt.SX, (size_t)bf_rows((int)nd) * Pn * 2 * C);
    dalloc(P, &bt.SXb, (size_t)bf_rows((int)nd) * Pn * 2 * C);
    dalloc(P, &bt.HX, bf_q_elems((int)nd, 64) * C);
    dalloc(P, &bt.DX, (size_t)bf_rows((int)nd) * Pn * 2 * C);
    dalloc(P, &bt.mref, nd * C);
    dalloc(P, &bt.gref, nd * C);
    dalloc(P, &bt.lref, C);
    // rows past ndim of the centred copy are read by the likelihood's matrix products and never written
    if (!P.alloc_rc && !sizing &&
        hipMemset(bt.DX, 0, sizeof(uint16_t) * (size_t)bf_rows((int)nd) * Pn * 2 * C) != hipSuccess) {
      set_error("hipMemset failed");
      P.alloc_rc = SMCF_EHIP;
    }
  }
  dalloc(P, &bt.qdiag, 4 * nd * C);
  dalloc(P, &bt.S, nd * Pn * C);
  dalloc(P, &bt.S2, nd * Pn * C);
  bt.Y = nullptr;
  if (P.mover != SMCF_MOVER_GIBBS) dalloc(P, &bt.Y, nd * Pn * C);  // H X of the leapfrog (HMC movers only)
  dalloc(P, &bt.like, Pn * C);
  dalloc(P, &bt.prior, Pn * C);
  dalloc(P, &bt.like2, Pn * C);
  dalloc(P, &bt.prior2, Pn * C);
  dalloc(P, &bt.w, Pn * C);
  dalloc(P, &bt.parent, Pn * C);
  dalloc(P, &bt.assigned, Pn * C);
  dalloc(P, &bt.mean, nd * C);
  dalloc(P, &bt.cov, nd * C);
  dalloc(P, &bt.key, C);
  dalloc(P, &bt.gamma, C);
  dalloc(P, &bt.neglog, C);
  dalloc(P, &bt.iter, C);
  dalloc(P, &bt.active, C);
  dalloc(P, &bt.done, C);
  dalloc(P, &bt.sel, 2 * C);
  dalloc(P, &bt.par, C);
  dalloc(P, &bt.gtrace, 202 * C);
  dalloc(P, &P.mdev, 1);
  if (P.mover != SMCF_MOVER_GIBBS) {
    HmcDev& h = P.hm;
    size_t nn2 = 2 * (size_t)nnode;
    dalloc(P, &h.Xc, nd * Pn * C);
    dalloc(P, &h.Xt, nd * Pn * C);
    dalloc(P, &h.Xt2, nd * Pn * C);
    dalloc(P, &h.Qg, nd * nd * C);
    dalloc(P, &h.accf, Pn * C);
    dalloc(P, &h.wla, 2 * hmc_work_entries((int)C, (int)Pn));
    dalloc(P, &h.wlw, 2 * hmc_work_entries((int)C, (int)Pn));
    dalloc(P, &h.wln, 2 * hmc_work_entries((int)C, (int)Pn));
    dalloc(P, &h.part, (size_t)32 * 6 * Pn * C);
    h.QgX = h.XXa = h.XXb = nullptr;
    if (P.precision == SMCF_PRECISION_MIXED_BF16) {
      uint4 *q = nullptr, *xa = nullptr, *xb = nullptr;  // 16-byte fragments
      size_t xf = hmc_xx_rows((int)nd) / 8 * Pn * 2 * C;
      dalloc(P, &q, hmc_qgx_frags((int)nd) * C);
      dalloc(P, &xa, xf);
      dalloc(P, &xb, xf);
      h.QgX = q;
      h.XXa = xa;
      h.XXb = xb;
      // rows past ndim are read by the products and never written
      if (!P.alloc_rc && !sizing &&
          (hipMemset(xa, 0, sizeof(uint4) * xf) != hipSuccess || hipMemset(xb, 0, sizeof(uint4) * xf) != hipSuccess)) {
        set_error("hipMemset failed");
        P.alloc_rc = SMCF_EHIP;
      }
    }
    dalloc(P, &h.Pm, nd * Pn * C);
    dalloc(P, &h.LS, nn2 * Pn * C);
    dalloc(P, &h.Lc, Pn * C);
    dalloc(P, &h.Rc, Pn * C);
    dalloc(P, &h.postc, Pn * C);
    dalloc(P, &h.hcur, Pn * C);
    dalloc(P, &h.dtau, Pn * C);
    dalloc(P, &h.dtmp, Pn * C);
    dalloc(P, &h.score, Pn * C);
    dalloc(P, &h.ldlo, C);
    dalloc(P, &h.ldhi, C);
    dalloc(P, &h.wptr, (Pn + 1) * C);
    dalloc(P, &h.tptr, (Pn + 1) * C);
    dalloc(P, &h.wcols, Pn * C);
    dalloc(P, &h.tcols, Pn * C);
    dalloc(P, &h.wamax, C);
    dalloc(P, &h.tamax, C);
    dalloc(P, &h.ncol, C);
    dalloc(P, &h.cpar, Pn * C);
    dalloc(P, &h.cchild, Pn * C);
    dalloc(P, &h.dead, Pn * C);
    dalloc(P, &h.idstart, Pn * C);
    dalloc(P, &h.flag, C);
  }
  if (P.alloc_rc) return P.alloc_rc;
  if (sizing) {
    *bytes = P.bytes;
    return SMCF_OK;
  }
  SMCF_HIP(hipMemset(bt.parent, 0, sizeof(int) * Pn * C));
  SMCF_HIP(hipMemset(bt.gtrace, 0, sizeof(double) * 202 * C));
  {
    // the problem as a one-entry member table of the stage kernels
    MemberDev md = member_of(P);
    SMCF_HIP(hipMemcpy(P.mdev, &md, sizeof(md), hipMemcpyHostToDevice));
  }
  guard.p = nullptr;
  *out = holder.release();
  return SMCF_OK;
}

int smcf_problem_create(smcf_ctx* ctx, const smcf_problem_desc* d, int max_batch, smcf_problem** out) {
  return problem_build(ctx, d, max_batch, false, out, nullptr);
}

int smcf_problem_bytes(const smcf_problem_desc* d, int max_batch, size_t* bytes) {
  if (!bytes) return SMCF_EINVAL;
  return problem_build(nullptr, d, max_batch, true, nullptr, bytes);
}

int smcf_batch_times(const smcf_problem* pb, int B, double* seconds) {
  if (!pb || !seconds || B < 0 || (size_t)B > pb->p.t_done.size()) return SMCF_EINVAL;
  std::copy(pb->p.t_done.begin(), pb->p.t_done.begin() + B, seconds);
  return SMCF_OK;
}

int smcf_batch_admit_times(const smcf_problem* pb, int B, double* seconds) {
  if (!pb || !seconds || B < 0 || (size_t)B > pb->p.t_admit.size()) return SMCF_EINVAL;
  std::copy(pb->p.t_admit.begin(), pb->p.t_admit.begin() + B, seconds);
  return SMCF_OK;
}

int smcf_problem_device_bytes(const smcf_problem* pb, size_t* bytes) {
  if (!pb || !bytes) return SMCF_EINVAL;
  *bytes = pb->p.bytes;
  return SMCF_OK;
}

smcf_ctx* smcf_problem_context(const smcf_problem* pb) { return pb ? pb->p.ctx : nullptr; }

int smcf_gibbs_block_rows(int precision, int ndim) {
  if (ndim < 1) return SMCF_EINVAL;
  return gibbs_kb(precision, ndim);
}

int smcf_set_gibbs_block_rows(int precision, int ndim_lo, int ndim_hi, int rows) {
  if (ndim_lo < 1 || ndim_hi < ndim_lo || !gibbs_rows_valid(precision, rows)) {
    set_error("set_gibbs_block_rows: rows 16 or 32 (FP64), 16, 32 or 48 (split BF16), and 1 <= ndim_lo <= ndim_hi");
    return SMCF_EINVAL;
  }
  set_gibbs_rows(precision, ndim_lo, ndim_hi, rows);
  return SMCF_OK;
}

int smcf_problem_destroy(smcf_problem* pb) {
  if (!pb) return SMCF_OK;
  if (pb->p.in_group || pb->p.engine) {
    set_error("smcf_problem_destroy: the problem is a member of a group (destroy the group first)");
    return SMCF_ESTATE;
  }
  (void)hipSetDevice(pb->p.ctx->device);
  (void)hipStreamSynchronize(pb->p.ctx->stream);
  for (void* p : pb->p.allocs) (void)hipFree(p);
  (void)smcf_fault_smc_set_output_dir(pb, nullptr);  // a later problem at the same address starts without dumps
  (void)smcf_fault_smc_set_outer_stage(pb, 0);         // ... and with the host outer stage
  delete pb;
  return SMCF_OK;
}

int smcf_problem_dims(const smcf_problem* pb, int* nnode, int* ndof, int* ndim_slip, int* ndim_fault, int* nobs) {
  if (!pb) return SMCF_EINVAL;
  if (nnode) *nnode = pb->p.m.nnode;
  if (ndof) *ndof = pb->p.m.ndof;
  if (ndim_slip) *ndim_slip = pb->p.m.ndim;
  if (ndim_fault) *ndim_fault = pb->p.m.ndim_fault;
  if (nobs) *nobs = pb->p.m.nobs;
  return SMCF_OK;
}

int smcf_dc3d0(smcf_ctx* ctx, int n, const double* in, double* u, int* iret) {
  if (!ctx || n < 0 || (n > 0 && (!in || !u || !iret))) return SMCF_EINVAL;
  if (n == 0) return SMCF_OK;
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  double *din = nullptr, *du = nullptr;
  int* dir = nullptr;
  TmpBufs tb;
  SMCF_HIP(tb.alloc(&din, 10 * (size_t)n));
  SMCF_HIP(tb.alloc(&du, 12 * (size_t)n));
  SMCF_HIP(tb.alloc(&dir, (size_t)n));
  SMCF_HIP(hipMemcpyAsync(din, in, sizeof(double) * 10 * n, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dir, iret, sizeof(int) * n, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_dc3d0_batch(n, din, du, dir, s));
  SMCF_HIP(hipMemcpyAsync(u, du, sizeof(double) * 12 * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipMemcpyAsync(iret, dir, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_call_dc3d0(smcf_ctx* ctx, int n, const double* in, double* uret) {
  if (!ctx || n < 0 || (n > 0 && (!in || !uret))) return SMCF_EINVAL;
  if (n == 0) return SMCF_OK;
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  TmpBufs tb;
  double *din, *dout;
  SMCF_HIP(tb.alloc(&din, 10 * (size_t)n));
  SMCF_HIP(tb.alloc(&dout, 3 * (size_t)n));
  SMCF_HIP(hipMemcpyAsync(din, in, sizeof(double) * 10 * n, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_call_dc3d0(n, din, dout, s));
  SMCF_HIP(hipMemcpyAsync(uret, dout, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_calc_responce(smcf_ctx* ctx, int npts, const double* pts, int npatch, const int* cny, int nnode,
                       const double* coor, const double* slip_dist, double xf, double yf, double zf, double strike,
                       double dip, int ntarget, const int* target_id, double* uobs) {
  if (!ctx || npts < 0 || npatch < 1 || nnode < 1 || ntarget < 0 || !cny || !coor || !slip_dist ||
      (ntarget > 0 && !target_id) || (npts > 0 && (!pts || !uobs)))
    return SMCF_EINVAL;
  for (int t = 0; t < ntarget; ++t)
    if (target_id[t] < 1 || target_id[t] > npatch) {
      set_error("smcf_calc_responce: target patch id out of range");
      return SMCF_EINVAL;
    }
  for (int k = 0; k < 4 * npatch; ++k)
    if (cny[k] < 1 || cny[k] > nnode) {
      set_error("smcf_calc_responce: node id out of range in cny");
      return SMCF_EINVAL;
    }
  if (npts == 0) return SMCF_OK;
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  TmpBufs tb;
  double *dp, *dcoor, *dslip, *du;
  int *dcny, *dt;
  SMCF_HIP(tb.alloc(&dp, 2 * (size_t)npts));
  SMCF_HIP(tb.alloc(&dcoor, 2 * (size_t)nnode));
  SMCF_HIP(tb.alloc(&dslip, 2 * (size_t)nnode));
  SMCF_HIP(tb.alloc(&du, 3 * (size_t)npts));
  SMCF_HIP(tb.alloc(&dcny, 4 * (size_t)npatch));
  SMCF_HIP(tb.alloc(&dt, (size_t)std::max(ntarget, 1)));
  SMCF_HIP(hipMemcpyAsync(dp, pts, sizeof(double) * 2 * npts, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dcoor, coor, sizeof(double) * 2 * nnode, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dslip, slip_dist, sizeof(double) * 2 * nnode, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dcny, cny, sizeof(int) * 4 * npatch, hipMemcpyHostToDevice, s));
  if (ntarget > 0) SMCF_HIP(hipMemcpyAsync(dt, target_id, sizeof(int) * ntarget, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_calc_responce(npts, dp, dcny, dcoor, dslip, xf, yf, zf, strike, dip, ntarget, dt, du, s));
  SMCF_HIP(hipMemcpyAsync(uobs, du, sizeof(double) * 3 * npts, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_calc_greens_func(smcf_problem* pb, int B, const double* theta, double* gmat) {
  if (!pb || !theta || !gmat) return SMCF_EINVAL;
  Problem& P = pb->p;
  if (B < 1 || B > P.bt.cap) { set_error("batch size outside capacity"); return SMCF_EINVAL; }
  hipStream_t s = P.ctx->stream;
  SMCF_HIP(hipMemcpyAsync(P.bt.theta, theta, sizeof(double) * P.m.ndim_fault * B, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_geometry(P.m, P.bt, B, s));
  CHECK_LAUNCH(launch_greens(P.m, P.bt, B, s));
  SMCF_HIP(hipMemcpyAsync(gmat, P.bt.G, sizeof(double) * P.m.nobs * P.m.ndim * B, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_laplacian(smcf_problem* pb, const double* theta, int* li, double* lv, int* lti, double* ltv, double* ll) {
  if (!pb || !theta) return SMCF_EINVAL;
  Problem& P = pb->p;
  MeshDev& m = P.m;
  hipStream_t s = P.ctx->stream;
  SMCF_HIP(hipMemcpyAsync(P.bt.theta, theta, sizeof(double) * m.ndim_fault, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_geometry(m, P.bt, 1, s));
  CHECK_LAUNCH(launch_laplacian(m, P.bt, 1, s));
  if (li) SMCF_HIP(hipMemcpyAsync(li, P.bt.li, sizeof(int) * 18 * m.nnode, hipMemcpyDeviceToHost, s));
  if (lv) SMCF_HIP(hipMemcpyAsync(lv, P.bt.lv, sizeof(double) * 18 * m.nnode, hipMemcpyDeviceToHost, s));
  if (lti) SMCF_HIP(hipMemcpyAsync(lti, P.bt.lti, sizeof(int) * 18 * m.ndof, hipMemcpyDeviceToHost, s));
  if (ltv) SMCF_HIP(hipMemcpyAsync(ltv, P.bt.ltv, sizeof(double) * 18 * m.ndof, hipMemcpyDeviceToHost, s));
  std::vector<int> ix;
  std::vector<double> vx;
  if (ll) {
    ix.resize(25 * (size_t)m.ndim);
    vx.resize(25 * (size_t)m.ndim);
    SMCF_HIP(hipMemcpyAsync(ix.data(), P.bt.llc_idx, sizeof(int) * ix.size(), hipMemcpyDeviceToHost, s));
    SMCF_HIP(hipMemcpyAsync(vx.data(), P.bt.llc_val, sizeof(double) * vx.size(), hipMemcpyDeviceToHost, s));
  }
  SMCF_HIP(hipStreamSynchronize(s));
  if (ll) {
    size_t nd = m.ndim;
    std::fill(ll, ll + nd * nd, 0.0);
    for (size_t i = 0; i < nd; ++i)
      for (int w = 0; w < 25; ++w) {
        int r = ix[i * 25 + w];
        if (r < 0) break;
        ll[i * nd + r] = vx[i * 25 + w];
      }
  }
  return SMCF_OK;
}

int smcf_discretize_fault(smcf_problem* pb, const double* theta, int* cny, double* coor, int* n2e, int* n2es,
                          int* id_dof, double* luni, double* lmat) {
  if (!pb || !theta) return SMCF_EINVAL;
  Problem& P = pb->p;
  MeshDev& m = P.m;
  SMCF_HIP(hipSetDevice(P.ctx->device));
  hipStream_t s = P.ctx->stream;
  const int nn = m.nnode, np = m.nplane;
  std::vector<int> nk(nn), nl(nn), npl(nn), pn0(np), pnx(np), pne(np), idd(m.ndof);
  std::vector<double> geo(16 * (size_t)np), slots(9 * (size_t)nn);
  SMCF_HIP(hipMemcpyAsync(P.bt.theta, theta, sizeof(double) * m.ndim_fault, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_geometry(m, P.bt, 1, s));
  CHECK_LAUNCH(launch_laplacian(m, P.bt, 1, s));
  SMCF_HIP(hipMemcpyAsync(geo.data(), P.bt.geo, sizeof(double) * geo.size(), hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipMemcpyAsync(slots.data(), P.bt.luni, sizeof(double) * slots.size(), hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipMemcpyAsync(nk.data(), m.node_k, sizeof(int) * nn, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipMemcpyAsync(nl.data(), m.node_l, sizeof(int) * nn, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipMemcpyAsync(npl.data(), m.node_plane, sizeof(int) * nn, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipMemcpyAsync(pn0.data(), m.plane_node0, sizeof(int) * np, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipMemcpyAsync(pnx.data(), m.plane_nxi, sizeof(int) * np, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipMemcpyAsync(pne.data(), m.plane_neta, sizeof(int) * np, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipMemcpyAsync(idd.data(), m.id_dof, sizeof(int) * m.ndof, hipMemcpyDeviceToHost, s));
  if (cny) SMCF_HIP(hipMemcpyAsync(cny, m.cny, sizeof(int) * 4 * m.npatch, hipMemcpyDeviceToHost, s));
  if (n2e) SMCF_HIP(hipMemcpyAsync(n2e, m.n2e, sizeof(int) * 4 * nn, hipMemcpyDeviceToHost, s));
  if (n2es) SMCF_HIP(hipMemcpyAsync(n2es, m.n2es, sizeof(int) * nn, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  if (id_dof) std::copy(idd.begin(), idd.end(), id_dof);
  if (coor)  // init.f90:240-244: xi = (i - 1) dxi - lxi / 2, eta = (j - 1) deta - leta / 2
    for (int n = 0; n < nn; ++n) {
      const double* g = &geo[16 * (size_t)npl[n]];
      coor[2 * (size_t)n] = (nk[n] - 1) * g[9] - g[11] / 2.0;
      coor[2 * (size_t)n + 1] = (nl[n] - 1) * g[10] - g[12] / 2.0;
    }
  if (luni || lmat) {
    std::vector<double> lu((size_t)nn * nn, 0.0);  // dense luni, column-major
    for (int n = 0; n < nn; ++n) {
      int ip = npl[n];
      for (int dl = -1; dl <= 1; ++dl)
        for (int dk = -1; dk <= 1; ++dk) {
          int k = nk[n] + dk, l = nl[n] + dl;
          if (k < 1 || k > pnx[ip] + 1 || l < 1 || l > pne[ip] + 1) continue;
          int j = pn0[ip] + (k - 1) + (l - 1) * (pnx[ip] + 1);
          lu[(size_t)j * nn + n] = slots[9 * (size_t)n + (dk + 1) + 3 * (dl + 1)];
        }
    }
    if (luni) std::copy(lu.begin(), lu.end(), luni);
    if (lmat) {  // init.f90:461-473: the rows of both slip components over the DOF columns
      size_t rows = 2 * (size_t)nn;
      std::fill(lmat, lmat + rows * 2 * m.ndof, 0.0);
      for (int n = 0; n < nn; ++n)
        for (int q = 0; q < m.ndof; ++q) {
          int j = idd[q] - 1;
          lmat[(size_t)(2 * q) * rows + 2 * n] = lu[(size_t)j * nn + n];
          lmat[(size_t)(2 * q + 1) * rows + 2 * n + 1] = lu[(size_t)j * nn + n];
        }
    }
  }
  return SMCF_OK;
}

int smcf_slip_calc_gsd_gsg(smcf_problem* pb, int B, const double* theta, double* gsgmat, double* gsdvec, double* dsd) {
  if (!pb || !theta) return SMCF_EINVAL;
  Problem& P = pb->p;
  MeshDev& m = P.m;
  if (B < 1 || B > P.bt.cap) { set_error("batch size outside capacity"); return SMCF_EINVAL; }
  hipStream_t s = P.ctx->stream;
  SMCF_HIP(hipMemcpyAsync(P.bt.theta, theta, sizeof(double) * m.ndim_fault * B, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_geometry(m, P.bt, B, s));
  CHECK_LAUNCH(launch_greens(m, P.bt, B, s));
  CHECK_LAUNCH(launch_gram(P, m, P.bt, B, s));
  size_t nd = m.ndim;
  if (gsgmat) SMCF_HIP(hipMemcpyAsync(gsgmat, P.bt.H, sizeof(double) * nd * nd * B, hipMemcpyDeviceToHost, s));
  if (gsdvec) SMCF_HIP(hipMemcpyAsync(gsdvec, P.bt.bvec, sizeof(double) * nd * B, hipMemcpyDeviceToHost, s));
  if (dsd) SMCF_HIP(hipMemcpyAsync(dsd, P.bt.dsd, sizeof(double) * B, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

// plain mean over the P final particles of instances 0 .. B-1 (the loop of calc_slip_map, main.f90:620-633:
// ascending particle order, then one division)
__global__ void k_plain_mean(int nd, int P, int B, BatchDev bt, double* __restrict__ mean) {
  int i = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (i >= nd || b >= B) return;
  const double* S = (bt.par[b] ? bt.S2 : bt.S) + ((size_t)b * nd + i) * P;
  double acc = 0.0;
  for (int p = 0; p < P; ++p) acc += S[p];
  mean[(size_t)b * nd + i] = acc / P;
}

int smcf_slip_mean_batch(smcf_problem* pb, int B, const double* theta, const uint64_t* keys, double* mean,
                         double* neglog) {
  if (!pb || !theta || !keys || !mean || B < 1) return SMCF_EINVAL;
  Problem& P = pb->p;
  if (B > P.bt.cap) {
    set_error("smcf_slip_mean_batch: B exceeds the problem's max_batch");
    return SMCF_EINVAL;
  }
  SMCF_HIP(hipSetDevice(P.ctx->device));
  std::vector<double> nl(B);
  int rc = run_stream(P, B, theta, keys, false, nl.data(), nullptr);  // B <= cap: slot b ends holding geometry b
  if (rc) return rc;
  hipStream_t s = P.ctx->stream;
  size_t nd = P.m.ndim;
  double* dm = nullptr;
  SMCF_HIP(hipMalloc(&dm, sizeof(double) * nd * B));
  hipLaunchKernelGGL(k_plain_mean, dim3((unsigned)((nd + 127) / 128), B), dim3(128), 0, s, (int)nd, P.bt.P, B, P.bt,
                     dm);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(mean, dm, sizeof(double) * nd * B, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(dm);
  if (e != hipSuccess) {
    set_error(std::string("smcf_slip_mean_batch: ") + hipGetErrorString(e));
    return SMCF_EHIP;
  }
  if (neglog) std::copy(nl.begin(), nl.end(), neglog);
  return SMCF_OK;
}

int smcf_slip_smc_exec(smcf_problem* pb, const double* theta, uint64_t key, double* particles, double* likelihood,
                       double* prior, double* mean, int* nstage, double* neglog, double* gamma_trace) {
  if (!pb || !theta) return SMCF_EINVAL;
  Problem& P = pb->p;
  SMCF_HIP(hipSetDevice(P.ctx->device));
  int rc = run_stream(P, 1, theta, &key, mean != nullptr, nullptr, nullptr);
  if (rc) return rc;
  hipStream_t s = P.ctx->stream;
  MeshDev& m = P.m;
  BatchDev& bt = P.bt;
  if (prior && P.mover == SMCF_MOVER_GIBBS) {  // the Gibbs stage loop leaves the prior of the final particles to here
    CHECK_LAUNCH(launch_prior(m, bt, 1, s));
    SMCF_HIP(hipStreamSynchronize(s));
  }
  size_t nd = m.ndim, Pn = bt.P;
  int par = 0, it = 0;
  SMCF_HIP(hipMemcpy(&par, bt.par, sizeof(int), hipMemcpyDeviceToHost));
  SMCF_HIP(hipMemcpy(&it, bt.iter, sizeof(int), hipMemcpyDeviceToHost));
  if (particles) {
    std::vector<double> dm(nd * Pn);
    SMCF_HIP(hipMemcpy(dm.data(), par ? bt.S2 : bt.S, sizeof(double) * nd * Pn, hipMemcpyDeviceToHost));
    for (size_t p = 0; p < Pn; ++p)
      for (size_t i = 0; i < nd; ++i) particles[p * nd + i] = dm[i * Pn + p];
  }
  if (likelihood)
    SMCF_HIP(hipMemcpy(likelihood, par ? bt.like2 : bt.like, sizeof(double) * Pn, hipMemcpyDeviceToHost));
  if (prior) SMCF_HIP(hipMemcpy(prior, par ? bt.prior2 : bt.prior, sizeof(double) * Pn, hipMemcpyDeviceToHost));
  if (mean) SMCF_HIP(hipMemcpy(mean, bt.mean, sizeof(double) * nd, hipMemcpyDeviceToHost));
  if (neglog) SMCF_HIP(hipMemcpy(neglog, bt.neglog, sizeof(double), hipMemcpyDeviceToHost));
  if (gamma_trace) SMCF_HIP(hipMemcpy(gamma_trace, bt.gtrace, sizeof(double) * 202, hipMemcpyDeviceToHost));
  if (nstage) *nstage = std::min(it - 1, 201);
  (void)s;
  return SMCF_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Component procedures of one geometry (slot 0 of the problem).  Particles are passed in the reference layout
// (ndim x n, one column per particle, n <= nparticle_slip) and transposed into the dimension-major device layout.
// ---------------------------------------------------------------------------------------------------------------
static int slot0_geometry(Problem& P, const double* theta, hipStream_t s) {
  MeshDev& m = P.m;
  BatchDev& bt = P.bt;
  SMCF_HIP(hipMemcpyAsync(bt.theta, theta, sizeof(double) * m.ndim_fault, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, s, bt, 1, (const int*)nullptr, 0);
  CHECK_LAUNCH(launch_geometry(m, bt, 1, s));
  CHECK_LAUNCH(launch_greens(m, bt, 1, s));
  CHECK_LAUNCH(launch_laplacian(m, bt, 1, s));
  CHECK_LAUNCH(launch_gram(P, m, bt, 1, s));
  return SMCF_OK;
}

static int check_component(smcf_problem* pb, const double* theta, int n) {
  if (!pb || !theta || n < 1) {
    set_error("component call: missing problem or theta, or n < 1");
    return SMCF_EINVAL;
  }
  if (n > pb->p.P) {
    set_error("component call: n = " + std::to_string(n) + " exceeds nparticle_slip = " + std::to_string(pb->p.P));
    return SMCF_EINVAL;
  }
  return SMCF_OK;
}

// particles (ndim x n, column per particle) -> dimension-major rows of length P at dst
static int upload_columns(Problem& P, const double* cols, int n, double* dst, hipStream_t s) {
  size_t nd = P.m.ndim, Pn = P.P;
  std::vector<double> t(nd * Pn, 0.0);
  for (int p = 0; p < n; ++p)
    for (size_t i = 0; i < nd; ++i) t[i * Pn + p] = cols[(size_t)p * nd + i];
  SMCF_HIP(hipMemcpyAsync(dst, t.data(), sizeof(double) * nd * Pn, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipStreamSynchronize(s));  // t is a stack buffer
  return SMCF_OK;
}

static int download_columns(Problem& P, const double* src, int n, double* cols, hipStream_t s) {
  size_t nd = P.m.ndim, Pn = P.P;
  std::vector<double> t(nd * Pn);
  SMCF_HIP(hipMemcpyAsync(t.data(), src, sizeof(double) * nd * Pn, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  for (int p = 0; p < n; ++p)
    for (size_t i = 0; i < nd; ++i) cols[(size_t)p * nd + i] = t[i * Pn + p];
  return SMCF_OK;
}

int smcf_slip_calc_likelihood(smcf_problem* pb, const double* theta, int n, const double* particles, double* out) {
  int rc = check_component(pb, theta, n);
  if (rc) return rc;
  if (!particles || !out) return SMCF_EINVAL;
  Problem& P = pb->p;
  SMCF_HIP(hipSetDevice(P.ctx->device));
  hipStream_t s = P.ctx->stream;
  if ((rc = slot0_geometry(P, theta, s))) return rc;
  if ((rc = upload_columns(P, particles, n, P.bt.S, s))) return rc;  // k_state_init left par = 0: S is current
  CHECK_LAUNCH(launch_like(P.m, P.bt, 1, 0, s));
  SMCF_HIP(hipMemcpyAsync(out, P.bt.like, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_slip_calc_prior(smcf_problem* pb, const double* theta, int n, const double* particles, double* out) {
  int rc = check_component(pb, theta, n);
  if (rc) return rc;
  if (!particles || !out) return SMCF_EINVAL;
  Problem& P = pb->p;
  SMCF_HIP(hipSetDevice(P.ctx->device));
  hipStream_t s = P.ctx->stream;
  if ((rc = slot0_geometry(P, theta, s))) return rc;
  if ((rc = upload_columns(P, particles, n, P.bt.S, s))) return rc;
  CHECK_LAUNCH(launch_prior(P.m, P.bt, 1, s));
  SMCF_HIP(hipMemcpyAsync(out, P.bt.prior, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

static int hmc_component_state(Problem& P, double gamma, int n, const double* cov, hipStream_t s) {
  if (P.mover == SMCF_MOVER_GIBBS) {
    set_error("component call needs a problem with an HMC mover (its trajectory buffers)");
    return SMCF_EUNSUPPORTED;
  }
  int one = 1;
  SMCF_HIP(hipMemcpyAsync(P.bt.gamma, &gamma, sizeof(double), hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(P.bt.active, &one, sizeof(int), hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(P.hm.ncol, &n, sizeof(int), hipMemcpyHostToDevice, s));
  if (cov) SMCF_HIP(hipMemcpyAsync(P.bt.cov, cov, sizeof(double) * P.m.ndim, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipStreamSynchronize(s));  // the host values above live on this stack frame
  return SMCF_OK;
}

int smcf_slip_calc_grad(smcf_problem* pb, const double* theta, double gamma, int n, const double* particles,
                        double* grad) {
  int rc = check_component(pb, theta, n);
  if (rc) return rc;
  if (!particles || !grad) return SMCF_EINVAL;
  Problem& P = pb->p;
  SMCF_HIP(hipSetDevice(P.ctx->device));
  hipStream_t s = P.ctx->stream;
  if ((rc = slot0_geometry(P, theta, s))) return rc;
  if ((rc = hmc_component_state(P, gamma, n, nullptr, s))) return rc;
  if ((rc = upload_columns(P, particles, n, P.hm.Xt, s))) return rc;
  TmpBufs tb;
  double* dg = nullptr;
  SMCF_HIP(tb.alloc(&dg, (size_t)P.m.ndim * n));
  CHECK_LAUNCH(launch_grad_cols(P.m, P.bt, P.hm, n, gamma, dg, s));
  SMCF_HIP(hipMemcpyAsync(grad, dg, sizeof(double) * P.m.ndim * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_slip_leapfrog(smcf_problem* pb, const double* theta, double gamma, int n, const double* cur,
                       const double* p0, const double* cov_diag, double dtau, int ntau, double* cand, double* p_out,
                       double* ham_cur, double* ham_cand, double* likelihood_cand, double* prior_cand) {
  int rc = check_component(pb, theta, n);
  if (rc) return rc;
  if (!cur || !p0 || !cov_diag || ntau < 0) return SMCF_EINVAL;
  Problem& P = pb->p;
  SMCF_HIP(hipSetDevice(P.ctx->device));
  hipStream_t s = P.ctx->stream;
  if ((rc = slot0_geometry(P, theta, s))) return rc;
  if ((rc = hmc_component_state(P, gamma, n, cov_diag, s))) return rc;
  if ((rc = upload_columns(P, cur, n, P.hm.Xt, s))) return rc;
  if ((rc = upload_columns(P, p0, n, P.hm.Pm, s))) return rc;
  TmpBufs tb;
  double *dhc = nullptr, *dhn = nullptr, *dlk = nullptr, *dpr = nullptr;
  SMCF_HIP(tb.alloc(&dhc, (size_t)n));
  SMCF_HIP(tb.alloc(&dhn, (size_t)n));
  SMCF_HIP(tb.alloc(&dlk, (size_t)n));
  SMCF_HIP(tb.alloc(&dpr, (size_t)n));
  CHECK_LAUNCH(launch_leapfrog_cols(P.m, P.bt, P.hm, n, gamma, dtau, ntau, P.max_slip, dhc, dhn, dlk, dpr, s));
  if (likelihood_cand) SMCF_HIP(hipMemcpyAsync(likelihood_cand, dlk, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  if (prior_cand) SMCF_HIP(hipMemcpyAsync(prior_cand, dpr, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  if (cand && (rc = download_columns(P, P.hm.Xt, n, cand, s))) return rc;
  if (p_out && (rc = download_columns(P, P.hm.Pm, n, p_out, s))) return rc;
  if (ham_cur) SMCF_HIP(hipMemcpyAsync(ham_cur, dhc, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  if (ham_cand) SMCF_HIP(hipMemcpyAsync(ham_cand, dhn, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_slip_gen_init_particles(smcf_problem* pb, const double* theta, uint64_t key, double* particles,
                                 double* likelihood, double* prior) {
  if (!pb || !theta) return SMCF_EINVAL;
  Problem& P = pb->p;
  SMCF_HIP(hipSetDevice(P.ctx->device));
  hipStream_t s = P.ctx->stream;
  SMCF_HIP(hipMemcpyAsync(P.bt.theta, theta, sizeof(double) * P.m.ndim_fault, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(P.bt.key, &key, sizeof(uint64_t), hipMemcpyHostToDevice, s));
  int rc = setup_pre(P, 1, nullptr, s);
  if (rc) return rc;
  CHECK_LAUNCH(launch_init_chain(P.mdev, nullptr, nullptr, nullptr, 1, P.m.chain_w, s));
  if ((rc = setup_post(P, 1, nullptr, s))) return rc;
  if (particles && (rc = download_columns(P, P.bt.S, P.P, particles, s))) return rc;
  if (likelihood) SMCF_HIP(hipMemcpyAsync(likelihood, P.bt.like, sizeof(double) * P.P, hipMemcpyDeviceToHost, s));
  if (prior) SMCF_HIP(hipMemcpyAsync(prior, P.bt.prior, sizeof(double) * P.P, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

// particles (ndim x n, column per particle) -> dimension-major, leading dimension ld
static std::vector<double> to_dim_major(const double* cols, int nd, int n, int ld) {
  std::vector<double> t((size_t)nd * ld, 0.0);
  for (int p = 0; p < n; ++p)
    for (int i = 0; i < nd; ++i) t[(size_t)i * ld + p] = cols[(size_t)p * nd + i];
  return t;
}

static int check_ell(const int* idx, int rows, int maxcol, const char* what) {
  for (int k = 0; k < 9 * rows; ++k)
    if (idx[k] < 1 || idx[k] > maxcol) {
      set_error(std::string(what) + ": ELL index out of range");
      return SMCF_EINVAL;
    }
  return SMCF_OK;
}

int smcf_slip_calc_gsd_gsg_gmat(smcf_problem* pb, const double* gmat, double sigma_sar2, double sigma_gnss2,
                                double* gsgmat, double* gsdvec, double* dsd) {
  if (!pb || !gmat) return SMCF_EINVAL;
  Problem& P = pb->p;
  SMCF_HIP(hipSetDevice(P.ctx->device));
  hipStream_t s = P.ctx->stream;
  size_t nd = P.m.ndim, nobs = P.m.nobs;
  double scal[8] = {0, 0, 0, 0, 0, sigma_gnss2, sigma_sar2, 0};
  SMCF_HIP(hipMemcpyAsync(P.bt.G, gmat, sizeof(double) * nobs * nd, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(P.bt.scal, scal, sizeof(scal), hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_gram(P, P.m, P.bt, 1, s, nullptr, true));
  if (gsgmat) SMCF_HIP(hipMemcpyAsync(gsgmat, P.bt.H, sizeof(double) * nd * nd, hipMemcpyDeviceToHost, s));
  if (gsdvec) SMCF_HIP(hipMemcpyAsync(gsdvec, P.bt.bvec, sizeof(double) * nd, hipMemcpyDeviceToHost, s));
  if (dsd) SMCF_HIP(hipMemcpyAsync(dsd, P.bt.dsd, sizeof(double), hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_slip_likelihood_gram(smcf_ctx* ctx, int ndim, int n, const double* particles, const double* gsgmat,
                              const double* gsdvec, double dsd, int nsar_total, int ngnss, double log_sigma_sar2,
                              double log_sigma_gnss2, double* out) {
  if (!ctx || ndim < 1 || n < 1 || !particles || !gsgmat || !gsdvec || !out) return SMCF_EINVAL;
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  int ld = (n + 63) / 64 * 64;
  std::vector<double> t = to_dim_major(particles, ndim, n, ld);
  double scal[8] = {0, 0, 0, nsar_total * log_sigma_sar2 / 2.0, 3.0 * ngnss * log_sigma_gnss2 / 2.0, 0, 0, 0};
  int zero = 0;
  TmpBufs tb;
  double *dS, *dH, *db, *dd, *dsc, *dout;
  int* dz;
  SMCF_HIP(tb.alloc(&dS, t.size()));
  SMCF_HIP(tb.alloc(&dH, (size_t)ndim * ndim));
  SMCF_HIP(tb.alloc(&db, (size_t)ndim));
  SMCF_HIP(tb.alloc(&dd, 1));
  SMCF_HIP(tb.alloc(&dsc, 8));
  SMCF_HIP(tb.alloc(&dout, (size_t)ld));
  SMCF_HIP(tb.alloc(&dz, 1));
  SMCF_HIP(hipMemcpyAsync(dS, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dH, gsgmat, sizeof(double) * ndim * ndim, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(db, gsdvec, sizeof(double) * ndim, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dd, &dsd, sizeof(double), hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dsc, scal, sizeof(scal), hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dz, &zero, sizeof(int), hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_like_raw(ndim, ld, dH, db, dd, dsc, dS, dz, dout, s));
  SMCF_HIP(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_slip_prior_ell(smcf_ctx* ctx, int nnode, int ndim, int n, const double* particles, const double* alpha2_full,
                        const int* lmat_index, const double* lmat_val, double* out) {
  if (!ctx || nnode < 1 || ndim < 1 || n < 1 || !particles || !alpha2_full || !lmat_index || !lmat_val || !out)
    return SMCF_EINVAL;
  int rc = check_ell(lmat_index, 2 * nnode, ndim, "smcf_slip_prior_ell");
  if (rc) return rc;
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  std::vector<double> t = to_dim_major(particles, ndim, n, n);
  int nn2 = 2 * nnode;
  TmpBufs tb;
  double *dS, *da, *dv, *dout;
  int* di;
  SMCF_HIP(tb.alloc(&dS, t.size()));
  SMCF_HIP(tb.alloc(&da, (size_t)nn2));
  SMCF_HIP(tb.alloc(&di, 9 * (size_t)nn2));
  SMCF_HIP(tb.alloc(&dv, 9 * (size_t)nn2));
  SMCF_HIP(tb.alloc(&dout, (size_t)n));
  SMCF_HIP(hipMemcpyAsync(dS, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(da, alpha2_full, sizeof(double) * nn2, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(di, lmat_index, sizeof(int) * 9 * nn2, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dv, lmat_val, sizeof(double) * 9 * nn2, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_prior_raw(nn2, n, n, dS, da, di, dv, dout, s));
  SMCF_HIP(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_slip_grad_ell(smcf_ctx* ctx, int nnode, int ndim, int n, const double* particles, double gamma,
                       const double* gsgmat, const double* gsdvec, const double* alpha2_full, const int* lmat_index,
                       const double* lmat_val, const int* ltmat_index, const double* ltmat_val, double* grad) {
  if (!ctx || nnode < 1 || ndim < 1 || n < 1 || !particles || !gsgmat || !gsdvec || !alpha2_full || !lmat_index ||
      !lmat_val || !ltmat_index || !ltmat_val || !grad)
    return SMCF_EINVAL;
  int nn2 = 2 * nnode;
  int rc = check_ell(lmat_index, nn2, ndim, "smcf_slip_grad_ell");
  if (!rc) rc = check_ell(ltmat_index, ndim, nn2, "smcf_slip_grad_ell");
  if (rc) return rc;
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  int ld = (n + 63) / 64 * 64;
  std::vector<double> t = to_dim_major(particles, ndim, n, ld);
  int na[2] = {n, 1};
  TmpBufs tb;
  double *dS, *dH, *db, *da, *dv, *dtv, *dY, *dLS, *dG;
  int *di, *dti, *dna;
  SMCF_HIP(tb.alloc(&dS, t.size()));
  SMCF_HIP(tb.alloc(&dH, (size_t)ndim * ndim));
  SMCF_HIP(tb.alloc(&db, (size_t)ndim));
  SMCF_HIP(tb.alloc(&da, (size_t)nn2));
  SMCF_HIP(tb.alloc(&di, 9 * (size_t)nn2));
  SMCF_HIP(tb.alloc(&dv, 9 * (size_t)nn2));
  SMCF_HIP(tb.alloc(&dti, 9 * (size_t)ndim));
  SMCF_HIP(tb.alloc(&dtv, 9 * (size_t)ndim));
  SMCF_HIP(tb.alloc(&dY, t.size()));
  SMCF_HIP(tb.alloc(&dLS, (size_t)nn2 * ld));
  SMCF_HIP(tb.alloc(&dG, (size_t)ndim * n));
  SMCF_HIP(tb.alloc(&dna, 2));
  SMCF_HIP(hipMemcpyAsync(dS, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dH, gsgmat, sizeof(double) * ndim * ndim, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(db, gsdvec, sizeof(double) * ndim, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(da, alpha2_full, sizeof(double) * nn2, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(di, lmat_index, sizeof(int) * 9 * nn2, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dv, lmat_val, sizeof(double) * 9 * nn2, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dti, ltmat_index, sizeof(int) * 9 * ndim, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dtv, ltmat_val, sizeof(double) * 9 * ndim, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dna, na, sizeof(na), hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_grad_raw(ndim, nn2, n, ld, gamma, dH, dS, db, da, di, dv, dti, dtv, dna, dY, dLS, dG, s));
  SMCF_HIP(hipMemcpyAsync(grad, dG, sizeof(double) * ndim * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_slip_hmc_accept_state(const smcf_problem* pb, double* next_dtau, double* mean_accept) {
  if (!pb) return SMCF_EINVAL;
  const Problem& P = pb->p;
  if (P.mover != SMCF_MOVER_HMC_ACCEPT) {
    set_error("smcf_slip_hmc_accept_state: the problem's mover is not SMCF_MOVER_HMC_ACCEPT");
    return SMCF_EUNSUPPORTED;
  }
  SMCF_HIP(hipSetDevice(P.ctx->device));
  double ld = 0.0, ab = 0.0;
  SMCF_HIP(hipMemcpy(&ld, P.hm.ldlo, sizeof(double), hipMemcpyDeviceToHost));
  SMCF_HIP(hipMemcpy(&ab, P.hm.ldhi, sizeof(double), hipMemcpyDeviceToHost));
  if (next_dtau) *next_dtau = std::exp(ld);
  if (mean_accept) *mean_accept = ab;
  return SMCF_OK;
}

int smcf_slip_calc_moments(smcf_ctx* ctx, int ndim, int P, const double* particles, const double* weights,
                           double* mean, double* cov_diag) {
  if (!ctx || ndim < 1 || P < 1 || !particles || !weights) return SMCF_EINVAL;
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  std::vector<double> t((size_t)ndim * P);
  for (int p = 0; p < P; ++p)
    for (int i = 0; i < ndim; ++i) t[(size_t)i * P + p] = particles[(size_t)p * ndim + i];
  TmpBufs tb;
  double *dS, *dw, *dm, *dc;
  SMCF_HIP(tb.alloc(&dS, t.size()));
  SMCF_HIP(tb.alloc(&dw, (size_t)P));
  SMCF_HIP(tb.alloc(&dm, (size_t)ndim));
  SMCF_HIP(tb.alloc(&dc, (size_t)ndim));
  SMCF_HIP(hipMemcpyAsync(dS, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dw, weights, sizeof(double) * P, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_moments_raw(ndim, P, dS, dw, dm, dc, s));
  if (mean) SMCF_HIP(hipMemcpyAsync(mean, dm, sizeof(double) * ndim, hipMemcpyDeviceToHost, s));
  if (cov_diag) SMCF_HIP(hipMemcpyAsync(cov_diag, dc, sizeof(double) * ndim, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_fault_calc_likelihood_batch(smcf_problem* pb, int B, const double* theta, const uint64_t* keys,
                                     double* neglog, int* nstage) {
  if (!pb || !theta || !keys || !neglog) return SMCF_EINVAL;
  Problem& P = pb->p;
  SMCF_HIP(hipSetDevice(P.ctx->device));
  return run_stream(P, B, theta, keys, false, neglog, nstage);
}

int smcf_find_next_gamma(smcf_ctx* ctx, int P, double gprev, const double* L, double cv, double tol, int cap,
                         double* gamma, double* w, double* nev) {
  if (!ctx || P < 1 || !L) return SMCF_EINVAL;
  if (P > kMaxInnerParticles) { set_error("P exceeds the tempering workgroup's LDS"); return SMCF_EINVAL; }
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  double *dL, *dw, *dout;
  TmpBufs tb;
  SMCF_HIP(tb.alloc(&dL, (size_t)P));
  SMCF_HIP(tb.alloc(&dw, (size_t)P));
  SMCF_HIP(tb.alloc(&dout, 3));
  SMCF_HIP(hipMemcpyAsync(dL, L, sizeof(double) * P, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_find_next_gamma(P, gprev, dL, cv, tol, cap, dout, dw, s));
  double out3[3];
  SMCF_HIP(hipMemcpyAsync(out3, dout, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
  if (w) SMCF_HIP(hipMemcpyAsync(w, dw, sizeof(double) * P, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  if (gamma) *gamma = out3[0];
  if (nev) *nev = out3[1];
  return SMCF_OK;
}

int smcf_outer_reweight_resample(smcf_ctx* ctx, int ndim, int N, double gamma_prev, uint64_t seed, int stage,
                                 int device_sort, double* particles, double* likelihood, double* q, double* gamma,
                                 double* neglog_evidence, double* weights, double* cov_chol, int* info,
                                 int* assigned_num, int* sorted_idx) {
  if (!ctx || ndim < 1 || ndim > kMaxOuterDim || N < 1 || !particles || !likelihood || !gamma || !neglog_evidence ||
      !weights || !cov_chol || !assigned_num || !sorted_idx) {
    set_error("outer_reweight_resample: missing argument, or ndim outside 1..64, or N < 1");
    return SMCF_EINVAL;
  }
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  auto mix = [](uint64_t x) {  // the outer key derivation of the host stage functions
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
  };
  uint64_t okey = mix(seed ^ 0x5EEDF00Dull);
  size_t n = (size_t)N, nx = (size_t)N * ndim;
  TmpBufs tb;
  double *dX = nullptr, *dL = nullptr, *dq = nullptr, *dw = nullptr, *dscal = nullptr, *dcov = nullptr;
  double *dX2 = nullptr, *dL2 = nullptr, *dq2 = nullptr;
  int *da = nullptr, *da2 = nullptr, *didx = nullptr, *dtmp = nullptr, *dsi = nullptr;
  SMCF_HIP(tb.alloc(&dX, nx));
  SMCF_HIP(tb.alloc(&dL, n));
  SMCF_HIP(tb.alloc(&dw, n));
  SMCF_HIP(tb.alloc(&dscal, 4));
  SMCF_HIP(tb.alloc(&dcov, (size_t)ndim * ndim));
  SMCF_HIP(tb.alloc(&da, n));
  if (q) SMCF_HIP(tb.alloc(&dq, n));
  SMCF_HIP(hipMemcpyAsync(dX, particles, sizeof(double) * nx, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dL, likelihood, sizeof(double) * n, hipMemcpyHostToDevice, s));
  if (q) SMCF_HIP(hipMemcpyAsync(dq, q, sizeof(double) * n, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_outer_stage(N, ndim, gamma_prev, dL, dq, dX, okey, stage, dw, dscal, dcov, da, s));
  double sc[4];
  SMCF_HIP(hipMemcpyAsync(sc, dscal, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipMemcpyAsync(weights, dw, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipMemcpyAsync(cov_chol, dcov, sizeof(double) * ndim * ndim, hipMemcpyDeviceToHost, s));
  if (device_sort) {
    SMCF_HIP(tb.alloc(&dX2, nx));
    SMCF_HIP(tb.alloc(&dL2, n));
    SMCF_HIP(tb.alloc(&da2, n));
    SMCF_HIP(tb.alloc(&didx, n));
    SMCF_HIP(tb.alloc(&dtmp, n));
    SMCF_HIP(tb.alloc(&dsi, n));
    if (q) SMCF_HIP(tb.alloc(&dq2, n));
    CHECK_LAUNCH(launch_outer_sort(N, ndim, da, didx, dtmp, dX, dL, dq, da2, dX2, dL2, dq2, dsi, s));
    SMCF_HIP(hipMemcpyAsync(assigned_num, da2, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    SMCF_HIP(hipMemcpyAsync(sorted_idx, dsi, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    SMCF_HIP(hipMemcpyAsync(particles, dX2, sizeof(double) * nx, hipMemcpyDeviceToHost, s));
    SMCF_HIP(hipMemcpyAsync(likelihood, dL2, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    if (q) SMCF_HIP(hipMemcpyAsync(q, dq2, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    SMCF_HIP(hipStreamSynchronize(s));
  } else {  // the reference's order: qsort_ on the host (fault_reorder_to_send)
    SMCF_HIP(hipMemcpyAsync(assigned_num, da, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    SMCF_HIP(hipStreamSynchronize(s));
    int rc = smcf_fault_reorder_to_send(ndim, N, assigned_num, particles, likelihood, q, sorted_idx);
    if (rc) return rc;
  }
  *gamma = sc[0];
  *neglog_evidence = sc[1];
  if (info) *info = (int)sc[2];
  return SMCF_OK;
}

int smcf_find_next_gamma_noalloc(smcf_ctx* ctx, int P, double gprev, const double* L, double cv, double tol,
                                 int cap, double* gamma, double* w, double* nev) {
  if (!ctx || P < 1 || !L) return SMCF_EINVAL;
  if (P > kNoAllocMaxP) {
    set_error("smcf_find_next_gamma_noalloc: P = " + std::to_string(P) + " exceeds the scratch maximum of " +
              std::to_string(kNoAllocMaxP) + " particles");
    return SMCF_EINVAL;
  }
  if (!ctx->scratch) {
    set_error("smcf_find_next_gamma_noalloc: the context has no scratch address");
    return SMCF_ESTATE;
  }
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  double* dL = ctx->scratch;          // kNoAllocMaxP likelihoods
  double* dw = dL + kNoAllocMaxP;     // kNoAllocMaxP weights
  double* dout = dw + kNoAllocMaxP;   // gamma, -log evidence increment
  SMCF_HIP(hipMemcpyAsync(dL, L, sizeof(double) * P, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_find_next_gamma(P, gprev, dL, cv, tol, cap, dout, dw, s));
  double out2[2];
  SMCF_HIP(hipMemcpyAsync(out2, dout, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
  if (w) SMCF_HIP(hipMemcpyAsync(w, dw, sizeof(double) * P, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  if (gamma) *gamma = out2[0];
  if (nev) *nev = out2[1];
  return SMCF_OK;
}

int smcf_normalize_resample(smcf_ctx* ctx, int P, const double* w, double uh, double ul, double* wn, int* assigned) {
  if (!ctx || P < 1 || !w) return SMCF_EINVAL;
  if (P > kMaxInnerParticles) { set_error("P exceeds the tempering workgroup's LDS"); return SMCF_EINVAL; }
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  double *dw, *dwn;
  int* da;
  TmpBufs tb;
  SMCF_HIP(tb.alloc(&dw, (size_t)P));
  SMCF_HIP(tb.alloc(&dwn, (size_t)P));
  SMCF_HIP(tb.alloc(&da, 2 * (size_t)P));
  SMCF_HIP(hipMemcpyAsync(dw, w, sizeof(double) * P, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_normalize_resample(P, dw, uh, ul, dwn, da, s));
  if (wn) SMCF_HIP(hipMemcpyAsync(wn, dwn, sizeof(double) * P, hipMemcpyDeviceToHost, s));
  if (assigned) SMCF_HIP(hipMemcpyAsync(assigned, da, sizeof(int) * P, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_trunc_normal(smcf_ctx* ctx, int n, const double* mu, const double* sg, double lo, double hi, const double* u,
                      double* out) {
  if (!ctx || n < 0 || (n > 0 && (!mu || !sg || !u || !out))) return SMCF_EINVAL;
  if (n == 0) return SMCF_OK;
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  double *dm, *ds, *du, *dout;
  TmpBufs tb;
  SMCF_HIP(tb.alloc(&dm, (size_t)n));
  SMCF_HIP(tb.alloc(&ds, (size_t)n));
  SMCF_HIP(tb.alloc(&du, (size_t)n));
  SMCF_HIP(tb.alloc(&dout, (size_t)n));
  SMCF_HIP(hipMemcpyAsync(dm, mu, sizeof(double) * n, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(ds, sg, sizeof(double) * n, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(du, u, sizeof(double) * n, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_trunc_normal(n, dm, ds, lo, hi, du, dout, s));
  SMCF_HIP(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_trunc_normal_f32(smcf_ctx* ctx, int n, const double* mu, const double* sq, double lo, double hi,
                          const float* u, double* out) {
  if (!ctx || n < 0 || (n > 0 && (!mu || !sq || !u || !out))) return SMCF_EINVAL;
  if (n == 0) return SMCF_OK;
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  double *dm, *ds, *dout;
  float* du;
  TmpBufs tb;
  SMCF_HIP(tb.alloc(&dm, (size_t)n));
  SMCF_HIP(tb.alloc(&ds, (size_t)n));
  SMCF_HIP(tb.alloc(&du, (size_t)n));
  SMCF_HIP(tb.alloc(&dout, (size_t)n));
  SMCF_HIP(hipMemcpyAsync(dm, mu, sizeof(double) * n, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(ds, sq, sizeof(double) * n, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(du, u, sizeof(float) * n, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_trunc_normal_f32(n, dm, ds, lo, hi, du, dout, s));
  SMCF_HIP(hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

int smcf_philox(smcf_ctx* ctx, int n, const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  if (!ctx || n < 0 || (n > 0 && (!ctr || !key || !out))) return SMCF_EINVAL;
  if (n == 0) return SMCF_OK;
  SMCF_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  uint32_t *dc, *dk, *dout;
  TmpBufs tb;
  SMCF_HIP(tb.alloc(&dc, 4 * (size_t)n));
  SMCF_HIP(tb.alloc(&dk, 2 * (size_t)n));
  SMCF_HIP(tb.alloc(&dout, 4 * (size_t)n));
  SMCF_HIP(hipMemcpyAsync(dc, ctr, sizeof(uint32_t) * 4 * n, hipMemcpyHostToDevice, s));
  SMCF_HIP(hipMemcpyAsync(dk, key, sizeof(uint32_t) * 2 * n, hipMemcpyHostToDevice, s));
  CHECK_LAUNCH(launch_philox(n, dc, dk, dout, s));
  SMCF_HIP(hipMemcpyAsync(out, dout, sizeof(uint32_t) * 4 * n, hipMemcpyDeviceToHost, s));
  SMCF_HIP(hipStreamSynchronize(s));
  return SMCF_OK;
}

}  // extern "C"
