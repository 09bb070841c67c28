// Stage engine: the inner-SMC stage loop of slip_smc_exec (smc_slip.f90:1445-1663) over the slots of one or more
// problems at once, and the problem groups built on it.
//
// Every stage kernel runs over a *stage set*: a device table of members (problems: per-slot buffers, dimensions,
// sampler settings) and the list of (member, slot) instances, so one launch covers instances whose meshes,
// observation sets and particle counts differ.  The chunked kernels (Gibbs sweeps, likelihoods) take item lists
// (member, slot, 256-particle chunk) in XCD-aware order; the Gibbs sweeps launch once per kernel class (form and
// block rows from the tuning table), so every instance keeps the block rows - and therefore the arithmetic - it has
// when it runs alone.  The per-instance kernels skip the instances without a running geometry; the item lists are
// rewritten with the running instances only whenever that set changes, spread evenly over the XCDs.
//
// After every stage the host reads the instances' status back in one copy and harvests the finished ones; once
// enough slots of a member are free its next geometries are set up in them on the set-up stream (each member's
// Green's, Laplacian and Gram kernels, then one prior-chain launch over every instance admitted together),
// overlapping the stage loop, and released into it at the first stage boundary after their set-up completed.
// Every instance keeps its own stage counter and RNG key, so results do not depend on the slot, the batch, the
// other members or the moment an instance ran.  With graphs on, the stage's launches are captured once per engine
// run into a HIP graph and each stage is one graph launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "smcf_internal.h"

namespace smcf {

namespace {

using Clock = std::chrono::steady_clock;

// one batch call's geometries (thetas nf x n, keys) and where their results go
struct Request {
  int n = 0;
  const double* theta = nullptr;
  const uint64_t* keys = nullptr;
  double* neglog = nullptr;    // may be null
  int* nstage = nullptr;       // may be null
  double* t_done = nullptr;    // seconds since t0 to each geometry's result (may be null)
  double* t_admit = nullptr;   // seconds since t0 to each geometry's set-up launch (may be null)
  Clock::time_point t0;
  int remaining = 0;
  bool finished = false;
  int rc = SMCF_OK;
  std::string err;
};

struct Geo {
  Request* r = nullptr;
  int i = 0;
};

// one admitted instance as k_admit reads it
struct AdmitEntry {
  int member, slot, k, pad;
  uint64_t key;
  double theta[kMaxOuterDim];
};

// every instance of the set held as done (skipped by the stage kernels) until an admission releases it
__global__ void k_idle(StageSet ss) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ss.ninst) return;
  int2 is = ss.inst[g];
  const BatchDev& bt = ss.mem[is.x].bt;
  bt.done[is.y] = 1;
  bt.active[is.y] = 0;
}

// thetas and keys of the admitted instances into their slots, and each member's slot list (entry k of the member
// at sel[half cap + k]) for its set-up kernels
__global__ void k_admit(const MemberDev* __restrict__ mem, const AdmitEntry* __restrict__ e, int n, int half) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const AdmitEntry& a = e[k];
  const MemberDev& M = mem[a.member];
  for (int d = 0; d < M.nf; ++d) M.bt.theta[(size_t)a.slot * M.nf + d] = a.theta[d];
  M.bt.key[a.slot] = a.key;
  M.bt.sel[(size_t)half * M.bt.cap + a.k] = a.slot;
}

// the admitted instances join the stage loop
__global__ void k_release(const MemberDev* __restrict__ mem, const int2* __restrict__ list, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  mem[list[k].x].bt.done[list[k].y] = 0;
}

// status of every instance after the stage: -log Z, done, stage counter
__global__ void k_status(StageSet ss, double* __restrict__ out) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ss.ninst) return;
  int2 is = ss.inst[g];
  const BatchDev& bt = ss.mem[is.x].bt;
  out[3 * (size_t)g] = bt.neglog[is.y];
  out[3 * (size_t)g + 1] = (double)bt.done[is.y];
  out[3 * (size_t)g + 2] = (double)bt.iter[is.y];
}

struct MemberRun {
  Problem* P = nullptr;
  int W = 0;      // slots in use: 0 .. W-1
  int first = 0;  // its first instance in the stage set
  bool fresh = false;  // added since the last build: its instances are not yet marked idle
  int admit_min = 1;
  std::deque<Geo> pending;
  std::vector<Geo> job;  // per slot (r == nullptr: free)
  std::vector<int> freed;
  int busy = 0;
};

struct Admission {
  bool live = false;
  int half = 0;
  std::vector<AdmitEntry> ent;
  std::vector<int2> list;
  std::vector<int2> jobs;  // prior-chain wavefronts: (first entry of list, count), one member each
  std::vector<Geo> geo;
  hipEvent_t ready{};
};

// items of the chunked kernels for the given instances (member, slot, chunks): each instance in turn goes to the
// XCD whose queue is shortest (the lowest such XCD; with equal chunk counts XCD g % 8 for instance g), its chunks at
// consecutive positions of that queue; item 8r + x is entry r of XCD x's queue (workgroup w runs on XCD w % 8), and
// the queues are padded to equal length with empty items.  The longest queue is at most ceil(T / 8) + c items for T
// chunks in all and c the most of one instance (xcd_items_bound).
void xcd_items(const std::vector<int4>& insts, std::vector<int4>& out) {
  std::vector<std::vector<int4>> q(8);
  for (size_t g = 0; g < insts.size(); ++g) {
    int x = 0;
    for (int y = 1; y < 8; ++y)
      if (q[y].size() < q[x].size()) x = y;
    for (int c = 0; c < insts[g].z; ++c) q[x].push_back(make_int4(insts[g].x, insts[g].y, c, 0));
  }
  size_t L = 0;
  for (auto& v : q) L = std::max(L, v.size());
  out.assign(8 * L, make_int4(-1, 0, 0, 0));
  for (int x = 0; x < 8; ++x)
    for (size_t r = 0; r < q[x].size(); ++r) out[8 * r + x] = q[x][r];
}

// length that xcd_items keeps within for any subset of these instances
size_t xcd_items_bound(const std::vector<int4>& insts) {
  size_t T = 0, c = 0;
  for (auto& e : insts) {
    T += (size_t)e.z;
    c = std::max(c, (size_t)e.z);
  }
  return insts.empty() ? 0 : 8 * ((T + 7) / 8 + c);
}

}  // namespace

struct Engine {
  smcf_ctx* ctx = nullptr;
  std::vector<MemberRun> mr;
  bool want_moments = false, graph_on = false, async = false;
  bool hmc = false, any_full = false, any_f32 = false, any_bf = false, use_graph = false;
  // device tables, grown as members join (never shrunk while the engine lives)
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  Buf b_mem, b_inst, b_items, b_stat, b_adm[2], b_list[2];
  MemberDev* dmem = nullptr;
  double* dstat = nullptr;
  AdmitEntry* dadm[2] = {nullptr, nullptr};
  int2* dlist[2] = {nullptr, nullptr};
  StageSet ss{};
  struct GibbsLaunch {
    int precision, kb;
    GibbsArgs args;
    ItemList it;
  };
  std::vector<GibbsLaunch> glaunch;
  ItemList like_fp{}, like_bf{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // The chunked kernels' item lists hold only the instances running a geometry, rewritten whenever that set changed
  // (an instance released or harvested), so the XCD-aware order spreads the running instances evenly over the XCDs:
  // with the finished and waiting instances left in place the busiest XCD set a launch's time.  The launches cover
  // each list's full length (the stage graph's grids are fixed); the rest is padding that returns at once.
  struct ListSrc {
    size_t off = 0;
    size_t len = 0;
    std::vector<int4> inst;   // (member in the launch's table, slot, chunks, 0) of every instance of the list
    std::vector<int> member;  // the instance's member in mr
  };
  std::vector<ListSrc> lists;
  std::vector<int4> items_h;
  long alive_version = 0, items_version = -1;
  Admission adm;
  std::vector<double> stat;
  int busy_total = 0;
  // asynchronous submissions: batch calls of client threads (a group's outer SMCs, a server's requests), served by
  // the thread in run(); with `persistent` (serving) members join and leave while it runs
  std::mutex mu;
  std::mutex capture_mu;  // held while the stage graph is captured and by threads that call HIP beside the engine
  std::condition_variable cv_engine, cv_done;
  std::deque<std::pair<Problem*, Request*>> inbox;
  std::vector<Problem*> add_q, leave_q;
  std::function<void(Problem*)> on_added, on_removed;
  bool persistent = false, stopping = false;
  int clients = 0;
  int rc = SMCF_OK;
  std::string err;

  ~Engine() {
    drop_graph();
    if (adm.ready) (void)hipEventDestroy(adm.ready);
    for (Buf* b : {&b_mem, &b_inst, &b_items, &b_stat, &b_adm[0], &b_adm[1], &b_list[0], &b_list[1]})
      if (b->p) (void)hipFree(b->p);
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }

  int grow(Buf& b, size_t bytes) {
    if (bytes <= b.bytes) return SMCF_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    size_t cap = std::max(bytes, (size_t)4096);
    hipError_t e = hipMalloc(&b.p, cap);
    if (e != hipSuccess) {
      set_error(std::string("stage engine: hipMalloc: ") + hipGetErrorString(e));
      return SMCF_ENOMEM;
    }
    b.bytes = cap;
    return SMCF_OK;
  }

  int init(smcf_ctx* c, bool moments, bool graph) {
    ctx = c;
    want_moments = moments;
    graph_on = graph;
    SMCF_HIP(hipEventCreateWithFlags(&adm.ready, hipEventDisableTiming));
    return SMCF_OK;
  }

  // a member with W slots (0 .. W-1) in use, all free; its instances are held idle until admissions release them
  void add(Problem* P, int W) {
    MemberRun m;
    m.P = P;
    m.W = W;
    m.admit_min = std::max(1, W / 16);
    m.job.assign(W, Geo{});
    for (int b = 0; b < W; ++b) m.freed.push_back(b);
    m.fresh = true;
    mr.push_back(std::move(m));
    P->engine = this;
  }

  // tables, item lists and (optionally) the stage graph of the current members; members added since the last build
  // have their instances marked idle
  int build() {
    drop_graph();
    int n = (int)mr.size(), ninst = 0;
    std::vector<MemberDev> md(n);
    std::vector<int2> inst;
    int nd_max = 1, P_max = 1, kb_max = 16;
    hmc = any_full = any_f32 = any_bf = false;
    for (int i = 0; i < n; ++i) {
      Problem& P = *mr[i].P;
      mr[i].first = ninst;
      md[i] = member_of(P);
      for (int b = 0; b < mr[i].W; ++b) inst.push_back(make_int2(i, b));
      ninst += mr[i].W;
      nd_max = std::max(nd_max, P.m.ndim);
      P_max = std::max(P_max, P.P);
      kb_max = std::max(kb_max, P.bt.kb);
      if (P.mover != SMCF_MOVER_GIBBS) hmc = true;
      if (P.precision == SMCF_PRECISION_MIXED) any_f32 = true;
      if (P.bt.QX) any_bf = true;
      if (P.bt.Q && !P.bt.QX) any_full = true;
    }
    if (hmc && n != 1) {
      set_error("stage engine: HMC movers run one problem at a time");
      return SMCF_EUNSUPPORTED;
    }
    // item lists: likelihoods (FP64 Gram form / centred split BF16; member = index in the device table) and the
    // Gibbs launches: per kernel class (form, block rows) groups of at most kGibbsArgMembers members whose table
    // travels as a kernel argument (member = index in that table)
    std::vector<int4> fp, bf;
    struct Cls {
      int precision, kb;
      std::vector<int> members;
    };
    std::vector<Cls> cls;
    for (int i = 0; i < n; ++i) {
      Problem& P = *mr[i].P;
      // 64 particles per likelihood workgroup (Gram form and centred split-BF16 form)
      int chunks = (P.P + 63) / 64;
      for (int b = 0; b < mr[i].W; ++b) (P.bt.QX ? bf : fp).push_back(make_int4(i, b, chunks, 0));
      if (P.mover != SMCF_MOVER_GIBBS || mr[i].W == 0) continue;
      auto it = std::find_if(cls.begin(), cls.end(),
                             [&](const Cls& c) { return c.precision == P.precision && c.kb == P.bt.kb; });
      if (it == cls.end()) {
        cls.push_back({P.precision, P.bt.kb, {}});
        it = cls.end() - 1;
      }
      it->members.push_back(i);
    }
    std::vector<int4> all, tmp;
    lists.clear();
    auto add_list = [&](const std::vector<int4>& v, const std::vector<int>& member, size_t& off, int& cnt) {
      // every instance's items; the region is as long as any subset's (the running instances') list can be
      xcd_items(v, tmp);
      off = all.size();
      size_t len = xcd_items_bound(v);
      cnt = (int)len;
      all.insert(all.end(), tmp.begin(), tmp.end());
      all.resize(off + len, make_int4(-1, 0, 0, 0));
      ListSrc L;
      L.off = off;
      L.len = len;
      L.inst = v;
      L.member = member;
      lists.push_back(std::move(L));
    };
    auto members_of = [](const std::vector<int4>& v) {
      std::vector<int> m;
      for (auto& e : v) m.push_back(e.x);
      return m;
    };
    size_t off_fp = 0, off_bf = 0;
    int n_fp = 0, n_bf = 0;
    add_list(fp, members_of(fp), off_fp, n_fp);
    add_list(bf, members_of(bf), off_bf, n_bf);
    glaunch.clear();
    std::vector<std::pair<size_t, int>> goff;
    for (auto& c : cls)
      for (size_t g0 = 0; g0 < c.members.size(); g0 += kGibbsArgMembers) {
        GibbsLaunch gl{};
        gl.precision = c.precision;
        gl.kb = c.kb;
        std::vector<int4> v;
        std::vector<int> vm;
        for (size_t k = g0; k < std::min(c.members.size(), g0 + kGibbsArgMembers); ++k) {
          int i = c.members[k];
          gl.args.m[k - g0] = gibbs_member_of(*mr[i].P);
          int chunks = (mr[i].P->P + 255) / 256;
          for (int b = 0; b < mr[i].W; ++b) {
            v.push_back(make_int4((int)(k - g0), b, chunks, 0));
            vm.push_back(i);
          }
        }
        goff.push_back({0, 0});
        add_list(v, vm, goff.back().first, goff.back().second);
        glaunch.push_back(gl);
      }
    int r;
    if ((r = grow(b_mem, sizeof(MemberDev) * std::max(n, 1))) || (r = grow(b_inst, sizeof(int2) * std::max(ninst, 1))) ||
        (r = grow(b_items, sizeof(int4) * std::max<size_t>(all.size(), 1))) ||
        (r = grow(b_stat, sizeof(double) * 3 * std::max(ninst, 1))))
      return r;
    for (int h = 0; h < 2; ++h)
      if ((r = grow(b_adm[h], sizeof(AdmitEntry) * std::max(ninst, 1))) ||
          (r = grow(b_list[h], sizeof(int2) * 2 * std::max(ninst, 1))))  // the list, then the chain jobs
        return r;
    dmem = (MemberDev*)b_mem.p;
    dstat = (double*)b_stat.p;
    for (int h = 0; h < 2; ++h) {
      dadm[h] = (AdmitEntry*)b_adm[h].p;
      dlist[h] = (int2*)b_list[h].p;
    }
    int2* dinst = (int2*)b_inst.p;
    int4* ditems = (int4*)b_items.p;
    hipStream_t s = ctx->stream;
    if (n) SMCF_HIP(hipMemcpyAsync(dmem, md.data(), sizeof(MemberDev) * n, hipMemcpyHostToDevice, s));
    if (ninst) SMCF_HIP(hipMemcpyAsync(dinst, inst.data(), sizeof(int2) * ninst, hipMemcpyHostToDevice, s));
    if (!all.empty()) SMCF_HIP(hipMemcpyAsync(ditems, all.data(), sizeof(int4) * all.size(), hipMemcpyHostToDevice, s));
    items_h = all;
    items_version = -1;  // the running instances' lists are written before the next stage
    ss = StageSet{dmem, dinst, ninst, nd_max, P_max, kb_max};
    like_fp = ItemList{ditems + off_fp, n_fp};
    like_bf = ItemList{ditems + off_bf, n_bf};
    for (size_t k = 0; k < glaunch.size(); ++k) glaunch[k].it = ItemList{ditems + goff[k].first, goff[k].second};
    stat.assign(3 * (size_t)ninst, 0.0);
    for (auto& m : mr)
      if (m.fresh && m.W > 0) {
        StageSet sub = ss;
        sub.inst = dinst + m.first;
        sub.ninst = m.W;
        hipLaunchKernelGGL(k_idle, dim3((m.W + 255) / 256), dim3(256), 0, s, sub);
        SMCF_HIP(hipGetLastError());
        m.fresh = false;
      }
    // the host copies above are read before this call returns
    SMCF_HIP(hipStreamSynchronize(s));
    use_graph = graph_on && !hmc && !ctx->profiling && ninst > 0;
    if (use_graph) {
      // other threads of the process make no HIP call while the stage is captured (capture_lock)
      std::lock_guard<std::mutex> cl(capture_mu);
      SMCF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      int rl = stage_launch(s);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s, &g);
      if (rl) {
        if (g) (void)hipGraphDestroy(g);
        return rl;
      }
      if (e != hipSuccess) {
        set_error(std::string("stage engine: hipStreamEndCapture: ") + hipGetErrorString(e));
        return SMCF_EHIP;
      }
      graph = g;
      SMCF_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    }
    return SMCF_OK;
  }

  // the launches of one stage over the whole set, ending with the status copy kernel
  int stage_launch(hipStream_t s) {
    hipError_t e;
    {
      ProfScope ps(ctx, "gamma_resample", s);
      if ((e = launch_gamma_resample(ss, s)) != hipSuccess) return launch_failed("gamma_resample", e);
    }
    if ((want_moments || hmc) && (e = launch_moments(ss, s)) != hipSuccess) return launch_failed("moments", e);
    if (hmc) {
      ProfScope ps(ctx, "hmc", s);
      Problem& P = *mr[0].P;
      const HmcCfg hcfg{P.mover, P.hmc_ntau, P.hmc_accept, P.precision == SMCF_PRECISION_MIXED_BF16 ? 1 : 0};
      if ((e = launch_hmc_stage(P.m, P.bt, P.hm, P.hh, mr[0].W, P.max_slip, hcfg, s)) != hipSuccess)
        return launch_failed("hmc_stage", e);
    } else {
      if ((e = launch_gather(ss, s)) != hipSuccess) return launch_failed("gather", e);
      if ((e = launch_form_q(ss, any_full, any_f32, any_bf, s)) != hipSuccess) return launch_failed("form_q", e);
      if (any_bf && (e = launch_like_ref(ss, s)) != hipSuccess) return launch_failed("like_ref", e);
      {
        ProfScope ps(ctx, "gibbs", s);
        for (auto& g : glaunch)
          if ((e = launch_gibbs(g.args, g.it, g.precision, g.kb, s)) != hipSuccess) return launch_failed("gibbs", e);
      }
      {
        // the Gibbs mover reads no prior values; they are formed once for the final particles
        ProfScope ps(ctx, "likelihood", s);
        if ((e = launch_like_bf(ss, like_bf, s)) != hipSuccess || (e = launch_like_stage(ss, like_fp, s)) != hipSuccess)
          return launch_failed("likelihood", e);
      }
    }
    hipLaunchKernelGGL(k_status, dim3((ss.ninst + 255) / 256), dim3(256), 0, s, ss, dstat);
    if ((e = hipGetLastError()) != hipSuccess) return launch_failed("status", e);
    return SMCF_OK;
  }

  int launch_failed(const char* what, hipError_t e) {
    set_error(std::string("stage engine: launch of ") + what + ": " + hipGetErrorString(e));
    return SMCF_EHIP;
  }

  static double since(const Clock::time_point& t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  // the item lists of the instances now running a geometry (the host's stream synchronisation at the end of every
  // stage keeps items_h unchanged until the copy has been read)
  int upload_items(hipStream_t s) {
    std::vector<int4> sub, tmp;
    for (auto& L : lists) {
      sub.clear();
      for (size_t k = 0; k < L.inst.size(); ++k)
        if (mr[L.member[k]].job[L.inst[k].y].r) sub.push_back(L.inst[k]);
      xcd_items(sub, tmp);
      if (tmp.size() > L.len) {
        set_error("stage engine: an item list outgrew its region");
        return SMCF_ESTATE;
      }
      std::copy(tmp.begin(), tmp.end(), items_h.begin() + L.off);
      std::fill(items_h.begin() + L.off + tmp.size(), items_h.begin() + L.off + L.len, make_int4(-1, 0, 0, 0));
    }
    if (!items_h.empty())
      SMCF_HIP(hipMemcpyAsync(b_items.p, items_h.data(), sizeof(int4) * items_h.size(), hipMemcpyHostToDevice, s));
    items_version = alive_version;
    return SMCF_OK;
  }

  // one stage, its status read back, finished instances harvested
  int step() {
    hipStream_t s = ctx->stream;
    if (items_version != alive_version) {
      int r = upload_items(s);
      if (r) return r;
    }
    if (gexec) {
      SMCF_HIP(hipGraphLaunch(gexec, s));
    } else {
      int r = stage_launch(s);
      if (r) return r;
    }
    SMCF_HIP(hipMemcpyAsync(stat.data(), dstat, sizeof(double) * stat.size(), hipMemcpyDeviceToHost, s));
    SMCF_HIP(hipStreamSynchronize(s));
    for (auto& m : mr)
      for (int b = 0; b < m.W; ++b) {
        Geo& g = m.job[b];
        size_t k = 3 * (size_t)(m.first + b);
        if (!g.r || stat[k + 1] == 0.0) continue;
        Request& r = *g.r;
        if (r.neglog) r.neglog[g.i] = stat[k];
        if (r.nstage) r.nstage[g.i] = std::min((int)stat[k + 2] - 1, 201);
        if (r.t_done) r.t_done[g.i] = since(r.t0);
        g.r = nullptr;
        ++alive_version;
        --m.busy;
        --busy_total;
        m.freed.push_back(b);
        if (--r.remaining == 0) finish(r, SMCF_OK, "");
      }
    return SMCF_OK;
  }

  void finish(Request& r, int code, const std::string& msg) {
    std::unique_lock<std::mutex> lk(mu, std::defer_lock);
    if (async) lk.lock();
    r.rc = code;
    r.err = msg;
    r.finished = true;
    if (async) cv_done.notify_all();
  }

  // set up the next geometries of every member with enough free slots, in one admission on the set-up stream
  int start_admission() {
    adm.ent.clear();
    adm.list.clear();
    adm.geo.clear();
    std::vector<std::pair<int, int>> who;  // member, count
    for (int i = 0; i < (int)mr.size(); ++i) {
      MemberRun& m = mr[i];
      if (m.pending.empty() || m.freed.empty()) continue;
      if ((int)m.freed.size() < m.admit_min && m.busy > 0) continue;
      int n = (int)std::min(m.freed.size(), m.pending.size());
      const int nf = m.P->m.ndim_fault;
      for (int k = 0; k < n; ++k) {
        Geo g = m.pending.front();
        m.pending.pop_front();
        AdmitEntry e{};
        e.member = i;
        e.slot = m.freed[k];
        e.k = k;
        e.key = g.r->keys[g.i];
        std::copy(g.r->theta + (size_t)g.i * nf, g.r->theta + (size_t)(g.i + 1) * nf, e.theta);
        if (g.r->t_admit) g.r->t_admit[g.i] = since(g.r->t0);
        adm.ent.push_back(e);
        adm.list.push_back(make_int2(i, e.slot));
        adm.geo.push_back(g);
      }
      m.freed.erase(m.freed.begin(), m.freed.begin() + n);
      who.push_back({i, n});
    }
    if (who.empty()) return SMCF_OK;
    // prior chains: up to 32 / chain_lanes consecutive entries of one member per wavefront
    adm.jobs.clear();
    for (size_t k = 0; k < adm.list.size();) {
      const int mi = adm.list[k].x, per = kChainLanes / mr[mi].P->m.chain_lanes;
      size_t e = k;
      while (e < adm.list.size() && adm.list[e].x == mi && (int)(e - k) < per) ++e;
      adm.jobs.push_back(make_int2((int)k, (int)(e - k)));
      k = e;
    }
    adm.half = 1 - adm.half;
    const int h = adm.half, n = (int)adm.list.size(), nj = (int)adm.jobs.size();
    hipStream_t s2 = ctx->setup_stream;
    SMCF_HIP(hipMemcpyAsync(dadm[h], adm.ent.data(), sizeof(AdmitEntry) * n, hipMemcpyHostToDevice, s2));
    SMCF_HIP(hipMemcpyAsync(dlist[h], adm.list.data(), sizeof(int2) * n, hipMemcpyHostToDevice, s2));
    SMCF_HIP(hipMemcpyAsync(dlist[h] + n, adm.jobs.data(), sizeof(int2) * nj, hipMemcpyHostToDevice, s2));
    hipLaunchKernelGGL(k_admit, dim3((n + 63) / 64), dim3(64), 0, s2, dmem, dadm[h], n, h);
    SMCF_HIP(hipGetLastError());
    int chain_w = 1, r;
    for (auto& w : who) {
      Problem& P = *mr[w.first].P;
      if ((r = setup_pre(P, w.second, P.bt.sel + (size_t)h * P.bt.cap, s2))) return r;
      chain_w = std::max(chain_w, P.m.chain_w);
    }
    {
      ProfScope ps(ctx, "init_chain", s2);
      hipError_t e = launch_init_chain(dmem, dlist[h], nullptr, dlist[h] + n, nj, chain_w, s2);
      if (e != hipSuccess) {
        set_error(std::string("stage engine: launch of init_chain: ") + hipGetErrorString(e));
        return SMCF_EHIP;
      }
    }
    for (auto& w : who) {
      Problem& P = *mr[w.first].P;
      if ((r = setup_post(P, w.second, P.bt.sel + (size_t)h * P.bt.cap, s2))) return r;
    }
    SMCF_HIP(hipEventRecord(adm.ready, s2));
    adm.live = true;
    return SMCF_OK;
  }

  // the completed admission's instances join the stage loop
  int release_admission() {
    int n = (int)adm.list.size();
    hipLaunchKernelGGL(k_release, dim3((n + 63) / 64), dim3(64), 0, ctx->stream, dmem, dlist[adm.half], n);
    SMCF_HIP(hipGetLastError());
    for (int k = 0; k < n; ++k) {
      MemberRun& m = mr[adm.list[k].x];
      m.job[adm.list[k].y] = adm.geo[k];
      ++m.busy;
      ++busy_total;
    }
    ++alive_version;
    adm.live = false;
    return SMCF_OK;
  }

  bool nothing_pending() const {
    for (auto& m : mr)
      if (!m.pending.empty()) return false;
    return true;
  }

  int member_index(const Problem* P) const {
    for (int i = 0; i < (int)mr.size(); ++i)
      if (mr[i].P == P) return i;
    return -1;
  }

  void queue(int member, Request* r) {
    for (int i = 0; i < r->n; ++i) mr[member].pending.push_back(Geo{r, i});
    if (r->n == 0) r->finished = true;
  }

  // serve until every request has finished (asynchronously: until every client has left and, for a persistent
  // engine, the engine is stopped)
  int run() {
    int r = loop();
    if (r) {
      // the error ends every request still queued or in flight
      std::string msg = smcf_last_error();
      (void)hipStreamSynchronize(ctx->setup_stream);
      (void)hipStreamSynchronize(ctx->stream);
      std::vector<Request*> open;
      for (auto& m : mr) {
        for (auto& g : m.pending) open.push_back(g.r);
        for (auto& g : m.job)
          if (g.r) open.push_back(g.r);
      }
      for (auto& g : adm.geo) open.push_back(g.r);
      std::unique_lock<std::mutex> lk(mu, std::defer_lock);
      if (async) lk.lock();
      rc = r;
      err = msg;
      for (auto& p : inbox) open.push_back(p.second);
      inbox.clear();
      // problems that never joined have no client running
      clients -= (int)add_q.size();
      add_q.clear();
      for (Request* q : open) {
        q->rc = r;
        q->err = msg;
        q->finished = true;
      }
      if (async) {
        cv_done.notify_all();
        // clients still running their outer SMC submit again and get the error at once; wait until they left
        cv_engine.wait(lk, [&] { return clients == 0; });
      }
      return r;
    }
    SMCF_HIP(hipStreamSynchronize(ctx->stream));
    resolve_profiles(ctx);
    return SMCF_OK;
  }

  // members that joined or left since the last stage (only between admissions: an admission's lists index members)
  // (a joining member's client starts at once: its first batch call is routed at the next stage boundary, after
  // the tables include it; a leaving member's problem is released only after the tables no longer do)
  int change_members(std::vector<Problem*>& adds, std::vector<Problem*>& leaves) {
    for (Problem* P : leaves) {
      int i = member_index(P);
      if (i >= 0) mr.erase(mr.begin() + i);
      P->engine = nullptr;
    }
    for (Problem* P : adds) {
      add(P, P->bt.cap);
      if (on_added) on_added(P);
    }
    int r = build();
    if (r) return r;
    for (Problem* P : leaves)
      if (on_removed) on_removed(P);
    return SMCF_OK;
  }

  int loop() {
    for (;;) {
      std::vector<Problem*> adds, leaves;
      if (async) {
        std::unique_lock<std::mutex> lk(mu);
        while (!inbox.empty()) {
          int i = member_index(inbox.front().first);
          Request* q = inbox.front().second;
          inbox.pop_front();
          if (i < 0) {
            q->rc = SMCF_ESTATE;
            q->err = "stage engine: batch call of a problem that is not a member";
            q->finished = true;
            cv_done.notify_all();
          } else {
            queue(i, q);
          }
        }
        if (!adm.live) {
          adds.swap(add_q);
          // a member leaves once its client has finished and none of its geometries is left
          for (auto it = leave_q.begin(); it != leave_q.end();) {
            int i = member_index(*it);
            if (i < 0 || (mr[i].busy == 0 && mr[i].pending.empty())) {
              leaves.push_back(*it);
              it = leave_q.erase(it);
            } else {
              ++it;
            }
          }
        }
        if (adds.empty() && leaves.empty() && busy_total == 0 && !adm.live && nothing_pending()) {
          auto over = [&] { return clients == 0 && (!persistent || stopping) && add_q.empty() && leave_q.empty(); };
          if (over()) return SMCF_OK;
          cv_engine.wait(lk, [&] { return !inbox.empty() || !add_q.empty() || !leave_q.empty() || over(); });
          continue;
        }
      } else if (busy_total == 0 && !adm.live && nothing_pending()) {
        return SMCF_OK;
      }
      int r;
      if ((!adds.empty() || !leaves.empty()) && (r = change_members(adds, leaves))) return r;
      if (busy_total > 0 && (r = step())) return r;
      if (adm.live) {
        hipError_t q = busy_total > 0 ? hipEventQuery(adm.ready) : hipEventSynchronize(adm.ready);
        if (q == hipSuccess) {
          if ((r = release_admission())) return r;
        } else if (q != hipErrorNotReady) {
          set_error(std::string("stage engine: set-up: ") + hipGetErrorString(q));
          return SMCF_EHIP;
        }
      }
      if (!adm.live && (r = start_admission())) return r;
    }
  }
};

int run_stream(Problem& P, int N, const double* theta, const uint64_t* keys, bool want_moments, double* neglog,
               int* nstage) {
  if (N < 1) {
    set_error("no geometries");
    return SMCF_EINVAL;
  }
  if (P.engine) return engine_submit_wait(P, N, theta, keys, neglog, nstage);
  Engine E;
  int r = E.init(P.ctx, want_moments, false);
  if (r) return r;
  E.add(&P, std::min(N, P.bt.cap));
  struct Detach {  // the problem leaves this engine however the call ends
    Problem* p;
    ~Detach() { p->engine = nullptr; }
  } detach{&P};
  if ((r = E.build())) return r;
  P.t_done.assign(N, 0.0);
  P.t_admit.assign(N, 0.0);
  Request q;
  q.n = q.remaining = N;
  q.theta = theta;
  q.keys = keys;
  q.neglog = neglog;
  q.nstage = nstage;
  q.t_done = P.t_done.data();
  q.t_admit = P.t_admit.data();
  q.t0 = Clock::now();
  E.queue(0, &q);
  return E.run();
}

int engine_submit_wait(Problem& P, int N, const double* theta, const uint64_t* keys, double* neglog, int* nstage) {
  Engine* E = P.engine;
  P.t_done.assign(N, 0.0);
  P.t_admit.assign(N, 0.0);
  Request q;
  q.n = q.remaining = N;
  q.theta = theta;
  q.keys = keys;
  q.neglog = neglog;
  q.nstage = nstage;
  q.t_done = P.t_done.data();
  q.t_admit = P.t_admit.data();
  q.t0 = Clock::now();
  std::unique_lock<std::mutex> lk(E->mu);
  if (E->rc) {
    set_error(E->err);
    return E->rc;
  }
  E->inbox.push_back({&P, &q});
  E->cv_engine.notify_all();
  E->cv_done.wait(lk, [&] { return q.finished; });
  if (q.rc) set_error(q.err);
  return q.rc;
}

// Serving: one persistent engine on the server's context; requests' problems join it as members when admitted and
// leave it when their outer SMC has finished.
Engine* serve_engine_create(smcf_ctx* ctx, std::function<void(Problem*)> added, std::function<void(Problem*)> removed) {
  auto* E = new Engine();
  if (E->init(ctx, false, true)) {
    delete E;
    return nullptr;
  }
  E->async = true;
  E->persistent = true;
  E->on_added = std::move(added);
  E->on_removed = std::move(removed);
  return E;
}

int serve_engine_run(Engine* E) { return E->run(); }

void serve_engine_add(Engine* E, Problem* P) {
  std::lock_guard<std::mutex> lk(E->mu);
  E->add_q.push_back(P);
  ++E->clients;
  E->cv_engine.notify_all();
}

void serve_engine_leave(Engine* E, Problem* P) {
  std::lock_guard<std::mutex> lk(E->mu);
  E->leave_q.push_back(P);
  --E->clients;
  E->cv_engine.notify_all();
}

void serve_engine_stop(Engine* E) {
  std::lock_guard<std::mutex> lk(E->mu);
  E->stopping = true;
  E->cv_engine.notify_all();
}

void serve_engine_destroy(Engine* E) { delete E; }

std::mutex& serve_engine_capture_lock(Engine* E) { return E->capture_mu; }

int serve_engine_error(Engine* E, std::string* msg) {
  std::lock_guard<std::mutex> lk(E->mu);
  if (msg) *msg = E->err;
  return E->rc;
}

}  // namespace smcf

using namespace smcf;

struct smcf_group {
  smcf_ctx* ctx = nullptr;
  std::vector<smcf_problem*> members;
  bool graph = true;
};

extern "C" {

int smcf_group_create(smcf_ctx* ctx, int n, smcf_problem* const* members, smcf_group** out) {
  if (!ctx || n < 1 || !members || !out) return SMCF_EINVAL;
  for (int i = 0; i < n; ++i) {
    if (!members[i]) return SMCF_EINVAL;
    Problem& P = members[i]->p;
    if (P.ctx != ctx) {
      set_error("smcf_group_create: every member must be created on the group's context");
      return SMCF_EINVAL;
    }
    if (P.mover != SMCF_MOVER_GIBBS) {
      set_error("smcf_group_create: group members take the Gibbs mover");
      return SMCF_EUNSUPPORTED;
    }
    if (P.in_group) {
      set_error("smcf_group_create: a problem belongs to one group at a time");
      return SMCF_ESTATE;
    }
    for (int j = 0; j < i; ++j)
      if (members[j] == members[i]) {
        set_error("smcf_group_create: a problem appears twice");
        return SMCF_EINVAL;
      }
  }
  auto* g = new smcf_group();
  g->ctx = ctx;
  g->members.assign(members, members + n);
  for (auto* m : g->members) m->p.in_group = true;
  *out = g;
  return SMCF_OK;
}

int smcf_group_set_graph(smcf_group* g, int on) {
  if (!g) return SMCF_EINVAL;
  g->graph = on != 0;
  return SMCF_OK;
}

int smcf_group_destroy(smcf_group* g) {
  if (!g) return SMCF_OK;
  for (auto* m : g->members) m->p.in_group = false;
  delete g;
  return SMCF_OK;
}

int smcf_group_calc_likelihood_batch(smcf_group* g, const int* B, const double* const* theta,
                                     const uint64_t* const* keys, double* const* neglog, int* const* nstage) {
  if (!g || !B || !theta || !keys || !neglog) return SMCF_EINVAL;
  int n = (int)g->members.size(), total = 0;
  for (int i = 0; i < n; ++i) {
    if (B[i] < 0 || (B[i] > 0 && (!theta[i] || !keys[i] || !neglog[i]))) return SMCF_EINVAL;
    total += B[i];
  }
  if (total == 0) return SMCF_OK;
  SMCF_HIP(hipSetDevice(g->ctx->device));
  Engine E;
  int r = E.init(g->ctx, false, g->graph);
  if (r) return r;
  for (int i = 0; i < n; ++i) {
    Problem& P = g->members[i]->p;
    E.add(&P, std::min(B[i], P.bt.cap));
  }
  struct Detach {
    smcf_group* g;
    ~Detach() {
      for (auto* m : g->members) m->p.engine = nullptr;
    }
  } detach{g};
  if ((r = E.build())) return r;
  std::vector<Request> q(n);
  auto t0 = Clock::now();
  for (int i = 0; i < n; ++i) {
    Problem& P = g->members[i]->p;
    P.t_done.assign(B[i], 0.0);
    P.t_admit.assign(B[i], 0.0);
    q[i].n = q[i].remaining = B[i];
    q[i].theta = theta[i];
    q[i].keys = keys[i];
    q[i].neglog = neglog[i];
    q[i].nstage = nstage ? nstage[i] : nullptr;
    q[i].t_done = P.t_done.data();
    q[i].t_admit = P.t_admit.data();
    q[i].t0 = t0;
    E.queue(i, &q[i]);
  }
  return E.run();
}

int smcf_group_fault_smc_exec(smcf_group* g, const double* const* range, const int* N, const uint64_t* seed,
                              int max_stages, double* const* particles, double* const* likelihood, int* nstage,
                              double* evidence) {
  if (!g || !range || !N || !seed || !particles || !likelihood) return SMCF_EINVAL;
  int n = (int)g->members.size();
  for (int i = 0; i < n; ++i)
    if (!range[i] || N[i] < 1 || !particles[i] || !likelihood[i]) return SMCF_EINVAL;
  for (auto* m : g->members)
    if (outer_stage_mode(m) != 0) {
      // the device outer stage would launch on the context's stream from the members' threads, beside the engine
      set_error("smcf_group_fault_smc_exec: group members run the host outer stage (smcf_fault_smc_set_outer_stage 0)");
      return SMCF_EUNSUPPORTED;
    }
  SMCF_HIP(hipSetDevice(g->ctx->device));
  Engine E;
  int r = E.init(g->ctx, false, g->graph);
  if (r) return r;
  E.async = true;
  for (auto* m : g->members) E.add(&m->p, m->p.bt.cap);
  struct Detach {
    smcf_group* g;
    ~Detach() {
      for (auto* m : g->members) m->p.engine = nullptr;
    }
  } detach{g};
  if ((r = E.build())) return r;
  E.clients = n;
  std::vector<int> rcs(n, SMCF_OK);
  std::vector<std::string> errs(n);
  std::vector<std::thread> th;
  const int device = g->ctx->device;
  for (int i = 0; i < n; ++i)
    th.emplace_back([&, i] {
      (void)hipSetDevice(device);
      // every wave in one call: the engine streams it through the member's slots
      rcs[i] = smcf_fault_smc_exec(g->members[i], range[i], N[i], seed[i], max_stages, INT_MAX / 2, particles[i],
                                   likelihood[i], nstage ? &nstage[i] : nullptr, evidence ? &evidence[i] : nullptr,
                                   nullptr);
      if (rcs[i]) errs[i] = smcf_last_error();
      std::lock_guard<std::mutex> lk(E.mu);
      --E.clients;
      E.cv_engine.notify_all();
    });
  r = E.run();
  for (auto& t : th) t.join();
  if (r) return r;
  for (int i = 0; i < n; ++i)
    if (rcs[i]) {
      set_error(errs[i]);
      return rcs[i];
    }
  return SMCF_OK;
}

}  // extern "C"
