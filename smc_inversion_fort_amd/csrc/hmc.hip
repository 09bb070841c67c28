// Reference HMC mover on the device (smc_slip.f90:591-1214), batched over instances.
//
// The children of one parent form a sequential Metropolis chain; different parents are
// independent.  Work is therefore organised in waves: wave k holds the k-th child of every
// parent with more than k children (parents in ascending order, a compact column list per
// instance).  For each wave:
//   k_hmc_wave_init   candidate <- chain state, momenta p = z / sqrt(cov_diag), H(cur)
//   (ntau + 1) x [k_hs_mfma: Y = H X over the wave's columns ; k_hmc_step: gradient
//                 gamma (H x - b) + L^T (L x / alpha^2), kick, drift with escape and
//                 reflection]                                         (slip_leapfrog, 732-917)
//   k_hs_mfma + k_hmc_finish  likelihood, prior, Hamiltonian, ESJD score (tuning), MH test
// Tuning (parents 1, 6, 11, ... in 1-based numbering) runs first with log-uniform dtau per
// child, then k_tune_resample turns the scores into the dtau of every child and the new
// log-dtau range (slip_hmc_tuning, 919-1214); sampling then runs every parent (591-730).
//
// Mover SMCF_MOVER_HMC_ACCEPT (no reference counterpart) runs the same trajectories with ntau
// steps and one step size per instance: no tuning pass; after sampling the step size is adapted
// toward the target acceptance from the stage's mean acceptance probability (k_hmc_adapt).
// Y = H X runs on the FP64 matrix cores (k_hs_mfma).
#include "smcf_device.h"
#include "smcf_internal.h"

namespace smcf {

constexpr int kTuneFactor = 5;
constexpr int kNtauRef = 50;  // smc_slip.f90:669 (tuned ntau values are drawn but not used)
constexpr double kAdaptGain = 2.0;                       // mover 2: log dtau += gain * (abar - target)
constexpr double kLogDtau0 = -2.302585092994045684;      // mover 2: first stage's log dtau = log(0.1)

// per-instance wave lists: waves over all parents (mode 0) and over tuning parents (mode 1)
__global__ void __launch_bounds__(256) k_hmc_prep(int P, BatchDev bt, HmcDev h) {
  __shared__ int red[256];
  __shared__ int base;
  int b = blockIdx.x, t = threadIdx.x;
  if (!bt.active[b]) return;
  const int* a = bt.assigned + (size_t)b * P;
  for (int mode = 0; mode < 2; ++mode) {
    int* ptr = (mode ? h.tptr : h.wptr) + (size_t)b * (P + 1);
    int* cols = (mode ? h.tcols : h.wcols) + (size_t)b * P;
    // amax
    int am = 0;
    for (int i = t; i < P; i += 256)
      if (!mode || i % kTuneFactor == 0) am = max(am, a[i]);
    red[t] = am;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) red[t] = max(red[t], red[t + s]);
      __syncthreads();
    }
    int amax = red[0];
    __syncthreads();
    if (t == 0) {
      (mode ? h.tamax : h.wamax)[b] = amax;
      ptr[0] = 0;
      base = 0;
    }
    __syncthreads();
    for (int k = 0; k < amax; ++k) {
      // stable compaction of {i : a_i > k} in ascending i, 256 at a time
      for (int i0 = 0; i0 < P; i0 += 256) {
        int i = i0 + t;
        int f = (i < P && a[i] > k && (!mode || i % kTuneFactor == 0)) ? 1 : 0;
        red[t] = f;
        __syncthreads();
        for (int s = 1; s < 256; s <<= 1) {  // inclusive scan
          int v = t >= s ? red[t - s] : 0;
          __syncthreads();
          red[t] += v;
          __syncthreads();
        }
        if (f) cols[base + red[t] - 1] = i;
        __syncthreads();
        if (t == 255) base += red[255];
        __syncthreads();
      }
      if (t == 0) ptr[k + 1] = base;
      __syncthreads();
    }
  }
}

// chain state <- parents (source buffer); tuning children get their dtau draws
__global__ void k_hmc_chain_init(int nd, int P, BatchDev bt, HmcDev h, int tune) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (!bt.active[b] || i >= P) return;
  int cur = bt.par[b];
  const double* Ss = (cur ? bt.S : bt.S2) + (size_t)b * nd * P;
  const double* Ls = (cur ? bt.like : bt.like2) + (size_t)b * P;
  const double* Rs = (cur ? bt.prior : bt.prior2) + (size_t)b * P;
  double* Xc = h.Xc + (size_t)b * nd * P;
  for (int d = 0; d < nd; ++d) Xc[(size_t)d * P + i] = Ss[(size_t)d * P + i];
  double g = bt.gamma[b];
  h.Lc[(size_t)b * P + i] = Ls[i];
  h.Rc[(size_t)b * P + i] = Rs[i];
  h.postc[(size_t)b * P + i] = g * Ls[i] + Rs[i];
  if (tune && i % kTuneFactor == 0) {
    // dtau draws of this parent's children (smc_slip.f90:979-992)
    const int* a = bt.assigned + (size_t)b * P;
    int j0 = h.idstart[(size_t)b * P + i];
    double lo = h.ldlo[b], hi = h.ldhi[b];
    uint32_t stage = (uint32_t)(bt.iter[b] - 1);
    for (int j = j0; j < j0 + a[i]; ++j) {
      uint64_t w0, w1;
      philox_draw(bt.key[b], kSiteTuneDtau, stage, (uint32_t)j, 0, 0, w0, w1);
      double ld = lo + clamp_u(u01(w1)) * (hi - lo);
      h.dtau[(size_t)b * P + j] = exp(ld);
    }
  }
}

__global__ void k_hmc_wave_init(int nd, int P, int k, BatchDev bt, HmcDev h, int tune) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (!bt.active[b]) return;
  const int* ptr = (tune ? h.tptr : h.wptr) + (size_t)b * (P + 1);
  const int* cols = (tune ? h.tcols : h.wcols) + (size_t)b * P;
  int amax = (tune ? h.tamax : h.wamax)[b];
  int n = k < amax ? ptr[k + 1] - ptr[k] : 0;
  if (m == 0) h.ncol[b] = n;
  if (m >= n) return;
  int i = cols[ptr[k] + m];
  int j = h.idstart[(size_t)b * P + i] + k;  // child slot j = id_start(i) + k
  h.cpar[(size_t)b * P + m] = i;
  h.cchild[(size_t)b * P + m] = j;
  h.dead[(size_t)b * P + m] = 0;
  const double* Xc = h.Xc + (size_t)b * nd * P;
  double* Xt = h.Xt + (size_t)b * nd * P;
  double* Pm = h.Pm + (size_t)b * nd * P;
  const double* cov = bt.cov + (size_t)b * nd;
  uint32_t stage = (uint32_t)(bt.iter[b] - 1);
  uint32_t site = tune ? kSiteTuneMom : kSiteHmcMom;
  double ham = h.postc[(size_t)b * P + i];
  for (int d = 0; d < nd; ++d) {
    Xt[(size_t)d * P + m] = Xc[(size_t)d * P + i];
    double p = draw_normal(bt.key[b], site, stage, (uint32_t)j, (uint32_t)d) / sqrt(cov[d]);
    Pm[(size_t)d * P + m] = p;
    ham = ham + p * cov[d] * p / 2.0;
  }
  h.hcur[(size_t)b * P + m] = ham;
}

__global__ void k_idstart(int P, BatchDev bt, HmcDev h, int B) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || !bt.active[b]) return;
  const int* a = bt.assigned + (size_t)b * P;
  int* s = h.idstart + (size_t)b * P;
  int acc = 0;
  for (int i = 0; i < P; ++i) {
    s[i] = acc;
    acc += a[i];
  }
}

// Y(:, m) = H X(:, m) over the first ncol[b] columns, on the FP64 matrix cores.  One wavefront computes a
// 64-row x 64-column block (4 x 4 tiles of v_mfma_f64_16x16x4f64, operands streamed by mfma_panel_t:
// A(r, k) = H(i0 + r, k) column-major, B(k, c) = X(k, p0 + c) dimension-major); a workgroup's 4 wavefronts take
// 4 consecutive row blocks of the same 64 columns.  Lane l holds D(4t + l/16, l%16) in register t.
constexpr int kHsMT = 4;
__global__ void __launch_bounds__(256) k_hs_mfma(int nd, int P, const double* __restrict__ Hall,
                                                 const double* __restrict__ Xall, const int* __restrict__ ncol,
                                                 const int* __restrict__ active, double* __restrict__ Yall) {
  constexpr int MT = kHsMT, KB = 16 * MT;
  int b = blockIdx.z;
  if (!active[b]) return;
  int n = ncol[b];
  int p0 = blockIdx.x * 64;
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int i0 = (blockIdx.y * 4 + wave) * KB;
  if (p0 >= n || i0 >= nd) return;
  const double* H = Hall + (size_t)b * nd * nd;
  const double* X = Xall + (size_t)b * nd * P;
  double* Y = Yall + (size_t)b * nd * P;
  int fi = lane % 16, fk = lane / 16;
  int pcol[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) pcol[q] = min(p0 + 16 * q + fi, n - 1);
  uint32_t bo[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) bo[q] = (uint32_t)(fk * P + pcol[q]);
  uint32_t ao[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) ao[m] = (uint32_t)(fk * nd + min(i0 + 16 * m + fi, nd - 1));
  dbl4 acc[MT][4];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[m][q] = dbl4{0.0, 0.0, 0.0, 0.0};
  int nd4 = nd & ~3;
  mfma_panel_t<MT, kPanelBuffers>(H, ao, (uint32_t)(4 * nd), X, bo, (uint32_t)(4 * P), nd4 / 4, acc);
  if (nd4 < nd) {  // last contraction rows (nd = 2 mod 4): rows past nd contribute zero
    bool in = nd4 + fk < nd;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double bn = in ? X[(size_t)(nd4 + fk) * P + pcol[q]] : 0.0;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        int r = i0 + 16 * m + fi;
        double av = (in && r < nd) ? H[(size_t)(nd4 + fk) * nd + r] : 0.0;
        acc[m][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bn, acc[m][q], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      int row = i0 + 16 * m + 4 * t + fk;
      if (row >= nd) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int col = p0 + 16 * q + fi;
        if (col < n) Y[(size_t)row * P + col] = acc[m][q][t];
      }
    }
}

static void launch_hs(int nd, int P, int B, const double* H, const double* X, const int* ncol, const int* active,
                      double* Y, hipStream_t s) {
  int nblk = (nd + 16 * kHsMT - 1) / (16 * kHsMT);
  dim3 g((P + 63) / 64, (nblk + 3) / 4, B);
  hipLaunchKernelGGL(k_hs_mfma, g, dim3(256), 0, s, nd, P, H, X, ncol, active, Y);
}

// one leapfrog gradient + kick + drift (slip_leapfrog, smc_slip.f90:756-904)
//   s = 0: half kick, full drift; 0 < s < ntau: full kick, full drift; s = ntau: full kick, half drift
__global__ void k_hmc_step(MeshDev m, int P, int s, int ntau, double max_slip, BatchDev bt, HmcDev h, int tune) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (!bt.active[b] || col >= h.ncol[b]) return;
  size_t cb = (size_t)b * P + col;
  if (h.dead[cb]) return;
  int nd = m.ndim, nn2 = 2 * m.nnode;
  const double* Y = bt.Y + (size_t)b * nd * P;
  double* Xt = h.Xt + (size_t)b * nd * P;
  double* Pm = h.Pm + (size_t)b * nd * P;
  double* LS = h.LS + (size_t)b * nn2 * P;
  const double* bv = bt.bvec + (size_t)b * nd;
  const double* cov = bt.cov + (size_t)b * nd;
  const int* li = bt.li + (size_t)b * nn2 * 9;
  const double* lv = bt.lv + (size_t)b * nn2 * 9;
  const int* lti = bt.lti + (size_t)b * nd * 9;
  const double* ltv = bt.ltv + (size_t)b * nd * 9;
  double gamma = bt.gamma[b], a2 = bt.scal[8 * (size_t)b];
  double dtau = h.dtau[(size_t)b * P + h.cchild[cb]];
  for (int r = 0; r < nn2; ++r) {
    double tt = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) tt = tt + lv[9 * r + q] * Xt[(size_t)(li[9 * r + q] - 1) * P + col];
    LS[(size_t)r * P + col] = tt / a2;
  }
  for (int d = 0; d < nd; ++d) {
    double g1 = Y[(size_t)d * P + col];
    g1 = g1 - bv[d];
    g1 = g1 * gamma;
    double g2 = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) g2 = g2 + ltv[9 * d + q] * LS[(size_t)(lti[9 * d + q] - 1) * P + col];
    double g = g1 + g2;
    double p = Pm[(size_t)d * P + col];
    p = (s == 0) ? p - g * 5e-1 * dtau : p - g * dtau;
    Pm[(size_t)d * P + col] = p;
  }
  const double* Xc = h.Xc + (size_t)b * nd * P;
  int ip = h.cpar[cb];
  for (int d = 0; d < nd; ++d) {
    double x = Xt[(size_t)d * P + col], p = Pm[(size_t)d * P + col];
    x = (s < ntau) ? x + cov[d] * p * dtau : x + cov[d] * p * 5e-1 * dtau;
    if (fabs(x) > 1e3) {
      for (int e = 0; e < nd; ++e) Xt[(size_t)e * P + col] = Xc[(size_t)e * P + ip];
      h.dead[cb] = 1;
      return;
    }
    while (x < 0.0 || x > max_slip) {
      if (x > max_slip) { x = 2.0 * max_slip - x; p = -p; }
      if (x < 0.0) { x = -x; p = -p; }
    }
    Xt[(size_t)d * P + col] = x;
    Pm[(size_t)d * P + col] = p;
  }
}

// final Hamiltonian, ESJD score (tuning), Metropolis test, chain / child update
__global__ void k_hmc_finish(MeshDev m, int P, BatchDev bt, HmcDev h, int tune, int record_acc) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (!bt.active[b] || col >= h.ncol[b]) return;
  size_t cb = (size_t)b * P + col;
  int nd = m.ndim, nn2 = 2 * m.nnode;
  int i = h.cpar[cb], j = h.cchild[cb];
  double* Xc = h.Xc + (size_t)b * nd * P;
  const double* Xt = h.Xt + (size_t)b * nd * P;
  const double* Pm = h.Pm + (size_t)b * nd * P;
  const double* Y = bt.Y + (size_t)b * nd * P;
  const double* cov = bt.cov + (size_t)b * nd;
  const double* bv = bt.bvec + (size_t)b * nd;
  const double* sc = bt.scal + 8 * (size_t)b;
  double gamma = bt.gamma[b];
  double hc = h.hcur[cb], hn, lk = 0.0, pr = 0.0, post = 0.0;
  if (h.dead[cb]) {
    hn = 1e20;
  } else {
    double v = 0.0, bs = 0.0;
    for (int d = 0; d < nd; ++d) {
      double x = Xt[(size_t)d * P + col];
      v = v + x * Y[(size_t)d * P + col];
      bs = bs + bv[d] * x;
    }
    lk = (v - 2.0 * bs + bt.dsd[b]) / 2.0 + sc[3] + sc[4];
    const int* li = bt.li + (size_t)b * nn2 * 9;
    const double* lv = bt.lv + (size_t)b * nn2 * 9;
    double two_a2 = 2.0 * sc[0];
    for (int r = 0; r < nn2; ++r) {
      double tt = 0.0;
#pragma unroll
      for (int q = 0; q < 9; ++q) tt = tt + lv[9 * r + q] * Xt[(size_t)(li[9 * r + q] - 1) * P + col];
      pr = pr + (tt * tt) / two_a2;
    }
    post = gamma * lk + pr;
    hn = post;
    for (int d = 0; d < nd; ++d) {
      double p = Pm[(size_t)d * P + col];
      hn = hn + p * cov[d] * p / 2.0;
    }
  }
  uint32_t stage = (uint32_t)(bt.iter[b] - 1);
  if (record_acc) h.score[(size_t)b * P + j] = exp(fmin(0.0, hc - hn));  // acceptance probability (mover 2)
  if (tune) {
    double score = 0.0;
    for (int d = 0; d < nd; ++d) {
      double df = Xt[(size_t)d * P + col] - Xc[(size_t)d * P + i];
      score = score + (df * df) / cov[d];
    }
    score = score * exp(fmin(0.0, hc - hn));
    h.score[(size_t)b * P + j] = score;
  }
  uint64_t w0, w1;
  philox_draw(bt.key[b], tune ? kSiteTuneMH : kSiteHmcMH, stage, (uint32_t)j, 0, 0, w0, w1);
  double mh = clamp_u(u01(w0));
  size_t ib = (size_t)b * P + i;
  if (hc - hn > log(mh)) {
    for (int d = 0; d < nd; ++d) Xc[(size_t)d * P + i] = Xt[(size_t)d * P + col];
    h.Lc[ib] = lk;
    h.Rc[ib] = pr;
    h.postc[ib] = post;
  }
  if (!tune) {
    int cur = bt.par[b];
    double* Sd = (cur ? bt.S2 : bt.S) + (size_t)b * nd * P;
    for (int d = 0; d < nd; ++d) Sd[(size_t)d * P + j] = Xc[(size_t)d * P + i];
    (cur ? bt.like2 : bt.like)[(size_t)b * P + j] = h.Lc[ib];
    (cur ? bt.prior2 : bt.prior)[(size_t)b * P + j] = h.Rc[ib];
  }
}

// scores -> weights -> systematic residual counts (real(16) running u as double-double
// prefix sums) -> dtau of every child; new log-dtau range from the real(16) mean
__global__ void k_tune_resample(int P, BatchDev bt, HmcDev h, int B) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || !bt.active[b]) return;
  const int* a = bt.assigned + (size_t)b * P;
  const int* ids = h.idstart + (size_t)b * P;
  double* score = h.score + (size_t)b * P;
  double* dtau = h.dtau + (size_t)b * P;
  double* dnew = h.dtmp + (size_t)b * P;
  int cnt = 0;
  double ss = 0.0;
  for (int i = 0; i < P; i += kTuneFactor)
    for (int j = ids[i]; j < ids[i] + a[i]; ++j) { ss = ss + score[j]; cnt++; }
  if (cnt == 0) return;  // smc_slip.f90:993-995: dtau unchanged
  if (ss < 1e-5) {
    ss = 0.0;
    for (int i = 0; i < P; i += kTuneFactor)
      for (int j = ids[i]; j < ids[i] + a[i]; ++j) { score[j] = 1.0; ss = ss + 1.0; }
  }
  uint64_t w0, w1;
  philox_draw(bt.key[b], kSiteTuneResample, (uint32_t)(bt.iter[b] - 1), 0, 0, 0, w0, w1);
  dd u = two_sum((double)(w0 >> 11) * 0x1.0p-53, (double)(w1 >> 11) * 0x1.0p-106);
  if (dd_lt_d(u, 1e-5)) u = dd{1e-5, 0.0};
  dd u0 = dd_div_d(u, (double)P);
  dd c = {0.0, 0.0};
  int Aprev = 0, out = 0;
  for (int i = 0; i < P; i += kTuneFactor)
    for (int j = ids[i]; j < ids[i] + a[i]; ++j) {
      double wj = score[j] / ss;
      c = dd_add_d(c, wj);
      int A = (int)dd_floor(dd_mul_d(dd_add(c, dd_neg(u0)), (double)P)) + 1;
      A = max(Aprev, min(P, A));
      for (int q = Aprev; q < A; ++q) dnew[out++] = dtau[j];
      Aprev = A;
    }
  if (out != P) h.flag[b] = 1;  // the reference stops here (smc_slip.f90:1156-1173)
  for (int q = 0; q < out; ++q) dtau[q] = dnew[q];
  dd mean = {0.0, 0.0};
  for (int q = 0; q < P; ++q) mean = dd_add_d(mean, dtau[q]);
  mean = dd_div_d(mean, (double)P);
  // log of the double-double mean to first order, then + / - log(5)
  dd lm = two_sum(log(mean.hi), mean.lo / mean.hi);
  double l5 = log(5.0);
  h.ldhi[b] = dd_add_d(lm, l5).hi;
  h.ldlo[b] = dd_add_d(lm, -l5).hi;
}

// reference mover: log-dtau range [-10, 0] and dtau = exp(-5) before the first tuning (uninitialised in the
// reference, smc_slip.f90:993-995); mover 2: ldlo = log dtau of the first stage, ldhi = last mean acceptance
__global__ void k_hmc_tune_init(HmcDev h, int B, int P, const int* __restrict__ sel, int accept_mover) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  if (sel) b = sel[b];
  h.ldhi[b] = 0.0;
  h.ldlo[b] = accept_mover ? kLogDtau0 : -1e1;
  h.flag[b] = 0;
  for (int q = 0; q < P; ++q) h.dtau[(size_t)b * P + q] = exp(-5.0);
}

hipError_t launch_hmc_init(BatchDev& bt, HmcDev& h, int B, hipStream_t s, const int* sel, int mover) {
  hipLaunchKernelGGL(k_hmc_tune_init, dim3((B + 63) / 64), dim3(64), 0, s, h, B, bt.P, sel,
                     mover == SMCF_MOVER_HMC_ACCEPT ? 1 : 0);
  return hipGetLastError();
}

// mover 2: every child of the stage moves with dtau = exp(log dtau) of its instance
__global__ void k_hmc_dtau_fill(int P, BatchDev bt, HmcDev h) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (!bt.active[b] || j >= P) return;
  h.dtau[(size_t)b * P + j] = exp(h.ldlo[b]);
}

// mover 2: mean acceptance probability of the stage's P children (summed in child order) and the step update
__global__ void k_hmc_adapt(int P, BatchDev bt, HmcDev h, int B, double target) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || !bt.active[b]) return;
  const double* a = h.score + (size_t)b * P;
  double sa = 0.0;
  for (int j = 0; j < P; ++j) sa = sa + a[j];
  double abar = sa / (double)P;
  h.ldhi[b] = abar;
  h.ldlo[b] = h.ldlo[b] + kAdaptGain * (abar - target);
}

static hipError_t run_waves(const MeshDev& m, BatchDev& bt, HmcDev& h, int B, int tune, int amax, double max_slip,
                            int ntau, int record_acc, hipStream_t s) {
  int P = bt.P, nd = m.ndim;
  dim3 gc((P + 127) / 128, B);
  for (int k = 0; k < amax; ++k) {
    hipLaunchKernelGGL(k_hmc_wave_init, gc, dim3(128), 0, s, nd, P, k, bt, h, tune);
    for (int st = 0; st <= ntau; ++st) {
      launch_hs(nd, P, B, bt.H, h.Xt, h.ncol, bt.active, bt.Y, s);
      hipLaunchKernelGGL(k_hmc_step, gc, dim3(128), 0, s, m, P, st, ntau, max_slip, bt, h, tune);
    }
    launch_hs(nd, P, B, bt.H, h.Xt, h.ncol, bt.active, bt.Y, s);
    hipLaunchKernelGGL(k_hmc_finish, gc, dim3(128), 0, s, m, P, bt, h, tune, record_acc);
  }
  return hipGetLastError();
}

hipError_t launch_hmc_stage(const MeshDev& m, BatchDev& bt, HmcDev& h, int B, double max_slip, const HmcCfg& cfg,
                            hipStream_t s, int* host_amax) {
  int P = bt.P;
  hipLaunchKernelGGL(k_idstart, dim3((B + 63) / 64), dim3(64), 0, s, P, bt, h, B);
  hipLaunchKernelGGL(k_hmc_prep, dim3(B), dim3(256), 0, s, P, bt, h);
  // wave counts for the host loop bounds
  hipError_t e = hipMemcpyAsync(host_amax, h.wamax, sizeof(int) * B, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(host_amax + B, h.tamax, sizeof(int) * B, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  int wmax = 0, tmax = 0;
  for (int b = 0; b < B; ++b) {
    wmax = max(wmax, host_amax[b]);
    tmax = max(tmax, host_amax[B + b]);
  }
  if (cfg.mover == SMCF_MOVER_HMC_ACCEPT) {
    hipLaunchKernelGGL(k_hmc_dtau_fill, dim3((P + 127) / 128, B), dim3(128), 0, s, P, bt, h);
    hipLaunchKernelGGL(k_hmc_chain_init, dim3((P + 127) / 128, B), dim3(128), 0, s, m.ndim, P, bt, h, 0);
    e = run_waves(m, bt, h, B, 0, wmax, max_slip, cfg.ntau, 1, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_hmc_adapt, dim3((B + 63) / 64), dim3(64), 0, s, P, bt, h, B, cfg.accept);
    return hipGetLastError();
  }
  // tuning
  hipLaunchKernelGGL(k_hmc_chain_init, dim3((P + 127) / 128, B), dim3(128), 0, s, m.ndim, P, bt, h, 1);
  e = run_waves(m, bt, h, B, 1, tmax, max_slip, kNtauRef, 0, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_tune_resample, dim3((B + 63) / 64), dim3(64), 0, s, P, bt, h, B);
  // sampling
  hipLaunchKernelGGL(k_hmc_chain_init, dim3((P + 127) / 128, B), dim3(128), 0, s, m.ndim, P, bt, h, 0);
  return run_waves(m, bt, h, B, 0, wmax, max_slip, kNtauRef, 0, s);
}

// ---------------------------------------------------------------------------
// Component kernels on one instance (slot 0) for the C-ABI counterparts of slip_calc_grad and slip_leapfrog
// ---------------------------------------------------------------------------
// gradient gamma (H x - b) + L^T (L x / alpha^2) of columns 0 .. n-1 of Xt (Y = H Xt already formed)
__global__ void k_grad_cols(MeshDev m, int P, int n, double gamma, BatchDev bt, HmcDev h, double* __restrict__ G) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n) return;
  int nd = m.ndim, nn2 = 2 * m.nnode;
  double a2 = bt.scal[0];
  for (int r = 0; r < nn2; ++r) {
    double tt = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) tt = tt + bt.lv[9 * r + q] * h.Xt[(size_t)(bt.li[9 * r + q] - 1) * P + col];
    h.LS[(size_t)r * P + col] = tt / a2;
  }
  for (int d = 0; d < nd; ++d) {
    double g1 = bt.Y[(size_t)d * P + col];
    g1 = g1 - bt.bvec[d];
    g1 = g1 * gamma;
    double g2 = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) g2 = g2 + bt.ltv[9 * d + q] * h.LS[(size_t)(bt.lti[9 * d + q] - 1) * P + col];
    G[(size_t)col * nd + d] = g1 + g2;
  }
}

hipError_t launch_grad_cols(const MeshDev& m, BatchDev& bt, HmcDev& h, int n, double gamma, double* G, hipStream_t s) {
  launch_hs(m.ndim, bt.P, 1, bt.H, h.Xt, h.ncol, bt.active, bt.Y, s);
  hipLaunchKernelGGL(k_grad_cols, dim3((n + 127) / 128), dim3(128), 0, s, m, bt.P, n, gamma, bt, h, G);
  return hipGetLastError();
}

// Raw-array forms (the reference's argument lists): particles dimension-major with leading dimension ld.
// prior  sum_r (L x)_r^2 / (2 alpha2_r)  (slip_calc_prior, smc_slip.f90:106-137)
__global__ void k_prior_raw(int nn2, int n, int ld, const double* __restrict__ X, const double* __restrict__ a2,
                            const int* __restrict__ li, const double* __restrict__ lv, double* __restrict__ out) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n) return;
  double pr = 0.0;
  for (int r = 0; r < nn2; ++r) {
    double tt = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) tt = tt + lv[9 * r + q] * X[(size_t)(li[9 * r + q] - 1) * ld + col];
    pr = pr + (tt * tt) / (2.0 * a2[r]);
  }
  out[col] = pr;
}

// gradient gamma (Y - b) + L^T (L x / alpha2) with Y = H x formed (slip_calc_grad, smc_slip.f90:139-182); LS is
// nn2 x ld scratch; G ndim x n, column per particle
__global__ void k_grad_raw(int nd, int nn2, int n, int ld, double gamma, const double* __restrict__ X,
                           const double* __restrict__ Y, const double* __restrict__ bv, const double* __restrict__ a2,
                           const int* __restrict__ li, const double* __restrict__ lv, const int* __restrict__ lti,
                           const double* __restrict__ ltv, double* __restrict__ LS, double* __restrict__ G) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n) return;
  for (int r = 0; r < nn2; ++r) {
    double tt = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) tt = tt + lv[9 * r + q] * X[(size_t)(li[9 * r + q] - 1) * ld + col];
    LS[(size_t)r * ld + col] = tt / a2[r];
  }
  for (int d = 0; d < nd; ++d) {
    double g1 = Y[(size_t)d * ld + col];
    g1 = g1 - bv[d];
    g1 = g1 * gamma;
    double g2 = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) g2 = g2 + ltv[9 * d + q] * LS[(size_t)(lti[9 * d + q] - 1) * ld + col];
    G[(size_t)col * nd + d] = g1 + g2;
  }
}

hipError_t launch_prior_raw(int nn2, int n, int ld, const double* X, const double* a2, const int* li, const double* lv,
                            double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_prior_raw, dim3((n + 127) / 128), dim3(128), 0, s, nn2, n, ld, X, a2, li, lv, out);
  return hipGetLastError();
}

// ncol_active: device ints {n, 1} (column count and active flag of the single instance)
hipError_t launch_grad_raw(int nd, int nn2, int n, int ld, double gamma, const double* H, const double* X,
                           const double* bv, const double* a2, const int* li, const double* lv, const int* lti,
                           const double* ltv, const int* ncol_active, double* Y, double* LS, double* G,
                           hipStream_t s) {
  launch_hs(nd, ld, 1, H, X, ncol_active, ncol_active + 1, Y, s);
  hipLaunchKernelGGL(k_grad_raw, dim3((n + 127) / 128), dim3(128), 0, s, nd, nn2, n, ld, gamma, X, Y, bv, a2, li, lv,
                     lti, ltv, LS, G);
  return hipGetLastError();
}

// trajectory set-up of n independent columns (column m is its own parent and child): H(cur) from the
// likelihood and prior of Xt (Y = H Xt formed) and the momenta in Pm
__global__ void k_leapfrog_begin(MeshDev m, int P, int n, double gamma, double dtau, BatchDev bt, HmcDev h) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n) return;
  int nd = m.ndim, nn2 = 2 * m.nnode;
  const double* sc = bt.scal;
  double v = 0.0, bs = 0.0;
  for (int d = 0; d < nd; ++d) {
    double x = h.Xt[(size_t)d * P + col];
    v = v + x * bt.Y[(size_t)d * P + col];
    bs = bs + bt.bvec[d] * x;
  }
  double lk = (v - 2.0 * bs + bt.dsd[0]) / 2.0 + sc[3] + sc[4];
  double pr = 0.0, two_a2 = 2.0 * sc[0];
  for (int r = 0; r < nn2; ++r) {
    double tt = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) tt = tt + bt.lv[9 * r + q] * h.Xt[(size_t)(bt.li[9 * r + q] - 1) * P + col];
    pr = pr + (tt * tt) / two_a2;
  }
  double ham = gamma * lk + pr;
  for (int d = 0; d < nd; ++d) {
    double p = h.Pm[(size_t)d * P + col];
    ham = ham + p * bt.cov[d] * p / 2.0;
    h.Xc[(size_t)d * P + col] = h.Xt[(size_t)d * P + col];
  }
  h.hcur[col] = ham;
  h.cpar[col] = col;
  h.cchild[col] = col;
  h.dead[col] = 0;
  h.dtau[col] = dtau;
  h.Lc[col] = lk;
  h.Rc[col] = pr;
}

// final Hamiltonian of the n trajectories (Y = H Xt formed): ham_cand (1e20 after an escape), the candidate and
// the momenta
__global__ void k_leapfrog_end(MeshDev m, int P, int n, double gamma, BatchDev bt, HmcDev h,
                               double* __restrict__ hn_out, double* __restrict__ lk_out, double* __restrict__ pr_out) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n) return;
  int nd = m.ndim, nn2 = 2 * m.nnode;
  if (h.dead[col]) {  // escaped: the candidate is the start point (smc_slip.f90:776-782)
    hn_out[col] = 1e20;
    lk_out[col] = h.Lc[col];
    pr_out[col] = h.Rc[col];
    return;
  }
  const double* sc = bt.scal;
  double v = 0.0, bs = 0.0;
  for (int d = 0; d < nd; ++d) {
    double x = h.Xt[(size_t)d * P + col];
    v = v + x * bt.Y[(size_t)d * P + col];
    bs = bs + bt.bvec[d] * x;
  }
  double lk = (v - 2.0 * bs + bt.dsd[0]) / 2.0 + sc[3] + sc[4];
  double pr = 0.0, two_a2 = 2.0 * sc[0];
  for (int r = 0; r < nn2; ++r) {
    double tt = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) tt = tt + bt.lv[9 * r + q] * h.Xt[(size_t)(bt.li[9 * r + q] - 1) * P + col];
    pr = pr + (tt * tt) / two_a2;
  }
  double hn = gamma * lk + pr;
  for (int d = 0; d < nd; ++d) {
    double p = h.Pm[(size_t)d * P + col];
    hn = hn + p * bt.cov[d] * p / 2.0;
  }
  hn_out[col] = hn;
  lk_out[col] = lk;
  pr_out[col] = pr;
}

hipError_t launch_leapfrog_cols(const MeshDev& m, BatchDev& bt, HmcDev& h, int n, double gamma, double dtau, int ntau,
                                double max_slip, double* hc_out, double* hn_out, double* lk_out, double* pr_out,
                                hipStream_t s) {
  int P = bt.P, nd = m.ndim;
  dim3 gc((n + 127) / 128);
  launch_hs(nd, P, 1, bt.H, h.Xt, h.ncol, bt.active, bt.Y, s);
  hipLaunchKernelGGL(k_leapfrog_begin, gc, dim3(128), 0, s, m, P, n, gamma, dtau, bt, h);
  for (int st = 0; st <= ntau; ++st) {
    launch_hs(nd, P, 1, bt.H, h.Xt, h.ncol, bt.active, bt.Y, s);
    hipLaunchKernelGGL(k_hmc_step, dim3((P + 127) / 128, 1), dim3(128), 0, s, m, P, st, ntau, max_slip, bt, h, 0);
  }
  launch_hs(nd, P, 1, bt.H, h.Xt, h.ncol, bt.active, bt.Y, s);
  hipLaunchKernelGGL(k_leapfrog_end, gc, dim3(128), 0, s, m, P, n, gamma, bt, h, hn_out, lk_out, pr_out);
  hipError_t e = hipMemcpyAsync(hc_out, h.hcur, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace smcf
