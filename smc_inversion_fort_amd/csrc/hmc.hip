// Reference HMC mover on the device (smc_slip.f90:591-1214), batched over instances.
//
// The children of one parent form a sequential Metropolis chain; different parents are
// independent.  Work is therefore organised in waves: wave k holds the k-th child of every
// parent with more than k children (parents in ascending order, a compact column list per
// instance).  Once per stage Q_g = gamma H + L^T L / alpha2 is formed (the Hessian of the tempered
// potential), so a leapfrog gradient is one matrix product.  For each wave:
//   k_hmc_wave_cols, k_hmc_wave_init, k_hmc_hcur   columns; candidate <- chain state, momenta
//                 p = z / sqrt(cov_diag) (element-parallel); H(cur)
//   (ntau + 1) x k_hmc_leap   G = Q_g X - gamma b on the FP64 matrix cores, kick and drift with
//                 escape and reflection fused into the product's epilogue        (slip_leapfrog, 732-917)
//   k_hs_work (Y = H X), k_hmc_sums (per-column partial sums), k_hmc_finish, k_hmc_commit   likelihood,
//                 prior, Hamiltonian, ESJD score (tuning) or acceptance probability, MH test, chain / child update
// Every per-(dimension, column) step is spread over threads: a thread per column looping over the
// dimensions waits one memory round trip per dimension.
// Tuning (parents 1, 6, 11, ... in 1-based numbering) runs first with log-uniform dtau per
// child, then k_tune_resample turns the scores into the dtau of every child and the new
// log-dtau range (slip_hmc_tuning, 919-1214); sampling then runs every parent (591-730).
//
// Mover SMCF_MOVER_HMC_ACCEPT (no reference counterpart) runs the same trajectories with ntau
// steps and one step size per instance: no tuning pass; after sampling the step size is adapted
// toward the target acceptance from the stage's mean acceptance probability (k_hmc_adapt).
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
  if (!bt.active[b]) {  // a finished instance has no waves (the host lists its work from these counts)
    if (t == 0) {
      h.wamax[b] = 0;
      h.tamax[b] = 0;
    }
    return;
  }
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

// the wave's columns: column m of wave k is the k-th child of the m-th listed parent
__global__ void k_hmc_wave_cols(int P, int k, BatchDev bt, HmcDev h, int tune) {
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
  size_t cb = (size_t)b * P + m;
  h.cpar[cb] = i;
  h.cchild[cb] = h.idstart[(size_t)b * P + i] + k;  // child slot j = id_start(i) + k
  h.dead[cb] = 0;
}

// Element-parallel kernels over (dimension, column) of a wave's listed 64-column tiles (items (instance, tile)
// built per stage on the host): blockIdx.x = item, threadIdx.x / 64 + 4 blockIdx.y strides the dimensions.  Sums
// over a column's dimensions are formed there as S = 4 gridDim.y partials (thread s adds dimensions s, s + S, ...
// in order) and the partials are added in order by the per-column kernel: fixed order, so a result does not depend
// on the batch it ran in.
constexpr int kRowGroups = 8;
constexpr int kSums = 6;  // partial sums per (column, s): x.Hx, b.x, prior rows, kinetic, tuning score, H(cur) kinetic

SMCF_DEV double* part_sum(const HmcDev& h, int P, int b, int s, int which, int col) {
  return h.part + (((size_t)b * 4 * kRowGroups + s) * kSums + which) * P + col;
}

SMCF_DEV bool wave_element(const int* __restrict__ work, const HmcDev& h, int& b, int& col) {
  b = work[2 * blockIdx.x];
  col = work[2 * blockIdx.x + 1] * 64 + threadIdx.x % 64;
  return col < h.ncol[b];
}

// candidate <- chain state and momenta p = z / sqrt(cov) (slip_hmc_sampling / slip_hmc_tuning before
// slip_leapfrog, smc_slip.f90:680-700, 1000-1020)
__global__ void __launch_bounds__(256) k_hmc_wave_init(int nd, int P, const int* __restrict__ work, BatchDev bt,
                                                       HmcDev h, int tune) {
  int b, m;
  if (!wave_element(work, h, b, m)) return;
  size_t cb = (size_t)b * P + m;
  int i = h.cpar[cb], j = h.cchild[cb];
  const double* Xc = h.Xc + (size_t)b * nd * P;
  double* Xt = h.Xt + (size_t)b * nd * P;
  double* Pm = h.Pm + (size_t)b * nd * P;
  const double* cov = bt.cov + (size_t)b * nd;
  uint32_t stage = (uint32_t)(bt.iter[b] - 1);
  uint32_t site = tune ? kSiteTuneMom : kSiteHmcMom;
  int s0 = blockIdx.y * 4 + threadIdx.x / 64;
  __bf16* XX = h.XXa ? (__bf16*)h.XXa + (size_t)b * hmc_xx_rows(nd) * P * 2 : nullptr;
  double kin = 0.0;
  for (int d = s0; d < nd; d += gridDim.y * 4) {
    double x = Xc[(size_t)d * P + i];
    Xt[(size_t)d * P + m] = x;
    if (XX) {  // the split copy read by the split-BF16 products: [d/8][column][hi 8 | lo 8]
      __bf16 hi, lo;
      split_bf16((float)x, hi, lo);
      size_t o = (((size_t)(d / 8) * P + m) * 2) * 8 + d % 8;
      XX[o] = hi;
      XX[o + 8] = lo;
    }
    double p = draw_normal(bt.key[b], site, stage, (uint32_t)j, (uint32_t)d) / sqrt(cov[d]);
    Pm[(size_t)d * P + m] = p;
    kin = kin + p * cov[d] * p / 2.0;
  }
  *part_sum(h, P, b, s0, 5, m) = kin;
}

// H(cur) = chain state's gamma L + prior, plus the S kinetic partials of k_hmc_wave_init in order
__global__ void k_hmc_hcur(int P, int S, BatchDev bt, HmcDev h) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (!bt.active[b] || m >= h.ncol[b]) return;
  size_t cb = (size_t)b * P + m;
  double kin = 0.0;
  for (int q = 0; q < S; ++q) kin = kin + *part_sum(h, P, b, q, 5, m);
  h.hcur[cb] = h.postc[(size_t)b * P + h.cpar[cb]] + kin;
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

// ---------------------------------------------------------------------------
// Products over the stage's work lists.  A wave's columns are listed per stage on the host as items (instance,
// column tile); a launch covers exactly the listed tiles.  (Grids over every (column tile, row group, instance)
// with early exits place the few live workgroups of a small wave on the same few compute units.)
//  * wide kernels (an instance with more than kNarrowCols columns in the wave): blockIdx.x = item * nrg + row
//    group, the 4 wavefronts of a workgroup take 4 consecutive 64-row blocks of one 64-column tile;
//  * narrow kernels (at most kNarrowCols columns): a workgroup forms one 64-row x 16-column tile with the
//    contraction split over its 4 wavefronts (a quarter of the per-wavefront chain of loads), the partials added
//    in LDS in wavefront order.
// The choice depends on the instance's own column count, so an instance's arithmetic does not depend on its batch.
// ---------------------------------------------------------------------------
constexpr int kNarrowCols = 48;  // an instance whose wave has at most this many columns takes the narrow kernels

SMCF_DEV int row_groups(int nd) { return ((nd + 16 * kHsMT - 1) / (16 * kHsMT) + 3) / 4; }

// acc[m][q] += A(16m + r, k) B(k, c_q) over the contraction, for MT row tiles and NT column tiles (the operand
// addressing and register pipeline of mfma_panel_t, for any NT)
template <int MT, int NT, int NB>
__device__ __forceinline__ void mfma_panel_nt(const double* __restrict__ abase, const uint32_t (&ao)[MT],
                                              uint32_t lda4, const double* __restrict__ bbase,
                                              const uint32_t (&bo)[NT], uint32_t ldb4, int nsteps,
                                              dbl4 (&acc)[MT][NT]) {
  if (nsteps <= 0) return;
  double x[NB][MT], y[NB][NT];
  auto load = [&](int k, int st) {
    uint32_t q = (uint32_t)min(st, nsteps - 1);
#pragma unroll
    for (int m = 0; m < MT; ++m) x[k][m] = abase[ao[m] + q * lda4];
#pragma unroll
    for (int n = 0; n < NT; ++n) y[k][n] = bbase[bo[n] + q * ldb4];
  };
  auto prod = [&](int k) {
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int m = 0; m < MT; ++m)
        acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[k][m], y[k][n], acc[m][n], 0, 0, 0);
  };
#pragma unroll
  for (int k = 0; k < NB; ++k) load(k, k);
  int full = nsteps - nsteps % NB;
  for (int st = 0; st < full; st += NB) {
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      prod(k);
      load(k, st + k + NB);
    }
  }
#pragma unroll
  for (int k = 0; k < NB - 1; ++k)
    if (k < nsteps - full) prod(k);
}

// rows i0 .. i0 + 63 x columns p0 .. p0 + 16 NT - 1 of A X (A symmetric nd x nd, X dimension-major with leading
// dimension P, columns past n - 1 read column n - 1); lane l holds D(4t + l/16, l%16) of tile (m, q) in acc[m][q][t]
// register buffers of the operand pipeline: deep for the narrow tile (a small wave's launch waits on one
// wavefront's chain of loads), the shared default for the wide one (its registers go to accumulators)
template <int NT>
struct TileDepth {
  static constexpr int value = NT == 1 ? 8 : kPanelBuffers;
};

template <int NT>
__device__ __forceinline__ void ax_tile(int nd, int P, int n, const double* __restrict__ A,
                                        const double* __restrict__ X, int p0, int i0, int lane,
                                        dbl4 (&acc)[kHsMT][NT]) {
  constexpr int MT = kHsMT;
  int fi = lane % 16, fk = lane / 16;
  int pcol[NT];  // n >= 1: the callers return for tiles at or past column n
#pragma unroll
  for (int q = 0; q < NT; ++q) pcol[q] = min(p0 + 16 * q + fi, n - 1);
  uint32_t bo[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) bo[q] = (uint32_t)(fk * P + pcol[q]);
  uint32_t ao[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) ao[m] = (uint32_t)(fk * nd + min(i0 + 16 * m + fi, nd - 1));
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int q = 0; q < NT; ++q) acc[m][q] = dbl4{0.0, 0.0, 0.0, 0.0};
  int nd4 = nd & ~3;
  mfma_panel_nt<MT, NT, TileDepth<NT>::value>(A, ao, (uint32_t)(4 * nd), X, bo, (uint32_t)(4 * P), nd4 / 4, acc);
  if (nd4 < nd) {  // last contraction rows (nd = 2 mod 4): rows past nd contribute zero
    bool in = nd4 + fk < nd;
#pragma unroll
    for (int q = 0; q < NT; ++q) {
      double bn = in ? X[(size_t)(nd4 + fk) * P + pcol[q]] : 0.0;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        int r = i0 + 16 * m + fi;
        double av = (in && r < nd) ? A[(size_t)(nd4 + fk) * nd + r] : 0.0;
        acc[m][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bn, acc[m][q], 0, 0, 0);
      }
    }
  }
}

// Y = H X over the listed tiles of one wave (the final positions' likelihood products)
template <int NT>
__global__ void __launch_bounds__(256) k_hs_work(int nd, int P, const double* __restrict__ Hall,
                                                 const double* __restrict__ Xall, const int* __restrict__ work,
                                                 HmcDev h, double* __restrict__ Yall) {
  int nrg = row_groups(nd);
  int item = blockIdx.x / nrg, rg = blockIdx.x % nrg;
  int b = work[2 * item], p0 = work[2 * item + 1] * 16 * NT;
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int i0 = (rg * 4 + wave) * 16 * kHsMT;
  int n = h.ncol[b];
  if (i0 >= nd || p0 >= n) return;
  dbl4 acc[kHsMT][NT];
  ax_tile<NT>(nd, P, n, Hall + (size_t)b * nd * nd, Xall + (size_t)b * nd * P, p0, i0, lane, acc);
  double* Y = Yall + (size_t)b * nd * P;
  int fi = lane % 16, fk = lane / 16;
#pragma unroll
  for (int m = 0; m < kHsMT; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      int row = i0 + 16 * m + 4 * t + fk;
      if (row >= nd) continue;
#pragma unroll
      for (int q = 0; q < NT; ++q) {
        int col = p0 + 16 * q + fi;
        if (col < n) Y[(size_t)row * P + col] = acc[m][q][t];
      }
    }
}

// Q_g = gamma H + (L^T L) / alpha2 of every active instance, column-major: the Hessian of gamma * likelihood +
// prior, so that Q_g x - gamma b is the gradient of slip_calc_grad (smc_slip.f90:139-182).  The dense gamma H,
// then the <= 25 entries of each column of L^T L, in the expression of the Gibbs precision.
__global__ void k_hmc_q_dense(int nd, BatchDev bt, HmcDev h) {
  int b = blockIdx.y;
  if (!bt.active[b]) return;
  size_t n = (size_t)nd * nd;
  double g = bt.gamma[b];
  const double* H = bt.H + (size_t)b * n;
  double* Q = h.Qg + (size_t)b * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    Q[e] = g * H[e];
}

__global__ void k_hmc_q_sparse(int nd, BatchDev bt, HmcDev h) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (!bt.active[b] || i >= nd) return;
  double g = bt.gamma[b], a2 = bt.scal[8 * (size_t)b];
  const double* H = bt.H + (size_t)b * nd * nd;
  double* Q = h.Qg + (size_t)b * nd * nd;
  const int* ix = bt.llc_idx + ((size_t)b * nd + i) * 25;
  const double* vx = bt.llc_val + ((size_t)b * nd + i) * 25;
  for (int w = 0; w < 25; ++w) {
    int r = ix[w];
    if (r < 0) break;
    Q[(size_t)i * nd + r] = g * H[(size_t)i * nd + r] + vx[w] / a2;
  }
}

static void launch_form_qg(int nd, BatchDev& bt, HmcDev& h, int B, hipStream_t s) {
  size_t n = (size_t)nd * nd;
  int blocks = (int)std::min<size_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(k_hmc_q_dense, dim3(blocks, B), dim3(256), 0, s, nd, bt, h);
  hipLaunchKernelGGL(k_hmc_q_sparse, dim3((nd + 127) / 128, B), dim3(128), 0, s, nd, bt, h);
}

// One leapfrog step of the wave's columns (slip_leapfrog, smc_slip.f90:756-904) fused with its gradient:
//   G = Q_g X_in - gamma b on the FP64 matrix cores (a listed tile per workgroup, as k_hs_work), then each lane
//   kicks and drifts the elements of G it holds:
//   s = 0: half kick, full drift; 0 < s < ntau: full kick, full drift; s = ntau: full kick, half drift.
// Momenta are updated in place (each element has one owner); positions go from X_in to X_out, since other
// workgroups still read X_in for their products.  A position past |x| > 1e3 ends its column's trajectory (dead;
// the candidate is then the start point, smc_slip.f90:776-782); dead columns are not updated.
template <int NT>
__global__ void __launch_bounds__(256) k_hmc_leap(int nd, int P, int s, int ntau, double max_slip,
                                                  const double* __restrict__ Xin_all, double* __restrict__ Xout_all,
                                                  const int* __restrict__ work, BatchDev bt, HmcDev h) {
  int nrg = row_groups(nd);
  int item = blockIdx.x / nrg, rg = blockIdx.x % nrg;
  int b = work[2 * item], p0 = work[2 * item + 1] * 16 * NT;
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int i0 = (rg * 4 + wave) * 16 * kHsMT;
  int n = h.ncol[b];
  if (i0 >= nd || p0 >= n) return;
  const double* X = Xin_all + (size_t)b * nd * P;
  dbl4 acc[kHsMT][NT];
  ax_tile<NT>(nd, P, n, h.Qg + (size_t)b * nd * nd, X, p0, i0, lane, acc);
  double* Xo = Xout_all + (size_t)b * nd * P;
  double* Pm = h.Pm + (size_t)b * nd * P;
  const double* bv = bt.bvec + (size_t)b * nd;
  const double* cov = bt.cov + (size_t)b * nd;
  double gam = bt.gamma[b];
  int fi = lane % 16, fk = lane / 16;
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    int col = p0 + 16 * q + fi;
    if (col >= n) continue;
    size_t cb = (size_t)b * P + col;
    if (h.dead[cb]) continue;
    double dt = h.dtau[(size_t)b * P + h.cchild[cb]];
#pragma unroll
    for (int m = 0; m < kHsMT; ++m)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        int row = i0 + 16 * m + 4 * t + fk;
        if (row >= nd) continue;
        size_t e = (size_t)row * P + col;
        double g = acc[m][q][t] - gam * bv[row];
        double p = Pm[e];
        p = (s == 0) ? p - g * 5e-1 * dt : p - g * dt;
        double x = X[e];
        x = (s < ntau) ? x + cov[row] * p * dt : x + cov[row] * p * 5e-1 * dt;
        if (fabs(x) > 1e3) {
          h.dead[cb] = 1;
          continue;
        }
        while (x < 0.0 || x > max_slip) {
          if (x > max_slip) { x = 2.0 * max_slip - x; p = -p; }
          if (x < 0.0) { x = -x; p = -p; }
        }
        Xo[e] = x;
        Pm[e] = p;
      }
  }
}

// ---------------------------------------------------------------------------
// Split-BF16 leapfrog gradients (SMCF_PRECISION_MIXED_BF16 with an HMC mover; no reference counterpart).  The wide
// leapfrog product G = Q_g X is formed from split operands on the BF16 matrix cores, Q_g X ~ Qhi Xhi + Qhi Xlo +
// Qlo Xhi with FP32 accumulation (about 2^-16 relative error per term), and the kick and drift follow in FP64.
// The trajectory then follows a slightly different force, but every leapfrog map is still reversible and volume
// preserving (kicks depend on position only), and the Metropolis test uses the FP64 Hamiltonian (H x of the final
// position in FP64, k_hs_work): the sampler stays exact, only its acceptance can change.  Small waves (the narrow
// kernel) stay FP64.
// ---------------------------------------------------------------------------
// Q_g in fragment order for v_mfma_f32_16x16x32_bf16: fragment f = (rt nks + st) 64 + l of lane l = r + 16g holds
// Q_g(16 rt + r, 32 st + 8g + e), e = 0..7 (zero past ndim), hi at 2f and lo at 2f + 1
__global__ void k_qg_to_bf(int nd, BatchDev bt, HmcDev h) {
  int b = blockIdx.y;
  if (!bt.active[b]) return;
  int nrt = (nd + 15) / 16, nks = (nd + 31) / 32;
  size_t nf = (size_t)nrt * nks * 64;
  const double* Q = h.Qg + (size_t)b * nd * nd;
  bf16x8* X = (bf16x8*)h.QgX + (size_t)b * nf * 2;
  for (size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += (size_t)gridDim.x * blockDim.x) {
    int l = (int)(f % 64);
    size_t rest = f / 64;
    int st = (int)(rest % nks), rt = (int)(rest / nks);
    int r = 16 * rt + l % 16, k0 = 32 * st + 8 * (l / 16);
    bf16x8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      double v = (r < nd && k0 + e < nd) ? Q[(size_t)(k0 + e) * nd + r] : 0.0;
      __bf16 a, c;
      split_bf16((float)v, a, c);
      hi[e] = a;
      lo[e] = c;
    }
    X[2 * f] = hi;
    X[2 * f + 1] = lo;
  }
}

// the wide leapfrog step with split-BF16 products: tiles, items and epilogue as k_hmc_leap<4>; the positions' split
// copy XX_in gives the B fragments (one 16-byte load per part), the epilogue writes the new positions to X_out and
// their split copy to XX_out; lane l holds D(4 (l/16) + t, l%16) of a tile in register t (the FP32-accumulating
// instruction's layout)
__global__ void __launch_bounds__(256) k_hmc_leap_bf(int nd, int P, int s, int ntau, double max_slip,
                                                     const double* __restrict__ Xin_all, double* __restrict__ Xout_all,
                                                     const bf16x8* __restrict__ XXin_all, __bf16* __restrict__ XXout_all,
                                                     const int* __restrict__ work, BatchDev bt, HmcDev h) {
  constexpr int MT = kHsMT, NT = 4;
  int nrg = row_groups(nd);
  int item = blockIdx.x / nrg, rg = blockIdx.x % nrg;
  int b = work[2 * item], p0 = work[2 * item + 1] * 16 * NT;
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int i0 = (rg * 4 + wave) * 16 * MT;
  int n = h.ncol[b];
  if (i0 >= nd || p0 >= n) return;
  int nrt = (nd + 15) / 16, nks = (nd + 31) / 32, rt0 = i0 / 16;
  size_t xrows = hmc_xx_rows(nd);
  const bf16x8* QX = (const bf16x8*)h.QgX + (size_t)b * nrt * nks * 64 * 2;
  const bf16x8* XXin = XXin_all + (size_t)b * (xrows / 8) * P * 2;
  const double* X = Xin_all + (size_t)b * nd * P;
  int fi = lane % 16, fk = lane / 16;
  int pcol[NT];
#pragma unroll
  for (int q = 0; q < NT; ++q) pcol[q] = min(p0 + 16 * q + fi, n - 1);
  flt4 acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int q = 0; q < NT; ++q) acc[m][q] = flt4{0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = {};
  bf16x8 ah[2][MT], al[2][MT], bh[2][NT], bl[2][NT];
  auto load = [&](int k, int st) {
    st = min(st, nks - 1);  // a prefetch past the end re-reads the last step, unused
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      bool in = rt0 + m < nrt;  // row tiles past the last one (wave-uniform)
      size_t f = ((size_t)(rt0 + m) * nks + st) * 64 + lane;
      ah[k][m] = in ? QX[2 * f] : zero8;
      al[k][m] = in ? QX[2 * f + 1] : zero8;
    }
#pragma unroll
    for (int q = 0; q < NT; ++q) {
      size_t o = ((size_t)(4 * st + fk) * P + pcol[q]) * 2;
      bh[k][q] = XXin[o];
      bl[k][q] = XXin[o + 1];
    }
  };
  auto prod = [&](int k) {
#pragma unroll
    for (int q = 0; q < NT; ++q)
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        acc[m][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[k][m], bh[k][q], acc[m][q], 0, 0, 0);
        acc[m][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[k][m], bl[k][q], acc[m][q], 0, 0, 0);
        acc[m][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[k][m], bh[k][q], acc[m][q], 0, 0, 0);
      }
  };
  load(0, 0);
  load(1, 1);
  int full = nks - nks % 2;
  for (int st = 0; st < full; st += 2) {
    prod(0);
    load(0, st + 2);
    prod(1);
    load(1, st + 3);
  }
  if (nks % 2) prod(0);
  double* Xo = Xout_all + (size_t)b * nd * P;
  __bf16* XXo = XXout_all + (size_t)b * xrows * P * 2;
  double* Pm = h.Pm + (size_t)b * nd * P;
  const double* bv = bt.bvec + (size_t)b * nd;
  const double* cov = bt.cov + (size_t)b * nd;
  double gam = bt.gamma[b];
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    int col = p0 + 16 * q + fi;
    if (col >= n) continue;
    size_t cb = (size_t)b * P + col;
    if (h.dead[cb]) continue;
    double dt = h.dtau[(size_t)b * P + h.cchild[cb]];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        int row = i0 + 16 * m + 4 * fk + t;
        if (row >= nd) continue;
        size_t e = (size_t)row * P + col;
        double g = (double)acc[m][q][t] - gam * bv[row];
        double p = Pm[e];
        p = (s == 0) ? p - g * 5e-1 * dt : p - g * dt;
        double x = X[e];
        x = (s < ntau) ? x + cov[row] * p * dt : x + cov[row] * p * 5e-1 * dt;
        if (fabs(x) > 1e3) {
          h.dead[cb] = 1;
          continue;
        }
        while (x < 0.0 || x > max_slip) {
          if (x > max_slip) { x = 2.0 * max_slip - x; p = -p; }
          if (x < 0.0) { x = -x; p = -p; }
        }
        Xo[e] = x;
        Pm[e] = p;
        __bf16 hi, lo;
        split_bf16((float)x, hi, lo);
        size_t o = (((size_t)(row / 8) * P + col) * 2) * 8 + row % 8;
        XXo[o] = hi;
        XXo[o + 8] = lo;
      }
  }
}

// The narrow tile: rows i0 .. i0 + 63, columns p0 .. p0 + 15 of A X with the contraction split over the 4
// wavefronts, then either Y (LEAP = false) or the leapfrog kick and drift of k_hmc_leap (LEAP = true).
template <bool LEAP>
__global__ void __launch_bounds__(256) k_narrow(int nd, int P, int s, int ntau, double max_slip,
                                                const double* __restrict__ Aall, const double* __restrict__ Xin_all,
                                                double* __restrict__ Out_all, const int* __restrict__ work,
                                                BatchDev bt, HmcDev h) {
  constexpr int MT = kHsMT;
  __shared__ double red[4][MT][4][64];
  int nrb = (nd + 16 * MT - 1) / (16 * MT);
  int item = blockIdx.x / nrb, i0 = (blockIdx.x % nrb) * 16 * MT;
  int b = work[2 * item], p0 = work[2 * item + 1] * 16;
  int n = h.ncol[b];
  if (p0 >= n) return;  // the whole workgroup (before its barrier)
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int fi = lane % 16, fk = lane / 16;
  const double* A = Aall + (size_t)b * nd * nd;
  const double* X = Xin_all + (size_t)b * nd * P;
  int pcol = min(p0 + fi, n - 1);
  uint32_t bo[1] = {(uint32_t)(fk * P + pcol)};
  uint32_t ao[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) ao[m] = (uint32_t)(fk * nd + min(i0 + 16 * m + fi, nd - 1));
  dbl4 acc[MT][1];
#pragma unroll
  for (int m = 0; m < MT; ++m) acc[m][0] = dbl4{0.0, 0.0, 0.0, 0.0};
  int nd4 = nd & ~3, nsteps = nd4 / 4, q = (nsteps + 3) / 4;
  int st0 = min(nsteps, wave * q), st1 = min(nsteps, st0 + q);
  mfma_panel_nt<MT, 1, TileDepth<1>::value>(A + (size_t)st0 * 4 * nd, ao, (uint32_t)(4 * nd), X + (size_t)st0 * 4 * P,
                                            bo, (uint32_t)(4 * P), st1 - st0, acc);
  if (wave == 3 && nd4 < nd) {  // last contraction rows (nd = 2 mod 4): rows past nd contribute zero
    bool in = nd4 + fk < nd;
    double bn = in ? X[(size_t)(nd4 + fk) * P + pcol] : 0.0;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      int r = i0 + 16 * m + fi;
      double av = (in && r < nd) ? A[(size_t)(nd4 + fk) * nd + r] : 0.0;
      acc[m][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bn, acc[m][0], 0, 0, 0);
    }
  }
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) red[wave][m][t][lane] = acc[m][0][t];
  __syncthreads();
  // wavefront w finishes row tile m = w
  int m = wave, col = p0 + fi;
  if (col >= n) return;
  size_t cb = (size_t)b * P + col;
  if (LEAP && h.dead[cb]) return;
  double* Out = Out_all + (size_t)b * nd * P;
  double* Pm = h.Pm + (size_t)b * nd * P;
  const double* bv = bt.bvec + (size_t)b * nd;
  const double* cov = bt.cov + (size_t)b * nd;
  double gam = bt.gamma[b];
  double dt = LEAP ? h.dtau[(size_t)b * P + h.cchild[cb]] : 0.0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    int row = i0 + 16 * m + 4 * t + fk;
    if (row >= nd) continue;
    double y = ((red[0][m][t][lane] + red[1][m][t][lane]) + red[2][m][t][lane]) + red[3][m][t][lane];
    size_t e = (size_t)row * P + col;
    if (!LEAP) {
      Out[e] = y;
      continue;
    }
    double g = y - gam * bv[row];
    double p = Pm[e];
    p = (s == 0) ? p - g * 5e-1 * dt : p - g * dt;
    double x = X[e];
    x = (s < ntau) ? x + cov[row] * p * dt : x + cov[row] * p * 5e-1 * dt;
    if (fabs(x) > 1e3) {
      h.dead[cb] = 1;
      continue;
    }
    while (x < 0.0 || x > max_slip) {
      if (x > max_slip) { x = 2.0 * max_slip - x; p = -p; }
      if (x < 0.0) { x = -x; p = -p; }
    }
    Out[e] = x;
    Pm[e] = p;
  }
}

// The per-column sums of the final positions as S partials (Y = H x formed): x.Y and b.x (the likelihood,
// slip_calc_likelihood, smc_slip.f90:56-104), the prior rows (L x)_r^2 / (2 alpha2) (slip_calc_prior, 106-137),
// the kinetic energy and, in tuning, the ESJD displacement sum (x - x_cur)^2 / cov (slip_hmc_tuning, 1020-1040).
__global__ void __launch_bounds__(256) k_hmc_sums(int nd, int nn2, int P, const double* __restrict__ Xall,
                                                  const int* __restrict__ work, BatchDev bt, HmcDev h, int tune) {
  int b, col;
  if (!wave_element(work, h, b, col)) return;
  size_t cb = (size_t)b * P + col;
  int i = h.cpar[cb];
  const double* X = Xall + (size_t)b * nd * P;
  const double* Y = bt.Y + (size_t)b * nd * P;
  const double* Pm = h.Pm + (size_t)b * nd * P;
  const double* Xc = h.Xc + (size_t)b * nd * P;
  const double* bv = bt.bvec + (size_t)b * nd;
  const double* cov = bt.cov + (size_t)b * nd;
  const int* li = bt.li + (size_t)b * nn2 * 9;
  const double* lv = bt.lv + (size_t)b * nn2 * 9;
  double two_a2 = 2.0 * bt.scal[8 * (size_t)b];
  int s0 = blockIdx.y * 4 + threadIdx.x / 64, S = gridDim.y * 4;
  double v = 0.0, bs = 0.0, kin = 0.0, sc = 0.0, pr = 0.0;
  for (int d = s0; d < nd; d += S) {
    double x = X[(size_t)d * P + col];
    v = v + x * Y[(size_t)d * P + col];
    bs = bs + bv[d] * x;
    double p = Pm[(size_t)d * P + col];
    kin = kin + p * cov[d] * p / 2.0;
    if (tune) {
      double df = x - Xc[(size_t)d * P + i];
      sc = sc + (df * df) / cov[d];
    }
  }
  for (int r = s0; r < nn2; r += S) {
    double tt = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) tt = tt + lv[9 * r + q] * X[(size_t)(li[9 * r + q] - 1) * P + col];
    pr = pr + (tt * tt) / two_a2;
  }
  *part_sum(h, P, b, s0, 0, col) = v;
  *part_sum(h, P, b, s0, 1, col) = bs;
  *part_sum(h, P, b, s0, 2, col) = pr;
  *part_sum(h, P, b, s0, 3, col) = kin;
  *part_sum(h, P, b, s0, 4, col) = sc;
}

// final Hamiltonian from the partials of k_hmc_sums (added in order), tuning score or acceptance probability,
// Metropolis test and chain-state values (slip_hmc_sampling / slip_hmc_tuning after slip_leapfrog,
// smc_slip.f90:700-730, 1020-1060); k_hmc_commit then moves the positions
__global__ void k_hmc_finish(int P, int S, BatchDev bt, HmcDev h, int tune, int record_acc) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (!bt.active[b] || col >= h.ncol[b]) return;
  size_t cb = (size_t)b * P + col;
  int i = h.cpar[cb], j = h.cchild[cb];
  const double* sc = bt.scal + 8 * (size_t)b;
  double gamma = bt.gamma[b];
  bool dead = h.dead[cb] != 0;
  double hc = h.hcur[cb], hn, lk = 0.0, pr = 0.0, post = 0.0, disp = 0.0;
  if (dead) {
    hn = 1e20;  // an escaped trajectory's candidate is its start point: no displacement either
  } else {
    double v = 0.0, bs = 0.0, kin = 0.0;
    for (int q = 0; q < S; ++q) {
      v = v + *part_sum(h, P, b, q, 0, col);
      bs = bs + *part_sum(h, P, b, q, 1, col);
      pr = pr + *part_sum(h, P, b, q, 2, col);
      kin = kin + *part_sum(h, P, b, q, 3, col);
      if (tune) disp = disp + *part_sum(h, P, b, q, 4, col);
    }
    lk = (v - 2.0 * bs + bt.dsd[b]) / 2.0 + sc[3] + sc[4];
    post = gamma * lk + pr;
    hn = post + kin;
  }
  uint32_t stage = (uint32_t)(bt.iter[b] - 1);
  if (record_acc) h.score[(size_t)b * P + j] = exp(fmin(0.0, hc - hn));  // acceptance probability (mover 2)
  if (tune) h.score[(size_t)b * P + j] = disp * exp(fmin(0.0, hc - hn));
  uint64_t w0, w1;
  philox_draw(bt.key[b], tune ? kSiteTuneMH : kSiteHmcMH, stage, (uint32_t)j, 0, 0, w0, w1);
  double mh = clamp_u(u01(w0));
  size_t ib = (size_t)b * P + i;
  int acc = hc - hn > log(mh) ? 1 : 0;
  h.accf[cb] = acc;
  if (acc) {
    h.Lc[ib] = lk;
    h.Rc[ib] = pr;
    h.postc[ib] = post;
  }
  if (!tune) {
    int cur = bt.par[b];
    (cur ? bt.like2 : bt.like)[(size_t)b * P + j] = h.Lc[ib];
    (cur ? bt.prior2 : bt.prior)[(size_t)b * P + j] = h.Rc[ib];
  }
}

// the chain state of each column's parent takes the accepted candidate; in sampling, child j <- chain state
__global__ void __launch_bounds__(256) k_hmc_commit(int nd, int P, const double* __restrict__ Xall,
                                                    const int* __restrict__ work, BatchDev bt, HmcDev h, int tune) {
  int b, col;
  if (!wave_element(work, h, b, col)) return;
  size_t cb = (size_t)b * P + col;
  int i = h.cpar[cb], j = h.cchild[cb];
  bool acc = h.accf[cb] != 0;
  const double* Xt = Xall + (size_t)b * nd * P;
  double* Xc = h.Xc + (size_t)b * nd * P;
  double* Sd = (bt.par[b] ? bt.S2 : bt.S) + (size_t)b * nd * P;
  for (int d = blockIdx.y * 4 + threadIdx.x / 64; d < nd; d += gridDim.y * 4) {
    double x = acc ? Xt[(size_t)d * P + col] : Xc[(size_t)d * P + i];
    if (acc) Xc[(size_t)d * P + i] = x;
    if (!tune) Sd[(size_t)d * P + j] = x;
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

// one pass (tuning or sampling) of a stage: the waves of the pass's work lists (device copies at wa: every
// instance's 64-column tiles; ww: the wide instances' 64-column tiles; wn: the narrow instances' 16-column tiles)
static hipError_t run_waves(const MeshDev& m, BatchDev& bt, HmcDev& h, const HmcPassHost& ph, const int* wa,
                            const int* ww, const int* wn, int B, int tune, double max_slip, int ntau, int record_acc,
                            int bf, hipStream_t s) {
  int P = bt.P, nd = m.ndim, nn2 = 2 * m.nnode;
  int nrg = (((nd + 16 * kHsMT - 1) / (16 * kHsMT)) + 3) / 4, nrb = (nd + 16 * kHsMT - 1) / (16 * kHsMT);
  dim3 gc((P + 127) / 128, B);
  int ry = std::min((nd + 3) / 4, kRowGroups), rl = std::min((std::max(nd, nn2) + 3) / 4, kRowGroups);
  int kmax = (int)ph.offa.size() - 1;
  for (int k = 0; k < kmax; ++k) {
    int na = ph.offa[k + 1] - ph.offa[k], nw = ph.offw[k + 1] - ph.offw[k], nn = ph.offn[k + 1] - ph.offn[k];
    if (na == 0) continue;  // no instance has a k-th child left
    const int* la = wa + 2 * (size_t)ph.offa[k];
    const int* lw = ww + 2 * (size_t)ph.offw[k];
    const int* ln = wn + 2 * (size_t)ph.offn[k];
    hipLaunchKernelGGL(k_hmc_wave_cols, gc, dim3(128), 0, s, P, k, bt, h, tune);
    hipLaunchKernelGGL(k_hmc_wave_init, dim3(na, ry), dim3(256), 0, s, nd, P, la, bt, h, tune);
    hipLaunchKernelGGL(k_hmc_hcur, gc, dim3(128), 0, s, P, 4 * ry, bt, h);
    const double* xin = h.Xt;
    double* xout = h.Xt2;
    for (int st = 0; st <= ntau; ++st) {
      if (nw && bf)
        hipLaunchKernelGGL(k_hmc_leap_bf, dim3(nw * nrg), dim3(256), 0, s, nd, P, st, ntau, max_slip, xin, xout,
                           (const bf16x8*)(st % 2 ? h.XXb : h.XXa), (__bf16*)(st % 2 ? h.XXa : h.XXb), lw, bt, h);
      else if (nw)
        hipLaunchKernelGGL(k_hmc_leap<4>, dim3(nw * nrg), dim3(256), 0, s, nd, P, st, ntau, max_slip, xin, xout, lw,
                           bt, h);
      if (nn)
        hipLaunchKernelGGL(k_narrow<true>, dim3(nn * nrb), dim3(256), 0, s, nd, P, st, ntau, max_slip, h.Qg, xin,
                           xout, ln, bt, h);
      double* t = const_cast<double*>(xin);
      xin = xout;
      xout = t;
    }
    if (nw) hipLaunchKernelGGL(k_hs_work<4>, dim3(nw * nrg), dim3(256), 0, s, nd, P, bt.H, xin, lw, h, bt.Y);
    if (nn)
      hipLaunchKernelGGL(k_narrow<false>, dim3(nn * nrb), dim3(256), 0, s, nd, P, 0, ntau, max_slip, bt.H, xin, bt.Y,
                         ln, bt, h);
    hipLaunchKernelGGL(k_hmc_sums, dim3(na, rl), dim3(256), 0, s, nd, nn2, P, xin, la, bt, h, tune);
    hipLaunchKernelGGL(k_hmc_finish, gc, dim3(128), 0, s, P, 4 * rl, bt, h, tune, record_acc);
    hipLaunchKernelGGL(k_hmc_commit, dim3(na, ry), dim3(256), 0, s, nd, P, xin, la, bt, h, tune);
  }
  return hipGetLastError();
}

// reads the pass's wave pointers (B x (kmax + 1)) back
static hipError_t build_pass(HmcPassHost& ph, const int* dptr, int B, int P, int kmax, hipStream_t s) {
  ph.ptr.assign((size_t)B * (kmax + 1), 0);
  return hipMemcpy2DAsync(ph.ptr.data(), sizeof(int) * (kmax + 1), dptr, sizeof(int) * (P + 1),
                          sizeof(int) * (kmax + 1), B, hipMemcpyDeviceToHost, s);
}

// the work items of every wave of the pass: for every instance with a k-th child left, its 64-column tiles (all,
// and as wide items) or its 16-column tiles (narrow items)
static void list_pass(HmcPassHost& ph, const int* amax, int B, int kmax) {
  ph.la.clear();
  ph.lw.clear();
  ph.ln.clear();
  ph.offa.assign(kmax + 1, 0);
  ph.offw.assign(kmax + 1, 0);
  ph.offn.assign(kmax + 1, 0);
  auto push = [](std::vector<int>& l, int b, int t) {
    l.push_back(b);
    l.push_back(t);
  };
  for (int k = 0; k < kmax; ++k) {
    ph.offa[k] = (int)ph.la.size() / 2;
    ph.offw[k] = (int)ph.lw.size() / 2;
    ph.offn[k] = (int)ph.ln.size() / 2;
    for (int b = 0; b < B; ++b) {
      if (k >= amax[b]) continue;
      const int* pp = ph.ptr.data() + (size_t)b * (kmax + 1);
      int n = pp[k + 1] - pp[k];
      for (int t = 0; t < (n + 63) / 64; ++t) push(ph.la, b, t);
      if (n > kNarrowCols)
        for (int t = 0; t < (n + 63) / 64; ++t) push(ph.lw, b, t);
      else
        for (int t = 0; t < (n + 15) / 16; ++t) push(ph.ln, b, t);
    }
  }
  ph.offa[kmax] = (int)ph.la.size() / 2;
  ph.offw[kmax] = (int)ph.lw.size() / 2;
  ph.offn[kmax] = (int)ph.ln.size() / 2;
}

static hipError_t upload_pass(const HmcPassHost& ph, int* wa, int* ww, int* wn, hipStream_t s) {
  const std::vector<int>* src[3] = {&ph.la, &ph.lw, &ph.ln};
  int* dst[3] = {wa, ww, wn};
  for (int q = 0; q < 3; ++q)
    if (!src[q]->empty()) {
      hipError_t e = hipMemcpyAsync(dst[q], src[q]->data(), sizeof(int) * src[q]->size(), hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return e;
    }
  return hipSuccess;
}

hipError_t launch_hmc_stage(const MeshDev& m, BatchDev& bt, HmcDev& h, HmcHost& hh, int B, double max_slip,
                            const HmcCfg& cfg, hipStream_t s) {
  int P = bt.P;
  bool accept = cfg.mover == SMCF_MOVER_HMC_ACCEPT;
  hipLaunchKernelGGL(k_idstart, dim3((B + 63) / 64), dim3(64), 0, s, P, bt, h, B);
  hipLaunchKernelGGL(k_hmc_prep, dim3(B), dim3(256), 0, s, P, bt, h);
  launch_form_qg(m.ndim, bt, h, B, s);  // the stage's gamma is set
  if (cfg.bf) {
    size_t nf = (size_t)((m.ndim + 15) / 16) * ((m.ndim + 31) / 32) * 64;
    hipLaunchKernelGGL(k_qg_to_bf, dim3((unsigned)std::min<size_t>((nf + 255) / 256, 1024), B), dim3(256), 0, s,
                       m.ndim, bt, h);
  }
  // wave counts, then the wave pointers, for the work lists of the passes
  hh.amax.resize(2 * (size_t)B);
  hipError_t e = hipMemcpyAsync(hh.amax.data(), h.wamax, sizeof(int) * B, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(hh.amax.data() + B, h.tamax, sizeof(int) * B, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  int wmax = 0, tmax = 0;
  for (int b = 0; b < B; ++b) {
    wmax = max(wmax, hh.amax[b]);
    tmax = max(tmax, hh.amax[B + b]);
  }
  size_t wl = hmc_work_entries(B, P);
  if ((e = build_pass(hh.pass[0], h.wptr, B, P, wmax, s)) != hipSuccess) return e;
  if (!accept && (e = build_pass(hh.pass[1], h.tptr, B, P, tmax, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  list_pass(hh.pass[0], hh.amax.data(), B, wmax);
  if ((e = upload_pass(hh.pass[0], h.wla, h.wlw, h.wln, s)) != hipSuccess) return e;
  if (!accept) {
    list_pass(hh.pass[1], hh.amax.data() + B, B, tmax);
    if ((e = upload_pass(hh.pass[1], h.wla + wl, h.wlw + wl, h.wln + wl, s)) != hipSuccess) return e;
  }
  if (accept) {
    hipLaunchKernelGGL(k_hmc_dtau_fill, dim3((P + 127) / 128, B), dim3(128), 0, s, P, bt, h);
    hipLaunchKernelGGL(k_hmc_chain_init, dim3((P + 127) / 128, B), dim3(128), 0, s, m.ndim, P, bt, h, 0);
    e = run_waves(m, bt, h, hh.pass[0], h.wla, h.wlw, h.wln, B, 0, max_slip, cfg.ntau, 1, cfg.bf, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_hmc_adapt, dim3((B + 63) / 64), dim3(64), 0, s, P, bt, h, B, cfg.accept);
    return hipGetLastError();
  }
  // tuning
  hipLaunchKernelGGL(k_hmc_chain_init, dim3((P + 127) / 128, B), dim3(128), 0, s, m.ndim, P, bt, h, 1);
  e = run_waves(m, bt, h, hh.pass[1], h.wla + wl, h.wlw + wl, h.wln + wl, B, 1, max_slip, kNtauRef, 0, cfg.bf, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_tune_resample, dim3((B + 63) / 64), dim3(64), 0, s, P, bt, h, B);
  // sampling
  hipLaunchKernelGGL(k_hmc_chain_init, dim3((P + 127) / 128, B), dim3(128), 0, s, m.ndim, P, bt, h, 0);
  return run_waves(m, bt, h, hh.pass[0], h.wla, h.wlw, h.wln, B, 0, max_slip, kNtauRef, 0, cfg.bf, s);
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

// final Hamiltonian of the n trajectories (Y = H Xf formed, Xf the final positions): ham_cand (1e20 after an
// escape), and the candidate into Xt (the start point after an escape, smc_slip.f90:776-782)
__global__ void k_leapfrog_end(MeshDev m, int P, int n, double gamma, const double* __restrict__ Xf, BatchDev bt,
                               HmcDev h, double* __restrict__ hn_out, double* __restrict__ lk_out,
                               double* __restrict__ pr_out) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n) return;
  int nd = m.ndim, nn2 = 2 * m.nnode;
  if (h.dead[col]) {
    for (int d = 0; d < nd; ++d) h.Xt[(size_t)d * P + col] = h.Xc[(size_t)d * P + col];
    hn_out[col] = 1e20;
    lk_out[col] = h.Lc[col];
    pr_out[col] = h.Rc[col];
    return;
  }
  const double* sc = bt.scal;
  double v = 0.0, bs = 0.0;
  for (int d = 0; d < nd; ++d) {
    double x = Xf[(size_t)d * P + col];
    v = v + x * bt.Y[(size_t)d * P + col];
    bs = bs + bt.bvec[d] * x;
  }
  double lk = (v - 2.0 * bs + bt.dsd[0]) / 2.0 + sc[3] + sc[4];
  double pr = 0.0, two_a2 = 2.0 * sc[0];
  for (int r = 0; r < nn2; ++r) {
    double tt = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) tt = tt + bt.lv[9 * r + q] * Xf[(size_t)(bt.li[9 * r + q] - 1) * P + col];
    pr = pr + (tt * tt) / two_a2;
  }
  if (Xf != h.Xt)
    for (int d = 0; d < nd; ++d) h.Xt[(size_t)d * P + col] = Xf[(size_t)d * P + col];
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
  launch_form_qg(nd, bt, h, 1, s);
  int nrg = (((nd + 16 * kHsMT - 1) / (16 * kHsMT)) + 3) / 4, nt = (n + 63) / 64;
  std::vector<int> list(2 * (size_t)nt, 0);
  for (int t = 0; t < nt; ++t) list[2 * t + 1] = t;  // instance 0, tile t
  hipError_t e = hipMemcpyAsync(h.wla, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  const double* xin = h.Xt;
  double* xout = h.Xt2;
  for (int st = 0; st <= ntau; ++st) {
    hipLaunchKernelGGL(k_hmc_leap<4>, dim3(nt * nrg), dim3(256), 0, s, nd, P, st, ntau, max_slip, xin, xout, h.wla,
                       bt, h);
    double* t = const_cast<double*>(xin);
    xin = xout;
    xout = t;
  }
  launch_hs(nd, P, 1, bt.H, xin, h.ncol, bt.active, bt.Y, s);
  hipLaunchKernelGGL(k_leapfrog_end, gc, dim3(128), 0, s, m, P, n, gamma, xin, bt, h, hn_out, lk_out, pr_out);
  e = hipMemcpyAsync(hc_out, h.hcur, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  // the list lives on this frame: the copy must have been consumed before return
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace smcf
