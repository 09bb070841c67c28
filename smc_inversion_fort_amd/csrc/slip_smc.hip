// Inner slip SMC kernels, batched over geometries (one instance = one outer particle).
//
//  k_init_chain      prior draw: one sequential Gibbs chain per instance, Newton-solved
//                    truncated normals (smc_slip.f90:184-264); one lane per instance
//  k_gamma_resample  one workgroup per instance: bisection on gamma for CV = 2, evidence,
//                    real(16)-grade normalisation and systematic residual resampling in
//                    double-double (smc_slip.f90:314-387, 465-483, 572-589)
//  k_moments         weighted mean / variance diagonal in double-double (smc_slip.f90:485-570)
//  k_gather          children <- parents (index map from the resampling)
//  k_gibbs           truncated-MVN Gibbs sweeps on the tempered posterior (benchmark mover)
#include <mutex>
#include <vector>

#include "smcf_device.h"
#include "smcf_internal.h"

namespace smcf {

// ---------------------------------------------------------------------------
// Prior draw: one sequential Gibbs chain over all P particles, one sweep per particle, started at
// s = 0, with Newton on the normal CDF for the truncated draw (slip_gen_init_particles,
// smc_slip.f90:184-287).
//
// The chain is pipelined without changing a single operation.  LL couples strike slip (even
// dimensions) only with strike slip and dip slip only with dip slip, so each component is its
// own chain.  Within a component, coordinate d of particle k reads particle k's coordinates
// d' < d and particle k-1's coordinates d < d' <= d + band.  Lane l (of the L = chain_lanes lanes
// of a component) runs particles k = l, l + L, ...; particle k starts at round k * lag with
// lag > band, so in every round particle k-1 is exactly lag coordinates ahead and has written
// everything particle k reads.  Each lane keeps its particle's last W coordinates in an LDS
// ring (W >= lag + 1); (L - 1) * lag >= ndof guarantees that a lane's ring is not reused while the
// next lane still reads it, and that particle 0 reads zeros from the ring of lane L - 1.
//
// A chain keeps only about ndof / lag particles in flight per component, so one instance leaves most of a
// wavefront's lanes idle; the 64 lanes hold 32 / L instances of one member instead (lane = q 2L + c L + l for
// instance q, component c).  The chain wavefronts run beside the stage kernels and hold their registers for the
// whole chain, so fewer of them leave more room to the sweep workgroups.
__global__ void __launch_bounds__(64) k_init_chain(const MemberDev* __restrict__ mem, const int2* __restrict__ list,
                                                   const int* __restrict__ sel, const int2* __restrict__ jobs) {
  extern __shared__ double ring[];
  const int2 job = jobs ? jobs[blockIdx.x] : make_int2((int)blockIdx.x, 1);
  auto entry = [&](int k) { return list ? list[k] : make_int2(0, sel ? sel[k] : k); };
  const MemberDev& M = mem[entry(job.x).x];
  const int L = M.chain_lanes;
  int lane = threadIdx.x, q = lane / (2 * L), c = (lane / L) % 2, l = lane % L;
  const bool live = q < job.y;  // lanes of the wavefront's unused instance places run the rounds idle
  int b = entry(job.x + (live ? q : 0)).y;
  int nd = M.nd, ndof = nd / 2, lag = M.chain_lag, W = M.chain_w, P = M.P;
  double max_slip = M.max_slip;
  uint64_t key = M.bt.key[b];
  double a2 = M.bt.scal[8 * (size_t)b];
  const int* IX = M.bt.llc_idx + (size_t)b * nd * 25;
  const double* VX = M.bt.llc_val + (size_t)b * nd * 25;
  double* S = M.bt.S + (size_t)b * nd * P;
  double* mine = ring + (size_t)lane * W;
  const double* prev = ring + (size_t)(q * 2 * L + c * L + (l + L - 1) % L) * W;
  for (int w = 0; w < W; ++w) mine[w] = 0.0;
  __syncthreads();
  long rounds = (long)(P - 1) * lag + ndof;
  for (long t = 0; t < rounds; ++t) {
    long kmax = t / lag;
    long k = kmax - (((kmax - l) % L) + L) % L;
    long dl = t - k * lag;
    if (live && k >= 0 && k < P && dl < ndof) {
      int d = (int)dl, i = 2 * d + c;
      const int* ix = IX + (size_t)i * 25;
      const double* vx = VX + (size_t)i * 25;
      // the column's 25 slots are loaded together (the -1 padding sits at the end, so skipping j < 0 is
      // the reference's exit at the first pad); the sums run in slot order as in the sequential chain
      int jj[25];
      double vv[25];
#pragma unroll
      for (int w = 0; w < 25; ++w) {
        jj[w] = ix[w];
        vv[w] = vx[w];
      }
      double lii = 0.0, mu = 0.0;
      bool found = false;
#pragma unroll
      for (int w = 0; w < 25; ++w)
        if (!found && jj[w] == i) {
          lii = vv[w];
          found = true;
        }
      double s2 = 1.0 / (lii / a2);
#pragma unroll
      for (int w = 0; w < 25; ++w) {
        int j = jj[w];
        if (j >= 0 && j != i) {
          int dj = j >> 1;  // same component as i
          double cj = j < i ? mine[dj & (W - 1)] : prev[dj & (W - 1)];
          mu = mu - vv[w] / a2 * cj * s2;
        }
      }
      double v1 = draw_normal(key, kSiteGibbsInit, 0, (uint32_t)k, (uint32_t)i);
      double x = mu + v1 * sqrt(s2);
      double fx = cdf_norm(x, mu, s2);
      double f1 = cdf_norm(max_slip, mu, s2);
      double f0 = cdf_norm(0.0, mu, s2);
      double target = (f1 - f0) * fx + f0;
      double y = mu, err = 1.0, yp = y + err;
      int cnt = 0;
      while (err > 1e-8 && cnt < 100) {
        y = y - (cdf_norm(y, mu, s2) - target) / pdf_norm(y, mu, s2);
        err = fabs(y - yp);
        yp = y;
        cnt++;
      }
      mine[d & (W - 1)] = y;
      S[(size_t)i * P + k] = y;
    }
    __syncthreads();
  }
}

hipError_t launch_init_chain(const MemberDev* mem, const int2* list, const int* sel, const int2* jobs, int nw,
                             int chain_w, hipStream_t s) {
  size_t sh = sizeof(double) * 2 * kChainLanes * (size_t)chain_w;
  if (sh > 64 * 1024) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(k_init_chain, dim3(nw), dim3(2 * kChainLanes), sh, s, mem, list, sel, jobs);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tempering step for one instance.  Layout: 256 threads; thread t owns the contiguous
// particle range [t*C, (t+1)*C) with C = ceil(P/256), so per-thread partial sums and
// prefix sums run in ascending particle order.
// ---------------------------------------------------------------------------
constexpr int GT = 256;

struct GammaOut {
  double gamma, neglog_ev, wsum_hi, wsum_lo;
};

// bisection on gamma (returns gamma, leaves w[] = exp(-dg (L - Lmin)) of the final gamma)
SMCF_DEV double gamma_search(const double* L, double* w, int P, double gprev, double cv_t, double tol, int cap,
                             double* red, double* neglog_ev) {
  int t = threadIdx.x;
  int C = (P + GT - 1) / GT, i0 = t * C, i1 = min(P, i0 + C);
  double lmin_t = i0 < P ? L[i0] : L[0];
  for (int i = i0; i < i1; ++i) lmin_t = fmin(lmin_t, L[i]);
  double lmin = block_min<GT>(lmin_t, red);
  double lower = gprev, upper = 1.0, err = 1.0, gamma = 1.0, mean = 1.0;
  int cnt = 0;
  while (err > tol) {
    gamma = (lower + upper) / 2.0;
    double dg = gamma - gprev;
    double s = 0.0;
    for (int i = i0; i < i1; ++i) {
      double v = exp(-dg * (L[i] - lmin));
      w[i] = v;
      s = s + v;
    }
    mean = block_sum<GT>(s, red) / P;
    double q = 0.0;
    for (int i = i0; i < i1; ++i) q = q + (w[i] - mean) * (w[i] - mean);
    double sd = sqrt(block_sum<GT>(q, red) / P);
    double cv = sd / mean;
    if (cv > cv_t) upper = gamma; else lower = gamma;
    err = fabs(cv - cv_t);
    if (fabs(gamma - 1.0) < tol) break;
    if (cnt > cap) break;
    cnt++;
  }
  double dg = gamma - gprev;
  *neglog_ev = -log(mean) + dg * lmin;
  return gamma;
}

// w <- w / sum(w) (double-double sum), prefix sums, resampling counts and parent map
SMCF_DEV void normalize_resample(double* w, int P, dd u, int* assigned, int* parent, double* red, dd* scan) {
  int t = threadIdx.x;
  int C = (P + GT - 1) / GT, i0 = t * C, i1 = min(P, i0 + C);
  dd s = {0.0, 0.0};
  for (int i = i0; i < i1; ++i) s = dd_add_d(s, w[i]);
  scan[t] = s;
  __syncthreads();
  if (t == 0) {
    dd run = {0.0, 0.0};
    for (int k = 0; k < GT; ++k) {
      dd v = scan[k];
      scan[k] = run;  // exclusive prefix of the thread totals
      run = dd_add(run, v);
    }
    scan[GT] = run;
  }
  __syncthreads();
  dd tot = scan[GT];
  for (int i = i0; i < i1; ++i) w[i] = d_div_dd(w[i], tot);
  __syncthreads();
  // prefix sums of the normalised weights
  dd c = {0.0, 0.0};
  for (int i = i0; i < i1; ++i) c = dd_add_d(c, w[i]);
  scan[t] = c;
  __syncthreads();
  if (t == 0) {
    dd run = {0.0, 0.0};
    for (int k = 0; k < GT; ++k) {
      dd v = scan[k];
      scan[k] = run;
      run = dd_add(run, v);
    }
  }
  __syncthreads();
  if (dd_lt_d(u, 1e-5)) u = dd{1e-5, 0.0};
  dd u0 = dd_div_d(u, (double)P);
  dd cum = scan[t];
  int Aprev;
  {
    // A(i0 - 1)
    if (i0 == 0) Aprev = 0;
    else {
      dd x = dd_mul_d(dd_add(cum, dd_neg(u0)), (double)P);
      Aprev = (int)dd_floor(x) + 1;
    }
  }
  Aprev = max(0, min(P, Aprev));
  for (int i = i0; i < i1; ++i) {
    cum = dd_add_d(cum, w[i]);
    dd x = dd_mul_d(dd_add(cum, dd_neg(u0)), (double)P);
    int A = (int)dd_floor(x) + 1;
    A = max(Aprev, min(P, A));
    assigned[i] = A - Aprev;
    for (int j = Aprev; j < A; ++j) parent[j] = i;
    Aprev = A;
  }
  // counts that fall short of P (only possible through rounding at the very end of the
  // cumulative sum) leave the last children on the last particle
  if (i1 == P && i0 < P)
    for (int j = Aprev; j < P; ++j) parent[j] = P - 1;
  __syncthreads();
}

__global__ void __launch_bounds__(GT) k_gamma_resample(StageSet ss) {
  extern __shared__ double smem[];
  int2 is = ss.inst[blockIdx.x];
  const MemberDev& M = ss.mem[is.x];
  const BatchDev& bt = M.bt;
  int P = M.P;
  double* Ls = smem;            // P
  double* ws = Ls + P;          // P
  double* red = ws + P;         // GT
  dd* scan = (dd*)(red + GT);   // GT + 1
  int b = is.y, t = threadIdx.x;
  if (bt.done[b]) {
    if (t == 0) bt.active[b] = 0;
    return;
  }
  // every thread reads the stage state before the first barrier; thread 0 updates it
  // only after the barriers inside gamma_search
  int pb = bt.par[b];
  int it = bt.iter[b];
  double gprev = bt.gamma[b];
  const double* L = (pb ? bt.like2 : bt.like) + (size_t)b * P;
  for (int i = t; i < P; i += GT) Ls[i] = L[i];
  __syncthreads();
  double nev;
  double gamma = gamma_search(Ls, ws, P, gprev, 2.0, 1e-10, 100000, red, &nev);
  if (t == 0) {
    double nl = bt.neglog[b] + nev;
    if (it <= 201) bt.gtrace[(size_t)b * 202 + it - 1] = gamma;
    bt.gamma[b] = gamma;
    bt.iter[b] = it + 1;
    if (it > 200) {
      nl = 1e10;
      bt.done[b] = 1;
      bt.active[b] = 0;
    } else {
      bt.active[b] = 1;
      bt.par[b] = 1 - pb;
      if (1.0 - gamma <= 10e-8) bt.done[b] = 1;
    }
    bt.neglog[b] = nl;
  }
  if (it > 200) return;
  uint64_t w0, w1;
  philox_draw(bt.key[b], kSiteResample, (uint32_t)it, 0, 0, 0, w0, w1);
  dd u = two_sum((double)(w0 >> 11) * 0x1.0p-53, (double)(w1 >> 11) * 0x1.0p-106);
  normalize_resample(ws, P, u, bt.assigned + (size_t)b * P, bt.parent + (size_t)b * P, red, scan);
  for (int i = t; i < P; i += GT) bt.w[(size_t)b * P + i] = ws[i];
}

hipError_t launch_gamma_resample(const StageSet& ss, hipStream_t s) {
  // the workgroup's LDS is sized for the largest P: the sums run over its own P only
  size_t sh = sizeof(double) * (2 * (size_t)ss.P_max + GT) + sizeof(dd) * (GT + 1);
  hipLaunchKernelGGL(k_gamma_resample, dim3(ss.ninst), dim3(GT), sh, s, ss);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// weighted mean and variance diagonal of coordinate i over P particles (S dimension-major, w normalised), in
// double-double for the reference's real(16) sums (slip_calc_mean_particles / slip_calc_cov_particles,
// smc_slip.f90:485-570; variances below 1e-3 are floored at 1e-3)
SMCF_DEV void moments_one(int i, int P, const double* __restrict__ S, const double* __restrict__ w, double& mean,
                          double& cv) {
  dd m = {0.0, 0.0};
  for (int p = 0; p < P; ++p) m = dd_add_d(m, w[p] * S[(size_t)i * P + p]);
  mean = m.hi + m.lo;
  dd c = {0.0, 0.0};
  for (int p = 0; p < P; ++p) {
    double di = S[(size_t)i * P + p] - mean;
    c = dd_add_d(c, w[p] * di * di);
  }
  cv = c.hi + c.lo;
  if (fabs(cv) < 1e-3) cv = 1e-3;
}

// weighted mean and variance diagonal of the pre-resampling particles (source buffer)
__global__ void k_moments(StageSet ss) {
  int2 is = ss.inst[blockIdx.y];
  const MemberDev& M = ss.mem[is.x];
  const BatchDev& bt = M.bt;
  int nd = M.nd, P = M.P, b = is.y;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (!bt.active[b] || i >= nd) return;
  const double* S = (bt.par[b] ? bt.S : bt.S2) + (size_t)b * nd * P;  // source = previous buffer
  double mean, cv;
  moments_one(i, P, S, bt.w + (size_t)b * P, mean, cv);
  bt.mean[(size_t)b * nd + i] = mean;
  bt.cov[(size_t)b * nd + i] = cv;
}

__global__ void k_moments_raw(int nd, int P, const double* __restrict__ S, const double* __restrict__ w,
                              double* __restrict__ mean, double* __restrict__ cov) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nd) return;
  moments_one(i, P, S, w, mean[i], cov[i]);
}

hipError_t launch_moments_raw(int nd, int P, const double* S, const double* w, double* mean, double* cov,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_moments_raw, dim3((nd + 63) / 64), dim3(64), 0, s, nd, P, S, w, mean, cov);
  return hipGetLastError();
}

hipError_t launch_moments(const StageSet& ss, hipStream_t s) {
  hipLaunchKernelGGL(k_moments, dim3((ss.nd_max + 63) / 64, ss.ninst), dim3(64), 0, s, ss);
  return hipGetLastError();
}

// children <- parents into the (new) current buffer; likelihood / prior follow
__global__ void k_gather(StageSet ss) {
  int2 is = ss.inst[blockIdx.z];
  const MemberDev& M = ss.mem[is.x];
  const BatchDev& bt = M.bt;
  int nd = M.nd, P = M.P, b = is.y;
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (!bt.active[b] || j >= P) return;
  int cur = bt.par[b];
  const double* Ss = (cur ? bt.S : bt.S2) + (size_t)b * nd * P;
  double* Sd = (cur ? bt.S2 : bt.S) + (size_t)b * nd * P;
  int pa = bt.parent[(size_t)b * P + j];
  if (bt.SX) {  // split-BF16 copy of the children as well: 16-value groups [hi 8 | lo 8]
    int ng = bf_rows(nd) / 8;
    const uint4* Xs = (const uint4*)(cur ? bt.SX : bt.SXb) + (size_t)b * ng * P * 2;
    uint4* Xd = (uint4*)(cur ? bt.SXb : bt.SX) + (size_t)b * ng * P * 2;
    for (int g8 = blockIdx.y; g8 < ng; g8 += gridDim.y) {
      Xd[((size_t)g8 * P + j) * 2] = Xs[((size_t)g8 * P + pa) * 2];
      Xd[((size_t)g8 * P + j) * 2 + 1] = Xs[((size_t)g8 * P + pa) * 2 + 1];
    }
  }
  if (bt.S32) {  // mixed precision: the FP32 copy of the children as well
    float* Td = (cur ? bt.S32b : bt.S32) + (size_t)b * nd * P;
    for (int i = blockIdx.y; i < nd; i += gridDim.y) {
      double v = Ss[(size_t)i * P + pa];
      Sd[(size_t)i * P + j] = v;
      Td[(size_t)i * P + j] = (float)v;
    }
  } else if (!bt.SX) {
    // (the split-BF16 sweep reads only the split copy and writes the FP64 particles in its last sweep, and the
    // centred likelihood takes its reference particle from the parents: the FP64 children are not gathered)
    for (int i = blockIdx.y; i < nd; i += gridDim.y) Sd[(size_t)i * P + j] = Ss[(size_t)i * P + pa];
  }
  if (blockIdx.y == 0) {
    const double* Ls = (cur ? bt.like : bt.like2) + (size_t)b * P;
    double* Ld = (cur ? bt.like2 : bt.like) + (size_t)b * P;
    const double* Rs = (cur ? bt.prior : bt.prior2) + (size_t)b * P;
    double* Rd = (cur ? bt.prior2 : bt.prior) + (size_t)b * P;
    Ld[j] = Ls[pa];
    Rd[j] = Rs[pa];
  }
}

hipError_t launch_gather(const StageSet& ss, hipStream_t s) {
  int gy = ss.nd_max < 64 ? ss.nd_max : 64;
  hipLaunchKernelGGL(k_gather, dim3((ss.P_max + 255) / 256, gy, ss.ninst), dim3(256), 0, s, ss);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gibbs sweeps (truncated-MVN Gibbs on the tempered posterior: precision Q = gamma H + LL / alpha^2,
// conditional mean of coordinate i from q_i = (Q s)_i, box [0, max_slip]).
// ---------------------------------------------------------------------------
// one coordinate redraw: Philox uniform + truncated-normal inversion (kept out of line so the
// unrolled block loop does not replicate it)
__device__ __noinline__ double gibbs_draw(double mu, double s, double is, double max_slip, double u) {
  return trunc_normal_prec(mu, s, is, 0.0, max_slip, u);
}

// The same draw (bit for bit) for everything except the rare regions -- a bound beyond 35 sigma
// (Mills-ratio solve) and p < 1.4e-11 (AS 241's far tail) -- where it returns NaN and the caller
// takes gibbs_draw.  Without those branches the function needs under 32 VGPRs, so a call clobbers
// only caller-saved registers and saves nothing to scratch.
__device__ __noinline__ double gibbs_draw_fast(double mu, double s, double is, double max_slip, double u) {
  double a = (0.0 - mu) * s, b = (max_slip - mu) * s;
  if (a > 35.0 || b < -35.0) return __builtin_nan("");
  const double r2 = 0.70710678118654752440;
  double ta = (a <= -10.0 && a * a > b * b + 100.0) ? 0.0 : 0.5 * erfc(fabs(a) * r2);
  double tb = (b >= 10.0 && b * b > a * a + 100.0) ? 0.0 : 0.5 * erfc(fabs(b) * r2);
  double pa = a < 0.0 ? ta : 1.0 - ta, qa = a < 0.0 ? 1.0 - ta : ta;
  double pb = b < 0.0 ? tb : 1.0 - tb, qb = b < 0.0 ? 1.0 - tb : tb;
  double m = a >= 0.0 ? qa - qb : pb - pa;
  double p = pa + u * m, q = qb + (1.0 - u) * m;
  double pq = fmin(p, q);  // <= 1/2: AS 241's central region or its lower tail
  double q0 = pq - 0.5;
  double z = ndtri_central(q0);
  if (q0 < -0.425) {
    double r = sqrt(-log(pq));
    if (r > 5.0) return __builtin_nan("");
    z = -ndtri_tail(r);
  }
  z = p <= q ? z : -z;
  z = fmin(fmax(z, a), b);
  double x = mu + z * is;
  return fmin(fmax(x, 0.0), max_slip);
}

// Likelihood of slip_calc_likelihood (smc_slip.f90:56-104) in Gram form,
//   L(s) = (s^T H s - 2 b^T s + dsd) / 2 + c,
// for the 64 particles p0 .. p0+63 of one wavefront.  Y = H S is formed in 64-row blocks on the matrix
// cores (A(r, k) = H(i0 + r, j0 + k) read column-major, B(k, c) = s_{j0+k}(p0 + c)); the epilogue of each
// block accumulates s^T (Y - 2 b) over its rows (lane l holds D(4t + l/16, l%16) in register t), and the
// four row groups of a particle are summed with lane shuffles.  Y never leaves the registers.
#ifndef SMCF_LIKE_MT
#define SMCF_LIKE_MT 4
#define SMCF_LIKE_NB 3
#define SMCF_LIKE_OCC 2
#endif
constexpr int kLikeMT = SMCF_LIKE_MT;  // 16-row tiles per block of H (64 rows: S is streamed nd / 64 times)

// s^T (H s - 2 b) over the rows of blocks rb0, rb0 + rbstep, ... for the wave's 64 particles p0 .. p0 + 63 (each
// particle's rows summed in ascending order; part[n] holds this lane's share of particle p0 + 16n + lane % 16)
__device__ __forceinline__ void wave_like_part(int nd, int P, const double* __restrict__ H,
                                               const double* __restrict__ S, const double* __restrict__ bv, int p0,
                                               int lane, int rb0, int rbstep, double (&part)[4]) {
  constexpr int MT = kLikeMT, KB = 16 * MT;
  int fi = lane % 16, fk = lane / 16;
  int pcol[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) pcol[n] = min(p0 + 16 * n + fi, P - 1);
  uint32_t bo[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) bo[n] = (uint32_t)(fk * P + pcol[n]);
  int nd4 = nd & ~3, nblk = (nd + KB - 1) / KB;
  for (int rb = rb0; rb < nblk; rb += rbstep) {
    int i0 = rb * KB;
    dbl4 acc[MT][4];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n) acc[m][n] = dbl4{0.0, 0.0, 0.0, 0.0};
    // rows past nd read row nd - 1; the epilogue skips them
    uint32_t ao[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) ao[m] = (uint32_t)(fk * nd + min(i0 + 16 * m + fi, nd - 1));
    mfma_panel_t<MT, SMCF_LIKE_NB>(H, ao, (uint32_t)(4 * nd), S, bo, (uint32_t)(4 * P), nd4 / 4, acc);
    if (nd4 < nd) {  // last rows (nd = 2 mod 4): rows past nd contribute zero
      bool in = nd4 + fk < nd;
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        double bn = in ? S[(size_t)(nd4 + fk) * P + pcol[n]] : 0.0;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          int r = i0 + 16 * m + fi;
          double a = (in && r < nd) ? H[(size_t)(nd4 + fk) * nd + r] : 0.0;
          acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bn, acc[m][n], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        int row = i0 + 16 * m + 4 * t + fk;
        if (row < nd) {
          double b2 = 2.0 * bv[row];
#pragma unroll
          for (int n = 0; n < 4; ++n) part[n] = fma(acc[m][n][t] - b2, S[(size_t)row * P + pcol[n]], part[n]);
        }
      }
  }
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    part[n] += __shfl_xor(part[n], 16);
    part[n] += __shfl_xor(part[n], 32);
  }
}

__device__ __forceinline__ void wave_like(int nd, int P, const double* __restrict__ H, const double* __restrict__ S,
                                          const double* __restrict__ bv, double dsd, double c, int p0, int lane,
                                          double* __restrict__ Lout) {
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  wave_like_part(nd, P, H, S, bv, p0, lane, 0, 1, part);
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    int p = p0 + 16 * n + lane % 16;
    if (lane < 16 && p < P) Lout[p] = (part[n] + dsd) / 2.0 + c;
  }
}

// one workgroup of 4 wavefronts per 256 particles of an instance; all chunks of an instance on one
// XCD (workgroup w runs on XCD w % 8), so that its H is read into one L2
__global__ void __launch_bounds__(256, SMCF_LIKE_OCC)
    k_like_mfma(int nd, int P, int B, const double* __restrict__ Hall, const double* __restrict__ bvec,
                const double* __restrict__ dsd, const double* __restrict__ scal, const double* __restrict__ S0,
                const double* __restrict__ S1, const int* __restrict__ par, const int* __restrict__ active,
                int use_active, const int* __restrict__ sel, double* __restrict__ L0, double* __restrict__ L1) {
  int chunks = (P + 255) / 256;
  int w = blockIdx.x, xcd = w % 8, r = w / 8;
  int bi = (r / chunks) * 8 + xcd, chunk = r % chunks;
  if (bi >= B) return;
  int b = sel ? sel[bi] : bi;
  if (use_active && !active[b]) return;
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int p0 = chunk * 256 + wave * 64;
  if (p0 >= P) return;
  int pb = par[b];
  const double* S = (pb ? S1 : S0) + (size_t)b * nd * P;
  const double* sc = scal + 8 * (size_t)b;
  wave_like(nd, P, Hall + (size_t)b * nd * nd, S, bvec + (size_t)b * nd, dsd[b], sc[3] + sc[4], p0, lane,
            (pb ? L1 : L0) + (size_t)b * P);
}

// The stage likelihood of the moved particles of the FP64 and FP32 forms: one workgroup per 64 particles of an
// instance (XCD-aware items), its four wavefronts taking the 64-row blocks of H round-robin for the same particles, so
// the particles are streamed from memory by one workgroup, its wavefronts sharing them through the caches, instead of
// once per block; a particle's four partial sums are added in wavefront order.
__global__ void __launch_bounds__(256, SMCF_LIKE_OCC)
    k_like_items(const int4* __restrict__ items, const MemberDev* __restrict__ mem) {
  __shared__ double red[4][64];
  int4 it = items[blockIdx.x];
  if (it.x < 0) return;
  const MemberDev& M = mem[it.x];
  const BatchDev& bt = M.bt;
  int b = it.y, nd = M.nd, P = M.P;
  if (!bt.active[b]) return;
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int p0 = it.z * 64;
  if (p0 >= P) return;
  int pb = bt.par[b];
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  wave_like_part(nd, P, bt.H + (size_t)b * nd * nd, (pb ? bt.S2 : bt.S) + (size_t)b * nd * P,
                 bt.bvec + (size_t)b * nd, p0, lane, wave, 4, part);
#pragma unroll
  for (int n = 0; n < 4; ++n)
    if (lane < 16) red[wave][16 * n + lane] = part[n];
  __syncthreads();
  if (threadIdx.x < 64 && p0 + (int)threadIdx.x < P) {
    int t = threadIdx.x;
    const double* sc = bt.scal + 8 * (size_t)b;
    double v = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
    ((pb ? bt.like2 : bt.like) + (size_t)b * P)[p0 + t] = (v + bt.dsd[b]) / 2.0 + (sc[3] + sc[4]);
  }
}

hipError_t launch_like_raw(int nd, int P, const double* H, const double* bv, const double* dsd, const double* scal,
                           const double* S, const int* zero, double* out, hipStream_t s) {
  int chunks = (P + 255) / 256;
  hipLaunchKernelGGL(k_like_mfma, dim3(8 * chunks), dim3(256), 0, s, nd, P, 1, H, bv, dsd, scal, S, S, zero, zero, 0,
                     (const int*)nullptr, out, out);
  return hipGetLastError();
}

hipError_t launch_like(const MeshDev& m, BatchDev& bt, int B, int use_active, hipStream_t s, const int* sel) {
  int chunks = (bt.P + 255) / 256;
  int groups = (B + 7) / 8;
  hipLaunchKernelGGL(k_like_mfma, dim3(groups * 8 * chunks), dim3(256), 0, s, m.ndim, bt.P, B, bt.H, bt.bvec, bt.dsd,
                     bt.scal, bt.S, bt.S2, bt.par, bt.active, use_active, sel, bt.like, bt.like2);
  return hipGetLastError();
}

hipError_t launch_like_stage(const StageSet& ss, const ItemList& it, hipStream_t s) {
  if (it.n > 0) hipLaunchKernelGGL(k_like_items, dim3(it.n), dim3(256), 0, s, it.items, ss.mem);
  return hipGetLastError();
}

// Q is stored per stage in blocks of KB columns, row-major inside the block:
//   Qb[(blk*nd + j)*KB + k] = Q(j, blk*KB + k)   (zero for columns past nd)
//
// One wavefront owns 64 particles; a workgroup of 4 wavefronts owns 256 particles of one instance.
// Per block of KB = 16 MT coordinates (MT 16-coordinate tiles m) and sweep:
//  * matrix phase: q(i0 + 16m + r, p) = sum_j Q(j, i0 + 16m + r) s_j(p) for the wave's 64 particles
//    (four 16-particle tiles n) as 4 MT accumulators of v_mfma_f64_16x16x4f64 over j in steps of 4,
//    software-pipelined (mfma_panel_t).  Operand fragments are plain coalesced loads: A(r, k) =
//    Q(j0 + k, col) at lane r + 16k, B(k, c) = s_{j0+k}(p0 + c) at lane c + 16k; lane l holds
//    D(4t + l/16, l%16) in register t.
//  * draw phase, per tile: the tile's constants are staged in LDS, its accumulators are transposed through
//    8 KB of wave-private LDS so that lane l finds q of its own particle; the 16 coordinates are redrawn in
//    order with left-looking in-tile corrections (VALU), and the tile's deltas, left in LDS, correct the
//    block's later tiles by 16 MFMAs each.
// While one wavefront draws (VALU) another can run its matrix phase on the matrix cores.
//
// Workgroup -> (instance, particle chunk) is XCD-aware: workgroup w runs on XCD w % 8, and all
// chunks of one instance get the same w % 8, so an instance's Q blocks are fetched into one L2.
constexpr int kXcd = 8;
#ifndef SMCF_GIBBS_MT
#define SMCF_GIBBS_MT 2
#endif
constexpr int kGibbsMT = SMCF_GIBBS_MT;  // 16-row tiles per block of the FP64 Gibbs kernel
constexpr int kGibbsKB = 16 * kGibbsMT;
#ifndef SMCF_GIBBS_OCC
#define SMCF_GIBBS_OCC 4
#endif
constexpr int kGibbsOcc = SMCF_GIBBS_OCC;  // waves per SIMD the register budget of k_gibbs is sized for

template <int MT>
__device__ __forceinline__ void gibbs_body(int nd, int P, int b, int chunk, int sweeps, double max_slip,
                                           const double* __restrict__ Qball, double* __restrict__ S0,
                                           double* __restrict__ S1, const double* __restrict__ cdall,
                                           const uint64_t* __restrict__ keys, const int* __restrict__ iters,
                                           const int* __restrict__ active, const int* __restrict__ par) {
  constexpr int KB = 16 * MT;
  __shared__ double T[4][16][65];
  // per wave, the constants of the tile being drawn (read with wave-uniform addresses, so broadcast):
  //   CQ[kk (kk - 1) / 2 + k] = Q(i0 + 16m + k, i0 + 16m + kk), k < kk: the strictly lower part of the tile's
  //   diagonal block (in-tile corrections), packed so that four workgroups still fit one CU's LDS
  //   CD[c][kk] = cd[c nd + i0 + 16m + kk]: gamma b_i, 1 / Q_ii, sqrt(Q_ii), 1 / sqrt(Q_ii)
  __shared__ double CQ[4][120];
  __shared__ double CD[4][4][16];
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int pw0 = chunk * 256 + wave * 64;
  int p = pw0 + lane;
  bool valid = p < P;
  int pc = valid ? p : P - 1;
  double* S = (par[b] ? S1 : S0) + (size_t)b * nd * P;
  const double* Qb = Qball + (size_t)b * nd * ((nd + KB - 1) / KB) * KB;
  // per-coordinate constants of the stage: gamma b_i, 1 / Q_ii, sqrt(Q_ii), 1 / sqrt(Q_ii)
  const double* cd = cdall + (size_t)b * 4 * nd;
  uint64_t key = keys[b];
  uint32_t stage = (uint32_t)(iters[b] - 1);
  int nblk = (nd + KB - 1) / KB;
  int fi = lane % 16, fk = lane / 16;
  // fragment columns: particles of the four 16-particle tiles (clamped into range; unused if invalid)
  int pcol[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) pcol[n] = min(pw0 + 16 * n + fi, P - 1);
  uint32_t bo[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) bo[n] = (uint32_t)(fk * P + pcol[n]);
  int nd4 = nd & ~3;
  uint32_t ao[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) ao[m] = (uint32_t)(fk * KB + 16 * m + fi);
  double (*Tw)[65] = T[wave];
  double* CQw = CQ[wave];
  double (*CDw)[16] = CD[wave];
  // the packed entries e = lane and lane + 64 (< 120) this lane stages: (row k, column kk) of the block
  int ek[2], ekk[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    int e = lane + 64 * h, kk = 1;
    while (kk < 16 && kk * (kk + 1) / 2 <= e) ++kk;
    ekk[h] = kk;
    ek[h] = e - kk * (kk - 1) / 2;
  }
  for (int sw = 0; sw < sweeps; ++sw) {
    for (int blk = 0; blk < nblk; ++blk) {
      int i0 = blk * KB;
      const double* Qr = Qb + (size_t)blk * nd * KB;
      dbl4 acc[MT][4];
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n) acc[m][n] = dbl4{0.0, 0.0, 0.0, 0.0};
      const double* qa = Qr + (size_t)fk * KB + fi;
      const double* sb = S + (size_t)fk * P;
      mfma_panel_t<MT, kPanelBuffers>(Qr, ao, (uint32_t)(4 * KB), S, bo, (uint32_t)(4 * P), nd4 / 4, acc);
      if (nd4 < nd) {  // last rows (nd = 2 mod 4): rows past nd contribute zero
        bool in = nd4 + fk < nd;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          double bn = in ? sb[(size_t)nd4 * P + pcol[n]] : 0.0;
#pragma unroll
          for (int m = 0; m < MT; ++m) {
            double am = in ? qa[(size_t)nd4 * KB + 16 * m] : 0.0;
            acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(am, bn, acc[m][n], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        if (i0 + 16 * m >= nd) break;  // tiles past the last coordinate (wave-uniform)
        {  // the tile's constants into LDS
          int ik = i0 + 16 * m + fi;
          CDw[fk][fi] = ik < nd ? cd[(size_t)fk * nd + ik] : 0.0;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            int e = lane + 64 * h, r = i0 + 16 * m + ek[h];
            if (e < 120) CQw[e] = r < nd ? Qr[(size_t)r * KB + 16 * m + ekk[h]] : 0.0;
          }
        }
        // transpose tile m through LDS: Tw[row][particle]
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
          for (int t = 0; t < 4; ++t) Tw[4 * t + fk][16 * n + fi] = acc[m][n][t];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint64_t wodd = 0;
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
          int i = i0 + 16 * m + kk;
          if (i >= nd) {
            Tw[kk][lane] = 0.0;
            continue;
          }
          // coordinates 2m' and 2m'+1 share one Philox block (word i & 1); i0 and the tile offset are even
          double u;
          if ((kk & 1) == 0) {
            uint64_t w0;
            philox_draw(key, kSiteGibbsMove, stage, (uint32_t)pc, (uint32_t)(i >> 1), (uint32_t)sw, w0, wodd);
            u = u01_open(w0);
          } else {
            u = u01_open(wodd);
          }
          // left-looking corrections by the tile's coordinates redrawn before it (deltas in Tw[k][lane]):
          //   qi = T(kk) + Q(i0 + 16m + k, i) delta_k, k = 0 .. kk - 1, fused in ascending k (the operations and
          //   order of updating every later entry of the tile after each draw; Q read from row i0 + 16m + k)
          double qi = Tw[kk][lane];
#pragma unroll
          for (int k = 0; k < kk; ++k) qi = fma(CQw[kk * (kk - 1) / 2 + k], Tw[k][lane], qi);
          double so = S[(size_t)i * P + pc];
          double mu = so - (qi - CDw[0][kk]) * CDw[1][kk];
          // lean draw for the common regions; the full one (tails beyond 10 sigma) when it declines
          double sn = gibbs_draw_fast(mu, CDw[2][kk], CDw[3][kk], max_slip, u);
          if (sn != sn) sn = gibbs_draw(mu, CDw[2][kk], CDw[3][kk], max_slip, u);
          double dlt = sn - so;
          if (valid) S[(size_t)i * P + p] = sn;
          Tw[kk][lane] = dlt;  // the delta, read by the later coordinates of the tile
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // corrections of the later tiles by this tile's deltas (in Tw[k][p]):
        //   acc[m2](r, p) += sum_k Q(i0 + 16m + k, i0 + 16m2 + r) delta_k(p)
#pragma unroll
        for (int m2 = m + 1; m2 < MT; ++m2) {
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) {
            int kq = 4 * s4 + fk;
            double a = (i0 + 16 * m + kq < nd) ? Qr[(size_t)(i0 + 16 * m + kq) * KB + 16 * m2 + fi] : 0.0;
#pragma unroll
            for (int n = 0; n < 4; ++n)
              acc[m2][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Tw[kq][16 * n + fi], acc[m2][n], 0, 0, 0);
          }
        }
        __builtin_amdgcn_wave_barrier();  // every lane has read the tile's deltas before the next tile overwrites them
      }
    }
  }
}

// one workgroup per item (instance, 256-particle chunk) of the FP64 form's class with KB = 16 MT rows
template <int MT>
__global__ void __launch_bounds__(256, kGibbsOcc) k_gibbs(const int4* __restrict__ items, GibbsArgs a) {
  int4 it = items[blockIdx.x];
  if (it.x < 0) return;
  const GibbsMember& M = a.m[it.x];
  if (!M.active[it.y]) return;
  gibbs_body<MT>(M.nd, M.P, it.y, it.z, M.sweeps, M.max_slip, M.Q, M.S, M.S2, M.qdiag, M.key, M.iter, M.active,
                 M.par);
}



// ---------------------------------------------------------------------------
// Mixed-precision Gibbs sweeps (SMCF_PRECISION_MIXED).  The sweep of k_gibbs, with
//  * the matrix phase on v_mfma_f32_16x16x4f32 from FP32 copies of Q and S: half the bytes of S per
//    block and twice the FP64 matrix rate, over blocks of KB = 64 coordinates (four 16-row tiles; an
//    FP32 accumulator tile takes 4 registers where FP64 takes 8).  FP32 fragments: A(r, k) at lane
//    r + 16k, B(k, c) at lane c + 16k; lane l holds D(4 (l/16) + t, l % 16) in register t (checked on
//    the device before first use; the FP64 instruction orders its rows differently);
//  * q of each tile converted to double in LDS; the conditional mean, the in-tile corrections and the
//    old and new values in FP64; the draw in single precision (trunc_normal_prec_f32), with the FP64
//    draw where that declines (bounds beyond 8 sigma on the far side of the mean);
//  * the corrections of the later tiles by a tile's deltas on the FP32 matrix cores;
//  * every redrawn value stored to S and to its FP32 copy.
// ---------------------------------------------------------------------------
#ifndef SMCF_MIXED_MT
#define SMCF_MIXED_MT 3
#define SMCF_MIXED_NB 3
#define SMCF_MIXED_OCC 3
#endif
constexpr int kMixedMT = SMCF_MIXED_MT;    // 16-row tiles per block
constexpr int kMixedNB = SMCF_MIXED_NB;    // register buffers of the matrix-phase pipeline
constexpr int kMixedOcc = SMCF_MIXED_OCC;  // waves per SIMD the register budget is sized for

// mfma_panel_t in FP32: acc[m][n] += sum_j A(16m + r, j) B(j, c_n), same operand addressing and pipeline
template <int MT, int NB>
__device__ __forceinline__ void mfma_panel_f32(const float* __restrict__ abase, const uint32_t (&ao)[MT],
                                               uint32_t lda4, const float* __restrict__ bbase,
                                               const uint32_t (&bo)[4], uint32_t ldb4, int nsteps,
                                               flt4 (&acc)[MT][4]) {
  if (nsteps <= 0) return;
  float x[NB][MT], y[NB][4];
  auto load = [&](int k, int st) {
    uint32_t q = (uint32_t)min(st, nsteps - 1);
#pragma unroll
    for (int m = 0; m < MT; ++m) x[k][m] = abase[ao[m] + q * lda4];
#pragma unroll
    for (int n = 0; n < 4; ++n) y[k][n] = bbase[bo[n] + q * ldb4];
  };
  auto prod = [&](int k) {
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int m = 0; m < MT; ++m)
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[k][m], y[k][n], acc[m][n], 0, 0, 0);
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

// one coordinate redraw in single precision; NaN where trunc_normal_prec_f32 declines
__device__ __noinline__ double gibbs_draw_f32(double mu, double s, double is, double max_slip, float u) {
  return trunc_normal_prec_f32(mu, s, is, 0.0, max_slip, u);
}

__device__ __forceinline__ void gibbs_mixed_body(int nd, int P, int b, int chunk, int sweeps, double max_slip,
                                                 const double* __restrict__ Qball, const float* __restrict__ Q32all,
                                                 double* __restrict__ S0, double* __restrict__ S1,
                                                 float* __restrict__ T0, float* __restrict__ T1,
                                                 const double* __restrict__ cdall, const uint64_t* __restrict__ keys,
                                                 const int* __restrict__ iters, const int* __restrict__ active,
                                                 const int* __restrict__ par) {
  constexpr int MT = kMixedMT, KB = 16 * MT;
  __shared__ double T[4][16][65];
  // per wave, the constants of the tile being drawn (read with wave-uniform addresses, so broadcast):
  //   CQ[kk][k] = Q(i0 + 16m + k, i0 + 16m + kk), the tile's diagonal block (in-tile corrections)
  //   CD[c][kk] = cd[c nd + i0 + 16m + kk]: gamma b_i, 1 / Q_ii, sqrt(Q_ii), 1 / sqrt(Q_ii)
  __shared__ double CQ[4][16][16];
  __shared__ double CD[4][4][16];
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int pw0 = chunk * 256 + wave * 64;
  int p = pw0 + lane;
  bool valid = p < P;
  int pc = valid ? p : P - 1;
  int pb = par[b];
  double* S = (pb ? S1 : S0) + (size_t)b * nd * P;
  float* S32 = (pb ? T1 : T0) + (size_t)b * nd * P;
  size_t qstride = (size_t)nd * ((nd + KB - 1) / KB) * KB;
  const double* Qb = Qball + (size_t)b * qstride;
  const float* Qb32 = Q32all + (size_t)b * qstride;
  // per-coordinate constants of the stage: gamma b_i, 1 / Q_ii, sqrt(Q_ii), 1 / sqrt(Q_ii)
  const double* cd = cdall + (size_t)b * 4 * nd;
  uint64_t key = keys[b];
  uint32_t stage = (uint32_t)(iters[b] - 1);
  int nblk = (nd + KB - 1) / KB;
  int fi = lane % 16, fk = lane / 16;
  int pcol[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) pcol[n] = min(pw0 + 16 * n + fi, P - 1);
  uint32_t bo[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) bo[n] = (uint32_t)(fk * P + pcol[n]);
  uint32_t ao[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) ao[m] = (uint32_t)(fk * KB + 16 * m + fi);
  int nd4 = nd & ~3;
  double (*Tw)[65] = T[wave];
  double (*CQw)[16] = CQ[wave];
  double (*CDw)[16] = CD[wave];
  for (int sw = 0; sw < sweeps; ++sw) {
    for (int blk = 0; blk < nblk; ++blk) {
      int i0 = blk * KB;
      const double* Qr = Qb + (size_t)blk * nd * KB;
      const float* Qr32 = Qb32 + (size_t)blk * nd * KB;
      flt4 acc[MT][4];
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n) acc[m][n] = flt4{0.f, 0.f, 0.f, 0.f};
      mfma_panel_f32<MT, kMixedNB>(Qr32, ao, (uint32_t)(4 * KB), S32, bo, (uint32_t)(4 * P), nd4 / 4, acc);
      if (nd4 < nd) {  // last rows (nd = 2 mod 4): rows past nd contribute zero
        bool in = nd4 + fk < nd;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          float bn = in ? S32[(size_t)(nd4 + fk) * P + pcol[n]] : 0.f;
#pragma unroll
          for (int m = 0; m < MT; ++m) {
            float a = in ? Qr32[(size_t)(nd4 + fk) * KB + 16 * m + fi] : 0.f;
            acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bn, acc[m][n], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        if (i0 + 16 * m >= nd) break;  // tiles past the last coordinate (wave-uniform)
        {  // the tile's constants into LDS: lane (k, c4) = (lane % 16, lane / 16) brings 4 + 1 values
          int ik = i0 + 16 * m + fi;
          const double* qsrc = Qr + (size_t)min(ik, nd - 1) * KB + 16 * m + 4 * fk;  // Q(ik, i0 + 16m + 4c4 + t)
#pragma unroll
          for (int t = 0; t < 4; ++t) CQw[4 * fk + t][fi] = ik < nd ? qsrc[t] : 0.0;
          CDw[fk][fi] = ik < nd ? cd[(size_t)fk * nd + ik] : 0.0;
        }
        // transpose tile m through LDS in double: Tw[row][particle]
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
          for (int t = 0; t < 4; ++t) Tw[4 * fk + t][16 * n + fi] = (double)acc[m][n][t];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint64_t wodd = 0;
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
          int i = i0 + 16 * m + kk;
          if (i >= nd) {
            Tw[kk][lane] = 0.0;
            continue;
          }
          // coordinates 2m' and 2m'+1 share one Philox block (word i & 1); i0 and the tile offset are even
          uint64_t wc;
          if ((kk & 1) == 0) {
            philox_draw(key, kSiteGibbsMove, stage, (uint32_t)pc, (uint32_t)(i >> 1), (uint32_t)sw, wc, wodd);
          } else {
            wc = wodd;
          }
          // left-looking corrections by the tile's coordinates redrawn before it (deltas in Tw[k][lane]):
          //   qi = T(kk) + Q(i0 + 16m + k, i) delta_k, k = 0 .. kk - 1, fused in ascending k (the operations and
          //   order of updating every later entry of the tile after each draw; Q read from row i0 + 16m + k)
          double qi = Tw[kk][lane];
#pragma unroll
          for (int k = 0; k < kk; ++k) qi = fma(CQw[kk][k], Tw[k][lane], qi);
          double so = S[(size_t)i * P + pc];
          double mu = so - (qi - CDw[0][kk]) * CDw[1][kk];
          double sn = gibbs_draw_f32(mu, CDw[2][kk], CDw[3][kk], max_slip, u01_open_f32(wc));
          if (sn != sn) sn = gibbs_draw(mu, CDw[2][kk], CDw[3][kk], max_slip, u01_open(wc));
          double dlt = sn - so;
          if (valid) {
            S[(size_t)i * P + p] = sn;
            S32[(size_t)i * P + p] = (float)sn;
          }
          Tw[kk][lane] = dlt;  // the delta, read by the later coordinates of the tile
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // corrections of the later tiles by this tile's deltas (in Tw[k][p]), FP32 matrix cores:
        //   acc[m2](r, p) += sum_k Q(i0 + 16m + k, i0 + 16m2 + r) delta_k(p)
#pragma unroll
        for (int m2 = m + 1; m2 < MT; ++m2) {
#pragma unroll
          for (int s4 = 0; s4 < 4; ++s4) {
            int kq = 4 * s4 + fk;
            float a = (i0 + 16 * m + kq < nd) ? Qr32[(size_t)(i0 + 16 * m + kq) * KB + 16 * m2 + fi] : 0.f;
#pragma unroll
            for (int n = 0; n < 4; ++n)
              acc[m2][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, (float)Tw[kq][16 * n + fi], acc[m2][n], 0, 0, 0);
          }
        }
        __builtin_amdgcn_wave_barrier();  // every lane has read the deltas before the next tile overwrites them
      }
    }
  }
}

// one workgroup per item (instance, 256-particle chunk) of the FP32 form
__global__ void __launch_bounds__(256, kMixedOcc) k_gibbs_mixed(const int4* __restrict__ items, GibbsArgs a) {
  int4 it = items[blockIdx.x];
  if (it.x < 0) return;
  const GibbsMember& M = a.m[it.x];
  if (!M.active[it.y]) return;
  gibbs_mixed_body(M.nd, M.P, it.y, it.z, M.sweeps, M.max_slip, M.Q, M.Q32, M.S, M.S2, M.S32, M.S32b, M.qdiag, M.key,
                   M.iter, M.active, M.par);
}


// ---------------------------------------------------------------------------
// Mixed-precision Gibbs sweeps with split-BF16 products (SMCF_PRECISION_MIXED_BF16).  The sweep of
// k_gibbs_mixed, with the products q = Q S formed on the BF16 matrix cores (v_mfma_f32_16x16x32_bf16):
// every operand is split as x = hi + lo with hi = bf16(x), lo = bf16(x - hi), and
//   Q S ~ Qhi Shi + Qhi Slo + Qlo Shi        (FP32 accumulation),
// which leaves about 2^-16 relative error per term (the dropped Qlo Slo and the split residues are
// below 2^-17 each) at three BF16 products per FP32-equivalent one, each 16x the FP32 matrix rate.
//  * Operands in matrix-core fragment order, 16 bytes per lane and part: QX holds, for block blk,
//    32-column step st, 16-row tile m and lane l = r + 16g, the 8 values Q(32 st + 8g + e, blk KB +
//    16m + r), e = 0..7 (hi, then lo); SX holds the particles as [j/8][p][hi 8 | lo 8], so lane
//    c + 16g of a particle tile reads coordinates 32 st + 8g .. +7 of particle p0 + c.  A and B
//    assign each contraction index to the same (lane group, element) slot, so the products do not
//    depend on the instruction's internal k order; lane l holds D(4 (l/16) + t, l % 16) in register t
//    (the FP32 accumulator layout of k_gibbs_mixed).
//  * Draw phase as in k_gibbs_mixed (FP64 conditional means and in-tile corrections, single-precision
//    draw); each lane reads its tile's 16 current values from SX (hi + lo, exact in double: the values
//    the matrix phase multiplied, so the diagonal term Q_ii s_i of q_i cancels exactly in the
//    conditional mean), keeps them in registers and writes the new ones to SX as two 16-value groups
//    at the end of the tile.  The FP64 particles S are written in the last sweep only: between sweeps
//    the chain's state is its split copy (16 significant bits, below the 2^-16 error the products
//    already carry), and S, SX and DX agree when the kernel ends.
//  * Corrections of the later tiles of the block by a tile's deltas: the same BF16 products with the
//    half of the 32-wide contraction that the tile does not cover set to zero.
// ---------------------------------------------------------------------------
#ifndef SMCF_BF_MT
#define SMCF_BF_MT 3
#define SMCF_BF_NB 2
#define SMCF_BF_OCC 2
#endif
constexpr int kBfMT = SMCF_BF_MT;    // 16-row tiles per block
constexpr int kBfNB = SMCF_BF_NB;    // register buffers of the matrix-phase pipeline
constexpr int kBfOcc = SMCF_BF_OCC;  // waves per SIMD the register budget is sized for

// acc[m][n] += sum over nst 32-wide steps of A_m B_n in split BF16.  A fragment (st, m) at
// qa[ao[m] + st astep] (hi) and +1 (lo); B fragment (st, n) at sb[bo[n] + st bstep] (hi) and +1 (lo);
// offsets in 16-byte units.  Same software pipeline as mfma_panel_f32.  Each load is a wave-uniform base
// for the step plus the lane's fixed 32-bit byte offset, the scalar-base form of the global load, so the
// steps spend no vector instructions on addresses.
template <int MT, int NB>
__device__ __forceinline__ void mfma_panel_bf3(const bf16x8* __restrict__ qa, const uint32_t (&ao)[MT], uint32_t astep,
                                               const bf16x8* __restrict__ sb, const uint32_t (&bo)[4],
                                               uint32_t bstep, int nst, flt4 (&acc)[MT][4]) {
  if (nst <= 0) return;
  bf16x8 ah[NB][MT], al[NB][MT], bh[NB][4], bl[NB][4];
  uint32_t aoff[MT], boff[4];  // byte offsets of the lane's fragments
#pragma unroll
  for (int m = 0; m < MT; ++m) aoff[m] = ao[m] * 16u;
#pragma unroll
  for (int n = 0; n < 4; ++n) boff[n] = bo[n] * 16u;
  auto load = [&](int k, int st) {
    uint32_t q = (uint32_t)min(st, nst - 1);
    const char* qs = (const char*)(qa + (size_t)q * astep);
    const char* ss = (const char*)(sb + (size_t)q * bstep);
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      ah[k][m] = *(const bf16x8*)(qs + aoff[m]);
      al[k][m] = *(const bf16x8*)(qs + aoff[m] + 16u);
    }
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      bh[k][n] = *(const bf16x8*)(ss + boff[n]);
      bl[k][n] = *(const bf16x8*)(ss + boff[n] + 16u);
    }
  };
  auto prod = [&](int k) {
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int m = 0; m < MT; ++m)
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[k][m], bh[k][n], acc[m][n], 0, 0, 0);
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int m = 0; m < MT; ++m)
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[k][m], bl[k][n], acc[m][n], 0, 0, 0);
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int m = 0; m < MT; ++m)
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[k][m], bh[k][n], acc[m][n], 0, 0, 0);
  };
#pragma unroll
  for (int k = 0; k < NB; ++k) load(k, k);
  int full = nst - nst % NB;
  for (int st = 0; st < full; st += NB) {
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      prod(k);
      load(k, st + k + NB);
    }
  }
#pragma unroll
  for (int k = 0; k < NB - 1; ++k)
    if (k < nst - full) prod(k);
}

template <int MT>
__device__ __forceinline__ void gibbs_bf_body(int nd, int P, int b, int chunk, int sweeps, double max_slip,
                                              const double* __restrict__ Qball, const bf16x8* __restrict__ QXall,
                                              double* __restrict__ S0, double* __restrict__ S1,
                                              bf16x8* __restrict__ X0, bf16x8* __restrict__ X1,
                                              bf16x8* __restrict__ DXall, const double* __restrict__ mrefall,
                                              const double* __restrict__ cdall, const uint64_t* __restrict__ keys,
                                              const int* __restrict__ iters, const int* __restrict__ active,
                                              const int* __restrict__ par) {
  constexpr int KB = 16 * MT;
  __shared__ double T[4][16][65];
  // per wave, the constants of the tile being drawn (read with wave-uniform addresses, so broadcast):
  //   CQ[kk][k] = Q(i0 + 16m + k, i0 + 16m + kk), the tile's diagonal block (in-tile corrections)
  //   CD[c][kk] = cd[c nd + i0 + 16m + kk]: gamma b_i, 1 / Q_ii, sqrt(Q_ii), 1 / sqrt(Q_ii); CD[4][kk] the
  //   reference particle mref(i0 + 16m + kk) of the centred likelihood
  __shared__ double CQ[4][16][16];
  __shared__ double CD[4][5][16];
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int pw0 = chunk * 256 + wave * 64;
  int p = pw0 + lane;
  bool valid = p < P;
  int pc = valid ? p : P - 1;
  int pb = par[b];
  int ndp = bf_rows(nd), nst = ndp / 32;
  int nblk = (nd + KB - 1) / KB;
  double* S = (pb ? S1 : S0) + (size_t)b * nd * P;
  bf16x8* SX = (pb ? X1 : X0) + (size_t)b * (ndp / 8) * P * 2;
  bf16x8* DX = DXall + (size_t)b * (ndp / 8) * P * 2;
  const double* mref = mrefall + (size_t)b * nd;
  const double* Qb = Qball + (size_t)b * nd * nblk * KB;
  const bf16x8* QX = QXall + (size_t)b * (bf_q_elems(nd, KB) / 8);
  // per-coordinate constants of the stage: gamma b_i, 1 / Q_ii, sqrt(Q_ii), 1 / sqrt(Q_ii)
  const double* cd = cdall + (size_t)b * 4 * nd;
  uint64_t key = keys[b];
  uint32_t stage = (uint32_t)(iters[b] - 1);
  int fi = lane % 16, fk = lane / 16;
  uint32_t bo[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) bo[n] = (uint32_t)((fk * P + min(pw0 + 16 * n + fi, P - 1)) * 2);
  uint32_t ao[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) ao[m] = (uint32_t)((m * 64 + lane) * 2);
  const uint32_t astep = MT * 128, bstep = 8 * (uint32_t)P;
  double (*Tw)[65] = T[wave];
  double (*CQw)[16] = CQ[wave];
  double (*CDw)[16] = CD[wave];
  const bf16x8 zero8 = {};
  for (int sw = 0; sw < sweeps; ++sw) {
    for (int blk = 0; blk < nblk; ++blk) {
      int i0 = blk * KB;
      const double* Qr = Qb + (size_t)blk * nd * KB;
      const bf16x8* QXb = QX + (size_t)blk * nst * MT * 128;
      flt4 acc[MT][4];
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n) acc[m][n] = flt4{0.f, 0.f, 0.f, 0.f};
      mfma_panel_bf3<MT, kBfNB>(QXb, ao, astep, SX, bo, bstep, nst, acc);
      // What a tile needs from memory is requested one tile ahead, while the previous tile is drawn: its
      // constants (this lane's 4 entries of the diagonal block, one per-coordinate constant, the reference
      // particle's entry) and its current values in split form.  The draw is an out-of-line call, whose
      // entry waits for every outstanding memory operation of the wave, so each tile then waits for
      // memory once instead of once per dependent request.
      double cq[4], cdv, mrv;
      bf16x8 xv[4];  // hi, lo of coordinate group (i0 + 16m) / 8, then of the next
      auto fetch = [&](int m) {
        int ik = i0 + 16 * m + fi;
        const double* qsrc = Qr + (size_t)min(ik, nd - 1) * KB + 16 * m + 4 * fk;  // Q(ik, i0 + 16m + 4c4 + t)
#pragma unroll
        for (int t = 0; t < 4; ++t) cq[t] = ik < nd ? qsrc[t] : 0.0;
        cdv = ik < nd ? cd[(size_t)fk * nd + ik] : 0.0;
        mrv = (fk == 0 && ik < nd) ? mref[ik] : 0.0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          size_t o = ((size_t)((i0 + 16 * m) / 8 + h) * P + pc) * 2;
          xv[2 * h] = SX[o];
          xv[2 * h + 1] = SX[o + 1];
        }
      };
      fetch(0);
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        if (i0 + 16 * m >= nd) break;  // tiles past the last coordinate (wave-uniform)
        // the tile's constants into LDS: lane (k, c4) = (lane % 16, lane / 16) brings 4 + 1 values
#pragma unroll
        for (int t = 0; t < 4; ++t) CQw[4 * fk + t][fi] = cq[t];
        CDw[fk][fi] = cdv;
        if (fk == 0) CDw[4][fi] = mrv;
        // the tile's 16 current values (hi + lo is exact in double)
        double sv[16];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int e = 0; e < 8; ++e) sv[8 * h + e] = (double)(float)xv[2 * h][e] + (double)(float)xv[2 * h + 1][e];
        // transpose tile m through LDS in double: Tw[row][particle]
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
          for (int t = 0; t < 4; ++t) Tw[4 * fk + t][16 * n + fi] = (double)acc[m][n][t];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the next tile's requests, and the Q fragments of this tile's corrections of the later tiles: the
        // tile's coordinates are columns 16 hh .. 16 hh + 15 of 32-wide step stc, held by lane groups 2 hh
        // and 2 hh + 1 of a fragment; the other two groups contribute zero
        if (m + 1 < MT && i0 + 16 * (m + 1) < nd) fetch(m + 1);
        const int jt = i0 + 16 * m, stc = jt / 32, hh = (jt / 16) & 1;
        const bool mine = (fk >> 1) == hh;
        bf16x8 cqh[MT], cql[MT];
#pragma unroll
        for (int m2 = m + 1; m2 < MT; ++m2) {
          uint32_t o = (uint32_t)(((stc * MT + m2) * 64 + lane) * 2);
          bool use = mine && i0 + 16 * m2 < nd;
          cqh[m2] = use ? QXb[o] : zero8;
          cql[m2] = use ? QXb[o + 1] : zero8;
        }
        uint64_t wnext = 0;  // the second word of the current Philox block
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
          int i = i0 + 16 * m + kk;
          if (i >= nd) {
            Tw[kk][lane] = 0.0;
            continue;
          }
          // coordinates 2m' and 2m'+1 share one Philox block (word i & 1); i0 and the tile offset are even
          uint64_t wc;
          if ((kk & 1) == 0) {
            philox_draw(key, kSiteGibbsMove, stage, (uint32_t)pc, (uint32_t)(i >> 1), (uint32_t)sw, wc, wnext);
          } else {
            wc = wnext;
          }
          // left-looking corrections by the tile's coordinates redrawn before it (deltas in Tw[k][lane]):
          //   qi = T(kk) + Q(i0 + 16m + k, i) delta_k, k = 0 .. kk - 1, fused in ascending k (the operations and
          //   order of updating every later entry of the tile after each draw; Q read from row i0 + 16m + k)
          double qi = Tw[kk][lane];
#pragma unroll
          for (int k = 0; k < kk; ++k) qi = fma(CQw[kk][k], Tw[k][lane], qi);
          double so = sv[kk];
          double mu = so - (qi - CDw[0][kk]) * CDw[1][kk];
          double sn = gibbs_draw_f32(mu, CDw[2][kk], CDw[3][kk], max_slip, u01_open_f32(wc));
          if (sn != sn) sn = gibbs_draw(mu, CDw[2][kk], CDw[3][kk], max_slip, u01_open(wc));
          double dlt = sn - so;
          sv[kk] = sn;  // the new value, stored with the tile's others below
          Tw[kk][lane] = dlt;  // the delta, read by the later coordinates of the tile
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the tile's new values into the split copy (coordinate groups (i0 + 16m) / 8 and the next) and, in
        // the last sweep, into S, stored together at the end of the tile rather than one after each draw
        if (valid) {
          if (sw == sweeps - 1) {
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
              if (i0 + 16 * m + kk < nd) S[(size_t)(i0 + 16 * m + kk) * P + p] = sv[kk];
          }
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            bf16x8 hi, lo;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
              __bf16 a, c;
              split_bf16((float)sv[8 * h + e], a, c);
              hi[e] = a;
              lo[e] = c;
            }
            size_t o = ((size_t)((i0 + 16 * m) / 8 + h) * P + p) * 2;
            SX[o] = hi;
            SX[o + 1] = lo;
          }
          if (sw == sweeps - 1) {  // the moved particles relative to the reference, for the likelihood
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              bf16x8 hi, lo;
#pragma unroll
              for (int e = 0; e < 8; ++e) {
                __bf16 a, c;
                split_bf16((float)(sv[8 * h + e] - CDw[4][8 * h + e]), a, c);
                hi[e] = a;
                lo[e] = c;
              }
              size_t o = ((size_t)((i0 + 16 * m) / 8 + h) * P + p) * 2;
              DX[o] = hi;
              DX[o + 1] = lo;
            }
          }
        }
        // corrections of the later tiles by this tile's deltas (in Tw[k][p]):
        //   acc[m2](r, p) += sum_k Q(i0 + 16m + k, i0 + 16m2 + r) delta_k(p)
        if (m + 1 < MT) {
          int kb8 = 8 * (fk - 2 * hh);
          bf16x8 dh[4], dl[4];
#pragma unroll
          for (int n = 0; n < 4; ++n) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
              __bf16 a, c;
              split_bf16(mine ? (float)Tw[kb8 + e][16 * n + fi] : 0.f, a, c);
              dh[n][e] = a;
              dl[n][e] = c;
            }
          }
#pragma unroll
          for (int m2 = m + 1; m2 < MT; ++m2) {
            if (i0 + 16 * m2 >= nd) break;
#pragma unroll
            for (int n = 0; n < 4; ++n) {
              acc[m2][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cqh[m2], dh[n], acc[m2][n], 0, 0, 0);
              acc[m2][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cqh[m2], dl[n], acc[m2][n], 0, 0, 0);
              acc[m2][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cql[m2], dh[n], acc[m2][n], 0, 0, 0);
            }
          }
        }
        __builtin_amdgcn_wave_barrier();  // every lane has read the deltas and values before the next tile
      }
    }
  }
}

// one workgroup per item (instance, 256-particle chunk) of the split-BF16 form's class with KB = 16 MT rows
template <int MT>
__global__ void __launch_bounds__(256, kBfOcc) k_gibbs_bf(const int4* __restrict__ items, GibbsArgs a) {
  int4 it = items[blockIdx.x];
  if (it.x < 0) return;
  const GibbsMember& M = a.m[it.x];
  if (!M.active[it.y]) return;
  gibbs_bf_body<MT>(M.nd, M.P, it.y, it.z, M.sweeps, M.max_slip, M.Q, (const bf16x8*)M.QX, M.S, M.S2,
                    (bf16x8*)M.SX, (bf16x8*)M.SXb, (bf16x8*)M.DX, M.mref, M.qdiag, M.key, M.iter, M.active, M.par);
}


// split-BF16 copy of the current particles of the given slots: [j/8][p][hi 8 | lo 8], zero past nd
__global__ void k_s_to_bf(int nd, int P, BatchDev bt, const int* __restrict__ sel) {
  int b = sel ? sel[blockIdx.y] : (int)blockIdx.y;
  int cur = bt.par[b];
  int ndp = bf_rows(nd);
  const double* S = (cur ? bt.S2 : bt.S) + (size_t)b * nd * P;
  bf16x8* X = (bf16x8*)(cur ? bt.SXb : bt.SX) + (size_t)b * (ndp / 8) * P * 2;
  size_t n = (size_t)(ndp / 8) * P;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    int p = (int)(e % P), g8 = (int)(e / P);
    bf16x8 hi, lo;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      int j = 8 * g8 + k;
      float x = j < nd ? (float)S[(size_t)j * P + p] : 0.f;
      __bf16 a, c;
      split_bf16(x, a, c);
      hi[k] = a;
      lo[k] = c;
    }
    X[2 * e] = hi;
    X[2 * e + 1] = lo;
  }
}

hipError_t launch_sbf(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel) {
  hipLaunchKernelGGL(k_s_to_bf, dim3(128, B), dim3(256), 0, s, m.ndim, bt.P, bt, sel);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Likelihood of the split-BF16 form (SMCF_PRECISION_MIXED_BF16), centred on a reference particle m of the
// stage (the first particle after resampling, FP64):
//   L(s) = L(m) + sum_i D_i (g_i + Y_i / 2),   D = s - m,  g = H m - b,  Y = H D,
// which is slip_calc_likelihood's (s^T H s - 2 b^T s + dsd) / 2 + c (smc_slip.f90:56-104) rearranged exactly.
// L(m) and g are FP64 (k_like_ref); D and the epilogue are FP64; only Y = H D is formed from split-BF16
// operands on the matrix cores (HX: H in fragment order, formed once per instance; DX: D, written by the last
// Gibbs sweep of the stage).  Centring keeps the products' 2^-16 relative error on the spread of the particle
// cloud, not on s itself: |error in L| ~ 2^-16 sum |D_i H_ij D_j|, which tempering multiplies by the stage's
// increment of gamma (the increment is chosen so that gamma times the spread of L is O(1)).
// ---------------------------------------------------------------------------
constexpr int kLikeBfMT = 4;  // 16-row tiles per block of H (64 rows)

// m = child 0 after the resampling (its parent's FP64 particle: the children's FP64 copy is written by the last
// sweep); g = H m - b; L(m) (one workgroup of GT threads per instance)
__global__ void __launch_bounds__(GT) k_like_ref(StageSet ss) {
  __shared__ double red[GT];
  int2 is = ss.inst[blockIdx.x];
  const MemberDev& M = ss.mem[is.x];
  const BatchDev& bt = M.bt;
  int b = is.y, nd = M.nd, P = M.P;
  if (!bt.HX || !bt.active[b]) return;
  const double* S = (bt.par[b] ? bt.S : bt.S2) + (size_t)b * nd * P + bt.parent[(size_t)b * P];
  const double* H = bt.H + (size_t)b * nd * nd;
  const double* bv = bt.bvec + (size_t)b * nd;
  double* mref = bt.mref + (size_t)b * nd;
  double* g = bt.gref + (size_t)b * nd;
  for (int i = threadIdx.x; i < nd; i += GT) mref[i] = S[(size_t)i * P];
  __syncthreads();
  double part = 0.0;
  for (int i = threadIdx.x; i < nd; i += GT) {
    double hm = 0.0;
    for (int j = 0; j < nd; ++j) hm = fma(H[(size_t)j * nd + i], mref[j], hm);  // (H m)_i = sum_j H(i, j) m_j
    g[i] = hm - bv[i];
    part = fma(mref[i], 0.5 * hm - bv[i], part);
  }
  double tot = block_sum<GT>(part, red);
  if (threadIdx.x == 0) {
    const double* sc = bt.scal + 8 * (size_t)b;
    bt.lref[b] = tot + bt.dsd[b] / 2.0 + (sc[3] + sc[4]);
  }
}

// HX: split-BF16 copy of H in fragment order, fragment f = ((blk nst + st) MT + m) 64 + lane of lane = r + 16g
// holding H(blk KB + 16m + r, 32 st + 8g + e), e = 0..7 (hi at HX[2f], lo at HX[2f + 1]; zero past nd)
__global__ void k_h_to_bf(int nd, const double* __restrict__ Hall, bf16x8* __restrict__ HXall,
                          const int* __restrict__ sel) {
  constexpr int MT = kLikeBfMT, KB = 16 * MT;
  int b = sel ? sel[blockIdx.y] : (int)blockIdx.y;
  int nblk = (nd + KB - 1) / KB, nst = bf_rows(nd) / 32;
  size_t nfrag = (size_t)nblk * nst * MT * 64;
  const double* H = Hall + (size_t)b * nd * nd;
  bf16x8* HX = HXall + (size_t)b * (bf_q_elems(nd, KB) / 8);
  for (size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x; f < nfrag; f += (size_t)gridDim.x * blockDim.x) {
    int lane = (int)(f % 64);
    size_t rest = f / 64;
    int m = (int)(rest % MT);
    rest /= MT;
    int st = (int)(rest % nst), blk = (int)(rest / nst);
    int i = blk * KB + 16 * m + lane % 16, g = lane / 16;
    bf16x8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      int j = 32 * st + 8 * g + e;
      float x = (i < nd && j < nd) ? (float)H[(size_t)j * nd + i] : 0.f;
      __bf16 a, c;
      split_bf16(x, a, c);
      hi[e] = a;
      lo[e] = c;
    }
    HX[2 * f] = hi;
    HX[2 * f + 1] = lo;
  }
}

// One workgroup per 64 particles of an instance (XCD-aware items as k_like_mfma): its four wavefronts take the
// 64-row blocks of H round-robin (wave w: blocks w, w + 4, ...) for the same 64 particles and step through the
// contraction together, so each 32-row step of the particles' split copy D is loaded from memory once per
// workgroup into LDS (double-buffered, one barrier per step) and read from there by the four wavefronts; H's
// fragments come from memory per wavefront.  The products accumulate in the same order as a wavefront alone would
// (per step hi.hi, hi.lo, lo.hi), each wavefront sums its blocks' rows in ascending order, and the four partial
// sums of a particle are added in wavefront order.
__global__ void __launch_bounds__(256, 2)
    k_like_bf(const int4* __restrict__ items, const MemberDev* __restrict__ mem) {
  constexpr int MT = kLikeBfMT, KB = 16 * MT;
  __shared__ bf16x8 Dt[2][4][64][2];  // [buffer][row group of the step][particle][hi, lo]
  __shared__ double red[4][64];
  int4 it = items[blockIdx.x];
  if (it.x < 0) return;
  const MemberDev& M = mem[it.x];
  const BatchDev& bt = M.bt;
  int b = it.y, chunk = it.z, nd = M.nd, P = M.P;
  if (!bt.active[b]) return;
  int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  int p0 = chunk * 64;
  if (p0 >= P) return;
  int pb = bt.par[b];
  int ndp = bf_rows(nd), nst = ndp / 32, nblk = (nd + KB - 1) / KB;
  const double* S = (pb ? bt.S2 : bt.S) + (size_t)b * nd * P;
  const bf16x8* DX = (const bf16x8*)bt.DX + (size_t)b * (ndp / 8) * P * 2;
  const bf16x8* HX = (const bf16x8*)bt.HX + (size_t)b * (bf_q_elems(nd, KB) / 8);
  const double* mref = bt.mref + (size_t)b * nd;
  const double* g = bt.gref + (size_t)b * nd;
  int fi = lane % 16, fk = lane / 16;
  int pcol[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) pcol[n] = min(p0 + 16 * n + fi, P - 1);
  // this thread's share of a step's D tile: row group wave, particle p0 + lane (clamped; unused past P)
  const bf16x8* dsrc = DX + ((size_t)wave * P + min(p0 + lane, P - 1)) * 2;
  const size_t dstep = (size_t)4 * P * 2;
  uint32_t aoff[MT];
#pragma unroll
  for (int m = 0; m < MT; ++m) aoff[m] = (uint32_t)((m * 64 + lane) * 2) * 16u;
  double part[4] = {0.0, 0.0, 0.0, 0.0};
  const int npass = (nblk + 3) / 4;
  for (int q = 0; q < npass; ++q) {
    const int blk = 4 * q + wave;
    const bool mine = blk < nblk;  // wave-uniform; a wave without a block still loads and waits with the others
    const char* qa = (const char*)(HX + (size_t)min(blk, nblk - 1) * nst * MT * 128);
    flt4 acc[MT][4];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int n = 0; n < 4; ++n) acc[m][n] = flt4{0.f, 0.f, 0.f, 0.f};
    bf16x8 dh = dsrc[0], dl = dsrc[1];
    // H's fragments of the step, requested one step ahead (a wave without a block reads its clamped block's)
    bf16x8 ah[MT], al[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      ah[m] = *(const bf16x8*)(qa + aoff[m]);
      al[m] = *(const bf16x8*)(qa + aoff[m] + 16u);
    }
    __syncthreads();  // the previous pass has finished reading both buffers
    Dt[0][wave][lane][0] = dh;
    Dt[0][wave][lane][1] = dl;
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
      const int cur = st & 1;
      const bool more = st + 1 < nst;
      bf16x8 nah[MT], nal[MT];
      if (more) {
        dh = dsrc[(size_t)(st + 1) * dstep];
        dl = dsrc[(size_t)(st + 1) * dstep + 1];
        const char* qn = qa + (size_t)(st + 1) * MT * 128 * 16;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          nah[m] = *(const bf16x8*)(qn + aoff[m]);
          nal[m] = *(const bf16x8*)(qn + aoff[m] + 16u);
        }
      }
      if (mine) {
        bf16x8 bh[4], bl[4];
#pragma unroll
        for (int n = 0; n < 4; ++n) {
          bh[n] = Dt[cur][fk][16 * n + fi][0];
          bl[n] = Dt[cur][fk][16 * n + fi][1];
        }
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
          for (int m = 0; m < MT; ++m)
            acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[m], bh[n], acc[m][n], 0, 0, 0);
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
          for (int m = 0; m < MT; ++m)
            acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[m], bl[n], acc[m][n], 0, 0, 0);
#pragma unroll
        for (int n = 0; n < 4; ++n)
#pragma unroll
          for (int m = 0; m < MT; ++m)
            acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[m], bh[n], acc[m][n], 0, 0, 0);
      }
      if (more) {
        Dt[cur ^ 1][wave][lane][0] = dh;
        Dt[cur ^ 1][wave][lane][1] = dl;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          ah[m] = nah[m];
          al[m] = nal[m];
        }
      }
      __syncthreads();
    }
    if (mine) {
      const int i0 = blk * KB;
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          int row = i0 + 16 * m + 4 * fk + t;
          if (row < nd) {
            double mr = mref[row], gr = g[row];
#pragma unroll
            for (int n = 0; n < 4; ++n) {
              double d = S[(size_t)row * P + pcol[n]] - mr;
              part[n] = fma(d, fma(0.5, (double)acc[m][n][t], gr), part[n]);
            }
          }
        }
    }
  }
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    double v = part[n];
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    if (fk == 0) red[wave][16 * n + fi] = v;
  }
  __syncthreads();
  if (threadIdx.x < 64 && p0 + (int)threadIdx.x < P) {
    int t = threadIdx.x;
    double* Lout = (pb ? bt.like2 : bt.like) + (size_t)b * P;
    Lout[p0 + t] = bt.lref[b] + (((red[0][t] + red[1][t]) + red[2][t]) + red[3][t]);
  }
}

hipError_t launch_h_bf(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel) {
  size_t nfrag = bf_q_elems(m.ndim, 16 * kLikeBfMT) / 16;
  int blocks = (int)std::min<size_t>((nfrag + 255) / 256, 1024);
  hipLaunchKernelGGL(k_h_to_bf, dim3(blocks, B), dim3(256), 0, s, m.ndim, bt.H, (bf16x8*)bt.HX, sel);
  return hipGetLastError();
}

hipError_t launch_like_ref(const StageSet& ss, hipStream_t s) {
  hipLaunchKernelGGL(k_like_ref, dim3(ss.ninst), dim3(GT), 0, s, ss);
  return hipGetLastError();
}

hipError_t launch_like_bf(const StageSet& ss, const ItemList& it, hipStream_t s) {
  if (it.n > 0) hipLaunchKernelGGL(k_like_bf, dim3(it.n), dim3(256), 0, s, it.items, ss.mem);
  return hipGetLastError();
}

// Block rows of the Gibbs sweep kernels per (form, ndim): the FP64 and split-BF16 kernels come in 16-, 32- and (split
// BF16) 48-row variants.  The rows change the order in which a conditional mean's terms are added (the block's own
// earlier coordinates enter as corrections), so the choice comes from a fixed table, not from timing at run time:
// a problem's results are reproducible.  The table is the output of the bench's --autotune mode on the MI355X
// (fastest variant per ndim range); smcf_set_gibbs_block_rows replaces entries for problems created afterwards.
namespace {
struct BlockRule {
  int precision, ndim_lo, ndim_hi, rows;
};
std::mutex g_tile_mu;
// measured by the bench's --autotune mode: 16 rows win up to ndim 112 in both forms, 32 (FP64) and 48 (split
// BF16) from ndim 264 on; the boundary between the measured sizes is put at 160
std::vector<BlockRule> g_tile_rules = {
    {SMCF_PRECISION_FP64, 1, 160, 16},
    {SMCF_PRECISION_FP64, 161, 1 << 30, kGibbsKB},
    {SMCF_PRECISION_MIXED_BF16, 1, 160, 16},
    {SMCF_PRECISION_MIXED_BF16, 161, 1 << 30, 16 * kBfMT},
};
}  // namespace

int gibbs_kb(int precision, int nd) {
  if (precision == SMCF_PRECISION_MIXED) return 16 * kMixedMT;
  std::lock_guard<std::mutex> lk(g_tile_mu);
  for (auto it = g_tile_rules.rbegin(); it != g_tile_rules.rend(); ++it)  // the latest matching rule
    if (it->precision == precision && nd >= it->ndim_lo && nd <= it->ndim_hi) return it->rows;
  return precision == SMCF_PRECISION_MIXED_BF16 ? 16 * kBfMT : kGibbsKB;
}

bool gibbs_rows_valid(int precision, int rows) {
  if (precision == SMCF_PRECISION_FP64) return rows == 16 || rows == 32;
  if (precision == SMCF_PRECISION_MIXED_BF16) return rows == 16 || rows == 32 || rows == 48;
  return false;
}

void set_gibbs_rows(int precision, int lo, int hi, int rows) {
  std::lock_guard<std::mutex> lk(g_tile_mu);
  g_tile_rules.push_back({precision, lo, hi, rows});
}

hipError_t launch_gibbs(const GibbsArgs& a, const ItemList& it, int precision, int kb, hipStream_t s) {
  if (it.n < 1) return hipSuccess;
  dim3 g(it.n);
  if (precision == SMCF_PRECISION_MIXED_BF16) {
    auto k = kb == 16 ? k_gibbs_bf<1> : kb == 32 ? k_gibbs_bf<2> : kb == 48 ? k_gibbs_bf<3> : nullptr;
    if (!k) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, g, dim3(256), 0, s, it.items, a);
  } else if (precision == SMCF_PRECISION_MIXED) {
    if (kb != 16 * kMixedMT) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gibbs_mixed, g, dim3(256), 0, s, it.items, a);
  } else {
    auto k = kb == 16 ? k_gibbs<1> : kb == 32 ? k_gibbs<2> : nullptr;
    if (!k) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, g, dim3(256), 0, s, it.items, a);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// standalone component kernels (single problem, for parity checks)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(GT) k_find_next_gamma(int P, double gprev, const double* __restrict__ L,
                                                        double cv, double tol, int cap, double* __restrict__ out3,
                                                        double* __restrict__ w) {
  extern __shared__ double smem[];
  double* Ls = smem;
  double* ws = Ls + P;
  double* red = ws + P;
  for (int i = threadIdx.x; i < P; i += GT) Ls[i] = L[i];
  __syncthreads();
  double nev;
  double g = gamma_search(Ls, ws, P, gprev, cv, tol, cap, red, &nev);
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += GT) w[i] = ws[i];
  if (threadIdx.x == 0) { out3[0] = g; out3[1] = nev; }
}

hipError_t launch_find_next_gamma(int P, double gprev, const double* L, double cv, double tol, int cap,
                                  double* out3, double* w, hipStream_t s) {
  size_t sh = sizeof(double) * (2 * (size_t)P + GT);
  hipLaunchKernelGGL(k_find_next_gamma, dim3(1), dim3(GT), sh, s, P, gprev, L, cv, tol, cap, out3, w);
  return hipGetLastError();
}

__global__ void __launch_bounds__(GT) k_normalize_resample(int P, const double* __restrict__ w, double uh, double ul,
                                                           double* __restrict__ wn, int* __restrict__ assigned,
                                                           int* __restrict__ parent) {
  extern __shared__ double smem[];
  double* ws = smem;
  double* red = ws + P;
  dd* scan = (dd*)(red + GT);
  for (int i = threadIdx.x; i < P; i += GT) ws[i] = w[i];
  __syncthreads();
  normalize_resample(ws, P, two_sum(uh, ul), assigned, parent, red, scan);
  for (int i = threadIdx.x; i < P; i += GT) wn[i] = ws[i];
}

hipError_t launch_normalize_resample(int P, const double* w, double uh, double ul, double* wn, int* assigned,
                                     hipStream_t s) {
  size_t sh = sizeof(double) * ((size_t)P + GT) + sizeof(dd) * (GT + 1);
  int* parent = assigned + P;
  hipLaunchKernelGGL(k_normalize_resample, dim3(1), dim3(GT), sh, s, P, w, uh, ul, wn, assigned, parent);
  return hipGetLastError();
}

__global__ void k_trunc_normal(int n, const double* mu, const double* sg, double lo, double hi, const double* u,
                               double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = trunc_normal(mu[i], sg[i], lo, hi, u[i]);
}
__global__ void k_trunc_normal_f32(int n, const double* mu, const double* sq, double lo, double hi, const float* u,
                                   double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = trunc_normal_prec_f32(mu[i], sq[i], 1.0 / sq[i], lo, hi, u[i]);
}

hipError_t launch_trunc_normal_f32(int n, const double* mu, const double* sq, double lo, double hi, const float* u,
                                   double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_trunc_normal_f32, dim3((n + 255) / 256), dim3(256), 0, s, n, mu, sq, lo, hi, u, out);
  return hipGetLastError();
}

hipError_t launch_trunc_normal(int n, const double* mu, const double* sig, double lo, double hi, const double* u,
                               double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_trunc_normal, dim3((n + 255) / 256), dim3(256), 0, s, n, mu, sig, lo, hi, u, out);
  return hipGetLastError();
}

__global__ void k_philox(int n, const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t* c = ctr + 4 * i;
  uint64_t k = ((uint64_t)key[2 * i + 1] << 32) | key[2 * i];
  uint64_t w0, w1;
  philox_draw(k, c[3] >> 24, c[3] & 0xFFFFFFu, c[2], c[1], c[0], w0, w1);
  out[4 * i] = (uint32_t)w0;
  out[4 * i + 1] = (uint32_t)(w0 >> 32);
  out[4 * i + 2] = (uint32_t)w1;
  out[4 * i + 3] = (uint32_t)(w1 >> 32);
}
hipError_t launch_philox(int n, const uint32_t* ctr, const uint32_t* key, uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_philox, dim3((n + 255) / 256), dim3(256), 0, s, n, ctr, key, out);
  return hipGetLastError();
}

}  // namespace smcf
