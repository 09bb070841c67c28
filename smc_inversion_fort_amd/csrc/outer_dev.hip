// Outer reweight and resample of one outer stage on the device (smc_fault.f90:372-637), one workgroup:
//
//   k_outer_stage  tempering: bisection on gamma for CV 0.2 (tol 1e-8, at most 1000 iterations; L - q with
//                  the gsmc reference density), the stage's -log evidence, normalised weights, weighted mean
//                  and lower covariance x 0.04 with its Cholesky factor, and the systematic residual
//                  resampling counts (fault_find_next_gamma, fault_normalize_weights, fault_calc_mean_particles,
//                  fault_calc_cov_particles, fault_resample_particles)
//   k_outer_sort   a stable LSD radix sort of the parents by descending count (4-bit digits), the device
//                  form of the qsort_ call of fault_reorder_to_send (sort.cpp:3-7)
//   k_outer_reorder  particles, likelihood, q and counts into the sorted order
//
// Thread t owns the contiguous particles [t C, (t + 1) C), so per-thread sums and prefix sums run in ascending
// particle order and the block reductions combine them in a fixed tree: results do not depend on timing.  The
// reference sums sequentially in FP64; these sums differ from it by rounding only.  The resampling counts follow
// the closed form A_i = floor(N (c_i - u0)) + 1 of the real(16) running u, with the prefix sums c_i in
// double-double (as the inner resampling).  The radix sort is stable (ties in ascending particle index), where
// std::sort leaves tied parents in its own order: the same moves, a different order of tied parents.
#include "smcf_device.h"
#include "smcf_internal.h"

namespace smcf {

constexpr int OT = 1024;

// exclusive prefix over the block of one int per thread; *total receives the sum
SMCF_DEV int block_exscan(int v, int* tmp, int* total) {
  int t = threadIdx.x;
  tmp[t] = v;
  __syncthreads();
  for (int s = 1; s < OT; s <<= 1) {
    int a = t >= s ? tmp[t - s] : 0;
    __syncthreads();
    tmp[t] += a;
    __syncthreads();
  }
  int incl = tmp[t];
  *total = tmp[OT - 1];
  __syncthreads();
  return incl - v;
}

__global__ void __launch_bounds__(OT) k_outer_stage(int N, int nd, double gprev, const double* __restrict__ L,
                                                    const double* __restrict__ q, const double* __restrict__ X,
                                                    uint64_t okey, int stage, double* __restrict__ w,
                                                    double* __restrict__ scal, double* __restrict__ cov,
                                                    int* __restrict__ assigned) {
  __shared__ double red[OT];
  __shared__ dd scan[OT + 1];
  __shared__ double mean[kMaxOuterDim];
  int t = threadIdx.x, C = (N + OT - 1) / OT, i0 = min(N, t * C), i1 = min(N, i0 + C);
  auto R = [&](int i) { return q ? L[i] - q[i] : L[i]; };
  double lm = i0 < i1 ? R(i0) : R(0);
  for (int i = i0; i < i1; ++i) lm = fmin(lm, R(i));
  double lmin = block_min<OT>(lm, red);
  // bisection on gamma (fault_find_next_gamma, smc_fault.f90:392-485)
  double lower = gprev, upper = 1.0, err = 1.0, gamma = 1.0;
  int cnt = 0;
  while (err > 1e-8) {
    gamma = (lower + upper) / 2.0;
    double dg = gamma - gprev, s = 0.0;
    for (int i = i0; i < i1; ++i) {
      double v = exp(-dg * (R(i) - lmin));
      w[i] = v;
      s = s + v;
    }
    double m = block_sum<OT>(s, red) / N;
    double s2 = 0.0;
    for (int i = i0; i < i1; ++i) s2 = s2 + (w[i] - m) * (w[i] - m);
    double cv = sqrt(block_sum<OT>(s2, red) / N) / m;
    if (cv > 0.2) upper = gamma; else lower = gamma;
    err = fabs(cv - 0.2);
    if (fabs(gamma - 1.0) < 1e-8) break;
    if (cnt > 1000) break;
    cnt++;
  }
  double dg = gamma - gprev, s = 0.0;
  for (int i = i0; i < i1; ++i) {
    double v = exp(-dg * (R(i) - lmin));
    w[i] = v;
    s = s + v;
  }
  double wsum = block_sum<OT>(s, red);
  double nev = -log(wsum / N) + dg * lmin;
  // normalised weights (fault_normalize_weights)
  for (int i = i0; i < i1; ++i) w[i] = w[i] / wsum;
  __syncthreads();
  // weighted mean and lower covariance x 0.04 (fault_calc_mean_particles, fault_calc_cov_particles)
  for (int d = 0; d < nd; ++d) {
    double a = 0.0;
    for (int i = i0; i < i1; ++i) a = a + w[i] * X[(size_t)i * nd + d];
    double v = block_sum<OT>(a, red);
    if (t == 0) mean[d] = v;
  }
  __syncthreads();
  for (int j = 0; j < nd; ++j)
    for (int r = j; r < nd; ++r) {
      double a = 0.0;
      for (int i = i0; i < i1; ++i)
        a = a + w[i] * (X[(size_t)i * nd + r] - mean[r]) * (X[(size_t)i * nd + j] - mean[j]);
      double v = block_sum<OT>(a, red);
      if (t == 0) {
        cov[j * nd + r] = v * 0.04;
        if (r != j) cov[r * nd + j] = 0.0;
      }
    }
  __syncthreads();
  if (t == 0) {  // unblocked lower Cholesky; info = the first non-positive pivot (dpotrf's), 0 if none
    int info = 0;
    for (int j = 0; j < nd; ++j) {
      double ajj = cov[j * nd + j];
      for (int k = 0; k < j; ++k) ajj -= cov[k * nd + j] * cov[k * nd + j];
      if (ajj <= 0.0 || ajj != ajj) {
        cov[j * nd + j] = ajj;
        info = j + 1;
        break;
      }
      ajj = sqrt(ajj);
      cov[j * nd + j] = ajj;
      for (int r = j + 1; r < nd; ++r) {
        double tt = cov[j * nd + r];
        for (int k = 0; k < j; ++k) tt -= cov[k * nd + r] * cov[k * nd + j];
        cov[j * nd + r] = tt * (1.0 / ajj);
      }
    }
    scal[0] = gamma;
    scal[1] = nev;
    scal[2] = (double)info;
  }
  // systematic residual resampling (fault_resample_particles): u from the stage's keyed draw, clamped at 1e-5
  dd c = {0.0, 0.0};
  for (int i = i0; i < i1; ++i) c = dd_add_d(c, w[i]);
  scan[t] = c;
  __syncthreads();
  if (t == 0) {
    dd run = {0.0, 0.0};
    for (int k = 0; k < OT; ++k) {
      dd v = scan[k];
      scan[k] = run;
      run = dd_add(run, v);
    }
  }
  __syncthreads();
  uint64_t w0, w1;
  philox_draw(okey, kSiteOuterResample, (uint32_t)stage, 0, 0, 0, w0, w1);
  dd u = two_sum((double)(w0 >> 11) * 0x1.0p-53, (double)(w1 >> 11) * 0x1.0p-106);
  if (dd_lt_d(u, 1e-5)) u = dd{1e-5, 0.0};
  dd u0 = dd_div_d(u, (double)N);
  dd cum = scan[t];
  int Aprev = 0;
  if (i0 > 0) Aprev = max(0, min(N, (int)dd_floor(dd_mul_d(dd_add(cum, dd_neg(u0)), (double)N)) + 1));
  for (int i = i0; i < i1; ++i) {
    cum = dd_add_d(cum, w[i]);
    int A = (int)dd_floor(dd_mul_d(dd_add(cum, dd_neg(u0)), (double)N)) + 1;
    A = max(Aprev, min(N, A));
    assigned[i] = A - Aprev;
    Aprev = A;
  }
  if (i1 == N && i0 < N && Aprev < N) assigned[N - 1] += N - Aprev;  // rounding shortfall at the very end
}

// stable LSD radix sort of 0 .. N-1 by descending assigned count (key = amax - count, 4-bit digits), one
// workgroup; idx receives the order, tmp is N ints of scratch
__global__ void __launch_bounds__(OT) k_outer_sort(int N, const int* __restrict__ assigned, int* __restrict__ idx,
                                                   int* __restrict__ tmp) {
  __shared__ int sc[OT];
  __shared__ int amax_s;
  int t = threadIdx.x, C = (N + OT - 1) / OT, i0 = min(N, t * C), i1 = min(N, i0 + C);
  int am = 0;
  for (int i = i0; i < i1; ++i) {
    am = max(am, assigned[i]);
    idx[i] = i;
  }
  sc[t] = am;
  __syncthreads();
  for (int s = OT / 2; s > 0; s >>= 1) {
    if (t < s) sc[t] = max(sc[t], sc[t + s]);
    __syncthreads();
  }
  if (t == 0) amax_s = sc[0];
  __syncthreads();
  int amax = amax_s;
  int* src = idx;
  int* dst = tmp;
  for (int shift = 0; (amax >> shift) > 0; shift += 4) {
    int cnt[16];
#pragma unroll
    for (int d = 0; d < 16; ++d) cnt[d] = 0;
    for (int i = i0; i < i1; ++i) cnt[((amax - assigned[src[i]]) >> shift) & 15]++;
    int off[16], base = 0;
#pragma unroll
    for (int d = 0; d < 16; ++d) {
      int tot;
      off[d] = base + block_exscan(cnt[d], sc, &tot);
      base += tot;
    }
    for (int i = i0; i < i1; ++i) {
      int k = src[i];
      int d = ((amax - assigned[k]) >> shift) & 15;
      dst[off[d]++] = k;
    }
    __syncthreads();
    int* x = src;
    src = dst;
    dst = x;
  }
  if (src != idx)
    for (int i = i0; i < i1; ++i) idx[i] = src[i];
}

// particles, likelihood, q and counts into the order idx; sorted_idx (1-based, qsort_'s convention)
__global__ void k_outer_reorder(int N, int nd, const int* __restrict__ idx, const int* __restrict__ a_in,
                                const double* __restrict__ X_in, const double* __restrict__ L_in,
                                const double* __restrict__ q_in, int* __restrict__ a_out, double* __restrict__ X_out,
                                double* __restrict__ L_out, double* __restrict__ q_out, int* __restrict__ sorted_idx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int o = idx[i];
  a_out[i] = a_in[o];
  for (int d = 0; d < nd; ++d) X_out[(size_t)i * nd + d] = X_in[(size_t)o * nd + d];
  L_out[i] = L_in[o];
  if (q_in) q_out[i] = q_in[o];
  sorted_idx[i] = o + 1;
}

hipError_t launch_outer_stage(int N, int nd, double gprev, const double* L, const double* q, const double* X,
                              uint64_t okey, int stage, double* w, double* scal, double* cov, int* assigned,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_outer_stage, dim3(1), dim3(OT), 0, s, N, nd, gprev, L, q, X, okey, stage, w, scal, cov,
                     assigned);
  return hipGetLastError();
}

hipError_t launch_outer_sort(int N, int nd, const int* assigned, int* idx, int* tmp, const double* X_in,
                             const double* L_in, const double* q_in, int* a_out, double* X_out, double* L_out,
                             double* q_out, int* sorted_idx, hipStream_t s) {
  hipLaunchKernelGGL(k_outer_sort, dim3(1), dim3(OT), 0, s, N, assigned, idx, tmp);
  hipLaunchKernelGGL(k_outer_reorder, dim3((N + 255) / 256), dim3(256), 0, s, N, nd, idx, assigned, X_in, L_in, q_in,
                     a_out, X_out, L_out, q_out, sorted_idx);
  return hipGetLastError();
}

}  // namespace smcf
