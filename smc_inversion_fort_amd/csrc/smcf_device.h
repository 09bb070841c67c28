// Device helpers shared by the kernels of libsmcf: keyed counter-based RNG,
// double-double arithmetic and the truncated-normal inversion of the Gibbs mover.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SMCF_DEV __device__ __forceinline__

namespace smcf {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  A draw is addressed by
//   key     = 64-bit instance key
//   counter = (c0, c1, c2, site << 24 | stage)
// and yields two 64-bit words.  Every random number of the inner and outer SMC is
// one such draw, so results do not depend on launch order or thread count.
// ---------------------------------------------------------------------------
enum Site : uint32_t {
  kSiteGibbsInit = 1,
  kSiteResample = 2,
  kSiteTuneDtau = 3,
  kSiteTuneMom = 4,
  kSiteTuneMH = 5,
  kSiteTuneResample = 6,
  kSiteHmcMom = 7,
  kSiteHmcMH = 8,
  kSiteGibbsMove = 9,
  kSiteOuterInit = 16,
  kSiteOuterResample = 17,
  kSiteOuterProp = 18,
  kSiteOuterMH = 19,
};

SMCF_DEV void philox_draw(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2, uint32_t c1,
                          uint32_t c0, uint64_t& w0, uint64_t& w1) {
  uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = (site << 24) | (stage & 0xFFFFFFu);
  uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32 x 32 -> 64 product per multiplier (a single v_mad_u64_u32 each, rather than mul_lo + mul_hi)
    uint64_t p0 = (uint64_t)0xD2511F53u * x0, p1 = (uint64_t)0xCD9E8D57u * x2;
    uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    uint32_t y0 = hi1 ^ x1 ^ k0, y2 = hi0 ^ x3 ^ k1;
    x0 = y0; x1 = lo1; x2 = y2; x3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  w0 = ((uint64_t)x1 << 32) | x0;
  w1 = ((uint64_t)x3 << 32) | x2;
}

SMCF_DEV double u01(uint64_t w) { return (double)(w >> 11) * 0x1.0p-53; }
// strictly inside (0, 1): used by the inversion sampler
SMCF_DEV double u01_open(uint64_t w) { return ((double)(w >> 11) + 0.5) * 0x1.0p-53; }
// reference clamp of the intrinsic uniform at 1e-5 (smc_slip.f90:8-15)
SMCF_DEV double clamp_u(double u) { return u < 1e-5 ? 1e-5 : u; }
// Box-Muller, sine branch (smc_slip.f90:26-38)
SMCF_DEV double box_muller_sin(double r, double s) {
  const double pi = 3.141592653589793;  // 2*asin(1) in double
  return sqrt(-2.0 * log(r)) * sin(2.0 * pi * s);
}
SMCF_DEV double draw_normal(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2, uint32_t c1) {
  uint64_t w0, w1;
  philox_draw(key, site, stage, c2, c1, 0, w0, w1);
  return box_muller_sin(clamp_u(u01(w0)), clamp_u(u01(w1)));
}

// ---------------------------------------------------------------------------
// double-double (Dekker / Knuth error-free transformations)
// ---------------------------------------------------------------------------
struct dd {
  double hi, lo;
};
SMCF_DEV dd two_sum(double a, double b) {
#pragma clang fp contract(off)
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}
SMCF_DEV dd quick_two_sum(double a, double b) {
#pragma clang fp contract(off)
  double s = a + b;
  return {s, b - (s - a)};
}
SMCF_DEV dd dd_add(dd a, dd b) {
#pragma clang fp contract(off)
  dd s = two_sum(a.hi, b.hi);
  dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}
SMCF_DEV dd dd_add_d(dd a, double b) {
#pragma clang fp contract(off)
  dd s = two_sum(a.hi, b);
  s.lo += a.lo;
  return quick_two_sum(s.hi, s.lo);
}
SMCF_DEV dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
SMCF_DEV dd dd_mul_d(dd a, double b) {
#pragma clang fp contract(off)
  double p = a.hi * b;
  double e = fma(a.hi, b, -p);
  e = fma(a.lo, b, e);
  return quick_two_sum(p, e);
}
SMCF_DEV dd dd_div_d(dd a, double b) {
#pragma clang fp contract(off)
  double q1 = a.hi / b;
  dd r = dd_add(a, dd_neg(dd_mul_d(dd{q1, 0.0}, b)));
  double q2 = r.hi / b;
  return quick_two_sum(q1, q2);
}
// round(x / s) for double x and double-double s.  The residual x - q1 * s.hi is exact in one FMA
// (q1 = fl(x / s.hi)); it is written with explicit FMAs so that contraction cannot change it.
SMCF_DEV double d_div_dd(double x, dd s) {
#pragma clang fp contract(off)
  double q1 = x / s.hi;
  double r = fma(-q1, s.hi, x);
  r = fma(-q1, s.lo, r);
  return q1 + r / s.hi;
}
SMCF_DEV double dd_floor(dd a) {
  double f = floor(a.hi);
  if (f == a.hi && a.lo < 0.0) f -= 1.0;
  return f;
}
SMCF_DEV bool dd_lt_d(dd a, double b) { return a.hi < b || (a.hi == b && a.lo < 0.0); }

// ---------------------------------------------------------------------------
// Standard normal quantile, Wichura (1988) AS 241 PPND16 (relative error ~1e-16)
// ---------------------------------------------------------------------------
// AS 241: central region |q| <= 0.425, tail regions with r = sqrt(-log(min(p, 1 - p))) <= 5 and
// r > 5 (p < 1.4e-11).  The central rational is evaluated on every lane; the log / sqrt and the
// tail rational only in waves with a tail lane, and the far tail in its own (rare) branch.
SMCF_DEV double ndtri_central(double q) {
  double r = 0.180625 - q * q;
  return q * (((((((2509.0809287301226727 * r + 33430.575583588128105) * r + 67265.770927008700853) * r +
                  45921.953931549871457) * r + 13731.693765509461125) * r + 1971.5909503065514427) * r +
               133.14166789178437745) * r + 3.387132872796366608) /
         (((((((5226.495278852545925 * r + 28729.085735721942674) * r + 39307.89580009271061) * r +
              21213.794301586595867) * r + 5394.1960214247511077) * r + 687.1870074920579083) * r +
           42.313330701600911252) * r + 1.0);
}
SMCF_DEV double ndtri_tail(double r) {  // r <= 5
  r -= 1.6;
  return (((((((7.7454501427834140764e-4 * r + 0.0227238449892691845833) * r + 0.24178072517745061177) * r +
              1.27045825245236838258) * r + 3.64784832476320460504) * r + 5.7694972214606914055) * r +
           4.6303378461565452959) * r + 1.42343711074968357734) /
         (((((((1.05075007164441684324e-9 * r + 5.475938084995344946e-4) * r + 0.0151986665636164571966) * r +
              0.14810397642748007459) * r + 0.68976733498510000455) * r + 1.6763848301838038494) * r +
           2.05319162663775882187) * r + 1.0);
}
__device__ __noinline__ double ndtri_far_tail(double r) {  // r > 5
  r -= 5.0;
  return (((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r + 0.0012426609473880784386) * r +
              0.026532189526576123093) * r + 0.29656057182850489123) * r + 1.7848265399172913358) * r +
           5.4637849111641143699) * r + 6.6579046435011037772) /
         (((((((2.04426310338993978564e-15 * r + 1.4215117583164458887e-7) * r + 1.8463183175100546818e-5) * r +
              7.868691311456132591e-4) * r + 0.0148753612908506148525) * r + 0.13692988092273580531) * r +
           0.59983220655588793769) * r + 1.0);
}
SMCF_DEV double ndtri(double p) {
  double q = p - 0.5;
  double z = ndtri_central(q);
  if (fabs(q) > 0.425) {
    double r = sqrt(-log(q < 0 ? p : 1.0 - p));
    double t = r <= 5.0 ? ndtri_tail(r) : ndtri_far_tail(r);
    z = q < 0 ? -t : t;
  }
  return z;
}

// Standard normal truncated to [a, b] by inversion; the upper tail works with
// complementary probabilities so that far-tail truncation keeps full precision.
// Mills ratio series R(x) x = 1 - x^-2 + 3x^-4 - ... and log Q(x), valid for x >= 35
SMCF_DEV double mills_series(double x) {
  double r = 1.0 / (x * x);
  return 1.0 + r * (-1.0 + r * (3.0 + r * (-15.0 + r * (105.0 + r * (-945.0 + r * (10395.0 + r * -135135.0))))));
}
SMCF_DEV double logq_far(double x) { return -0.5 * x * x - log(x) - 0.91893853320467274178 + log(mills_series(x)); }
// beyond 35 sigma: Newton on log Q(z) = log Q(a) + log(1 - u (1 - Q(b)/Q(a)))
// (out of line: rarely taken, and it would otherwise dominate the draw's code and registers)
__device__ __noinline__ double tn_far(double a, double b, double u) {
  double la = logq_far(a), lb = logq_far(b);
  double T = la + log1p(-u * (-expm1(lb - la)));
  double z = a - log1p(-u * (-expm1(-a * (b - a)))) / a;
#pragma unroll
  for (int it = 0; it < 4; ++it) {
    double g = logq_far(z) - T;
    z = z + g * (mills_series(z) / z);
  }
  return fmin(fmax(z, a), b);
}
// Standard normal truncated to [a, b] by inversion.  One erfc per bound gives t = Q(|x|), from
// which Phi(x) and Q(x) follow without cancellation; a bound whose tail is below 1e-20 of the
// interval's mass is not evaluated.  p = Phi(a) + u m and q = Q(b) + (1 - u) m are both formed
// directly and the smaller is inverted, so every lane runs the same straight-line path except for
// the Mills-ratio solve beyond 35 sigma.
SMCF_DEV double tn_std(double a, double b, double u) {
  if (a > 35.0) return tn_far(a, b, u);
  if (b < -35.0) return -tn_far(-b, -a, 1.0 - u);  // mirror: quantile u of z is 1-u of -z
  const double r2 = 0.70710678118654752440;
  double ta = (a <= -10.0 && a * a > b * b + 100.0) ? 0.0 : 0.5 * erfc(fabs(a) * r2);
  double tb = (b >= 10.0 && b * b > a * a + 100.0) ? 0.0 : 0.5 * erfc(fabs(b) * r2);
  double pa = a < 0.0 ? ta : 1.0 - ta, qa = a < 0.0 ? 1.0 - ta : ta;
  double pb = b < 0.0 ? tb : 1.0 - tb, qb = b < 0.0 ? 1.0 - tb : tb;
  double m = a >= 0.0 ? qa - qb : pb - pa;
  double p = pa + u * m, q = qb + (1.0 - u) * m;
  double z = ndtri(fmin(p, q));
  z = p <= q ? z : -z;
  return fmin(fmax(z, a), b);
}
SMCF_DEV double trunc_normal(double mu, double sigma, double lo, double hi, double u) {
  double a = (lo - mu) / sigma, b = (hi - mu) / sigma;
  double x = mu + sigma * tn_std(a, b, u);
  return fmin(fmax(x, lo), hi);
}
// the Gibbs mover's parameterisation: s = sqrt(precision), is = 1 / s
SMCF_DEV double trunc_normal_prec(double mu, double s, double is, double lo, double hi, double u) {
  double a = (lo - mu) * s, b = (hi - mu) * s;
  double x = mu + tn_std(a, b, u) * is;
  return fmin(fmax(x, lo), hi);
}

// ---------------------------------------------------------------------------
// Single-precision truncated-normal draw of the mixed-precision Gibbs mover.  The conditional mean
// and the returned value stay double; the standardised bounds, the two tail masses and the inverse
// CDF are single precision.  Inverse CDF: Giles' single-precision erfinv polynomial for p > 0.0017, and
// below that the tail of AS 241 PPND7 (Wichura 1988, Appl. Statist. 37, 477-484), the 7-digit version of
// the algorithm used in double precision above, its quotients formed with the hardware reciprocal (1 ulp,
// below the 7-digit accuracy of the rational forms).
SMCF_DEV float ndtri7_lower(float p) {  // p in (0, 1/2]: z <= 0 with Phi(z) = p
  // p > 0.0017 (w = -ln(4 p (1 - p)) < 5): z = sqrt(2) erfinv(2p - 1) from the central polynomial of Giles'
  // single-precision erfinv (M. Giles, "Approximating the erfinv function", GPU Computing Gems Jade Edition,
  // 2011; relative error below 1.3e-7), one branch where AS 241 needs two (p < 0.075 is its tail), so fewer
  // waves run both.  2p - 1 is exact for p >= 1/4.
  float w = -0.69314718056f * __builtin_amdgcn_logf(4.0f * p * (1.0f - p));
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
  // p <= 0.0017: AS 241 PPND7's tail (r = sqrt(-ln p) > 2.5)
  // r = sqrt(-ln p) from the hardware log2 and square root (1 ulp each): p >= 2^-126 wherever a draw is
  // accepted (the interval's mass is at least 1e-30), so neither needs the library's denormal scaling
  float r = __builtin_amdgcn_sqrtf(-0.69314718056f * __builtin_amdgcn_logf(p));
  float z;
  if (r <= 5.0f) {
    r -= 1.6f;
    z = (((0.17023821103f * r + 1.3067284816f) * r + 2.7568153900f) * r + 1.4234372777f) *
        __builtin_amdgcn_rcpf((0.12021132975f * r + 0.73700164250f) * r + 1.0f);
  } else {
    r -= 5.0f;
    z = (((0.017337203997f * r + 0.42868294337f) * r + 3.0812263860f) * r + 6.6579051150f) *
        __builtin_amdgcn_rcpf((0.012258202635f * r + 0.24197894225f) * r + 1.0f);
  }
  return -z;
}
// e^y for y <= 1 in single precision: y log2(e) carried as ph + pl (pl from an FMA and the low part of
// log2(e)), e = rint(ph), then 2^((ph - e) + pl) from the hardware exp2 (1 ulp on [-1/2, 1/2]) scaled by
// 2^e.  Results below 2^-149 flush to zero.  The library expf adds overflow and underflow tests that
// cannot trigger for y <= 1.
SMCF_DEV float exp_le1_f32(float y) {
  const float l2e_hi = 1.44269502163f, l2e_lo = 1.92596299e-8f;
  float ph = y * l2e_hi;
  float pl = fmaf(y, l2e_hi, -ph) + y * l2e_lo;
  float e = __builtin_rintf(ph);
  return __builtin_ldexpf(__builtin_amdgcn_exp2f((ph - e) + pl), (int)e);
}
// erfc(x) for x >= 0 in single precision: the Chebyshev-fitted form of Numerical Recipes in C (2nd ed.,
// section 6.2, erfcc), t exp(-x^2 + c(t)) with t = 1 / (1 + x/2) and c a degree-9 polynomial in t,
// fractional error below 1.2e-7 for every x >= 0 (one reciprocal, one exponential, nine FMAs).
// c(t) <= 3e-8 on (0, 1] (its largest value, at t = 1, where erfc(0) = 1), so the exponent is below 1.
SMCF_DEV float erfc_cheb32(float x) {
  float t = __builtin_amdgcn_rcpf(1.0f + 0.5f * x);
  float c = -1.26551223f +
            t * (1.00002368f +
                 t * (0.37409196f +
                      t * (0.09678418f +
                           t * (-0.18628806f +
                                t * (0.27886807f +
                                     t * (-1.13520398f + t * (1.48851587f + t * (-0.82215223f + t * 0.17087277f))))))));
  return t * exp_le1_f32(c - x * x);
}
// uniform strictly inside (0, 1) from the top 23 bits h of a Philox word: (h + 1/2) 2^-23, exact in single
// precision.  Formed as the float with mantissa h and exponent 0, minus 1, plus 2^-24 (each step exact),
// which avoids the compiler's 64-bit integer conversion sequence.
SMCF_DEV float u01_open_f32(uint64_t w) {
  uint32_t h = (uint32_t)(w >> 41);
  return (__uint_as_float(0x3F800000u | h) - 1.0f) + 0x1.0p-24f;
}

// Same construction as tn_std: p = Phi(a) + u m and q = Q(b) + (1 - u) m, the smaller inverted.  NaN
// (the caller then takes the double-precision draw) when a bound lies beyond 8 sigma on the far side of
// the mean, or when the interval holds less than 1e-30 of the mass.
SMCF_DEV double trunc_normal_prec_f32(double mu, double s, double is, double lo, double hi, float u) {
  double a64 = (lo - mu) * s, b64 = (hi - mu) * s;
  if (a64 > 8.0 || b64 < -8.0) return __builtin_nan("");
  float a = (float)fmax(a64, -40.0), b = (float)fmin(b64, 40.0);
  const float r2 = 0.70710678118654752f;
  // A bound 10 or more standard deviations from the mean holds below 7.7e-24 of the mass beyond it, under
  // 2^-24 of the interval's mass (the other bound lies within 8 sigma, so the interval holds 6.2e-16 or
  // more): its tail is taken as zero, and a wave whose bounds all lie that far skips the erfc.
  float ta = 0.0f, tb = 0.0f;
  if (fabsf(a) < 10.0f) ta = 0.5f * erfc_cheb32(fabsf(a) * r2);
  if (fabsf(b) < 10.0f) tb = 0.5f * erfc_cheb32(fabsf(b) * r2);
  float pa = a < 0.0f ? ta : 1.0f - ta, qa = a < 0.0f ? 1.0f - ta : ta;
  float pb = b < 0.0f ? tb : 1.0f - tb, qb = b < 0.0f ? 1.0f - tb : tb;
  float m = a >= 0.0f ? qa - qb : pb - pa;
  if (!(m >= 1e-30f)) return __builtin_nan("");
  float p = pa + u * m, q = qb + (1.0f - u) * m;
  float z = ndtri7_lower(fminf(p, q));
  z = p <= q ? z : -z;
  z = fminf(fmaxf(z, a), b);
  double x = mu + (double)z * is;
  return fmin(fmax(x, lo), hi);
}

// cdf_norm / pdf_norm of smc_slip.f90:40-54
SMCF_DEV double cdf_norm(double x, double mu, double sigma2) {
  return (1.0 + erf((x - mu) / sqrt(2.0 * sigma2))) / 2.0;
}
SMCF_DEV double pdf_norm(double x, double mu, double sigma2) {
  const double pi = 3.141592653589793;
  return exp(-((x - mu) * (x - mu)) / (2.0 * sigma2)) / sqrt(2.0 * pi * sigma2);
}

// ---------------------------------------------------------------------------
// block reductions (deterministic: fixed tree for a fixed block size)
// ---------------------------------------------------------------------------
template <int NT>
SMCF_DEV double block_sum(double v, double* red) {
  int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) red[t] = red[t] + red[t + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}
template <int NT>
SMCF_DEV double block_min(double v, double* red) {
  int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) red[t] = fmin(red[t], red[t + s]);
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// FP64 matrix-core panel shared by the Gibbs, likelihood and HMC kernels
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// x = hi + lo with hi = bf16(x), lo = bf16(x - hi): the operands of the split-BF16 products
SMCF_DEV void split_bf16(float x, __bf16& h, __bf16& l) {
  h = (__bf16)x;
  l = (__bf16)(x - (float)h);
}

#ifndef SMCF_PANEL_BUFFERS
#define SMCF_PANEL_BUFFERS 2
#endif
constexpr int kPanelBuffers = SMCF_PANEL_BUFFERS;

// One block of MT 16-row tiles of a product Y = A B on the matrix cores (v_mfma_f64_16x16x4f64), 64 columns
// per wavefront:  acc[m][n] += sum_j A(16m + r, j) B(j, c_n) over j = 0 .. 4 nsteps - 1, where lane
// (r, k) = (lane % 16, lane / 16) reads A(16m + r, 4st + k) at abase[ao[m] + st lda4] and B(4st + k, c_n) at
// bbase[bo[n] + st ldb4].  Bases are wave-uniform and offsets 32-bit, so the loads take the scalar-base +
// 32-bit-offset form.  The loop is software-pipelined over NB register buffers: the operands of step st + NB
// are requested as soon as the products of step st are issued.  The main loop runs whole groups of NB steps and
// has no branches, so the wait before each group's products covers only that buffer's loads; the last
// nsteps % NB steps follow it.  (The compiler does not unroll a loop holding MFMA instructions when that needs a
// remainder loop, so a plain loop would wait for its own loads on every step.)
template <int MT, int NB>
__device__ __forceinline__ void mfma_panel_t(const double* __restrict__ abase, const uint32_t (&ao)[MT], uint32_t lda4,
                                           const double* __restrict__ bbase, const uint32_t (&bo)[4], uint32_t ldb4,
                                           int nsteps, dbl4 (&acc)[MT][4]) {
  if (nsteps <= 0) return;
  double x[NB][MT], y[NB][4];
  // loads of step st (clamped to the last step: a group's prefetch past the end re-reads it, unused)
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

}  // namespace smcf
