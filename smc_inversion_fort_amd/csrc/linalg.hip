// Dense FP64 linear algebra of the inner SMC, batched over geometries.
//
//  k_scal      theta -> alpha^2, sigma^2 and the likelihood constant (init.f90:175-196,
//              smc_fault.f90:148-173, smc_slip.f90:84-87)
//  k_sar_weight  SAR rows of Sigma^-1 G: (1/sigma_sar^2) Sinv_p^T G_p per path (smc_slip.f90:1390-1398)
//  k_sd        sd = Sigma^-1 d (SAR: column-oriented Sinv d; GNSS: d / (sigma_gnss^2 sigma_i^2))
//                                                                    (smc_slip.f90:1365-1377)
//  k_gram      H = (Sigma^-1 G)^T G                                  (smc_slip.f90:1390-1409)
//  k_gram_vec  b = G^T sd, dsd = d^T sd                              (smc_slip.f90:1379-1388)
//  k_prior     ELL prior per particle                                (smc_slip.f90:106-137)
//  k_form_q    Q = gamma H + LL / alpha^2 for the Gibbs mover
// The two GEMMs use 64x64 LDS tiles with a 4x4 register block per thread.
#include <algorithm>

#include "smcf_device.h"
#include "smcf_internal.h"

namespace smcf {

__global__ void k_scal(MeshDev m, const double* __restrict__ theta, double* __restrict__ scal, int B,
                       const int* __restrict__ sel) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  if (sel) b = sel[b];
  const double* th = theta + (size_t)b * m.ndim_fault;
  double la2, lss2, lsg2;
  if (m.fix_xbend) {
    la2 = th[2 * m.nplane + 1]; lss2 = th[2 * m.nplane + 2]; lsg2 = th[2 * m.nplane + 3];
  } else {
    la2 = th[3 * m.nplane]; lss2 = th[3 * m.nplane + 1]; lsg2 = th[3 * m.nplane + 2];
  }
  double* s = scal + 8 * (size_t)b;
  s[0] = exp(la2);
  s[1] = lss2;
  s[2] = lsg2;
  s[3] = m.nsar * lss2 / 2.0;
  s[4] = 3.0 * m.ngnss * lsg2 / 2.0;
  s[5] = exp(lsg2);
  s[6] = exp(lss2);
  s[7] = 0.0;
}

constexpr int TM = 64, TN = 64, TK = 16;

// SAR rows of the weighted design matrix for path p (rows r0 .. r0+n-1):
//   Gw(i, r0 + c) = (1 / sigma_sar^2) sum_l G(r0 + l, i) Sinv_p(l, c)
// (the reference's gtsmat columns of the path, dgemm('t', 'n'), smc_slip.f90:1390-1398), as a tiled GEMM over
// (i, c) with K = l; Gw is stored like G, column-major over the nsar SAR rows.
__global__ void __launch_bounds__(256) k_sar_weight(MeshDev m, const double* __restrict__ Gall,
                                                    const double* __restrict__ scal, double* __restrict__ Gwall,
                                                    int r0, int n, long off, const int* __restrict__ sel) {
  __shared__ double As[TK][TM + 1];
  __shared__ double Bs[TK][TN + 1];
  int b = sel ? sel[blockIdx.z] : (int)blockIdx.z;
  int i0 = blockIdx.x * TM, c0 = blockIdx.y * TN;
  int nobs = m.nobs, nd = m.ndim, nsar = m.nsar;
  const double* G = Gall + (size_t)b * nobs * nd;
  const double* A = m.sinv + off;
  double rs2 = 1.0 / scal[8 * (size_t)b + 6];
  int t = threadIdx.x, ty = t / 16, tx = t % 16;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
  for (int k0 = 0; k0 < n; k0 += TK) {
#pragma unroll
    for (int rep = 0; rep < 4; ++rep) {
      int kk = t % 16, ii = t / 16 + rep * 16;
      int l = k0 + kk;
      int i = i0 + ii, c = c0 + ii;
      As[kk][ii] = (l < n && i < nd) ? G[(size_t)i * nobs + r0 + l] : 0.0;
      Bs[kk][ii] = (l < n && c < n) ? A[(size_t)c * n + l] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      double a4[4], b4[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) a4[a] = As[kk][ty * 4 + a];
#pragma unroll
      for (int c = 0; c < 4; ++c) b4[c] = Bs[kk][tx * 4 + c];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] += a4[a] * b4[c];
    }
    __syncthreads();
  }
  double* Gw = Gwall + (size_t)b * nsar * nd;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int i = i0 + ty * 4 + a, cc = c0 + tx * 4 + c;
      if (i < nd && cc < n) Gw[(size_t)i * nsar + r0 + cc] = rs2 * acc[a][c];
    }
}

// sd = Sigma^-1 d per observation row: SAR rows of path p as dgemv('n') with alpha = 1/sigma_sar^2
// accumulated column by column (sd(r0 + i) += (alpha d(r0 + j)) Sinv_p(i, j), j ascending); GNSS rows
// d / (sigma_gnss^2 sigma_i^2) (smc_slip.f90:1365-1377)
__global__ void k_sd(MeshDev m, const double* __restrict__ scal, double* __restrict__ sdall,
                     const int* __restrict__ sel) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int b = sel ? sel[blockIdx.y] : (int)blockIdx.y;
  if (k >= m.nobs) return;
  double v;
  if (k < m.nsar) {
    int p = m.row_path[k];
    int r0 = m.sar_r0[p], n = m.sar_n[p], i = k - r0;
    const double* A = m.sinv + m.sar_off[p];
    double alpha = 1.0 / scal[8 * (size_t)b + 6];
    v = 0.0;
    for (int j = 0; j < n; ++j) v += (alpha * m.dvec[r0 + j]) * A[(size_t)j * n + i];
  } else {
    double sk = m.obs_sigma[k];
    v = m.dvec[k] / (scal[8 * (size_t)b + 5] * (sk * sk));
  }
  sdall[(size_t)b * m.nobs + k] = v;
}

// H(i, j) = sum_k W(k, i) G(k, j) with W = Sigma^-1 G: SAR rows from Gw, GNSS rows G(k, i) / c_k,
// c_k = obs_sigma_k^2 * sigma_gnss^2
__global__ void __launch_bounds__(256) k_gram(MeshDev m, const double* __restrict__ Gall,
                                              const double* __restrict__ Gwall, const double* __restrict__ scal,
                                              double* __restrict__ Hall, const int* __restrict__ sel) {
  __shared__ double As[TK][TM + 1];
  __shared__ double Bs[TK][TN + 1];
  int b = sel ? sel[blockIdx.z] : (int)blockIdx.z;
  int i0 = blockIdx.x * TM, j0 = blockIdx.y * TN;
  int nobs = m.nobs, nd = m.ndim;
  const double* G = Gall + (size_t)b * nobs * nd;
  const double* Gw = Gwall ? Gwall + (size_t)b * m.nsar * nd : nullptr;
  double sg2 = scal[8 * (size_t)b + 5];
  int t = threadIdx.x, ty = t / 16, tx = t % 16;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
  for (int k0 = 0; k0 < nobs; k0 += TK) {
#pragma unroll
    for (int rep = 0; rep < 4; ++rep) {
      int kk = t % 16, ii = t / 16 + rep * 16;
      int k = k0 + kk;
      int i = i0 + ii, j = j0 + ii;
      double av = 0.0, bv = 0.0;
      if (k < nobs) {
        if (k < m.nsar) {
          if (i < nd) av = Gw[(size_t)i * m.nsar + k];
        } else {
          double sk = m.obs_sigma[k];
          double ck = (sk * sk) * sg2;
          if (i < nd) av = G[(size_t)i * nobs + k] / ck;
        }
        if (j < nd) bv = G[(size_t)j * nobs + k];
      }
      As[kk][ii] = av;
      Bs[kk][ii] = bv;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      double a4[4], b4[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) a4[a] = As[kk][ty * 4 + a];
#pragma unroll
      for (int c = 0; c < 4; ++c) b4[c] = Bs[kk][tx * 4 + c];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] += a4[a] * b4[c];
    }
    __syncthreads();
  }
  double* H = Hall + (size_t)b * nd * nd;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int i = i0 + ty * 4 + a, j = j0 + tx * 4 + c;
      if (i < nd && j < nd) H[(size_t)j * nd + i] = acc[a][c];
    }
}

__global__ void k_gram_vec(MeshDev m, const double* __restrict__ Gall, const double* __restrict__ sdall,
                           double* __restrict__ bvec, double* __restrict__ dsd, int B, const int* __restrict__ sel) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = sel ? sel[blockIdx.y] : (int)blockIdx.y;
  int nobs = m.nobs, nd = m.ndim;
  const double* sd = sdall + (size_t)b * nobs;
  if (i < nd) {
    const double* g = Gall + ((size_t)b * nd + i) * nobs;
    double s = 0.0;
    for (int k = 0; k < nobs; ++k) s = s + g[k] * sd[k];
    bvec[(size_t)b * nd + i] = s;
  } else if (i == nd) {
    double s = 0.0;
    for (int k = 0; k < nobs; ++k) s = s + m.dvec[k] * sd[k];
    dsd[b] = s;
  }
}

hipError_t launch_gram(const Problem& P, const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel,
                       bool scal_given) {
  if (!scal_given) hipLaunchKernelGGL(k_scal, dim3((B + 63) / 64), dim3(64), 0, s, m, bt.theta, bt.scal, B, sel);
  for (int p = 0; p < m.npath && m.nsar > 0; ++p) {
    int n = P.sar_n[p];
    if (n == 0) continue;
    dim3 gw((m.ndim + TM - 1) / TM, (n + TN - 1) / TN, B);
    hipLaunchKernelGGL(k_sar_weight, gw, dim3(256), 0, s, m, bt.G, bt.scal, bt.Gw, P.sar_r0[p], n, P.sar_off[p],
                       sel);
  }
  hipLaunchKernelGGL(k_sd, dim3((m.nobs + 127) / 128, B), dim3(128), 0, s, m, bt.scal, bt.sdv, sel);
  dim3 g((m.ndim + TM - 1) / TM, (m.ndim + TN - 1) / TN, B);
  hipLaunchKernelGGL(k_gram, g, dim3(256), 0, s, m, bt.G, m.nsar > 0 ? bt.Gw : nullptr, bt.scal, bt.H, sel);
  hipLaunchKernelGGL(k_gram_vec, dim3((m.ndim + 1 + 127) / 128, B), dim3(128), 0, s, m, bt.G, bt.sdv, bt.bvec,
                     bt.dsd, B, sel);
  return hipGetLastError();
}

// prior of the current particles (slip_calc_prior, smc_slip.f90:106-137): sum over the rows of L s
// (ELL, 9 entries per row) of (L s)^2 / (2 alpha^2)
__global__ void k_prior(MeshDev m, int P, const double* __restrict__ S0, const double* __restrict__ S1,
                        const int* __restrict__ par, const double* __restrict__ scal, const int* __restrict__ liall,
                        const double* __restrict__ lvall, double* __restrict__ R0, double* __restrict__ R1,
                        const int* __restrict__ sel) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  int b = sel ? sel[blockIdx.y] : (int)blockIdx.y;
  if (p >= P) return;
  int nd = m.ndim;
  int pb = par[b];
  const double* S = (pb ? S1 : S0) + (size_t)b * nd * P;
  const double* sc = scal + 8 * (size_t)b;
  const int* li = liall + (size_t)b * 2 * m.nnode * 9;
  const double* lv = lvall + (size_t)b * 2 * m.nnode * 9;
  double two_a2 = 2.0 * sc[0];
  double pr = 0.0;
  for (int r = 0; r < 2 * m.nnode; ++r) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) t = t + lv[9 * r + q] * S[(size_t)(li[9 * r + q] - 1) * P + p];
    pr = pr + (t * t) / two_a2;
  }
  (pb ? R1 : R0)[(size_t)b * P + p] = pr;
}

hipError_t launch_prior(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel) {
  int P = bt.P;
  hipLaunchKernelGGL(k_prior, dim3((P + 127) / 128, B), dim3(128), 0, s, m, P, bt.S, bt.S2, bt.par, bt.scal, bt.li,
                     bt.lv, bt.prior, bt.prior2, sel);
  return hipGetLastError();
}

// Gibbs precision Q = gamma H + LL / alpha^2, written in the blocked layout of the Gibbs
// kernel: Qb[(blk*nd + j)*KB + k] = Q(j, blk*KB + k) (zero past the last column).

__global__ void k_form_qb(StageSet ss) {
  int2 is = ss.inst[blockIdx.y];
  const MemberDev& M = ss.mem[is.x];
  const BatchDev& bt = M.bt;
  int b = is.y, nd = M.nd, KB = M.kb;
  if (!bt.Q || bt.QX || !bt.active[b]) return;  // split-BF16 members: k_form_qx
  int nblk = (nd + KB - 1) / KB;
  size_t n = (size_t)nblk * nd * KB;
  double g = bt.gamma[b];
  const double* H = bt.H + (size_t)b * nd * nd;
  double* Qb = bt.Q + (size_t)b * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    int k = (int)(e % KB);
    size_t rest = e / KB;
    int j = (int)(rest % nd), blk = (int)(rest / nd);
    int i = blk * KB + k;
    Qb[e] = i < nd ? g * H[(size_t)i * nd + j] : 0.0;
  }
}

// one thread per (column i, slot w) of the sparse columns of LL: the slot's entry (row r) of column i
__global__ void k_form_qb_sparse(StageSet ss) {
  int2 is = ss.inst[blockIdx.y];
  const MemberDev& M = ss.mem[is.x];
  const BatchDev& bt = M.bt;
  int b = is.y, nd = M.nd, KB = M.kb;
  int t = blockIdx.x * blockDim.x + threadIdx.x, i = t / 25;
  if (!bt.Q || bt.QX || !bt.active[b] || i >= nd) return;  // split-BF16 members: k_form_qx_sparse
  int r = bt.llc_idx[(size_t)b * nd * 25 + t];
  if (r < 0) return;  // padding (after the column's entries)
  int nblk = (nd + KB - 1) / KB;
  double g = bt.gamma[b], a2 = bt.scal[8 * (size_t)b];
  const double* H = bt.H + (size_t)b * nd * nd;
  double* Qb = bt.Q + (size_t)b * nblk * nd * KB;
  Qb[((size_t)(i / KB) * nd + r) * KB + i % KB] = g * H[(size_t)i * nd + r] + bt.llc_val[(size_t)b * nd * 25 + t] / a2;
}

// Split-BF16 members: the stage's precision straight into the fragment order of the sweep kernel's products
// (fragment f = ((blk nst + st) MT + m) 64 + lane of lane = r + 16g holds Q(32 st + 8g + e, blk KB + 16m + r),
// e = 0..7, hi at QX[2f], lo at QX[2f + 1]), each value the same double as in the FP64 blocked layout rounded to
// single and split.  k_form_qx writes gamma H; k_form_qx_sparse then rewrites the entries of the sparse columns of
// LL (gamma H + LL / alpha^2), as k_form_qb and k_form_qb_sparse do for the FP64 layout.  Of the FP64 blocked layout
// only the diagonal KB x KB blocks, which the sweep's in-tile corrections and the draw constants read, are written.
SMCF_DEV void qx_store(bf16x8* __restrict__ QX, int KB, int nst, int i, int j, double v) {
  int blk = i / KB, k = i % KB;
  int lane = k % 16 + 16 * ((j % 32) / 8);
  size_t f = (((size_t)blk * nst + j / 32) * (KB / 16) + k / 16) * 64 + lane;
  __bf16 h, l;
  split_bf16((float)v, h, l);
  ((__bf16*)(QX + 2 * f))[j % 8] = h;
  ((__bf16*)(QX + 2 * f + 1))[j % 8] = l;
}

__global__ void k_form_qx(StageSet ss) {
  int2 is = ss.inst[blockIdx.y];
  const MemberDev& M = ss.mem[is.x];
  const BatchDev& bt = M.bt;
  int b = is.y, nd = M.nd, KB = M.kb;
  if (!bt.QX || !bt.active[b]) return;
  int MT = KB / 16, nblk = (nd + KB - 1) / KB, nst = bf_rows(nd) / 32;
  size_t nfrag = (size_t)nblk * nst * MT * 64;
  double g = bt.gamma[b];
  const double* H = bt.H + (size_t)b * nd * nd;
  double* Qb = bt.Q + (size_t)b * nblk * nd * KB;
  bf16x8* QX = (bf16x8*)bt.QX + (size_t)b * (bf_q_elems(nd, KB) / 8);
  for (size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x; f < nfrag; f += (size_t)gridDim.x * blockDim.x) {
    int lane = (int)(f % 64);
    size_t rest = f / 64;
    int m = (int)(rest % MT);
    rest /= MT;
    int st = (int)(rest % nst), blk = (int)(rest / nst);
    int k = 16 * m + lane % 16, i = blk * KB + k, j0 = 32 * st + 8 * (lane / 16);
    bf16x8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      int j = j0 + e;
      double q = (i < nd && j < nd) ? g * H[(size_t)i * nd + j] : 0.0;
      __bf16 h, l;
      split_bf16((float)q, h, l);
      hi[e] = h;
      lo[e] = l;
      if (i < nd && j < nd && j >= blk * KB && j < blk * KB + KB) Qb[((size_t)blk * nd + j) * KB + k] = q;
    }
    QX[2 * f] = hi;
    QX[2 * f + 1] = lo;
  }
}

__global__ void k_form_qx_sparse(StageSet ss) {
  int2 is = ss.inst[blockIdx.y];
  const MemberDev& M = ss.mem[is.x];
  const BatchDev& bt = M.bt;
  int b = is.y, nd = M.nd, KB = M.kb;
  int t = blockIdx.x * blockDim.x + threadIdx.x, i = t / 25;
  if (!bt.QX || !bt.active[b] || i >= nd) return;
  int r = bt.llc_idx[(size_t)b * nd * 25 + t];
  if (r < 0) return;  // padding (after the column's entries)
  int nblk = (nd + KB - 1) / KB, nst = bf_rows(nd) / 32;
  double g = bt.gamma[b], a2 = bt.scal[8 * (size_t)b];
  const double* H = bt.H + (size_t)b * nd * nd;
  double v = g * H[(size_t)i * nd + r] + bt.llc_val[(size_t)b * nd * 25 + t] / a2;
  qx_store((bf16x8*)bt.QX + (size_t)b * (bf_q_elems(nd, KB) / 8), KB, nst, i, r, v);
  int blk = i / KB;
  if (r >= blk * KB && r < blk * KB + KB) bt.Q[(size_t)b * nblk * nd * KB + ((size_t)blk * nd + r) * KB + i % KB] = v;
}

// per-coordinate constants of the stage for the Gibbs draws: gamma b_i, 1 / Q_ii, sqrt(Q_ii), 1 / sqrt(Q_ii)
__global__ void k_form_qdiag(StageSet ss) {
  int2 is = ss.inst[blockIdx.y];
  const MemberDev& M = ss.mem[is.x];
  const BatchDev& bt = M.bt;
  int b = is.y, nd = M.nd, KB = M.kb;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (!bt.Q || !bt.active[b] || i >= nd) return;
  int nblk = (nd + KB - 1) / KB;
  const double* Qb = bt.Q + (size_t)b * nblk * nd * KB;
  double qii = Qb[((size_t)(i / KB) * nd + i) * KB + i % KB];
  double* cd = bt.qdiag + (size_t)b * 4 * nd;
  double sq = sqrt(qii);
  cd[i] = bt.gamma[b] * bt.bvec[(size_t)b * nd + i];
  cd[nd + i] = 1.0 / qii;
  cd[2 * nd + i] = sq;
  cd[3 * nd + i] = 1.0 / sq;
}

// FP32 copy of the blocked precision (mixed precision)
__global__ void k_q_to_f32(StageSet ss) {
  int2 is = ss.inst[blockIdx.y];
  const MemberDev& M = ss.mem[is.x];
  const BatchDev& bt = M.bt;
  int b = is.y, nd = M.nd, KB = M.kb;
  if (!bt.Q32 || !bt.active[b]) return;
  size_t n = (size_t)((nd + KB - 1) / KB) * nd * KB;
  const double* Q = bt.Q + (size_t)b * n;
  float* Q32 = bt.Q32 + (size_t)b * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    Q32[e] = (float)Q[e];
}

// FP32 copy of the current particles of the given slots (mixed precision)
__global__ void k_s_to_f32(int nd, int P, BatchDev bt, const int* __restrict__ sel) {
  int b = sel ? sel[blockIdx.y] : (int)blockIdx.y;
  int cur = bt.par[b];
  size_t n = (size_t)nd * P;
  const double* S = (cur ? bt.S2 : bt.S) + (size_t)b * n;
  float* S32 = (cur ? bt.S32b : bt.S32) + (size_t)b * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    S32[e] = (float)S[e];
}

hipError_t launch_s32(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel) {
  hipLaunchKernelGGL(k_s_to_f32, dim3(256, B), dim3(256), 0, s, m.ndim, bt.P, bt, sel);
  return hipGetLastError();
}

hipError_t launch_form_q(const StageSet& ss, bool any_full, bool any_f32, bool any_bf, hipStream_t s) {
  int nd = ss.nd_max, KB = ss.kb_max;
  size_t n = (size_t)((nd + KB - 1) / KB) * nd * KB;
  int blocks = (int)std::min<size_t>((n + 255) / 256, 2048);
  if (any_full) {
    hipLaunchKernelGGL(k_form_qb, dim3(blocks, ss.ninst), dim3(256), 0, s, ss);
    hipLaunchKernelGGL(k_form_qb_sparse, dim3((nd * 25 + 255) / 256, ss.ninst), dim3(256), 0, s, ss);
  }
  if (any_bf) {
    size_t nfrag = bf_q_elems(nd, KB) / 16;
    hipLaunchKernelGGL(k_form_qx, dim3((int)std::min<size_t>((nfrag + 255) / 256, 1024), ss.ninst), dim3(256), 0, s,
                       ss);
    hipLaunchKernelGGL(k_form_qx_sparse, dim3((nd * 25 + 255) / 256, ss.ninst), dim3(256), 0, s, ss);
  }
  hipLaunchKernelGGL(k_form_qdiag, dim3((nd + 127) / 128, ss.ninst), dim3(128), 0, s, ss);
  if (any_f32) hipLaunchKernelGGL(k_q_to_f32, dim3(blocks, ss.ninst), dim3(256), 0, s, ss);
  return hipGetLastError();
}

}  // namespace smcf
