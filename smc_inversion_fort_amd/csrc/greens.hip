// Green's matrix assembly on the device.
//
//  k_geometry   theta -> per-plane fault frame (init.f90:121-173, gfunc.f90:283-299)
//  k_greens_p1  one thread per (observation point, patch, geometry): the 2x2 Gauss
//               point sources of the patch (gfunc.f90:136-192), Okada DC3D0 surface
//               displacement for unit strike/dip potency (DC3Dfortran.f90:2-442, z = 0),
//               accumulated per (patch node, slip direction).
//  k_greens_p2  one thread per (observation point, slip DOF, geometry): sums the partial
//               responses of the patches around the DOF node in node_to_elem order and
//               projects on the line-of-sight vectors (gfunc.f90:300-330).
//  k_dc3d0_full all 12 DC3D0 outputs for arbitrary (x, y, z), DC3Dfortran.f90:2-607.
#include "smcf_device.h"
#include "smcf_internal.h"

namespace smcf {

static constexpr double kPI2 = 6.283185307179586;

__global__ void k_geometry(MeshDev m, const double* __restrict__ theta, double* __restrict__ geo, int B,
                           const int* __restrict__ sel) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B * m.nplane) return;
  int b = t / m.nplane, ip = t % m.nplane + 1;  // 1-based plane
  if (sel) b = sel[b];
  const double* th = theta + (size_t)b * m.ndim_fault;
  const double pi = 3.141592653589793;  // 4*atan(1)
  int np = m.nplane;
  double xl, yl, xr, yr, dip;
  double ymin = th[0], ymax = th[1];
  if (m.fix_xbend) {
    if (ip == 1) { xl = m.xmin; yl = ymin; } else { xl = m.xbend[ip - 2]; yl = th[ip]; }
    if (ip == np) { xr = m.xmax; yr = ymax; } else { xr = m.xbend[ip - 1]; yr = th[ip + 1]; }
    dip = th[np + ip];
  } else {
    if (ip == 1) { xl = m.xmin; yl = ymin; } else { xl = th[2 * ip - 2]; yl = th[2 * ip - 1]; }
    if (ip == np) { xr = m.xmax; yr = ymax; } else { xr = th[2 * ip]; yr = th[2 * ip + 1]; }
    dip = th[2 * np + ip - 1];
  }
  double dip_rad = dip * pi / 180.0;
  double lxi = sqrt((xr - xl) * (xr - xl) + (yr - yl) * (yr - yl));
  double leta = -m.zmin / sin(dip_rad);
  double strike_rad = atan2(xr - xl, yr - yl);
  double strike = strike_rad / pi * 180.0;
  double xf = (xl + xr) / 2.0 - (m.zmin * cos(strike_rad)) / (2.0 * tan(dip_rad));
  double yf = (yl + yr) / 2.0 + (m.zmin * sin(strike_rad)) / (2.0 * tan(dip_rad));
  double zf = m.zmin / 2.0;
  // calc_responce / call_dc3d0 convert degrees back with 2*asin(1) (same double as pi)
  double sr = strike / 180.0 * pi, dr = dip / 180.0 * pi;
  double sd = sin(dip * (kPI2 / 360.0)), cd = cos(dip * (kPI2 / 360.0));
  if (fabs(cd) < 1.0e-6) {
    cd = 0.0;
    if (sd > 0.0) sd = 1.0;
    if (sd < 0.0) sd = -1.0;
  }
  double* g = geo + ((size_t)b * m.nplane + (ip - 1)) * 16;
  int nxi = m.plane_nxi[ip - 1], neta = m.plane_neta[ip - 1];
  g[0] = xf; g[1] = yf; g[2] = zf;
  g[3] = sin(sr); g[4] = cos(sr);
  g[5] = sin(dr); g[6] = cos(dr);
  g[7] = sd; g[8] = cd;
  g[9] = lxi / nxi; g[10] = leta / neta;
  g[11] = lxi; g[12] = leta; g[13] = dip; g[14] = strike; g[15] = 0.0;
}

// Displacement terms of UA0 / UB0 at z = 0 for unit POT/PI2 scaling left to the caller.
struct SurfTerms {
  double as[3], ad[3];  // UA0 strike, dip
  double bs[3], bd[3];  // UB0 strike, dip
};

SMCF_DEV void ua0_disp(double x, double y, double d, double sd, double cd, double* us, double* ud) {
  const double alp1 = (1.0 - 2.0 / 3.0) / 2.0, alp2 = (2.0 / 3.0) / 2.0;
  double p = y * cd + d * sd, q = y * sd - d * cd;
  double s = p * sd + q * cd, t = p * cd - q * sd;
  double r = sqrt(x * x + y * y + d * d);
  double r3 = r * r * r, r5 = r3 * r * r;
  double qr = 3.0 * q / r5;
  double x2 = x * x, xy = x * y;
  us[0] = alp1 * q / r3 + alp2 * x2 * qr;
  us[1] = alp1 * x / r3 * sd + alp2 * xy * qr;
  us[2] = -alp1 * x / r3 * cd + alp2 * x * d * qr;
  ud[0] = alp2 * x * p * qr;
  ud[1] = alp1 * s / r3 + alp2 * y * p * qr;
  ud[2] = -alp1 * t / r3 + alp2 * d * p * qr;
}

SMCF_DEV void ub0_disp_z0(double x, double y, double d, double sd, double cd, double* us, double* ud) {
  const double alp3 = (1.0 - 2.0 / 3.0) / (2.0 / 3.0);
  double p = y * cd + d * sd, q = y * sd - d * cd;
  double r = sqrt(x * x + y * y + d * d);
  double r2 = r * r, r3 = r2 * r, r5 = r3 * r2;
  double qr = 3.0 * q / r5;
  double x2 = x * x, y2 = y * y, xy = x * y;
  double c = d;  // D + Z with Z = 0
  double rd = r + d;
  double d12 = 1.0 / (r * rd * rd);
  double d32 = d12 * (2.0 * r + d) / r2;
  double d33 = d12 * (3.0 * r + d) / (r2 * rd);
  double fi1 = y * (d12 - x2 * d33);
  double fi2 = x * (d12 - y2 * d33);
  double fi3 = x / r3 - fi2;
  double fi4 = -xy * d32;
  double fi5 = 1.0 / (r * rd) - x2 * d32;
  us[0] = -x2 * qr - alp3 * fi1 * sd;
  us[1] = -xy * qr - alp3 * fi2 * sd;
  us[2] = -c * x * qr - alp3 * fi4 * sd;
  ud[0] = -x * p * qr + alp3 * fi3 * sd * cd;
  ud[1] = -y * p * qr + alp3 * fi1 * sd * cd;
  ud[2] = -c * p * qr + alp3 * fi5 * sd * cd;
}

__global__ void __launch_bounds__(128) k_greens_p1(MeshDev m, const double* __restrict__ geo,
                                                   double* __restrict__ gscr, int B, const int* __restrict__ sel) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;  // observation point
  int e = blockIdx.y;                              // patch (0-based)
  int b = sel ? sel[blockIdx.z] : (int)blockIdx.z;
  if (k >= m.npts) return;
  int ip = m.patch_plane[e];
  const double* g = geo + ((size_t)b * m.nplane + ip) * 16;
  double xf = g[0], yf = g[1], zf = g[2], ss = g[3], cs = g[4], sdr = g[5], cdr = g[6];
  double SD = g[7], CD = g[8], dxi = g[9], deta = g[10], lxi = g[11], leta = g[12];
  double xin[4], etan[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    int node = m.cny[4 * e + a] - 1;
    xin[a] = (m.node_k[node] - 1) * dxi - lxi / 2.0;
    etan[a] = (m.node_l[node] - 1) * deta - leta / 2.0;
  }
  double area4 = (xin[1] - xin[0]) * (etan[2] - etan[1]) / 4.0;
  double xo = m.pts[2 * k], yo = m.pts[2 * k + 1];
  double acc[24];
#pragma unroll
  for (int q = 0; q < 24; ++q) acc[q] = 0.0;
  const double gp = 1.0 / sqrt(3.0);
#pragma unroll
  for (int i1 = 0; i1 < 2; ++i1) {
#pragma unroll
    for (int i2 = 0; i2 < 2; ++i2) {
      double r1 = i1 ? gp : -gp, r2 = i2 ? gp : -gp;
      double nv[4] = {(1.0 - r1) * (1.0 - r2) / 4.0, (1.0 + r1) * (1.0 - r2) / 4.0,
                      (1.0 + r1) * (1.0 + r2) / 4.0, (1.0 - r1) * (1.0 + r2) / 4.0};
      double xi = 0.0, eta = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        xi = xi + nv[a] * xin[a];
        eta = eta + nv[a] * etan[a];
      }
      double xs = xf - eta * cdr * cs + xi * ss;
      double ys = yf + eta * cdr * ss + xi * cs;
      double zs = zf + eta * sdr;
      double x = ss * (xo - xs) + cs * (yo - ys);
      double y = -cs * (xo - xs) + ss * (yo - ys);
      double depth = -zs;
      if (depth < 0.0) continue;  // DC3D0 returns zeros
      double xx = fabs(x) < 1.0e-6 ? 0.0 : x;
      double yy = fabs(y) < 1.0e-6 ? 0.0 : y;
      double dd = fabs(depth) < 1.0e-6 ? 0.0 : depth;
      if (sqrt(xx * xx + yy * yy + dd * dd) == 0.0) continue;  // IRET = 1, zeros
      double ars[3], ard[3], ais[3], aid[3], bs[3], bd[3];
      ua0_disp(xx, yy, depth, SD, CD, ais, aid);  // image source: DD = depth - z
      if (dd != depth) {
        ua0_disp(xx, yy, dd, SD, CD, ars, ard);   // real source with snapped depth
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) { ars[c] = ais[c]; ard[c] = aid[c]; }
      }
      ub0_disp_z0(xx, yy, depth, SD, CD, bs, bd);
#pragma unroll
      for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int dir = 0; dir < 2; ++dir) {
          double pot = dir == 0 ? (-nv[a]) * area4 : nv[a] * area4;
          double f = pot / kPI2;
          const double* Ar = dir == 0 ? ars : ard;
          const double* Ai = dir == 0 ? ais : aid;
          const double* Bb = dir == 0 ? bs : bd;
          double u[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            double ar = f * Ar[c], ai = f * Ai[c], bb = f * Bb[c];
            u[c] = -ar + (ai + bb);
          }
          double uxr = ss * u[0] - cs * u[1];
          double uyr = cs * u[0] + ss * u[1];
          double* o = acc + (a * 2 + dir) * 3;
          o[0] = o[0] + 100.0 * uxr;
          o[1] = o[1] + 100.0 * uyr;
          o[2] = o[2] + 100.0 * u[2];
        }
      }
    }
  }
  double* out = gscr + ((size_t)b * m.npatch + e) * 24 * m.npts + k;
#pragma unroll
  for (int q = 0; q < 24; ++q) out[(size_t)q * m.npts] = acc[q];
}

__global__ void __launch_bounds__(128) k_greens_p2(MeshDev m, const double* __restrict__ gscr,
                                                   double* __restrict__ G, int B, const int* __restrict__ sel) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int d = blockIdx.y;
  int b = sel ? sel[blockIdx.z] : (int)blockIdx.z;
  if (k >= m.npts) return;
  int node = m.id_dof[d] - 1;
  double resp[6] = {0, 0, 0, 0, 0, 0};
  int nt = m.n2es[node];
  for (int t = 0; t < nt; ++t) {
    int e = m.n2e[4 * node + t] - 1;
    int a = m.n2e_slot[4 * node + t];
    const double* src = gscr + ((size_t)b * m.npatch + e) * 24 * m.npts + k;
#pragma unroll
    for (int q = 0; q < 6; ++q) resp[q] = resp[q] + src[(size_t)(a * 6 + q) * m.npts];
  }
  int r0 = m.pt_row0[k], nr = m.pt_nrow[k];
  double* Gb = G + (size_t)b * m.nobs * m.ndim;
  for (int r = r0; r < r0 + nr; ++r) {
    double e0 = m.unitvec[3 * r], e1 = m.unitvec[3 * r + 1], e2 = m.unitvec[3 * r + 2];
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
      double v = 0.0;
      v = v + resp[dir * 3 + 0] * e0;
      v = v + resp[dir * 3 + 1] * e1;
      v = v + resp[dir * 3 + 2] * e2;
      Gb[(size_t)(2 * d + dir) * m.nobs + r] = v;
    }
  }
}

hipError_t launch_geometry(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel) {
  int n = B * m.nplane;
  hipLaunchKernelGGL(k_geometry, dim3((n + 63) / 64), dim3(64), 0, s, m, bt.theta, bt.geo, B, sel);
  return hipGetLastError();
}

hipError_t launch_greens(const MeshDev& m, BatchDev& bt, int B, hipStream_t s, const int* sel) {
  dim3 g1((m.npts + 127) / 128, m.npatch, B);
  hipLaunchKernelGGL(k_greens_p1, g1, dim3(128), 0, s, m, bt.geo, bt.gscr, B, sel);
  dim3 g2((m.npts + 127) / 128, m.ndof, B);
  hipLaunchKernelGGL(k_greens_p2, g2, dim3(128), 0, s, m, bt.gscr, bt.G, B, sel);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Full DC3D0 (all 12 outputs, any z <= 0), DC3Dfortran.f90:2-607 / 1236-1286 / 1400-1414
// ---------------------------------------------------------------------------
struct PtC {
  double p, q, s, t, r, r2, r3, r5, qr, qrx, xy, x2, y2, d2, a3, a5, b3, c3, uy, uz, vy, vz, wy, wz;
};
SMCF_DEV void dccon1(double x, double y, double d, double sd, double cd, PtC& k) {
  k.p = y * cd + d * sd;
  k.q = y * sd - d * cd;
  k.s = k.p * sd + k.q * cd;
  k.t = k.p * cd - k.q * sd;
  k.r = sqrt(x * x + y * y + d * d);
  k.xy = x * y; k.x2 = x * x; k.y2 = y * y; k.d2 = d * d;
  k.r2 = k.r * k.r; k.r3 = k.r2 * k.r; k.r5 = k.r3 * k.r2;
  k.qr = 3.0 * k.q / k.r5;
  k.qrx = 5.0 * k.qr * x / k.r2;
  k.a3 = 1.0 - 3.0 * k.x2 / k.r2;
  k.a5 = 1.0 - 5.0 * k.x2 / k.r2;
  k.b3 = 1.0 - 3.0 * k.y2 / k.r2;
  k.c3 = 1.0 - 3.0 * k.d2 / k.r2;
  k.uy = sd - 5.0 * y * k.q / k.r2;
  k.uz = cd + 5.0 * d * k.q / k.r2;
  k.vy = k.s - 5.0 * y * k.p * k.q / k.r2;
  k.vz = k.t + 5.0 * d * k.p * k.q / k.r2;
  k.wy = k.uy + sd;
  k.wz = k.uz + cd;
}
SMCF_DEV void acc12(double* u, double pot, const double* du) {
#pragma unroll
  for (int i = 0; i < 12; ++i) u[i] = u[i] + pot / kPI2 * du[i];
}

SMCF_DEV void ua0_full(double x, double y, double d, const double* pot, double sd, double cd, const double* al,
                       double* u) {
  PtC k;
  double du[12];
  for (int i = 0; i < 12; ++i) u[i] = 0.0;
  dccon1(x, y, d, sd, cd, k);
  double s2d = 2.0 * sd * cd, c2d = cd * cd - sd * sd;
  if (pot[0] != 0.0) {
    du[0] = al[0] * k.q / k.r3 + al[1] * k.x2 * k.qr;
    du[1] = al[0] * x / k.r3 * sd + al[1] * k.xy * k.qr;
    du[2] = -al[0] * x / k.r3 * cd + al[1] * x * d * k.qr;
    du[3] = x * k.qr * (-al[0] + al[1] * (1.0 + k.a5));
    du[4] = al[0] * k.a3 / k.r3 * sd + al[1] * y * k.qr * k.a5;
    du[5] = -al[0] * k.a3 / k.r3 * cd + al[1] * d * k.qr * k.a5;
    du[6] = al[0] * (sd / k.r3 - y * k.qr) + al[1] * 3.0 * k.x2 / k.r5 * k.uy;
    du[7] = 3.0 * x / k.r5 * (-al[0] * y * sd + al[1] * (y * k.uy + k.q));
    du[8] = 3.0 * x / k.r5 * (al[0] * y * cd + al[1] * d * k.uy);
    du[9] = al[0] * (cd / k.r3 + d * k.qr) + al[1] * 3.0 * k.x2 / k.r5 * k.uz;
    du[10] = 3.0 * x / k.r5 * (al[0] * d * sd + al[1] * y * k.uz);
    du[11] = 3.0 * x / k.r5 * (-al[0] * d * cd + al[1] * (d * k.uz - k.q));
    acc12(u, pot[0], du);
  }
  if (pot[1] != 0.0) {
    du[0] = al[1] * x * k.p * k.qr;
    du[1] = al[0] * k.s / k.r3 + al[1] * y * k.p * k.qr;
    du[2] = -al[0] * k.t / k.r3 + al[1] * d * k.p * k.qr;
    du[3] = al[1] * k.p * k.qr * k.a5;
    du[4] = -al[0] * 3.0 * x * k.s / k.r5 - al[1] * y * k.p * k.qrx;
    du[5] = al[0] * 3.0 * x * k.t / k.r5 - al[1] * d * k.p * k.qrx;
    du[6] = al[1] * 3.0 * x / k.r5 * k.vy;
    du[7] = al[0] * (s2d / k.r3 - 3.0 * y * k.s / k.r5) + al[1] * (3.0 * y / k.r5 * k.vy + k.p * k.qr);
    du[8] = -al[0] * (c2d / k.r3 - 3.0 * y * k.t / k.r5) + al[1] * 3.0 * d / k.r5 * k.vy;
    du[9] = al[1] * 3.0 * x / k.r5 * k.vz;
    du[10] = al[0] * (c2d / k.r3 + 3.0 * d * k.s / k.r5) + al[1] * 3.0 * y / k.r5 * k.vz;
    du[11] = al[0] * (s2d / k.r3 - 3.0 * d * k.t / k.r5) + al[1] * (3.0 * d / k.r5 * k.vz - k.p * k.qr);
    acc12(u, pot[1], du);
  }
  if (pot[2] != 0.0) {
    du[0] = al[0] * x / k.r3 - al[1] * x * k.q * k.qr;
    du[1] = al[0] * k.t / k.r3 - al[1] * y * k.q * k.qr;
    du[2] = al[0] * k.s / k.r3 - al[1] * d * k.q * k.qr;
    du[3] = al[0] * k.a3 / k.r3 - al[1] * k.q * k.qr * k.a5;
    du[4] = -al[0] * 3.0 * x * k.t / k.r5 + al[1] * y * k.q * k.qrx;
    du[5] = -al[0] * 3.0 * x * k.s / k.r5 + al[1] * d * k.q * k.qrx;
    du[6] = -al[0] * 3.0 * k.xy / k.r5 - al[1] * x * k.qr * k.wy;
    du[7] = al[0] * (c2d / k.r3 - 3.0 * y * k.t / k.r5) - al[1] * (y * k.wy + k.q) * k.qr;
    du[8] = al[0] * (s2d / k.r3 - 3.0 * y * k.s / k.r5) - al[1] * d * k.qr * k.wy;
    du[9] = al[0] * 3.0 * x * d / k.r5 - al[1] * x * k.qr * k.wz;
    du[10] = -al[0] * (s2d / k.r3 - 3.0 * d * k.t / k.r5) - al[1] * y * k.qr * k.wz;
    du[11] = al[0] * (c2d / k.r3 + 3.0 * d * k.s / k.r5) - al[1] * (d * k.wz - k.q) * k.qr;
    acc12(u, pot[2], du);
  }
  if (pot[3] != 0.0) {
    du[0] = -al[0] * x / k.r3;
    du[1] = -al[0] * y / k.r3;
    du[2] = -al[0] * d / k.r3;
    du[3] = -al[0] * k.a3 / k.r3;
    du[4] = al[0] * 3.0 * k.xy / k.r5;
    du[5] = al[0] * 3.0 * x * d / k.r5;
    du[6] = du[4];
    du[7] = -al[0] * k.b3 / k.r3;
    du[8] = al[0] * 3.0 * y * d / k.r5;
    du[9] = -du[5];
    du[10] = -du[8];
    du[11] = al[0] * k.c3 / k.r3;
    acc12(u, pot[3], du);
  }
}

SMCF_DEV void ub0_full(double x, double y, double d, double z, const double* pot, double sd, double cd,
                       const double* al, double* u) {
  PtC k;
  double du[12];
  for (int i = 0; i < 12; ++i) u[i] = 0.0;
  dccon1(x, y, d, sd, cd, k);
  double c = d + z, rd = k.r + d;
  double d12 = 1.0 / (k.r * rd * rd);
  double d32 = d12 * (2.0 * k.r + d) / k.r2;
  double d33 = d12 * (3.0 * k.r + d) / (k.r2 * rd);
  double d53 = d12 * (8.0 * k.r2 + 9.0 * k.r * d + 3.0 * k.d2) / (k.r2 * k.r2 * rd);
  double d54 = d12 * (5.0 * k.r2 + 4.0 * k.r * d + k.d2) / k.r3 * d12;
  double fi1 = y * (d12 - k.x2 * d33), fi2 = x * (d12 - k.y2 * d33), fi3 = x / k.r3 - fi2;
  double fi4 = -k.xy * d32, fi5 = 1.0 / (k.r * rd) - k.x2 * d32;
  double fj1 = -3.0 * k.xy * (d33 - k.x2 * d54);
  double fj2 = 1.0 / k.r3 - 3.0 * d12 + 3.0 * k.x2 * k.y2 * d54;
  double fj3 = k.a3 / k.r3 - fj2, fj4 = -3.0 * k.xy / k.r5 - fj1;
  double fk1 = -y * (d32 - k.x2 * d53), fk2 = -x * (d32 - k.y2 * d53), fk3 = -3.0 * x * d / k.r5 - fk2;
  double a3 = al[2];
  if (pot[0] != 0.0) {
    du[0] = -k.x2 * k.qr - a3 * fi1 * sd;
    du[1] = -k.xy * k.qr - a3 * fi2 * sd;
    du[2] = -c * x * k.qr - a3 * fi4 * sd;
    du[3] = -x * k.qr * (1.0 + k.a5) - a3 * fj1 * sd;
    du[4] = -y * k.qr * k.a5 - a3 * fj2 * sd;
    du[5] = -c * k.qr * k.a5 - a3 * fk1 * sd;
    du[6] = -3.0 * k.x2 / k.r5 * k.uy - a3 * fj2 * sd;
    du[7] = -3.0 * k.xy / k.r5 * k.uy - x * k.qr - a3 * fj4 * sd;
    du[8] = -3.0 * c * x / k.r5 * k.uy - a3 * fk2 * sd;
    du[9] = -3.0 * k.x2 / k.r5 * k.uz + a3 * fk1 * sd;
    du[10] = -3.0 * k.xy / k.r5 * k.uz + a3 * fk2 * sd;
    du[11] = 3.0 * x / k.r5 * (-c * k.uz + a3 * y * sd);
    acc12(u, pot[0], du);
  }
  if (pot[1] != 0.0) {
    du[0] = -x * k.p * k.qr + a3 * fi3 * sd * cd;
    du[1] = -y * k.p * k.qr + a3 * fi1 * sd * cd;
    du[2] = -c * k.p * k.qr + a3 * fi5 * sd * cd;
    du[3] = -k.p * k.qr * k.a5 + a3 * fj3 * sd * cd;
    du[4] = y * k.p * k.qrx + a3 * fj1 * sd * cd;
    du[5] = c * k.p * k.qrx + a3 * fk3 * sd * cd;
    du[6] = -3.0 * x / k.r5 * k.vy + a3 * fj1 * sd * cd;
    du[7] = -3.0 * y / k.r5 * k.vy - k.p * k.qr + a3 * fj2 * sd * cd;
    du[8] = -3.0 * c / k.r5 * k.vy + a3 * fk1 * sd * cd;
    du[9] = -3.0 * x / k.r5 * k.vz - a3 * fk3 * sd * cd;
    du[10] = -3.0 * y / k.r5 * k.vz - a3 * fk1 * sd * cd;
    du[11] = -3.0 * c / k.r5 * k.vz + a3 * k.a3 / k.r3 * sd * cd;
    acc12(u, pot[1], du);
  }
  if (pot[2] != 0.0) {
    du[0] = x * k.q * k.qr - a3 * fi3 * sd * sd;
    du[1] = y * k.q * k.qr - a3 * fi1 * sd * sd;
    du[2] = c * k.q * k.qr - a3 * fi5 * sd * sd;
    du[3] = k.q * k.qr * k.a5 - a3 * fj3 * sd * sd;
    du[4] = -y * k.q * k.qrx - a3 * fj1 * sd * sd;
    du[5] = -c * k.q * k.qrx - a3 * fk3 * sd * sd;
    du[6] = x * k.qr * k.wy - a3 * fj1 * sd * sd;
    du[7] = k.qr * (y * k.wy + k.q) - a3 * fj2 * sd * sd;
    du[8] = c * k.qr * k.wy - a3 * fk1 * sd * sd;
    du[9] = x * k.qr * k.wz + a3 * fk3 * sd * sd;
    du[10] = y * k.qr * k.wz + a3 * fk1 * sd * sd;
    du[11] = c * k.qr * k.wz - a3 * k.a3 / k.r3 * sd * sd;
    acc12(u, pot[2], du);
  }
  if (pot[3] != 0.0) {
    du[0] = a3 * x / k.r3;
    du[1] = a3 * y / k.r3;
    du[2] = a3 * d / k.r3;
    du[3] = a3 * k.a3 / k.r3;
    du[4] = -a3 * 3.0 * k.xy / k.r5;
    du[5] = -a3 * 3.0 * x * d / k.r5;
    du[6] = du[4];
    du[7] = a3 * k.b3 / k.r3;
    du[8] = -a3 * 3.0 * y * d / k.r5;
    du[9] = -du[5];
    du[10] = -du[8];
    du[11] = -a3 * k.c3 / k.r3;
    acc12(u, pot[3], du);
  }
}

SMCF_DEV void uc0_full(double x, double y, double d, double z, const double* pot, double sd, double cd,
                       const double* al, double* u) {
  PtC k;
  double du[12];
  for (int i = 0; i < 12; ++i) u[i] = 0.0;
  dccon1(x, y, d, sd, cd, k);
  double s2d = 2.0 * sd * cd, c2d = cd * cd - sd * sd;
  double c = d + z, q2 = k.q * k.q, r7 = k.r5 * k.r2;
  double a7 = 1.0 - 7.0 * k.x2 / k.r2, b5 = 1.0 - 5.0 * k.y2 / k.r2, b7 = 1.0 - 7.0 * k.y2 / k.r2;
  double c5 = 1.0 - 5.0 * k.d2 / k.r2, c7 = 1.0 - 7.0 * k.d2 / k.r2, d7 = 2.0 - 7.0 * q2 / k.r2;
  double qr5 = 5.0 * k.q / k.r2, qr7 = 7.0 * k.q / k.r2, dr5 = 5.0 * d / k.r2;
  double A4 = al[3], A5 = al[4];
  double R2 = k.r2, R3 = k.r3, R5 = k.r5, Q = k.q, P = k.p, S = k.s, T = k.t, a3 = k.a3, a5 = k.a5;
  double QR = k.qr, QRX = k.qrx;
  if (pot[0] != 0.0) {
    du[0] = -A4 * a3 / R3 * cd + A5 * c * QR * a5;
    du[1] = 3.0 * x / R5 * (A4 * y * cd + A5 * c * (sd - y * qr5));
    du[2] = 3.0 * x / R5 * (-A4 * y * sd + A5 * c * (cd + d * qr5));
    du[3] = A4 * 3.0 * x / R5 * (2.0 + a5) * cd - A5 * c * QRX * (2.0 + a7);
    du[4] = 3.0 / R5 * (A4 * y * a5 * cd + A5 * c * (a5 * sd - y * qr5 * a7));
    du[5] = 3.0 / R5 * (-A4 * y * a5 * sd + A5 * c * (a5 * cd + d * qr5 * a7));
    du[6] = du[4];
    du[7] = 3.0 * x / R5 * (A4 * b5 * cd - A5 * 5.0 * c / R2 * (2.0 * y * sd + Q * b7));
    du[8] = 3.0 * x / R5 * (-A4 * b5 * sd + A5 * 5.0 * c / R2 * (d * b7 * sd - y * c7 * cd));
    du[9] = 3.0 / R5 * (-A4 * d * a5 * cd + A5 * c * (a5 * cd + d * qr5 * a7));
    du[10] = 15.0 * x / r7 * (A4 * y * d * cd + A5 * c * (d * b7 * sd - y * c7 * cd));
    du[11] = 15.0 * x / r7 * (-A4 * y * d * sd + A5 * c * (2.0 * d * cd - Q * c7));
    acc12(u, pot[0], du);
  }
  if (pot[1] != 0.0) {
    du[0] = A4 * 3.0 * x * T / R5 - A5 * c * P * QRX;
    du[1] = -A4 / R3 * (c2d - 3.0 * y * T / R2) + A5 * 3.0 * c / R5 * (S - y * P * qr5);
    du[2] = -A4 * a3 / R3 * sd * cd + A5 * 3.0 * c / R5 * (T + d * P * qr5);
    du[3] = A4 * 3.0 * T / R5 * a5 - A5 * 5.0 * c * P * QR / R2 * a7;
    du[4] = 3.0 * x / R5 * (A4 * (c2d - 5.0 * y * T / R2) - A5 * 5.0 * c / R2 * (S - y * P * qr7));
    du[5] = 3.0 * x / R5 * (A4 * (2.0 + a5) * sd * cd - A5 * 5.0 * c / R2 * (T + d * P * qr7));
    du[6] = du[4];
    du[7] = 3.0 / R5 * (A4 * (2.0 * y * c2d + T * b5) + A5 * c * (s2d - 10.0 * y * S / R2 - P * qr5 * b7));
    du[8] = 3.0 / R5 * (A4 * y * a5 * sd * cd - A5 * c * ((3.0 + a5) * c2d + y * P * dr5 * qr7));
    du[9] = 3.0 * x / R5 * (-A4 * (s2d - T * dr5) - A5 * 5.0 * c / R2 * (T + d * P * qr7));
    du[10] = 3.0 / R5 * (-A4 * (d * b5 * c2d + y * c5 * s2d) - A5 * c * ((3.0 + a5) * c2d + y * P * dr5 * qr7));
    du[11] = 3.0 / R5 * (-A4 * d * a5 * sd * cd - A5 * c * (s2d - 10.0 * d * T / R2 + P * qr5 * c7));
    acc12(u, pot[1], du);
  }
  if (pot[2] != 0.0) {
    du[0] = 3.0 * x / R5 * (-A4 * S + A5 * (c * Q * qr5 - z));
    du[1] = A4 / R3 * (s2d - 3.0 * y * S / R2) + A5 * 3.0 / R5 * (c * (T - y + y * Q * qr5) - y * z);
    du[2] = -A4 / R3 * (1.0 - a3 * sd * sd) - A5 * 3.0 / R5 * (c * (S - d + d * Q * qr5) - d * z);
    du[3] = -A4 * 3.0 * S / R5 * a5 + A5 * (c * QR * qr5 * a7 - 3.0 * z / R5 * a5);
    du[4] = 3.0 * x / R5 * (-A4 * (s2d - 5.0 * y * S / R2) - A5 * 5.0 / R2 * (c * (T - y + y * Q * qr7) - y * z));
    du[5] = 3.0 * x / R5 * (A4 * (1.0 - (2.0 + a5) * sd * sd) + A5 * 5.0 / R2 * (c * (S - d + d * Q * qr7) - d * z));
    du[6] = du[4];
    du[7] = 3.0 / R5 * (-A4 * (2.0 * y * s2d + S * b5) -
                        A5 * (c * (2.0 * sd * sd + 10.0 * y * (T - y) / R2 - Q * qr5 * b7) + z * b5));
    du[8] = 3.0 / R5 * (A4 * y * (1.0 - a5 * sd * sd) + A5 * (c * (3.0 + a5) * s2d - y * dr5 * (c * d7 + z)));
    du[9] = 3.0 * x / R5 * (-A4 * (c2d + S * dr5) + A5 * (5.0 * c / R2 * (S - d + d * Q * qr7) - 1.0 - z * dr5));
    du[10] = 3.0 / R5 * (A4 * (d * b5 * s2d - y * c5 * c2d) +
                         A5 * (c * ((3.0 + a5) * s2d - y * dr5 * d7) - y * (1.0 + z * dr5)));
    du[11] = 3.0 / R5 * (-A4 * d * (1.0 - a5 * sd * sd) -
                         A5 * (c * (c2d + 10.0 * d * (S - d) / R2 - Q * qr5 * c7) + z * (1.0 + c5)));
    acc12(u, pot[2], du);
  }
  if (pot[3] != 0.0) {
    du[0] = A4 * 3.0 * x * d / R5;
    du[1] = A4 * 3.0 * y * d / R5;
    du[2] = A4 * k.c3 / R3;
    du[3] = A4 * 3.0 * d / R5 * a5;
    du[4] = -A4 * 15.0 * k.xy * d / r7;
    du[5] = -A4 * 3.0 * x / R5 * c5;
    du[6] = du[4];
    du[7] = A4 * 3.0 * d / R5 * b5;
    du[8] = -A4 * 3.0 * y / R5 * c5;
    du[9] = du[5];
    du[10] = du[8];
    du[11] = A4 * 3.0 * d / R5 * (2.0 + c5);
    acc12(u, pot[3], du);
  }
}

// DC3D0 (DC3Dfortran.f90:2-118): all 12 outputs into u; *iret as the reference sets it (left unchanged when
// depth < 0, where every output is zero)
SMCF_DEV void dc3d0_dev(double alpha, double x, double y, double z, double depth, double dip, const double* pot,
                        double* u, int* iret) {
  double dua[12], dub[12], duc[12];
  for (int c = 0; c < 12; ++c) u[c] = 0.0;
  double xx = x, yy = y, zz = z, dd = depth + z;
  if (depth < 0.0) return;
  if (fabs(xx) < 1e-6) xx = 0.0;
  if (fabs(yy) < 1e-6) yy = 0.0;
  if (fabs(dd) < 1e-6) dd = 0.0;
  int ir = 0;
  if (z > 0.0) ir = 2;
  else if (sqrt(xx * xx + yy * yy + dd * dd) == 0.0) ir = 1;
  *iret = ir;
  if (ir != 0) return;
  double al[5] = {(1.0 - alpha) / 2.0, alpha / 2.0, (1.0 - alpha) / alpha, 1.0 - alpha, alpha};
  double sd = sin(dip * (kPI2 / 360.0)), cd = cos(dip * (kPI2 / 360.0));
  if (fabs(cd) < 1e-6) {
    cd = 0.0;
    if (sd > 0.0) sd = 1.0;
    if (sd < 0.0) sd = -1.0;
  }
  ua0_full(xx, yy, dd, pot, sd, cd, al, dua);
  for (int c = 0; c < 12; ++c) u[c] = (c < 9) ? u[c] - dua[c] : u[c] + dua[c];
  dd = depth - z;
  ua0_full(xx, yy, dd, pot, sd, cd, al, dua);
  ub0_full(xx, yy, dd, zz, pot, sd, cd, al, dub);
  uc0_full(xx, yy, dd, zz, pot, sd, cd, al, duc);
  for (int c = 0; c < 12; ++c) {
    double du = dua[c] + dub[c] + zz * duc[c];
    if (c >= 9) du = du + duc[c - 9];
    u[c] = u[c] + du;
  }
}

__global__ void k_dc3d0_full(int n, const double* __restrict__ in, double* __restrict__ uo,
                             int* __restrict__ iret) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* a = in + 10 * (size_t)i;
  double pot[4] = {a[6], a[7], a[8], a[9]};
  double u[12];
  int ir = iret[i];
  dc3d0_dev(a[0], a[1], a[2], a[3], a[4], a[5], pot, u, &ir);
  iret[i] = ir;
  double* o = uo + 12 * (size_t)i;
  for (int c = 0; c < 12; ++c) o[c] = u[c];
}

// call_dc3d0 (gfunc.f90:65-100): station rotated into the strike frame, alpha = 2/3, z = 0, depth = -zsource,
// POT1 = uxi area, POT2 = ueta area; displacement rotated back and scaled by 100 (cm)
SMCF_DEV void call_dc3d0_dev(double xs, double ys, double zs, double xo, double yo, double uxi, double ueta,
                             double dip, double area, double strike, double* uret) {
  const double pi = 3.141592653589793;
  double sr = strike / 180.0 * pi;
  double x = sin(sr) * (xo - xs) + cos(sr) * (yo - ys);
  double y = -cos(sr) * (xo - xs) + sin(sr) * (yo - ys);
  double pot[4] = {uxi * area, ueta * area, 0.0, 0.0};
  double u[12];
  int ir = 0;
  dc3d0_dev(2.0 / 3.0, x, y, 0.0, -zs, dip, pot, u, &ir);
  double uxr = sin(sr) * u[0] - cos(sr) * u[1];
  double uyr = cos(sr) * u[0] + sin(sr) * u[1];
  uret[0] = 100.0 * uxr;
  uret[1] = 100.0 * uyr;
  uret[2] = 100.0 * u[2];
}

__global__ void k_call_dc3d0(int n, const double* __restrict__ in, double* __restrict__ uret) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* a = in + 10 * (size_t)i;
  call_dc3d0_dev(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], uret + 3 * (size_t)i);
}

// calc_responce (gfunc.f90:102-194) for one observation point per thread: over the target patches, the 2 x 2
// Gauss points r = +-1/sqrt(3) with bilinearly interpolated (xi, eta, uxi, ueta), each a point source of area / 4
__global__ void k_calc_responce(int npts, const double* __restrict__ pts, const int* __restrict__ cny,
                                const double* __restrict__ coor, const double* __restrict__ slip, double xf, double yf,
                                double zf, double strike, double dip, int ntarget, const int* __restrict__ target,
                                double* __restrict__ uobs) {
  int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= npts) return;
  const double pi = 3.141592653589793;
  double dr = dip / 180.0 * pi, sr = strike / 180.0 * pi;
  double rv[2] = {-1.0 / sqrt(3.0), 1.0 / sqrt(3.0)};
  double xo = pts[2 * (size_t)ip], yo = pts[2 * (size_t)ip + 1];
  double u0 = 0.0, u1 = 0.0, u2 = 0.0;
  for (int it = 0; it < ntarget; ++it) {
    int ie = target[it] - 1;
    double xn[4], en[4], uxn[4], uen[4];
    for (int k = 0; k < 4; ++k) {
      int jn = cny[4 * (size_t)ie + k] - 1;
      xn[k] = coor[2 * (size_t)jn];
      en[k] = coor[2 * (size_t)jn + 1];
      uxn[k] = slip[2 * (size_t)jn];
      uen[k] = slip[2 * (size_t)jn + 1];
    }
    double area = (xn[1] - xn[0]) * (en[2] - en[1]);
    for (int a1 = 0; a1 < 2; ++a1) {
      double r1 = rv[a1];
      for (int a2 = 0; a2 < 2; ++a2) {
        double r2 = rv[a2];
        double nv[4] = {(1.0 - r1) * (1.0 - r2) / 4.0, (1.0 + r1) * (1.0 - r2) / 4.0, (1.0 + r1) * (1.0 + r2) / 4.0,
                        (1.0 - r1) * (1.0 + r2) / 4.0};
        double xi = 0.0, eta = 0.0, uxi = 0.0, ueta = 0.0;
        for (int k = 0; k < 4; ++k) {
          xi = xi + nv[k] * xn[k];
          eta = eta + nv[k] * en[k];
          uxi = uxi + nv[k] * uxn[k];
          ueta = ueta + nv[k] * uen[k];
        }
        double xs = xf - eta * cos(dr) * cos(sr) + xi * sin(sr);
        double ys = yf + eta * cos(dr) * sin(sr) + xi * cos(sr);
        double zs = zf + eta * sin(dr);
        double ur[3];
        call_dc3d0_dev(xs, ys, zs, xo, yo, uxi, ueta, dip, area / 4.0, strike, ur);
        u0 = u0 + ur[0];
        u1 = u1 + ur[1];
        u2 = u2 + ur[2];
      }
    }
  }
  uobs[3 * (size_t)ip] = u0;
  uobs[3 * (size_t)ip + 1] = u1;
  uobs[3 * (size_t)ip + 2] = u2;
}

hipError_t launch_call_dc3d0(int n, const double* in, double* uret, hipStream_t s) {
  hipLaunchKernelGGL(k_call_dc3d0, dim3((n + 127) / 128), dim3(128), 0, s, n, in, uret);
  return hipGetLastError();
}

hipError_t launch_calc_responce(int npts, const double* pts, const int* cny, const double* coor, const double* slip,
                                double xf, double yf, double zf, double strike, double dip, int ntarget,
                                const int* target, double* uobs, hipStream_t s) {
  hipLaunchKernelGGL(k_calc_responce, dim3((npts + 63) / 64), dim3(64), 0, s, npts, pts, cny, coor, slip, xf, yf, zf,
                     strike, dip, ntarget, target, uobs);
  return hipGetLastError();
}

hipError_t launch_dc3d0_batch(int n, const double* in, double* u, int* iret, hipStream_t s) {
  hipLaunchKernelGGL(k_dc3d0_full, dim3((n + 127) / 128), dim3(128), 0, s, n, in, u, iret);
  return hipGetLastError();
}

}  // namespace smcf
