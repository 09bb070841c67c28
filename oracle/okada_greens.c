/*
 * Oracle part 1: keyed RNG, Okada (1992) point source, fault mesh, Laplacian
 * and Green's matrix.  Test infrastructure only (see oracle.h).
 *
 * Reference files followed:
 *   DC3Dfortran.f90:2-118    DC3D0 (real + image source combination)
 *   DC3Dfortran.f90:122-269  UA0      DC3Dfortran.f90:273-442  UB0
 *   DC3Dfortran.f90:446-607  UC0      DC3Dfortran.f90:1236-1286 DCCON0/DCCON1
 *   DC3Dfortran.f90:1400-1414 assign_alpha
 *   init.f90:121-196  get_geometry / get_sigma
 *   init.f90:198-293  discretize_fault
 *   init.f90:295-541  gen_laplacian / gen_sparse_lmat / calc_ll
 *   gfunc.f90:6-348   unit slip, call_dc3d0, calc_responce(_dist), calc_greens_func
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), restated from the paper.
 * ====================================================================== */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t h0, l0, h1, l1;
        mulhilo32(0xD2511F53u, c0, &h0, &l0);
        mulhilo32(0xCD9E8D57u, c2, &h1, &l1);
        uint32_t n0 = h1 ^ c1 ^ k0, n1 = l1, n2 = h0 ^ c3 ^ k1, n3 = l0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Draw layout: key = instance key; counter = (c0, c1, c2, site<<24 | stage). */
void orc_draw_u64x2(uint64_t key, uint32_t site, uint32_t stage, uint32_t c2, uint32_t c1,
                    uint32_t c0, uint64_t out[2]) {
    uint32_t ctr[4] = {c0, c1, c2, (site << 24) | (stage & 0xFFFFFFu)};
    uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    uint32_t w[4];
    orc_philox4x32(ctr, k, w);
    out[0] = ((uint64_t)w[1] << 32) | w[0];
    out[1] = ((uint64_t)w[3] << 32) | w[2];
}

double orc_uniform_raw(uint64_t word) { return (double)(word >> 11) * 0x1.0p-53; }

/* random_number_correction, smc_slip.f90:8-15 */
double orc_clamp_uniform(double u) { return u < 1e-5 ? 1e-5 : u; }

/* slip_BoxMuller sin branch, smc_slip.f90:26-38 */
double orc_box_muller_sin(double r, double s) {
    double pi = 2.0 * asin(1.0);
    return sqrt(-2.0 * log(r)) * sin(2.0 * pi * s);
}

uint64_t orc_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* Key of the inner SMC run for one outer evaluation (outer stage, role, slot). */
uint64_t orc_inner_key(uint64_t seed, uint32_t outer_stage, uint32_t role, uint32_t slot) {
    uint64_t a = orc_splitmix64(seed);
    uint64_t b = ((uint64_t)outer_stage << 33) ^ ((uint64_t)role << 32) ^ (uint64_t)slot;
    return orc_splitmix64(a ^ orc_splitmix64(b));
}

/* ======================================================================
 * DC3D0: Okada (1992) buried point source in a half-space.
 * u[0..2] displacement, u[3..5] x-derivatives, u[6..8] y-, u[9..11] z-derivatives.
 * ====================================================================== */
typedef struct {
    double p, q, s, t, r, r2, r3, r5, qr, qrx, xy, x2, y2, d2, a3, a5, b3, c3;
    double uy, uz, vy, vz, wy, wz;
} pt_consts;

static void dccon1(double x, double y, double d, double sd, double cd, pt_consts* k) {
    k->p = y * cd + d * sd;
    k->q = y * sd - d * cd;
    k->s = k->p * sd + k->q * cd;
    k->t = k->p * cd - k->q * sd;
    k->r = sqrt(x * x + y * y + d * d);
    k->xy = x * y; k->x2 = x * x; k->y2 = y * y; k->d2 = d * d;
    k->r2 = pow(k->r, 2.0); k->r3 = pow(k->r, 3.0); k->r5 = pow(k->r, 5.0);
    k->qr = 3.0 * k->q / k->r5;
    k->qrx = 5.0 * k->qr * x / k->r2;
    k->a3 = 1.0 - 3.0 * k->x2 / k->r2;
    k->a5 = 1.0 - 5.0 * k->x2 / k->r2;
    k->b3 = 1.0 - 3.0 * k->y2 / k->r2;
    k->c3 = 1.0 - 3.0 * k->d2 / k->r2;
    k->uy = sd - 5.0 * y * k->q / k->r2;
    k->uz = cd + 5.0 * d * k->q / k->r2;
    k->vy = k->s - 5.0 * y * k->p * k->q / k->r2;
    k->vz = k->t + 5.0 * d * k->p * k->q / k->r2;
    k->wy = k->uy + sd;
    k->wz = k->uz + cd;
}

static const double PI2 = 6.283185307179586;

static void accum(double* u, double pot, const double* du) {
    for (int i = 0; i < 12; ++i) u[i] = u[i] + pot / PI2 * du[i];
}

/* Part A (Okada 1992 Tab. 4/5 with eq. 8), DC3Dfortran.f90:122-269 */
static void ua0(double x, double y, double d, const double pot[4], double sd, double cd,
                const double alp[5], double u[12]) {
    pt_consts k; double du[12];
    for (int i = 0; i < 12; ++i) u[i] = 0.0;
    dccon1(x, y, d, sd, cd, &k);
    double s2d = 2.0 * sd * cd, c2d = cd * cd - sd * sd;
    if (pot[0] != 0.0) {
        du[0] = alp[0] * k.q / k.r3 + alp[1] * k.x2 * k.qr;
        du[1] = alp[0] * x / k.r3 * sd + alp[1] * k.xy * k.qr;
        du[2] = -alp[0] * x / k.r3 * cd + alp[1] * x * d * k.qr;
        du[3] = x * k.qr * (-alp[0] + alp[1] * (1.0 + k.a5));
        du[4] = alp[0] * k.a3 / k.r3 * sd + alp[1] * y * k.qr * k.a5;
        du[5] = -alp[0] * k.a3 / k.r3 * cd + alp[1] * d * k.qr * k.a5;
        du[6] = alp[0] * (sd / k.r3 - y * k.qr) + alp[1] * 3.0 * k.x2 / k.r5 * k.uy;
        du[7] = 3.0 * x / k.r5 * (-alp[0] * y * sd + alp[1] * (y * k.uy + k.q));
        du[8] = 3.0 * x / k.r5 * (alp[0] * y * cd + alp[1] * d * k.uy);
        du[9] = alp[0] * (cd / k.r3 + d * k.qr) + alp[1] * 3.0 * k.x2 / k.r5 * k.uz;
        du[10] = 3.0 * x / k.r5 * (alp[0] * d * sd + alp[1] * y * k.uz);
        du[11] = 3.0 * x / k.r5 * (-alp[0] * d * cd + alp[1] * (d * k.uz - k.q));
        accum(u, pot[0], du);
    }
    if (pot[1] != 0.0) {
        du[0] = alp[1] * x * k.p * k.qr;
        du[1] = alp[0] * k.s / k.r3 + alp[1] * y * k.p * k.qr;
        du[2] = -alp[0] * k.t / k.r3 + alp[1] * d * k.p * k.qr;
        du[3] = alp[1] * k.p * k.qr * k.a5;
        du[4] = -alp[0] * 3.0 * x * k.s / k.r5 - alp[1] * y * k.p * k.qrx;
        du[5] = alp[0] * 3.0 * x * k.t / k.r5 - alp[1] * d * k.p * k.qrx;
        du[6] = alp[1] * 3.0 * x / k.r5 * k.vy;
        du[7] = alp[0] * (s2d / k.r3 - 3.0 * y * k.s / k.r5) + alp[1] * (3.0 * y / k.r5 * k.vy + k.p * k.qr);
        du[8] = -alp[0] * (c2d / k.r3 - 3.0 * y * k.t / k.r5) + alp[1] * 3.0 * d / k.r5 * k.vy;
        du[9] = alp[1] * 3.0 * x / k.r5 * k.vz;
        du[10] = alp[0] * (c2d / k.r3 + 3.0 * d * k.s / k.r5) + alp[1] * 3.0 * y / k.r5 * k.vz;
        du[11] = alp[0] * (s2d / k.r3 - 3.0 * d * k.t / k.r5) + alp[1] * (3.0 * d / k.r5 * k.vz - k.p * k.qr);
        accum(u, pot[1], du);
    }
    if (pot[2] != 0.0) {
        du[0] = alp[0] * x / k.r3 - alp[1] * x * k.q * k.qr;
        du[1] = alp[0] * k.t / k.r3 - alp[1] * y * k.q * k.qr;
        du[2] = alp[0] * k.s / k.r3 - alp[1] * d * k.q * k.qr;
        du[3] = alp[0] * k.a3 / k.r3 - alp[1] * k.q * k.qr * k.a5;
        du[4] = -alp[0] * 3.0 * x * k.t / k.r5 + alp[1] * y * k.q * k.qrx;
        du[5] = -alp[0] * 3.0 * x * k.s / k.r5 + alp[1] * d * k.q * k.qrx;
        du[6] = -alp[0] * 3.0 * k.xy / k.r5 - alp[1] * x * k.qr * k.wy;
        du[7] = alp[0] * (c2d / k.r3 - 3.0 * y * k.t / k.r5) - alp[1] * (y * k.wy + k.q) * k.qr;
        du[8] = alp[0] * (s2d / k.r3 - 3.0 * y * k.s / k.r5) - alp[1] * d * k.qr * k.wy;
        du[9] = alp[0] * 3.0 * x * d / k.r5 - alp[1] * x * k.qr * k.wz;
        du[10] = -alp[0] * (s2d / k.r3 - 3.0 * d * k.t / k.r5) - alp[1] * y * k.qr * k.wz;
        du[11] = alp[0] * (c2d / k.r3 + 3.0 * d * k.s / k.r5) - alp[1] * (d * k.wz - k.q) * k.qr;
        accum(u, pot[2], du);
    }
    if (pot[3] != 0.0) {
        du[0] = -alp[0] * x / k.r3;
        du[1] = -alp[0] * y / k.r3;
        du[2] = -alp[0] * d / k.r3;
        du[3] = -alp[0] * k.a3 / k.r3;
        du[4] = alp[0] * 3.0 * k.xy / k.r5;
        du[5] = alp[0] * 3.0 * x * d / k.r5;
        du[6] = du[4];
        du[7] = -alp[0] * k.b3 / k.r3;
        du[8] = alp[0] * 3.0 * y * d / k.r5;
        du[9] = -du[5];
        du[10] = -du[8];
        du[11] = alp[0] * k.c3 / k.r3;
        accum(u, pot[3], du);
    }
}

/* Part B, DC3Dfortran.f90:273-442 */
static void ub0(double x, double y, double d, double z, const double pot[4], double sd, double cd,
                const double alp[5], double u[12]) {
    pt_consts k; double du[12];
    for (int i = 0; i < 12; ++i) u[i] = 0.0;
    dccon1(x, y, d, sd, cd, &k);
    double c = d + z, rd = k.r + d;
    double d12 = 1.0 / (k.r * rd * rd);
    double d32 = d12 * (2.0 * k.r + d) / k.r2;
    double d33 = d12 * (3.0 * k.r + d) / (k.r2 * rd);
    double d53 = d12 * (8.0 * k.r2 + 9.0 * k.r * d + 3.0 * k.d2) / (k.r2 * k.r2 * rd);
    double d54 = d12 * (5.0 * k.r2 + 4.0 * k.r * d + k.d2) / k.r3 * d12;
    double fi1 = y * (d12 - k.x2 * d33);
    double fi2 = x * (d12 - k.y2 * d33);
    double fi3 = x / k.r3 - fi2;
    double fi4 = -k.xy * d32;
    double fi5 = 1.0 / (k.r * rd) - k.x2 * d32;
    double fj1 = -3.0 * k.xy * (d33 - k.x2 * d54);
    double fj2 = 1.0 / k.r3 - 3.0 * d12 + 3.0 * k.x2 * k.y2 * d54;
    double fj3 = k.a3 / k.r3 - fj2;
    double fj4 = -3.0 * k.xy / k.r5 - fj1;
    double fk1 = -y * (d32 - k.x2 * d53);
    double fk2 = -x * (d32 - k.y2 * d53);
    double fk3 = -3.0 * x * d / k.r5 - fk2;
    double a3 = alp[2];
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
        accum(u, pot[0], du);
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
        accum(u, pot[1], du);
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
        accum(u, pot[2], du);
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
        accum(u, pot[3], du);
    }
}

/* Part C, DC3Dfortran.f90:446-607 */
static void uc0(double x, double y, double d, double z, const double pot[4], double sd, double cd,
                const double alp[5], double u[12]) {
    pt_consts k; double du[12];
    for (int i = 0; i < 12; ++i) u[i] = 0.0;
    dccon1(x, y, d, sd, cd, &k);
    double s2d = 2.0 * sd * cd, c2d = cd * cd - sd * sd;
    double c = d + z, q2 = k.q * k.q, r7 = k.r5 * k.r2;
    double a7 = 1.0 - 7.0 * k.x2 / k.r2;
    double b5 = 1.0 - 5.0 * k.y2 / k.r2;
    double b7 = 1.0 - 7.0 * k.y2 / k.r2;
    double c5 = 1.0 - 5.0 * k.d2 / k.r2;
    double c7 = 1.0 - 7.0 * k.d2 / k.r2;
    double d7 = 2.0 - 7.0 * q2 / k.r2;
    double qr5 = 5.0 * k.q / k.r2, qr7 = 7.0 * k.q / k.r2, dr5 = 5.0 * d / k.r2;
    double A4 = alp[3], A5 = alp[4];
    double R3 = k.r3, R5 = k.r5, R2 = k.r2, Q = k.q, P = k.p, S = k.s, T = k.t;
    double a3 = k.a3, a5 = k.a5, c3 = k.c3, QR = k.qr, QRX = k.qrx;
    (void)c3;
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
        accum(u, pot[0], du);
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
        accum(u, pot[1], du);
    }
    if (pot[2] != 0.0) {
        du[0] = 3.0 * x / R5 * (-A4 * S + A5 * (c * Q * qr5 - z));
        du[1] = A4 / R3 * (s2d - 3.0 * y * S / R2) + A5 * 3.0 / R5 * (c * (T - y + y * Q * qr5) - y * z);
        du[2] = -A4 / R3 * (1.0 - a3 * sd * sd) - A5 * 3.0 / R5 * (c * (S - d + d * Q * qr5) - d * z);
        du[3] = -A4 * 3.0 * S / R5 * a5 + A5 * (c * QR * qr5 * a7 - 3.0 * z / R5 * a5);
        du[4] = 3.0 * x / R5 * (-A4 * (s2d - 5.0 * y * S / R2) - A5 * 5.0 / R2 * (c * (T - y + y * Q * qr7) - y * z));
        du[5] = 3.0 * x / R5 * (A4 * (1.0 - (2.0 + a5) * sd * sd) + A5 * 5.0 / R2 * (c * (S - d + d * Q * qr7) - d * z));
        du[6] = du[4];
        du[7] = 3.0 / R5 * (-A4 * (2.0 * y * s2d + S * b5) - A5 * (c * (2.0 * sd * sd + 10.0 * y * (T - y) / R2 - Q * qr5 * b7) + z * b5));
        du[8] = 3.0 / R5 * (A4 * y * (1.0 - a5 * sd * sd) + A5 * (c * (3.0 + a5) * s2d - y * dr5 * (c * d7 + z)));
        du[9] = 3.0 * x / R5 * (-A4 * (c2d + S * dr5) + A5 * (5.0 * c / R2 * (S - d + d * Q * qr7) - 1.0 - z * dr5));
        du[10] = 3.0 / R5 * (A4 * (d * b5 * s2d - y * c5 * c2d) + A5 * (c * ((3.0 + a5) * s2d - y * dr5 * d7) - y * (1.0 + z * dr5)));
        du[11] = 3.0 / R5 * (-A4 * d * (1.0 - a5 * sd * sd) - A5 * (c * (c2d + 10.0 * d * (S - d) / R2 - Q * qr5 * c7) + z * (1.0 + c5)));
        accum(u, pot[2], du);
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
        accum(u, pot[3], du);
    }
}

void orc_dc3d0(double alpha, double x, double y, double z, double depth, double dip,
               double pot1, double pot2, double pot3, double pot4, double u[12], int* iret) {
    const double EPS = 1.0e-6;
    double pot[4] = {pot1, pot2, pot3, pot4};
    double dua[12], dub[12], duc[12];
    for (int i = 0; i < 12; ++i) u[i] = 0.0;
    double xx = x, yy = y, zz = z, dd = depth + z;
    if (depth < 0.0) return;          /* IRET left untouched, DC3Dfortran.f90:67-70 */
    if (fabs(xx) < EPS) xx = 0.0;
    if (fabs(yy) < EPS) yy = 0.0;
    if (fabs(dd) < EPS) dd = 0.0;
    *iret = 0;
    if (z > 0.0) { *iret = 2; return; }
    double r = sqrt(xx * xx + yy * yy + dd * dd);
    if (r == 0.0) { *iret = 1; return; }
    double alp[5] = {(1.0 - alpha) / 2.0, alpha / 2.0, (1.0 - alpha) / alpha, 1.0 - alpha, alpha};
    /* DCCON0 */
    const double P18 = PI2 / 360.0;
    double sd = sin(dip * P18), cd = cos(dip * P18);
    if (fabs(cd) < EPS) {
        cd = 0.0;
        if (sd > 0.0) sd = 1.0;
        if (sd < 0.0) sd = -1.0;
    }
    ua0(xx, yy, dd, pot, sd, cd, alp, dua);
    for (int i = 0; i < 12; ++i) u[i] = (i < 9) ? u[i] - dua[i] : u[i] + dua[i];
    dd = depth - z;
    ua0(xx, yy, dd, pot, sd, cd, alp, dua);
    ub0(xx, yy, dd, zz, pot, sd, cd, alp, dub);
    uc0(xx, yy, dd, zz, pot, sd, cd, alp, duc);
    for (int i = 0; i < 12; ++i) {
        double du = dua[i] + dub[i] + zz * duc[i];
        if (i >= 9) du = du + duc[i - 9];
        u[i] = u[i] + du;
    }
}

/* ======================================================================
 * Geometry, mesh, Laplacian (init.f90)
 * ====================================================================== */
void orc_get_geometry(int iplane, const orc_fault_desc* f, const double* theta, int* nxi,
                      int* neta, double* xl, double* yl, double* xr, double* yr, double* dip,
                      double* dip_rad, double* lxi, double* leta) {
    /* iplane is 1-based as in init.f90:121-173; theta is 0-based storage of theta(1:) */
    double pi = 4.0 * atan(1.0);
    int np = f->nplane;
    double ymin = theta[0], ymax = theta[1];
    *nxi = f->nxi_ls[iplane - 1];
    *neta = f->neta_ls[iplane - 1];
    if (f->fix_xbend) {
        if (iplane == 1) { *xl = f->xmin; *yl = ymin; }
        else { *xl = f->xbend[iplane - 2]; *yl = theta[iplane]; }
        if (iplane == np) { *xr = f->xmax; *yr = ymax; }
        else { *xr = f->xbend[iplane - 1]; *yr = theta[iplane + 1]; }
        *dip = theta[np + iplane];
    } else {
        if (iplane == 1) { *xl = f->xmin; *yl = ymin; }
        else { *xl = theta[2 * iplane - 2]; *yl = theta[2 * iplane - 1]; }
        if (iplane == np) { *xr = f->xmax; *yr = ymax; }
        else { *xr = theta[2 * iplane]; *yr = theta[2 * iplane + 1]; }
        *dip = theta[2 * np + iplane - 1];
    }
    *dip_rad = *dip * pi / 180.0;
    *lxi = sqrt(pow(*xr - *xl, 2.0) + pow(*yr - *yl, 2.0));
    *leta = -f->zmin / sin(*dip_rad);
}

void orc_get_sigma(const double* theta, int fix_xbend, int nplane, double* log_sigma_sar2,
                   double* log_sigma_gnss2, double* log_alpha2) {
    if (fix_xbend) {
        *log_alpha2 = theta[2 * nplane + 1];
        *log_sigma_sar2 = theta[2 * nplane + 2];
        *log_sigma_gnss2 = theta[2 * nplane + 3];
    } else {
        *log_alpha2 = theta[3 * nplane];
        *log_sigma_sar2 = theta[3 * nplane + 1];
        *log_sigma_gnss2 = theta[3 * nplane + 2];
    }
}

void orc_mesh_sizes(const orc_fault_desc* f, int* nnode, int* ndof, int* npatch, int* ndof_index) {
    /* main.f90:217-238 */
    int nn = 0, nd = 0, npt = 0;
    if (ndof_index) ndof_index[0] = 1;
    for (int ip = 0; ip < f->nplane; ++ip) {
        int nxi = f->nxi_ls[ip], neta = f->neta_ls[ip];
        int nnp = (nxi + 1) * (neta + 1);
        int ndp = nnp - (nxi + 1);
        if (ip == 0) ndp -= neta;
        if (ip == f->nplane - 1) ndp -= neta;
        nn += nnp; nd += ndp; npt += nxi * neta;
        if (ndof_index) ndof_index[ip + 1] = ndof_index[ip] + ndp;
    }
    *nnode = nn; *ndof = nd; *npatch = npt;
}

void orc_discretize_fault(const orc_fault_desc* f, const double* theta, int* cny, double* coor,
                          int* node_to_elem_val, int* node_to_elem_size, int* id_dof) {
    int cnt = 1, offset_node = 0, offset_patch = 0;
    for (int iplane = 1; iplane <= f->nplane; ++iplane) {
        int nxi, neta;
        double xl, yl, xr, yr, dip, dip_rad, lxi, leta;
        orc_get_geometry(iplane, f, theta, &nxi, &neta, &xl, &yl, &xr, &yr, &dip, &dip_rad, &lxi, &leta);
        double dxi = lxi / nxi, deta = leta / neta;
        for (int j = 1; j <= neta + 1; ++j) {
            for (int i = 1; i <= nxi + 1; ++i) {
                int node_id = i + (nxi + 1) * (j - 1) + offset_node;
                coor[2 * (node_id - 1) + 0] = (i - 1) * dxi - lxi / 2.0;
                coor[2 * (node_id - 1) + 1] = (j - 1) * deta - leta / 2.0;
                node_to_elem_size[node_id - 1] = 0;
                for (int k = 0; k < 4; ++k) node_to_elem_val[4 * (node_id - 1) + k] = -1;
                if (j == 1) continue;
                if (iplane == 1 && i == 1) continue;
                if (iplane == f->nplane && i == nxi + 1) continue;
                id_dof[cnt - 1] = node_id;
                cnt++;
            }
        }
        for (int j = 1; j <= neta; ++j) {
            for (int i = 1; i <= nxi; ++i) {
                int patch_id = offset_patch + i + nxi * (j - 1);
                int n1 = i + (nxi + 1) * (j - 1) + offset_node;
                int nodes[4] = {n1, n1 + 1, n1 + 1 + nxi + 1, n1 + nxi + 1};
                for (int k = 0; k < 4; ++k) {
                    cny[4 * (patch_id - 1) + k] = nodes[k];
                    int nid = nodes[k];
                    node_to_elem_size[nid - 1]++;
                    node_to_elem_val[4 * (nid - 1) + node_to_elem_size[nid - 1] - 1] = patch_id;
                }
            }
        }
        offset_patch += nxi * neta;
        offset_node += (nxi + 1) * (neta + 1);
    }
}

void orc_gen_laplacian(const orc_fault_desc* f, const double* theta, int nnode, const int* id_dof,
                       int ndof, double* luni, double* lmat) {
    memset(luni, 0, sizeof(double) * (size_t)nnode * nnode);
#define LU(a, b) luni[(size_t)((b) - 1) * nnode + ((a) - 1)]
    int offset = 0;
    for (int iplane = 1; iplane <= f->nplane; ++iplane) {
        int nxi, neta;
        double xl, yl, xr, yr, dip, dip_rad, lxi, leta;
        orc_get_geometry(iplane, f, theta, &nxi, &neta, &xl, &yl, &xr, &yr, &dip, &dip_rad, &lxi, &leta);
        double dxi = lxi / nxi, deta = leta / neta;
        double cx = 1.0 / pow(dxi, 2.0), ce = 1.0 / pow(deta, 2.0), cc = 1.0 / (4.0 * dxi * deta);
        int W = nxi + 1;
        for (int n = offset + 1; n <= offset + W * (neta + 1); ++n) {
            int k = (n - 1 - offset) % W + 1;
            int l = (n - 1 - offset) / W + 1;
            LU(n, n) = LU(n, n) - 2.0 / pow(dxi, 2.0);
            if (k == nxi + 1) LU(n, n - 1) += cx; else LU(n, n + 1) += cx;
            if (k == 1) LU(n, n + 1) += cx; else LU(n, n - 1) += cx;
            if (l == neta + 1) LU(n, n - W) += ce; else LU(n, n + W) += ce;
            if (l == 1) LU(n, n + W) += ce; else LU(n, n - W) += ce;
            /* cross terms, init.f90:356-394 */
            if (k == nxi + 1 && l == neta + 1) LU(n, n - W - 1) += cc;
            else if (k == nxi + 1) LU(n, n + W - 1) += cc;
            else if (l == neta + 1) LU(n, n - W + 1) += cc;
            else LU(n, n + W + 1) += cc;
            if (k == 1 && l == neta + 1) LU(n, n - W + 1) -= cc;
            else if (k == 1) LU(n, n + W + 1) -= cc;
            else if (l == neta + 1) LU(n, n - W - 1) -= cc;
            else LU(n, n + W - 1) -= cc;
            if (k == 1 && l == 1) LU(n, n + W + 1) += cc;
            else if (k == 1) LU(n, n - W + 1) += cc;
            else if (l == 1) LU(n, n + W - 1) += cc;
            else LU(n, n - W - 1) += cc;
            if (k == nxi + 1 && l == 1) LU(n, n + W - 1) -= cc;
            else if (k == nxi + 1) LU(n, n - W - 1) -= cc;
            else if (l == 1) LU(n, n + W + 1) -= cc;
            else LU(n, n - W + 1) -= cc;
        }
        offset += W * (neta + 1);
    }
    size_t ldl = 2 * (size_t)nnode;
    memset(lmat, 0, sizeof(double) * ldl * 2 * ndof);
    for (int n = 1; n <= nnode; ++n) {
        for (int idof = 1; idof <= ndof; ++idof) {
            int jn = id_dof[idof - 1];
            lmat[(size_t)(2 * (idof - 1)) * ldl + 2 * (n - 1)] = LU(n, jn);
            lmat[(size_t)(2 * (idof - 1) + 1) * ldl + 2 * (n - 1) + 1] = LU(n, jn);
        }
    }
#undef LU
}

void orc_gen_sparse_lmat(const double* lmat, int nnode, int ndof, int* lmat_index,
                         double* lmat_val, int* ltmat_index, double* ltmat_val) {
    int nr = 2 * nnode, nc = 2 * ndof;
    for (int i = 0; i < nr; ++i) {
        int cnt = 0;
        for (int j = 0; j < nc; ++j) {
            double v = lmat[(size_t)j * nr + i];
            if (fabs(v) > 1e-8) {
                if (cnt < 9) { lmat_index[9 * i + cnt] = j + 1; lmat_val[9 * i + cnt] = v; }
                cnt++;
            }
        }
        for (; cnt < 9; ++cnt) { lmat_index[9 * i + cnt] = 1; lmat_val[9 * i + cnt] = 0.0; }
    }
    for (int i = 0; i < nc; ++i) {
        int cnt = 0;
        for (int j = 0; j < nr; ++j) {
            double v = lmat[(size_t)i * nr + j];
            if (fabs(v) > 1e-8) {
                if (cnt < 9) { ltmat_index[9 * i + cnt] = j + 1; ltmat_val[9 * i + cnt] = v; }
                cnt++;
            }
        }
        for (; cnt < 9; ++cnt) { ltmat_index[9 * i + cnt] = 1; ltmat_val[9 * i + cnt] = 0.0; }
    }
}

/* L^T L by plain ascending-k dot products (the reference uses dgemm, order unspecified) */
void orc_calc_ll(const double* lmat, int nnode, int ndof, double* llmat) {
    int n = 2 * nnode, m = 2 * ndof;
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            const double* ci = lmat + (size_t)i * n;
            const double* cj = lmat + (size_t)j * n;
            for (int k = 0; k < n; ++k) s += ci[k] * cj[k];
            llmat[(size_t)j * m + i] = s;
        }
}

/* ======================================================================
 * Green's function (gfunc.f90)
 * ====================================================================== */
void orc_call_dc3d0(double xs, double ys, double zs, double xo, double yo, double uxi,
                    double ueta, double dip, double area, double strike, double uret[3]) {
    double M_PI_ = 2.0 * asin(1.0);
    double sr = strike / 180.0 * M_PI_;
    double x = sin(sr) * (xo - xs) + cos(sr) * (yo - ys);
    double y = -cos(sr) * (xo - xs) + sin(sr) * (yo - ys);
    double u[12]; int iret = 0;
    orc_dc3d0(2.0 / 3.0, x, y, 0.0, -zs, dip, uxi * area, ueta * area, 0.0, 0.0, u, &iret);
    double uxr = sin(sr) * u[0] - cos(sr) * u[1];
    double uyr = cos(sr) * u[0] + sin(sr) * u[1];
    uret[0] = 100.0 * uxr;
    uret[1] = 100.0 * uyr;
    uret[2] = 100.0 * u[2];
}

/* calc_responce for a unit slip on node `inode` direction `idir` (gfunc.f90:102-194) */
static void responce_unit(const int* cny, const double* coor, int inode, int idir,
                          double xo, double yo, double xf, double yf, double zf, double strike,
                          double dip, const int* targets, int ntarget, double uobs[3]) {
    double M_PI_ = 2.0 * asin(1.0);
    double dip_rad = dip / 180.0 * M_PI_, strike_rad = strike / 180.0 * M_PI_;
    double rv[2] = {-1.0 / sqrt(3.0), 1.0 / sqrt(3.0)};
    uobs[0] = uobs[1] = uobs[2] = 0.0;
    for (int it = 0; it < ntarget; ++it) {
        int ie = targets[it];
        double xin[4], etan[4], uxin[4], uetan[4];
        for (int a = 0; a < 4; ++a) {
            int jn = cny[4 * (ie - 1) + a];
            xin[a] = coor[2 * (jn - 1)];
            etan[a] = coor[2 * (jn - 1) + 1];
            uxin[a] = (jn == inode && idir == 1) ? -1.0 : 0.0;
            uetan[a] = (jn == inode && idir == 2) ? 1.0 : 0.0;
        }
        double area = (xin[1] - xin[0]) * (etan[2] - etan[1]);
        for (int i1 = 0; i1 < 2; ++i1) {
            double r1 = rv[i1];
            for (int i2 = 0; i2 < 2; ++i2) {
                double r2 = rv[i2];
                double nv[4] = {(1.0 - r1) * (1.0 - r2) / 4.0, (1.0 + r1) * (1.0 - r2) / 4.0,
                                (1.0 + r1) * (1.0 + r2) / 4.0, (1.0 - r1) * (1.0 + r2) / 4.0};
                double xi = 0.0, eta = 0.0, uxi = 0.0, ueta = 0.0;
                for (int a = 0; a < 4; ++a) {
                    xi = xi + nv[a] * xin[a];
                    eta = eta + nv[a] * etan[a];
                    uxi = uxi + nv[a] * uxin[a];
                    ueta = ueta + nv[a] * uetan[a];
                }
                double xs = xf - eta * cos(dip_rad) * cos(strike_rad) + xi * sin(strike_rad);
                double ys = yf + eta * cos(dip_rad) * sin(strike_rad) + xi * cos(strike_rad);
                double zs = zf + eta * sin(dip_rad);
                double ur[3];
                orc_call_dc3d0(xs, ys, zs, xo, yo, uxi, ueta, dip, area / 4.0, strike, ur);
                for (int d = 0; d < 3; ++d) uobs[d] = uobs[d] + ur[d];
            }
        }
    }
}

void orc_calc_greens_func(const orc_fault_desc* f, const double* theta, int nobs, int nsar_total,
                          int ngnss, const double* obs_points, const double* obs_unitvec,
                          double* gmat) {
    int nnode, ndof, npatch;
    int* ndof_index = (int*)malloc(sizeof(int) * (f->nplane + 1));
    orc_mesh_sizes(f, &nnode, &ndof, &npatch, ndof_index);
    int* cny = (int*)malloc(sizeof(int) * 4 * npatch);
    double* coor = (double*)malloc(sizeof(double) * 2 * nnode);
    int* n2e = (int*)malloc(sizeof(int) * 4 * nnode);
    int* n2es = (int*)malloc(sizeof(int) * nnode);
    int* id_dof = (int*)malloc(sizeof(int) * ndof);
    orc_discretize_fault(f, theta, cny, coor, n2e, n2es, id_dof);
    double* resp = (double*)malloc(sizeof(double) * 3 * nobs);
    memset(gmat, 0, sizeof(double) * (size_t)nobs * 2 * ndof);
    double pi = 4.0 * atan(1.0);
    for (int iplane = 1; iplane <= f->nplane; ++iplane) {
        int nxi, neta;
        double xl, yl, xr, yr, dip, dip_rad, lxi, leta;
        orc_get_geometry(iplane, f, theta, &nxi, &neta, &xl, &yl, &xr, &yr, &dip, &dip_rad, &lxi, &leta);
        double strike_rad = atan2(xr - xl, yr - yl);
        double strike = strike_rad / pi * 180.0;
        double xf = (xl + xr) / 2.0 - (f->zmin * cos(strike_rad)) / (2.0 * tan(dip_rad));
        double yf = (yl + yr) / 2.0 + (f->zmin * sin(strike_rad)) / (2.0 * tan(dip_rad));
        double zf = f->zmin / 2.0;
        for (int idof = ndof_index[iplane - 1]; idof <= ndof_index[iplane] - 1; ++idof) {
            int inode = id_dof[idof - 1];
            int nt = n2es[inode - 1];
            const int* targets = n2e + 4 * (inode - 1);
            for (int idir = 1; idir <= 2; ++idir) {
                for (int io = 0; io < nsar_total; ++io) {
                    double u[3];
                    responce_unit(cny, coor, inode, idir, obs_points[2 * io], obs_points[2 * io + 1],
                                  xf, yf, zf, strike, dip, targets, nt, u);
                    for (int d = 0; d < 3; ++d) resp[3 * io + d] = u[d];
                }
                for (int ig = 0; ig < ngnss; ++ig) {
                    int row = nsar_total + 3 * ig;
                    double u[3];
                    responce_unit(cny, coor, inode, idir, obs_points[2 * row], obs_points[2 * row + 1],
                                  xf, yf, zf, strike, dip, targets, nt, u);
                    for (int c = 0; c < 3; ++c)
                        for (int d = 0; d < 3; ++d) resp[3 * (row + c) + d] = u[d];
                }
                double* col = gmat + (size_t)(2 * (idof - 1) + idir - 1) * nobs;
                for (int io = 0; io < nobs; ++io)
                    for (int d = 0; d < 3; ++d)
                        col[io] = col[io] + resp[3 * io + d] * obs_unitvec[3 * io + d];
            }
        }
    }
    free(resp); free(id_dof); free(n2es); free(n2e); free(coor); free(cny); free(ndof_index);
}
