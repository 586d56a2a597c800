/*
 * oracle/okada92.c -- TEST INFRASTRUCTURE (the CPU oracle), never shipped.
 *
 * FP64 restatement of Okada (1992), "Internal deformation due to shear and
 * tensile faults in a half-space", BSSA 82(2), 1018-1040: displacement and
 * displacement gradients at (x, y, z <= 0) due to a finite rectangular
 * dislocation.  The reference calls this routine through GeoGreensFunctions.jl
 * (`dc3d`, un-vendored, compat "0.2", Project.toml:11,32) at
 * src/BEM/GF.jl:49-53 and GF.jl:156-160.
 *
 * Conventions (those of the paper and of the call sites):
 *   x along strike, y horizontal normal, z up (z <= 0 inside the medium);
 *   the fault spans xi in [al1, al2] along strike and eta in [aw1, aw2]
 *   along dip (eta up-dip), its reference point at depth `depth`;
 *   alpha = (lambda + mu) / (lambda + 2 mu); dip in degrees.
 *   Output u[12] = ux uy uz  uxx uyx uzx  uxy uyy uzy  uxz uyz uzz
 *   (u_ab = d u_a / d b), the order the reference indexes at GF.jl:77-86.
 *
 * Structure follows the paper: the half-space field is
 *   u = u_A(x,y,-z) - u_A(x,y,z) + u_B(x,y,z) + z u_C(x,y,z)
 * (paper eq. 5), each part a Chinnery corner sum f(xi,eta)|| over the four
 * fault corners (paper eq. 23).  The A, B and C functions are Tables 6-9 of
 * the paper; the singular-point conventions (q = 0, R + xi = 0, R + eta = 0,
 * receivers on a fault edge) are the paper's section 6.
 *
 * Parity status: no golden vectors exist for dc3d in the reference.  This
 * function is pinned by tests/test_oracle_okada.py against what the
 * mathematics fixes: gradients vs. finite differences of the displacement,
 * Navier equilibrium, a traction-free free surface, the unit displacement
 * jump across the fault, additivity when a fault is split, and the 2-D
 * antiplane long-fault limit.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

#define OQ_EPS 1.0e-6
#define OQ_PI2 6.283185307179586476925286766559

typedef struct {
    double alp1, alp2, alp3, alp4, alp5;
    double sd, cd, sdsd, cdcd, sdcd, s2d, c2d;
} consts_t;

typedef struct {
    double xi, et, q;
    double xi2, et2, q2, r, r2, r3, r5;
    double y, d, tt, alx, ale, x11, y11, x32, y32;
    double ey, ez, fy, fz, gy, gz, hy, hz;
    int zero_r;
} corner_t;

static void set_consts(consts_t *c, double alpha, double dip)
{
    c->alp1 = (1.0 - alpha) / 2.0;
    c->alp2 = alpha / 2.0;
    c->alp3 = (1.0 - alpha) / alpha;
    c->alp4 = 1.0 - alpha;
    c->alp5 = alpha;
    c->sd = oq_sind(dip);
    c->cd = oq_cosd(dip);
    if (fabs(c->cd) < OQ_EPS) {
        c->cd = 0.0;
        if (c->sd > 0.0) c->sd = 1.0;
        if (c->sd < 0.0) c->sd = -1.0;
    }
    c->sdsd = c->sd * c->sd;
    c->cdcd = c->cd * c->cd;
    c->sdcd = c->sd * c->cd;
    c->s2d = 2.0 * c->sdcd;
    c->c2d = c->cdcd - c->sdsd;
}

/* Per-corner geometric terms (paper eqs. 10-14 and Table 8 auxiliaries). */
static void set_corner(corner_t *k, const consts_t *c, double xi, double et,
                       double q, int kxi, int ket)
{
    if (fabs(xi) < OQ_EPS) xi = 0.0;
    if (fabs(et) < OQ_EPS) et = 0.0;
    if (fabs(q) < OQ_EPS) q = 0.0;
    k->xi = xi; k->et = et; k->q = q;
    k->xi2 = xi * xi; k->et2 = et * et; k->q2 = q * q;
    k->r2 = k->xi2 + k->et2 + k->q2;
    k->r = sqrt(k->r2);
    k->zero_r = (k->r == 0.0);
    if (k->zero_r) return;
    k->r3 = k->r * k->r2;
    k->r5 = k->r3 * k->r2;
    k->y = et * c->cd + q * c->sd;
    k->d = et * c->sd - q * c->cd;
    k->tt = (q == 0.0) ? 0.0 : atan(xi * et / (q * k->r));
    if (kxi) {
        k->alx = -log(k->r - xi);
        k->x11 = 0.0;
        k->x32 = 0.0;
    } else {
        double rxi = k->r + xi;
        k->alx = log(rxi);
        k->x11 = 1.0 / (k->r * rxi);
        k->x32 = (k->r + rxi) * k->x11 * k->x11 / k->r;
    }
    if (ket) {
        k->ale = -log(k->r - et);
        k->y11 = 0.0;
        k->y32 = 0.0;
    } else {
        double ret = k->r + et;
        k->ale = log(ret);
        k->y11 = 1.0 / (k->r * ret);
        k->y32 = (k->r + ret) * k->y11 * k->y11 / k->r;
    }
    k->ey = c->sd / k->r - k->y * q / k->r3;
    k->ez = c->cd / k->r + k->d * q / k->r3;
    k->fy = k->d / k->r3 + k->xi2 * k->y32 * c->sd;
    k->fz = k->y / k->r3 + k->xi2 * k->y32 * c->cd;
    k->gy = 2.0 * k->x11 * c->sd - k->y * q * k->x32;
    k->gz = 2.0 * k->x11 * c->cd + k->d * q * k->x32;
    k->hy = k->d * q * k->x32 + xi * q * k->y32 * c->sd;
    k->hz = k->y * q * k->x32 + xi * q * k->y32 * c->cd;
}

/* Part A (paper Table 6, f^A terms), 12 components in the fault frame. */
static void part_a(const corner_t *k, const consts_t *c, double ds, double dd,
                   double dt, double u[12])
{
    double du[12];
    const double xi = k->xi, et = k->et, q = k->q, r = k->r;
    const double xy = xi * k->y11, qx = q * k->x11, qy = q * k->y11;
    const double a1 = c->alp1, a2 = c->alp2, sd = c->sd, cd = c->cd;
    int i;
    memset(u, 0, 12 * sizeof(double));
    if (ds != 0.0) {
        du[0] = k->tt / 2.0 + a2 * xi * qy;
        du[1] = a2 * q / r;
        du[2] = a1 * k->ale - a2 * q * qy;
        du[3] = -a1 * qy - a2 * k->xi2 * q * k->y32;
        du[4] = -a2 * xi * q / k->r3;
        du[5] = a1 * xy + a2 * xi * k->q2 * k->y32;
        du[6] = a1 * xy * sd + a2 * xi * k->fy + k->d / 2.0 * k->x11;
        du[7] = a2 * k->ey;
        du[8] = a1 * (cd / r + qy * sd) - a2 * q * k->fy;
        du[9] = a1 * xy * cd + a2 * xi * k->fz + k->y / 2.0 * k->x11;
        du[10] = a2 * k->ez;
        du[11] = -a1 * (sd / r - qy * cd) - a2 * q * k->fz;
        for (i = 0; i < 12; ++i) u[i] += ds / OQ_PI2 * du[i];
    }
    if (dd != 0.0) {
        du[0] = a2 * q / r;
        du[1] = k->tt / 2.0 + a2 * et * qx;
        du[2] = a1 * k->alx - a2 * q * qx;
        du[3] = -a2 * xi * q / k->r3;
        du[4] = -qy / 2.0 - a2 * et * q / k->r3;
        du[5] = a1 / r + a2 * k->q2 / k->r3;
        du[6] = a2 * k->ey;
        du[7] = a1 * k->d * k->x11 + xy / 2.0 * sd + a2 * et * k->gy;
        du[8] = a1 * k->y * k->x11 - a2 * q * k->gy;
        du[9] = a2 * k->ez;
        du[10] = a1 * k->y * k->x11 + xy / 2.0 * cd + a2 * et * k->gz;
        du[11] = -a1 * k->d * k->x11 - a2 * q * k->gz;
        for (i = 0; i < 12; ++i) u[i] += dd / OQ_PI2 * du[i];
    }
    if (dt != 0.0) {
        du[0] = -a1 * k->ale - a2 * q * qy;
        du[1] = -a1 * k->alx - a2 * q * qx;
        du[2] = k->tt / 2.0 - a2 * (et * qx + xi * qy);
        du[3] = -a1 * xy + a2 * xi * k->q2 * k->y32;
        du[4] = -a1 / r + a2 * k->q2 / k->r3;
        du[5] = -a1 * qy - a2 * q * k->q2 * k->y32;
        du[6] = -a1 * (cd / r + qy * sd) - a2 * q * k->fy;
        du[7] = -a1 * k->y * k->x11 - a2 * q * k->gy;
        du[8] = a1 * (k->d * k->x11 + xy * sd) + a2 * q * k->hy;
        du[9] = a1 * (sd / r - qy * cd) - a2 * q * k->fz;
        du[10] = a1 * k->d * k->x11 - a2 * q * k->gz;
        du[11] = a1 * (k->y * k->x11 + xy * cd) + a2 * q * k->hz;
        for (i = 0; i < 12; ++i) u[i] += dt / OQ_PI2 * du[i];
    }
}

/* Part B (paper Table 6, f^B terms, with the I and J/K integrals of
 * Table 7 including the cos(dip) = 0 limits). */
static void part_b(const corner_t *k, const consts_t *c, double ds, double dd,
                   double dt, double u[12])
{
    double du[12];
    const double xi = k->xi, et = k->et, q = k->q, r = k->r;
    const double sd = c->sd, cd = c->cd, a3 = c->alp3;
    const double rd = r + k->d;
    const double d11 = 1.0 / (r * rd);
    const double aj2 = xi * k->y / rd * d11;
    const double aj5 = -(k->d + k->y * k->y / rd) * d11;
    double ai1, ai2, ai3, ai4, ak1, ak2, ak3, ak4, aj1, aj3, aj4, aj6;
    double xy, qx, qy;
    int i;
    if (cd != 0.0) {
        if (xi == 0.0) {
            ai4 = 0.0;
        } else {
            double x = sqrt(k->xi2 + k->q2);
            ai4 = 1.0 / c->cdcd *
                  (xi / rd * c->sdcd +
                   2.0 * atan((et * (x + q * cd) + x * (r + x) * sd) /
                              (xi * (r + x) * cd)));
        }
        ai3 = (k->y * cd / rd - k->ale + sd * log(rd)) / c->cdcd;
        ak1 = xi * (d11 - k->y11 * sd) / cd;
        ak3 = (q * k->y11 - k->y * d11) / cd;
        aj3 = (ak1 - aj2 * sd) / cd;
        aj6 = (ak3 - aj5 * sd) / cd;
    } else {
        double rd2 = rd * rd;
        ai3 = (et / rd + k->y * q / rd2 - k->ale) / 2.0;
        ai4 = xi * k->y / rd2 / 2.0;
        ak1 = xi * q / rd * d11;
        ak3 = sd / rd * (k->xi2 * d11 - 1.0);
        aj3 = -xi / rd2 * (k->q2 * d11 - 0.5);
        aj6 = -k->y / rd2 * (k->xi2 * d11 - 0.5);
    }
    xy = xi * k->y11;
    ai1 = -xi / rd * cd - ai4 * sd;
    ai2 = log(rd) + ai3 * sd;
    ak2 = 1.0 / r + ak3 * sd;
    ak4 = xy * cd - ak1 * sd;
    aj1 = aj5 * cd - aj6 * sd;
    aj4 = -xy - aj2 * cd + aj3 * sd;
    qx = q * k->x11;
    qy = q * k->y11;
    memset(u, 0, 12 * sizeof(double));
    if (ds != 0.0) {
        du[0] = -xi * qy - k->tt - a3 * ai1 * sd;
        du[1] = -q / r + a3 * k->y / rd * sd;
        du[2] = q * qy - a3 * ai2 * sd;
        du[3] = k->xi2 * q * k->y32 - a3 * aj1 * sd;
        du[4] = xi * q / k->r3 - a3 * aj2 * sd;
        du[5] = -xi * k->q2 * k->y32 - a3 * aj3 * sd;
        du[6] = -xi * k->fy - k->d * k->x11 + a3 * (xy + aj4) * sd;
        du[7] = -k->ey + a3 * (1.0 / r + aj5) * sd;
        du[8] = q * k->fy - a3 * (qy - aj6) * sd;
        du[9] = -xi * k->fz - k->y * k->x11 + a3 * ak1 * sd;
        du[10] = -k->ez + a3 * k->y * d11 * sd;
        du[11] = q * k->fz + a3 * ak2 * sd;
        for (i = 0; i < 12; ++i) u[i] += ds / OQ_PI2 * du[i];
    }
    if (dd != 0.0) {
        du[0] = -q / r + a3 * ai3 * c->sdcd;
        du[1] = -et * qx - k->tt - a3 * xi / rd * c->sdcd;
        du[2] = q * qx + a3 * ai4 * c->sdcd;
        du[3] = xi * q / k->r3 + a3 * aj4 * c->sdcd;
        du[4] = et * q / k->r3 + qy + a3 * aj5 * c->sdcd;
        du[5] = -k->q2 / k->r3 + a3 * aj6 * c->sdcd;
        du[6] = -k->ey + a3 * aj1 * c->sdcd;
        du[7] = -et * k->gy - xy * sd + a3 * aj2 * c->sdcd;
        du[8] = q * k->gy + a3 * aj3 * c->sdcd;
        du[9] = -k->ez - a3 * ak3 * c->sdcd;
        du[10] = -et * k->gz - xy * cd - a3 * xi * d11 * c->sdcd;
        du[11] = q * k->gz - a3 * ak4 * c->sdcd;
        for (i = 0; i < 12; ++i) u[i] += dd / OQ_PI2 * du[i];
    }
    if (dt != 0.0) {
        du[0] = q * qy - a3 * ai3 * c->sdsd;
        du[1] = q * qx + a3 * xi / rd * c->sdsd;
        du[2] = et * qx + xi * qy - k->tt - a3 * ai4 * c->sdsd;
        du[3] = -xi * k->q2 * k->y32 - a3 * aj4 * c->sdsd;
        du[4] = -k->q2 / k->r3 - a3 * aj5 * c->sdsd;
        du[5] = q * k->q2 * k->y32 - a3 * aj6 * c->sdsd;
        du[6] = q * k->fy - a3 * aj1 * c->sdsd;
        du[7] = q * k->gy - a3 * aj2 * c->sdsd;
        du[8] = -q * k->hy - a3 * aj3 * c->sdsd;
        du[9] = q * k->fz + a3 * ak3 * c->sdsd;
        du[10] = q * k->gz + a3 * xi * d11 * c->sdsd;
        du[11] = -q * k->hz + a3 * ak4 * c->sdsd;
        for (i = 0; i < 12; ++i) u[i] += dt / OQ_PI2 * du[i];
    }
}

/* Part C (paper Table 6, f^C terms, Table 9 auxiliaries); z is the
 * receiver's z. */
static void part_c(const corner_t *k, const consts_t *c, double z, double ds,
                   double dd, double dt, double u[12])
{
    double du[12];
    const double xi = k->xi, et = k->et, q = k->q, r = k->r;
    const double sd = c->sd, cd = c->cd, a4 = c->alp4, a5 = c->alp5;
    const double cc = k->d + z;
    const double x53 = (8.0 * k->r2 + 9.0 * r * xi + 3.0 * k->xi2) * k->x11 *
                       k->x11 * k->x11 / k->r2;
    const double y53 = (8.0 * k->r2 + 9.0 * r * et + 3.0 * k->et2) * k->y11 *
                       k->y11 * k->y11 / k->r2;
    const double h = q * cd - z;
    const double z32 = sd / k->r3 - h * k->y32;
    const double z53 = 3.0 * sd / k->r5 - h * y53;
    const double y0 = k->y11 - k->xi2 * k->y32;
    const double z0 = z32 - k->xi2 * z53;
    const double ppy = cd / k->r3 + q * k->y32 * sd;
    const double ppz = sd / k->r3 - q * k->y32 * cd;
    const double qq = z * k->y32 + z32 + z0;
    const double qqy = 3.0 * cc * k->d / k->r5 - qq * sd;
    const double qqz = 3.0 * cc * k->y / k->r5 - qq * cd + q * k->y32;
    const double xy = xi * k->y11, qx = q * k->x11, qy = q * k->y11;
    const double qr = 3.0 * q / k->r5;
    const double cdr = (cc + k->d) / k->r3;
    const double yy0 = k->y / k->r3 - y0 * cd;
    const double yv = k->y, dv = k->d;
    int i;
    (void)qx;
    memset(u, 0, 12 * sizeof(double));
    if (ds != 0.0) {
        du[0] = a4 * xy * cd - a5 * xi * q * z32;
        du[1] = a4 * (cd / r + 2.0 * qy * sd) - a5 * cc * q / k->r3;
        du[2] = a4 * qy * cd - a5 * (cc * et / k->r3 - z * k->y11 + k->xi2 * z32);
        du[3] = a4 * y0 * cd - a5 * q * z0;
        du[4] = -a4 * xi * (cd / k->r3 + 2.0 * q * k->y32 * sd) + a5 * cc * xi * qr;
        du[5] = -a4 * xi * q * k->y32 * cd + a5 * xi * (3.0 * cc * et / k->r5 - qq);
        du[6] = -a4 * xi * ppy * cd - a5 * xi * qqy;
        du[7] = a4 * 2.0 * (dv / k->r3 - y0 * sd) * sd - yv / k->r3 * cd -
                a5 * (cdr * sd - et / k->r3 - cc * yv * qr);
        du[8] = -a4 * q / k->r3 + yy0 * sd +
                a5 * (cdr * cd + cc * dv * qr - (y0 * cd + q * z0) * sd);
        du[9] = a4 * xi * ppz * cd - a5 * xi * qqz;
        du[10] = a4 * 2.0 * (yv / k->r3 - y0 * cd) * sd + dv / k->r3 * cd -
                 a5 * (cdr * cd + cc * dv * qr);
        du[11] = yy0 * cd - a5 * (cdr * sd - cc * yv * qr - y0 * c->sdsd + q * z0 * cd);
        for (i = 0; i < 12; ++i) u[i] += ds / OQ_PI2 * du[i];
    }
    if (dd != 0.0) {
        du[0] = a4 * cd / r - qy * sd - a5 * cc * q / k->r3;
        du[1] = a4 * yv * k->x11 - a5 * cc * et * q * k->x32;
        du[2] = -dv * k->x11 - xy * sd - a5 * cc * (k->x11 - k->q2 * k->x32);
        du[3] = -a4 * xi / k->r3 * cd + a5 * cc * xi * qr + xi * q * k->y32 * sd;
        du[4] = -a4 * yv / k->r3 + a5 * cc * et * qr;
        du[5] = dv / k->r3 - y0 * sd + a5 * cc / k->r3 * (1.0 - 3.0 * k->q2 / k->r2);
        du[6] = -a4 * et / k->r3 + y0 * c->sdsd - a5 * (cdr * sd - cc * yv * qr);
        du[7] = a4 * (k->x11 - yv * yv * k->x32) -
                a5 * cc * ((dv + 2.0 * q * cd) * k->x32 - yv * et * q * x53);
        du[8] = xi * ppy * sd + yv * dv * k->x32 +
                a5 * cc * ((yv + 2.0 * q * sd) * k->x32 - yv * k->q2 * x53);
        du[9] = -q / k->r3 + y0 * c->sdcd - a5 * (cdr * cd + cc * dv * qr);
        du[10] = a4 * yv * dv * k->x32 -
                 a5 * cc * ((yv - 2.0 * q * sd) * k->x32 + dv * et * q * x53);
        du[11] = -xi * ppz * sd + k->x11 - dv * dv * k->x32 -
                 a5 * cc * ((dv - 2.0 * q * cd) * k->x32 - dv * k->q2 * x53);
        for (i = 0; i < 12; ++i) u[i] += dd / OQ_PI2 * du[i];
    }
    if (dt != 0.0) {
        du[0] = -a4 * (sd / r + qy * cd) - a5 * (z * k->y11 - k->q2 * z32);
        du[1] = a4 * 2.0 * xy * sd + dv * k->x11 - a5 * cc * (k->x11 - k->q2 * k->x32);
        du[2] = a4 * (yv * k->x11 + xy * cd) + a5 * q * (cc * et * k->x32 + xi * z32);
        du[3] = a4 * xi / k->r3 * sd + xi * q * k->y32 * cd +
                a5 * xi * (3.0 * cc * et / k->r5 - 2.0 * z32 - z0);
        du[4] = a4 * 2.0 * y0 * sd - dv / k->r3 + a5 * cc / k->r3 * (1.0 - 3.0 * k->q2 / k->r2);
        du[5] = -a4 * yy0 - a5 * (cc * et * qr - q * z0);
        du[6] = a4 * (q / k->r3 + y0 * c->sdcd) + a5 * (z / k->r3 * cd + cc * dv * qr - q * z0 * sd);
        du[7] = -a4 * 2.0 * xi * ppy * sd - yv * dv * k->x32 +
                a5 * cc * ((yv + 2.0 * q * sd) * k->x32 - yv * k->q2 * x53);
        du[8] = -a4 * (xi * ppy * cd - k->x11 + yv * yv * k->x32) +
                a5 * (cc * ((dv + 2.0 * q * cd) * k->x32 - yv * et * q * x53) + xi * qqy);
        du[9] = -et / k->r3 + y0 * c->cdcd - a5 * (z / k->r3 * sd - cc * yv * qr - y0 * c->sdsd + q * z0 * cd);
        du[10] = a4 * 2.0 * xi * ppz * sd - k->x11 + dv * dv * k->x32 -
                 a5 * cc * ((dv - 2.0 * q * cd) * k->x32 - dv * k->q2 * x53);
        du[11] = a4 * (xi * ppz * cd + yv * dv * k->x32) +
                 a5 * (cc * ((yv - 2.0 * q * sd) * k->x32 + dv * et * q * x53) + xi * qqz);
        for (i = 0; i < 12; ++i) u[i] += dt / OQ_PI2 * du[i];
    }
}

/* Flags for the negative extension of a fault edge (paper section 6). */
static void edge_flags(const double xi[2], const double et[2], double q,
                       int kxi[2], int ket[2])
{
    double r12 = sqrt(xi[0] * xi[0] + et[1] * et[1] + q * q);
    double r21 = sqrt(xi[1] * xi[1] + et[0] * et[0] + q * q);
    double r22 = sqrt(xi[1] * xi[1] + et[1] * et[1] + q * q);
    kxi[0] = kxi[1] = ket[0] = ket[1] = 0;
    if (xi[0] < 0.0 && r21 + xi[1] < OQ_EPS) kxi[0] = 1;
    if (xi[0] < 0.0 && r22 + xi[1] < OQ_EPS) kxi[1] = 1;
    if (et[0] < 0.0 && r12 + et[1] < OQ_EPS) ket[0] = 1;
    if (et[0] < 0.0 && r22 + et[1] < OQ_EPS) ket[1] = 1;
}

static int on_edge(const double xi[2], const double et[2], double q)
{
    return q == 0.0 && ((xi[0] * xi[1] <= 0.0 && et[0] * et[1] == 0.0) ||
                        (et[0] * et[1] <= 0.0 && xi[0] * xi[1] == 0.0));
}

int oq_dc3d(double alpha, double x, double y, double z, double depth,
            double dip, double al1, double al2, double aw1, double aw2,
            double disl1, double disl2, double disl3, double u[12])
{
    consts_t c;
    double xi[2], et[2], q, d, p;
    int kxi[2], ket[2], j, kk, i;
    double dua[12], dub[12], duc[12], du[12];

    memset(u, 0, 12 * sizeof(double));
    if (z > 0.0) return 2;
    set_consts(&c, alpha, dip);
    xi[0] = x - al1;
    xi[1] = x - al2;
    if (fabs(xi[0]) < OQ_EPS) xi[0] = 0.0;
    if (fabs(xi[1]) < OQ_EPS) xi[1] = 0.0;

    /* real source */
    d = depth + z;
    p = y * c.cd + d * c.sd;
    q = y * c.sd - d * c.cd;
    et[0] = p - aw1;
    et[1] = p - aw2;
    if (fabs(q) < OQ_EPS) q = 0.0;
    if (fabs(et[0]) < OQ_EPS) et[0] = 0.0;
    if (fabs(et[1]) < OQ_EPS) et[1] = 0.0;
    if (on_edge(xi, et, q)) return 1;
    edge_flags(xi, et, q, kxi, ket);
    for (kk = 0; kk < 2; ++kk) {
        for (j = 0; j < 2; ++j) {
            corner_t k;
            set_corner(&k, &c, xi[j], et[kk], q, kxi[kk], ket[j]);
            if (k.zero_r) continue;
            part_a(&k, &c, disl1, disl2, disl3, dua);
            for (i = 0; i < 12; i += 3) {
                du[i] = -dua[i];
                du[i + 1] = -dua[i + 1] * c.cd + dua[i + 2] * c.sd;
                du[i + 2] = -dua[i + 1] * c.sd - dua[i + 2] * c.cd;
                if (i == 9) {
                    du[i] = -du[i];
                    du[i + 1] = -du[i + 1];
                    du[i + 2] = -du[i + 2];
                }
            }
            for (i = 0; i < 12; ++i) u[i] += (j + kk != 1) ? du[i] : -du[i];
        }
    }

    /* image source */
    d = depth - z;
    p = y * c.cd + d * c.sd;
    q = y * c.sd - d * c.cd;
    et[0] = p - aw1;
    et[1] = p - aw2;
    if (fabs(q) < OQ_EPS) q = 0.0;
    if (fabs(et[0]) < OQ_EPS) et[0] = 0.0;
    if (fabs(et[1]) < OQ_EPS) et[1] = 0.0;
    if (on_edge(xi, et, q)) {
        memset(u, 0, 12 * sizeof(double));
        return 1;
    }
    edge_flags(xi, et, q, kxi, ket);
    for (kk = 0; kk < 2; ++kk) {
        for (j = 0; j < 2; ++j) {
            corner_t k;
            set_corner(&k, &c, xi[j], et[kk], q, kxi[kk], ket[j]);
            if (k.zero_r) continue;
            part_a(&k, &c, disl1, disl2, disl3, dua);
            part_b(&k, &c, disl1, disl2, disl3, dub);
            part_c(&k, &c, z, disl1, disl2, disl3, duc);
            for (i = 0; i < 12; i += 3) {
                du[i] = dua[i] + dub[i] + z * duc[i];
                du[i + 1] = (dua[i + 1] + dub[i + 1] + z * duc[i + 1]) * c.cd -
                            (dua[i + 2] + dub[i + 2] + z * duc[i + 2]) * c.sd;
                du[i + 2] = (dua[i + 1] + dub[i + 1] - z * duc[i + 1]) * c.sd +
                            (dua[i + 2] + dub[i + 2] - z * duc[i + 2]) * c.cd;
                if (i == 9) {
                    du[9] += duc[0];
                    du[10] += duc[1] * c.cd - duc[2] * c.sd;
                    du[11] += -duc[1] * c.sd - duc[2] * c.cd;
                }
            }
            for (i = 0; i < 12; ++i) u[i] += (j + kk != 1) ? du[i] : -du[i];
        }
    }
    return 0;
}
