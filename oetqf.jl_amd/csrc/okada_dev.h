// okada_dev.h -- device Okada (1992) rectangular dislocation, FP64.
//
// Displacement gradients at (x, y, z <= 0) from a unit strike- or dip-slip
// on the rectangle xi in [al1, al2], eta in [aw1, aw2] of a fault whose
// reference point lies at `depth` (Okada 1992, BSSA 82(2): eq. 5 split
// u = uA(x,y,-z) - uA(x,y,z) + uB + z uC, Chinnery corner sums, Tables 6-9,
// singular-point rules of section 6).  The reference reaches the same
// closed form through GeoGreensFunctions.dc3d (src/BEM/GF.jl:49-53,156-160).
//
// GPU shape: the slip type is a template parameter, so one thread evaluates
// only the terms of the slip the builder needs; the 12 gradient components
// are accumulated in registers and only the 9 gradients are returned
// (the displacement is never used by the stress builders).
#pragma once
#include <hip/hip_runtime.h>

namespace oq {

struct OkConst {
    double a1, a2, a3, a4, a5;  // (1-α)/2, α/2, (1-α)/α, 1-α, α
    double sd, cd;
};

__host__ __device__ inline OkConst ok_const(double alpha, double sd, double cd)
{
    OkConst c;
    c.a1 = (1.0 - alpha) / 2.0;
    c.a2 = alpha / 2.0;
    c.a3 = (1.0 - alpha) / alpha;
    c.a4 = 1.0 - alpha;
    c.a5 = alpha;
    if (fabs(cd) < 1e-6) {
        cd = 0.0;
        sd = sd > 0.0 ? 1.0 : (sd < 0.0 ? -1.0 : 0.0);
    }
    c.sd = sd;
    c.cd = cd;
    return c;
}

#define OQ_OK_EPS 1.0e-6

// Per-corner terms; returns false when R == 0.
struct OkCorner {
    double xi, et, q, xi2, et2, q2, r, r2, r3, r5, yb, db;
    double tt, alx, ale, x11, y11, x32, y32;
};

__device__ __forceinline__ bool ok_corner(OkCorner &k, const OkConst &c,
                                          double xi, double et, double q,
                                          bool kxi, bool ket)
{
    if (fabs(xi) < OQ_OK_EPS) xi = 0.0;
    if (fabs(et) < OQ_OK_EPS) et = 0.0;
    if (fabs(q) < OQ_OK_EPS) q = 0.0;
    k.xi = xi; k.et = et; k.q = q;
    k.xi2 = xi * xi; k.et2 = et * et; k.q2 = q * q;
    k.r2 = k.xi2 + k.et2 + k.q2;
    k.r = sqrt(k.r2);
    if (k.r == 0.0) return false;
    k.r3 = k.r * k.r2;
    k.r5 = k.r3 * k.r2;
    k.yb = et * c.cd + q * c.sd;
    k.db = et * c.sd - q * c.cd;
    k.tt = (q == 0.0) ? 0.0 : atan(xi * et / (q * k.r));
    if (kxi) {
        k.alx = -log(k.r - xi);
        k.x11 = 0.0;
        k.x32 = 0.0;
    } else {
        const double rxi = k.r + xi;
        k.alx = log(rxi);
        k.x11 = 1.0 / (k.r * rxi);
        k.x32 = (k.r + rxi) * k.x11 * k.x11 / k.r;
    }
    if (ket) {
        k.ale = -log(k.r - et);
        k.y11 = 0.0;
        k.y32 = 0.0;
    } else {
        const double ret = k.r + et;
        k.ale = log(ret);
        k.y11 = 1.0 / (k.r * ret);
        k.y32 = (k.r + ret) * k.y11 * k.y11 / k.r;
    }
    return true;
}

// Fault-frame gradient part of f^A (9 values: d/dx, d/dy, d/dz of the three
// fault-frame displacement components) for unit slip SLIP.
template <int SLIP>
__device__ __forceinline__ void ok_part_a(const OkCorner &k, const OkConst &c, double g[9])
{
    const double xi = k.xi, et = k.et, q = k.q, r = k.r, sd = c.sd, cd = c.cd;
    const double xy = xi * k.y11, qy = q * k.y11;
    const double ey = sd / r - k.yb * q / k.r3, ez = cd / r + k.db * q / k.r3;
    const double s = 1.0 / 6.283185307179586476925286766559;
    if (SLIP == 0) {
        const double fy = k.db / k.r3 + k.xi2 * k.y32 * sd;
        const double fz = k.yb / k.r3 + k.xi2 * k.y32 * cd;
        g[0] = -c.a1 * qy - c.a2 * k.xi2 * q * k.y32;
        g[1] = -c.a2 * xi * q / k.r3;
        g[2] = c.a1 * xy + c.a2 * xi * k.q2 * k.y32;
        g[3] = c.a1 * xy * sd + c.a2 * xi * fy + k.db / 2.0 * k.x11;
        g[4] = c.a2 * ey;
        g[5] = c.a1 * (cd / r + qy * sd) - c.a2 * q * fy;
        g[6] = c.a1 * xy * cd + c.a2 * xi * fz + k.yb / 2.0 * k.x11;
        g[7] = c.a2 * ez;
        g[8] = -c.a1 * (sd / r - qy * cd) - c.a2 * q * fz;
    } else {
        const double gy = 2.0 * k.x11 * sd - k.yb * q * k.x32;
        const double gz = 2.0 * k.x11 * cd + k.db * q * k.x32;
        g[0] = -c.a2 * xi * q / k.r3;
        g[1] = -qy / 2.0 - c.a2 * et * q / k.r3;
        g[2] = c.a1 / r + c.a2 * k.q2 / k.r3;
        g[3] = c.a2 * ey;
        g[4] = c.a1 * k.db * k.x11 + xy / 2.0 * sd + c.a2 * et * gy;
        g[5] = c.a1 * k.yb * k.x11 - c.a2 * q * gy;
        g[6] = c.a2 * ez;
        g[7] = c.a1 * k.yb * k.x11 + xy / 2.0 * cd + c.a2 * et * gz;
        g[8] = -c.a1 * k.db * k.x11 - c.a2 * q * gz;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] *= s;
}

// f^B gradients (Table 6 with the I/J/K integrals of Table 7, including
// the cos(dip) = 0 limit).
template <int SLIP>
__device__ __forceinline__ void ok_part_b(const OkCorner &k, const OkConst &c, double g[9])
{
    const double xi = k.xi, et = k.et, q = k.q, r = k.r, sd = c.sd, cd = c.cd;
    const double rd = r + k.db, d11 = 1.0 / (r * rd);
    const double aj2 = xi * k.yb / rd * d11;
    const double aj5 = -(k.db + k.yb * k.yb / rd) * d11;
    double ak1, ak3, aj3, aj6;
    if (cd != 0.0) {
        ak1 = xi * (d11 - k.y11 * sd) / cd;
        ak3 = (q * k.y11 - k.yb * d11) / cd;
        aj3 = (ak1 - aj2 * sd) / cd;
        aj6 = (ak3 - aj5 * sd) / cd;
    } else {
        const double rd2 = rd * rd;
        ak1 = xi * q / rd * d11;
        ak3 = sd / rd * (k.xi2 * d11 - 1.0);
        aj3 = -xi / rd2 * (k.q2 * d11 - 0.5);
        aj6 = -k.yb / rd2 * (k.xi2 * d11 - 0.5);
    }
    const double xy = xi * k.y11, qy = q * k.y11;
    const double ak2 = 1.0 / r + ak3 * sd;
    const double ak4 = xy * cd - ak1 * sd;
    const double aj1 = aj5 * cd - aj6 * sd;
    const double aj4 = -xy - aj2 * cd + aj3 * sd;
    const double ey = sd / r - k.yb * q / k.r3, ez = cd / r + k.db * q / k.r3;
    const double a3 = c.a3;
    const double s = 1.0 / 6.283185307179586476925286766559;
    if (SLIP == 0) {
        const double fy = k.db / k.r3 + k.xi2 * k.y32 * sd;
        const double fz = k.yb / k.r3 + k.xi2 * k.y32 * cd;
        g[0] = k.xi2 * q * k.y32 - a3 * aj1 * sd;
        g[1] = xi * q / k.r3 - a3 * aj2 * sd;
        g[2] = -xi * k.q2 * k.y32 - a3 * aj3 * sd;
        g[3] = -xi * fy - k.db * k.x11 + a3 * (xy + aj4) * sd;
        g[4] = -ey + a3 * (1.0 / r + aj5) * sd;
        g[5] = q * fy - a3 * (qy - aj6) * sd;
        g[6] = -xi * fz - k.yb * k.x11 + a3 * ak1 * sd;
        g[7] = -ez + a3 * k.yb * d11 * sd;
        g[8] = q * fz + a3 * ak2 * sd;
    } else {
        const double sdcd = sd * cd;
        const double gy = 2.0 * k.x11 * sd - k.yb * q * k.x32;
        const double gz = 2.0 * k.x11 * cd + k.db * q * k.x32;
        g[0] = xi * q / k.r3 + a3 * aj4 * sdcd;
        g[1] = et * q / k.r3 + qy + a3 * aj5 * sdcd;
        g[2] = -k.q2 / k.r3 + a3 * aj6 * sdcd;
        g[3] = -ey + a3 * aj1 * sdcd;
        g[4] = -et * gy - xy * sd + a3 * aj2 * sdcd;
        g[5] = q * gy + a3 * aj3 * sdcd;
        g[6] = -ez - a3 * ak3 * sdcd;
        g[7] = -et * gz - xy * cd - a3 * xi * d11 * sdcd;
        g[8] = q * gz - a3 * ak4 * sdcd;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] *= s;
}

// f^C: displacement (3) and gradients (9); z is the receiver's z.
template <int SLIP>
__device__ __forceinline__ void ok_part_c(const OkCorner &k, const OkConst &c, double z,
                                          double uc[3], double g[9])
{
    const double xi = k.xi, et = k.et, q = k.q, r = k.r, sd = c.sd, cd = c.cd;
    const double a4 = c.a4, a5 = c.a5, yv = k.yb, dv = k.db;
    const double cc = dv + z;
    const double x53 = (8.0 * k.r2 + 9.0 * r * xi + 3.0 * k.xi2) * k.x11 * k.x11 * k.x11 / k.r2;
    const double y53 = (8.0 * k.r2 + 9.0 * r * et + 3.0 * k.et2) * k.y11 * k.y11 * k.y11 / k.r2;
    const double h = q * cd - z;
    const double z32 = sd / k.r3 - h * k.y32;
    const double z53 = 3.0 * sd / k.r5 - h * y53;
    const double y0 = k.y11 - k.xi2 * k.y32;
    const double z0 = z32 - k.xi2 * z53;
    const double ppy = cd / k.r3 + q * k.y32 * sd;
    const double ppz = sd / k.r3 - q * k.y32 * cd;
    const double qq = z * k.y32 + z32 + z0;
    const double qqy = 3.0 * cc * dv / k.r5 - qq * sd;
    const double qqz = 3.0 * cc * yv / k.r5 - qq * cd + q * k.y32;
    const double xy = xi * k.y11, qy = q * k.y11;
    const double qr = 3.0 * q / k.r5;
    const double cdr = (cc + dv) / k.r3;
    const double yy0 = yv / k.r3 - y0 * cd;
    const double s = 1.0 / 6.283185307179586476925286766559;
    if (SLIP == 0) {
        uc[0] = a4 * xy * cd - a5 * xi * q * z32;
        uc[1] = a4 * (cd / r + 2.0 * qy * sd) - a5 * cc * q / k.r3;
        uc[2] = a4 * qy * cd - a5 * (cc * et / k.r3 - z * k.y11 + k.xi2 * z32);
        g[0] = a4 * y0 * cd - a5 * q * z0;
        g[1] = -a4 * xi * (cd / k.r3 + 2.0 * q * k.y32 * sd) + a5 * cc * xi * qr;
        g[2] = -a4 * xi * q * k.y32 * cd + a5 * xi * (3.0 * cc * et / k.r5 - qq);
        g[3] = -a4 * xi * ppy * cd - a5 * xi * qqy;
        g[4] = a4 * 2.0 * (dv / k.r3 - y0 * sd) * sd - yv / k.r3 * cd -
               a5 * (cdr * sd - et / k.r3 - cc * yv * qr);
        g[5] = -a4 * q / k.r3 + yy0 * sd + a5 * (cdr * cd + cc * dv * qr - (y0 * cd + q * z0) * sd);
        g[6] = a4 * xi * ppz * cd - a5 * xi * qqz;
        g[7] = a4 * 2.0 * (yv / k.r3 - y0 * cd) * sd + dv / k.r3 * cd - a5 * (cdr * cd + cc * dv * qr);
        g[8] = yy0 * cd - a5 * (cdr * sd - cc * yv * qr - y0 * sd * sd + q * z0 * cd);
    } else {
        uc[0] = a4 * cd / r - qy * sd - a5 * cc * q / k.r3;
        uc[1] = a4 * yv * k.x11 - a5 * cc * et * q * k.x32;
        uc[2] = -dv * k.x11 - xy * sd - a5 * cc * (k.x11 - k.q2 * k.x32);
        g[0] = -a4 * xi / k.r3 * cd + a5 * cc * xi * qr + xi * q * k.y32 * sd;
        g[1] = -a4 * yv / k.r3 + a5 * cc * et * qr;
        g[2] = dv / k.r3 - y0 * sd + a5 * cc / k.r3 * (1.0 - 3.0 * k.q2 / k.r2);
        g[3] = -a4 * et / k.r3 + y0 * sd * sd - a5 * (cdr * sd - cc * yv * qr);
        g[4] = a4 * (k.x11 - yv * yv * k.x32) - a5 * cc * ((dv + 2.0 * q * cd) * k.x32 - yv * et * q * x53);
        g[5] = xi * ppy * sd + yv * dv * k.x32 + a5 * cc * ((yv + 2.0 * q * sd) * k.x32 - yv * k.q2 * x53);
        g[6] = -q / k.r3 + y0 * sd * cd - a5 * (cdr * cd + cc * dv * qr);
        g[7] = a4 * yv * dv * k.x32 - a5 * cc * ((yv - 2.0 * q * sd) * k.x32 + dv * et * q * x53);
        g[8] = -xi * ppz * sd + k.x11 - dv * dv * k.x32 - a5 * cc * ((dv - 2.0 * q * cd) * k.x32 - dv * k.q2 * x53);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) uc[i] *= s;
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] *= s;
}

__device__ __forceinline__ bool ok_on_edge(double xi0, double xi1, double et0, double et1, double q)
{
    return q == 0.0 && ((xi0 * xi1 <= 0.0 && et0 * et1 == 0.0) ||
                        (et0 * et1 <= 0.0 && xi0 * xi1 == 0.0));
}

// Displacement gradients du[9] = (uxx uyx uzx uxy uyy uzy uxz uyz uzz) of
// a unit SLIP dislocation; returns false (du = 0) on a fault edge, as the
// paper's singular-case rule.
template <int SLIP>
__device__ bool okada_grad(const OkConst &c, double x, double y, double z, double depth,
                           double al1, double al2, double aw1, double aw2, double du[9])
{
#pragma unroll
    for (int i = 0; i < 9; ++i) du[i] = 0.0;
    if (z > 0.0) return false;
    double xi[2] = {x - al1, x - al2};
    if (fabs(xi[0]) < OQ_OK_EPS) xi[0] = 0.0;
    if (fabs(xi[1]) < OQ_OK_EPS) xi[1] = 0.0;
#pragma unroll
    for (int img = 0; img < 2; ++img) {
        const double d = img ? depth - z : depth + z;
        const double p = y * c.cd + d * c.sd;
        double q = y * c.sd - d * c.cd;
        double et[2] = {p - aw1, p - aw2};
        if (fabs(q) < OQ_OK_EPS) q = 0.0;
        if (fabs(et[0]) < OQ_OK_EPS) et[0] = 0.0;
        if (fabs(et[1]) < OQ_OK_EPS) et[1] = 0.0;
        if (ok_on_edge(xi[0], xi[1], et[0], et[1], q)) {
#pragma unroll
            for (int i = 0; i < 9; ++i) du[i] = 0.0;
            return false;
        }
        const double r12 = sqrt(xi[0] * xi[0] + et[1] * et[1] + q * q);
        const double r21 = sqrt(xi[1] * xi[1] + et[0] * et[0] + q * q);
        const double r22 = sqrt(xi[1] * xi[1] + et[1] * et[1] + q * q);
        const bool kxi[2] = {xi[0] < 0.0 && r21 + xi[1] < OQ_OK_EPS,
                             xi[0] < 0.0 && r22 + xi[1] < OQ_OK_EPS};
        const bool ket[2] = {et[0] < 0.0 && r12 + et[1] < OQ_OK_EPS,
                             et[0] < 0.0 && r22 + et[1] < OQ_OK_EPS};
        for (int kk = 0; kk < 2; ++kk) {
            for (int j = 0; j < 2; ++j) {
                OkCorner k;
                if (!ok_corner(k, c, xi[j], et[kk], q, kxi[kk], ket[j])) continue;
                const double sg = (j + kk != 1) ? 1.0 : -1.0;
                double ga[9], g[9];
                ok_part_a<SLIP>(k, c, ga);
                if (!img) {
                    // real source: -uA(x,y,z), fault frame -> (x,y,z); the
                    // z-derivative row changes sign with z.
#pragma unroll
                    for (int i = 0; i < 9; i += 3) {
                        const double s3 = (i == 6) ? -1.0 : 1.0;
                        g[i] = -ga[i] * s3;
                        g[i + 1] = (-ga[i + 1] * c.cd + ga[i + 2] * c.sd) * s3;
                        g[i + 2] = (-ga[i + 1] * c.sd - ga[i + 2] * c.cd) * s3;
                    }
                } else {
                    double gb[9], gc[9], uc[3];
                    ok_part_b<SLIP>(k, c, gb);
                    ok_part_c<SLIP>(k, c, z, uc, gc);
#pragma unroll
                    for (int i = 0; i < 9; i += 3) {
                        g[i] = ga[i] + gb[i] + z * gc[i];
                        g[i + 1] = (ga[i + 1] + gb[i + 1] + z * gc[i + 1]) * c.cd -
                                   (ga[i + 2] + gb[i + 2] + z * gc[i + 2]) * c.sd;
                        g[i + 2] = (ga[i + 1] + gb[i + 1] - z * gc[i + 1]) * c.sd +
                                   (ga[i + 2] + gb[i + 2] - z * gc[i + 2]) * c.cd;
                    }
                    g[6] += uc[0];
                    g[7] += uc[1] * c.cd - uc[2] * c.sd;
                    g[8] += -uc[1] * c.sd - uc[2] * c.cd;
                }
#pragma unroll
                for (int i = 0; i < 9; ++i) du[i] += sg * g[i];
            }
        }
    }
    return true;
}

}  // namespace oq
