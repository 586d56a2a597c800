// hex8_dev.h -- device stress of a uniform eigenstrain in a cuboid of an
// elastic half-space (z <= 0, traction-free surface), FP64.
//
// Replaces GeoGreensFunctions.stress_vol_hex8! as called by the reference
// (src/BEM/GF.jl:215-221, 277-283): cuboid [qx-T/2, qx+T/2] x [qy, qy+L] x
// [qz-W, qz] ("T-L-W <=> x-y-z"), rotation 0, tensor strain components
// [xx xy xz yy yz zz], stress in the same order.
//
//   grad u = m : int_V d2G/dx dy,  m = C : e*,  G = G_Kelvin + G_image
//   - Kelvin part in closed form (Eshelby 1957 potentials phi = int 1/r,
//     psi = int r; their 2nd and 4th derivatives over a cuboid are corner
//     sums of log(R+X), atan and rational terms);
//   - image part of Mindlin's (1936) half-space solution (regular inside
//     the medium) by Gauss-Legendre quadrature on sub-boxes bisected along
//     their longest axis until each is >= 5 half-diagonals from the
//     receiver's image point.
//
// The image kernel D[i][l][p] = d2 G_image_ij / dx_l dy_k summed over the
// moment pair p = (j, k) (without 1/(16 pi mu (1-nu))) enters only through
// the combinations a stress (36) or a fault's shear traction (12) takes,
// generated and simplified by oracle/derive/mindlin_image.py.
//
// GPU shape: everything that does not depend on the eigenstrain (the
// Eshelby tensors and the quadrature sums of the image combinations) is
// accumulated once per receiver-cell pair; the 6 unit eigenstrains the
// builders need are then contractions.  The reference calls the closed form
// once per unit eigenstrain (6x the geometric work).
#pragma once
#include <hip/hip_runtime.h>

namespace oq {

#define OQ_IPOW(a, n) oq_ipow_dev((a), (n))
__device__ __forceinline__ double oq_ipow_dev(double a, int n)
{
    double r = 1.0;
    const int m = n < 0 ? -n : n;
    for (int i = 0; i < m; ++i) r *= a;
    return n < 0 ? 1.0 / r : r;
}

// The pair sums of the image kernel that the shear traction on a fault of
// type FT takes (P[p], P[6 + p] for eigenstrain pair p; see
// h8_image_traction), simplified as a whole before their common
// subexpressions are taken: about half the operations of the 24 kernel
// entries they are sums of.
template <int FT>
__device__ __forceinline__ void hex8_image_pairs(double X, double Y, double zr, double ys, double nu,
                                                 double *P)
{
    const double R = sqrt(X * X + Y * Y + (zr + ys) * (zr + ys));
    if constexpr (FT == 0) {
#include "hex8_image_trac0.inc"
    } else {
#include "hex8_image_trac1.inc"
    }
}

__device__ __forceinline__ double h8_logrx(double R, double xc, double rho2)
{
    if (xc >= 0.0) return log(R + xc);
    if (rho2 > 0.0) return log(rho2 / (R - xc));
    return -log(R - xc);
}

__device__ __forceinline__ double h8_atanq(double a, double b, double c, double R)
{
    const double num = a * b;
    if (c == 0.0) return num == 0.0 ? 0.0 : copysign(M_PI / 2.0, num);
    return atan(num / (c * R));
}

__device__ __forceinline__ double h8_ratio(double num, double den)
{
    return den == 0.0 ? 0.0 : num / den;
}

// Eshelby tensors summed over the 8 corners.
//   phi[ab] (a<=b, order 00 01 02 11 12 22)
//   psi by index counts: P4[a] = psi_aaaa, P31[a][b] = psi_aaab (a!=b),
//   P22[c] = psi_aabb with {a,b} = the pair excluding c, P211[a] = psi_aabc.
struct Hex8Kelvin {
    double phi[6];
    double P4[3], P31[3][3], P22[3], P211[3];
};

__device__ __forceinline__ void h8_kelvin_add(Hex8Kelvin &K, const double X[3], double sg)
{
    const double R = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    if (R == 0.0) return;
    double A[3], L[3], rho2[3], B[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int a = (c + 1) % 3, b = (c + 2) % 3;
        rho2[c] = X[a] * X[a] + X[b] * X[b];
        L[c] = h8_logrx(R, X[c], rho2[c]);
        A[c] = h8_atanq(X[a], X[b], X[c], R);
        B[c] = -X[a] * X[b] * h8_ratio(X[c], rho2[c] * R);  // psi_aabb
    }
    K.phi[0] += -sg * A[0];
    K.phi[3] += -sg * A[1];
    K.phi[5] += -sg * A[2];
    K.phi[1] += sg * L[2];
    K.phi[2] += sg * L[1];
    K.phi[4] += sg * L[0];
#pragma unroll
    for (int c = 0; c < 3; ++c) K.P22[c] += sg * B[c];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const int q = (p + 1) % 3, r = (p + 2) % 3;
        // psi_pppp = 2 phi_pp - psi_ppqq - psi_pprr
        K.P4[p] += sg * (-2.0 * A[p] - B[r] - B[q]);
        K.P211[p] += sg * (X[p] / R);
    }
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (p == q) continue;
            const int r = 3 - p - q;
            K.P31[p][q] += sg * (X[q] * X[q] * h8_ratio(X[r], rho2[r] * R) + L[r] - X[r] / R);
        }
}

__device__ __forceinline__ double h8_psi(const Hex8Kelvin &K, int i, int j, int k, int l)
{
    int cnt[3] = {0, 0, 0};
    cnt[i]++; cnt[j]++; cnt[k]++; cnt[l]++;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (cnt[a] == 4) return K.P4[a];
        if (cnt[a] == 3) {
            const int b = cnt[(a + 1) % 3] == 1 ? (a + 1) % 3 : (a + 2) % 3;
            return K.P31[a][b];
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (cnt[a] == 0) return K.P22[a];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (cnt[a] == 2) return K.P211[a];
    return 0.0;
}

__device__ __forceinline__ double h8_phi(const Hex8Kelvin &K, int a, int b)
{
    if (a > b) { const int t = a; a = b; b = t; }
    const int idx = a == 0 ? b : (a == 1 ? 2 + b : 5);
    return K.phi[idx];
}

__constant__ double kGLX[7][6] = {
    {0}, {0.0}, {-0.5773502691896257645, 0.5773502691896257645},
    {-0.7745966692414833770, 0.0, 0.7745966692414833770},
    {-0.8611363115940525752, -0.3399810435848562648, 0.3399810435848562648, 0.8611363115940525752},
    {-0.9061798459386639928, -0.5384693101056830910, 0.0, 0.5384693101056830910, 0.9061798459386639928},
    {-0.9324695142031520278, -0.6612093864662645137, -0.2386191860831969086, 0.2386191860831969086,
     0.6612093864662645137, 0.9324695142031520278}};
__constant__ double kGLW[7][6] = {
    {0}, {2.0}, {1.0, 1.0},
    {0.5555555555555555556, 0.8888888888888888889, 0.5555555555555555556},
    {0.3478548451374538574, 0.6521451548625461426, 0.6521451548625461426, 0.3478548451374538574},
    {0.2369268850561890875, 0.4786286704993664680, 0.5688888888888888889, 0.4786286704993664680,
     0.2369268850561890875},
    {0.1713244923791703450, 0.3607615730481386076, 0.4679139345726910474, 0.4679139345726910474,
     0.3607615730481386076, 0.1713244923791703450}};

#define OQ_H8_NQ 6
#define OQ_H8_KAPPA 5.0
#define OQ_H8_MAXDEPTH 48

// Sub-box of the cuboid reached by `path` (depth bits, most significant
// first; 0 = lower half along the current longest axis).
__device__ __forceinline__ void h8_subbox(const double lo0[3], const double hi0[3],
                                          unsigned long long path, int depth,
                                          double lo[3], double hi[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) { lo[a] = lo0[a]; hi[a] = hi0[a]; }
    for (int d = depth - 1; d >= 0; --d) {
        int amax = 0;
        double hm = hi[0] - lo[0];
        if (hi[1] - lo[1] > hm) { amax = 1; hm = hi[1] - lo[1]; }
        if (hi[2] - lo[2] > hm) { amax = 2; }
        const double mid = 0.5 * (hi[amax] + lo[amax]);
        if ((path >> d) & 1ull) lo[amax] = mid;
        else hi[amax] = mid;
    }
}

// The combinations of the image kernel that a stress takes: P[6 q + c] for
// moment pair q, c = D[0][0], D[1][1], D[2][2], D[0][1] + D[1][0],
// D[0][2] + D[2][0], D[1][2] + D[2][1] (the image part of grad u enters the
// stress only through its diagonal and symmetric sums), simplified as a
// whole before their common subexpressions are taken: about 60 % of the
// operations of the 54 kernel entries.
__device__ __forceinline__ void hex8_image_stress_combos(double X, double Y, double zr, double ys, double nu,
                                                         double *P)
{
    const double R = sqrt(X * X + Y * Y + (zr + ys) * (zr + ys));
#include "hex8_image_stress.inc"
}

// Accumulates the 36 image stress combinations integrated over [lo0, hi0]
// into acc (depth-first over the bisection tree, lower child first).
__device__ void h8_image_integral(double x, double y, double z, const double lo0[3],
                                  const double hi0[3], double nu, double *acc)
{
    const double xi[3] = {x, y, -z};
    unsigned long long path = 0;
    int depth = 0;
    double Dk[36];
    for (;;) {
        double lo[3], hi[3], h[3], ctr[3], dist2 = 0.0, hd2 = 0.0;
        h8_subbox(lo0, hi0, path, depth, lo, hi);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            h[a] = 0.5 * (hi[a] - lo[a]);
            ctr[a] = 0.5 * (hi[a] + lo[a]);
            const double dd = fabs(xi[a] - ctr[a]) - h[a];
            if (dd > 0.0) dist2 += dd * dd;
            hd2 += h[a] * h[a];
        }
        if (depth < OQ_H8_MAXDEPTH && dist2 < OQ_H8_KAPPA * OQ_H8_KAPPA * hd2) {
            ++depth;
            path <<= 1;
            continue;
        }
        const double vol = h[0] * h[1] * h[2];
        for (int i = 0; i < OQ_H8_NQ; ++i)
            for (int j = 0; j < OQ_H8_NQ; ++j)
                for (int k = 0; k < OQ_H8_NQ; ++k) {
                    const double ys1 = ctr[0] + h[0] * kGLX[OQ_H8_NQ][i];
                    const double ys2 = ctr[1] + h[1] * kGLX[OQ_H8_NQ][j];
                    const double ys3 = ctr[2] + h[2] * kGLX[OQ_H8_NQ][k];
                    const double w = kGLW[OQ_H8_NQ][i] * kGLW[OQ_H8_NQ][j] * kGLW[OQ_H8_NQ][k] * vol;
                    hex8_image_stress_combos(x - ys1, y - ys2, z, ys3, nu, Dk);
#pragma unroll
                    for (int m = 0; m < 36; ++m) acc[m] += w * Dk[m];
                }
        // next leaf in depth-first order
        while (depth > 0 && (path & 1ull)) {
            path >>= 1;
            --depth;
        }
        if (depth == 0) break;
        path += 1;
    }
}

// Eshelby corner sums of the cuboid [lo, hi] at the receiver (x, y, z);
// returns whether the receiver lies inside the cuboid.
__device__ __forceinline__ bool hex8_kelvin(double x, double y, double z, const double lo[3], const double hi[3],
                            Hex8Kelvin &K)
{
    const double rv[3] = {x, y, z};
#pragma unroll
    for (int i = 0; i < 6; ++i) K.phi[i] = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        K.P4[a] = K.P22[a] = K.P211[a] = 0.0;
#pragma unroll
        for (int b = 0; b < 3; ++b) K.P31[a][b] = 0.0;
    }
    for (int s = 0; s < 8; ++s) {
        double X[3], sg = 1.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int bit = (s >> a) & 1;
            X[a] = rv[a] - (bit ? hi[a] : lo[a]);
            if (bit) sg = -sg;
        }
        h8_kelvin_add(K, X, sg);
    }
    return x >= lo[0] && x < hi[0] && y >= lo[1] && y < hi[1] && z >= lo[2] && z < hi[2];
}

// The image integral contracted with the shear-traction functional of a
// fault receiver, for the 6 unit eigenstrains at once.  The traction is
// linear in grad u, tau = sum_il c_il (grad u)_il, and only four c_il are
// nonzero (FT 0, strike slip: tau = -s_xy sin d + s_xz cos d gives c_01 =
// c_10 = -mu sin d, c_02 = c_20 = mu cos d; FT 1, dip slip: tau = (s_zz -
// s_yy)/2 sin 2d + s_yz cos 2d gives c_22 = -c_11 = mu sin 2d, c_12 = c_21 =
// mu cos 2d).  So the quadrature accumulates E[q] = sum_il c_il D[i][l][q]
// (6 sums, from 24 of the 54 kernel terms; the other 30 and their common
// subexpressions are dead code) instead of all 54 integrals: ca, cb are the
// two coefficients above.  Same bisection and Gauss points as
// h8_image_integral.
template <int FT>
__device__ void h8_image_traction(double x, double y, double z, const double lo0[3], const double hi0[3],
                                  double nu, double ca, double cb, double E[6])
{
    const double xi[3] = {x, y, -z};
    unsigned long long path = 0;
    int depth = 0;
    double Pk[12];
    for (;;) {
        double lo[3], hi[3], h[3], ctr[3], dist2 = 0.0, hd2 = 0.0;
        h8_subbox(lo0, hi0, path, depth, lo, hi);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            h[a] = 0.5 * (hi[a] - lo[a]);
            ctr[a] = 0.5 * (hi[a] + lo[a]);
            const double dd = fabs(xi[a] - ctr[a]) - h[a];
            if (dd > 0.0) dist2 += dd * dd;
            hd2 += h[a] * h[a];
        }
        if (depth < OQ_H8_MAXDEPTH && dist2 < OQ_H8_KAPPA * OQ_H8_KAPPA * hd2) {
            ++depth;
            path <<= 1;
            continue;
        }
        const double vol = h[0] * h[1] * h[2];
        for (int i = 0; i < OQ_H8_NQ; ++i)
            for (int j = 0; j < OQ_H8_NQ; ++j)
                // One Gauss point per iteration: unrolled, the six points'
                // temporaries spill (512 registers + 264 B/lane of scratch
                // for strike slip) and one wavefront fits a SIMD; rolled, the
                // body fits 244 VGPRs and two wavefronts hide each other's
                // FP64 latency (k_gf_fm 61 -> 47 ms at 4096 x 2048 pairs).
#pragma unroll 1
                for (int k = 0; k < OQ_H8_NQ; ++k) {
                    const double ys1 = ctr[0] + h[0] * kGLX[OQ_H8_NQ][i];
                    const double ys2 = ctr[1] + h[1] * kGLX[OQ_H8_NQ][j];
                    const double ys3 = ctr[2] + h[2] * kGLX[OQ_H8_NQ][k];
                    const double w = kGLW[OQ_H8_NQ][i] * kGLW[OQ_H8_NQ][j] * kGLW[OQ_H8_NQ][k] * vol;
                    hex8_image_pairs<FT>(x - ys1, y - ys2, z, ys3, nu, Pk);
#pragma unroll
                    for (int q = 0; q < 6; ++q) E[q] += w * (ca * Pk[q] + cb * Pk[6 + q]);
                }
        while (depth > 0 && (path & 1ull)) {
            path >>= 1;
            --depth;
        }
        if (depth == 0) break;
        path += 1;
    }
}

// Per receiver-cell pair: everything independent of the eigenstrain.
struct Hex8Pair {
    Hex8Kelvin K;
    double D[36];  // image stress combinations, integrated (h8_image_integral)
    bool inside;
};

__device__ void hex8_pair(double x, double y, double z, double qx, double qy, double qz,
                          double T, double L, double W, double nu, Hex8Pair &P)
{
    const double lo[3] = {qx - T / 2.0, qy, qz - W};
    const double hi[3] = {qx + T / 2.0, qy + L, qz};
    // The image integral first: its quadrature loop then runs without the
    // 24 Kelvin sums live, which takes most of its register spills away.
#pragma unroll
    for (int m = 0; m < 36; ++m) P.D[m] = 0.0;
    h8_image_integral(x, y, z, lo, hi, nu, P.D);
    P.inside = hex8_kelvin(x, y, z, lo, hi, P.K);
}

// mv = the eigenstress C : e* of the unit eigenstrain p, as [xx xy xz yy yz zz]
__device__ __forceinline__ void hex8_moment(int p, double mu, double nu, double mv[6])
{
    const double lambda = 2.0 * mu * nu / (1.0 - 2.0 * nu);
#pragma unroll
    for (int q = 0; q < 6; ++q) mv[q] = 0.0;
    if (p == 0 || p == 3 || p == 5) {
        mv[0] = mv[3] = mv[5] = lambda;
        mv[p] += 2.0 * mu;
    } else {
        mv[p] = 2.0 * mu;
    }
}

// Stress s[6] for the unit eigenstrain component p from the Eshelby sums K
// and (IMG) the 54 image integrals D; without them, the full-space part
// plus the eigenstress inside the cuboid.
template <bool IMG>
__device__ __forceinline__ void hex8_stress_unit_body(const Hex8Kelvin &K, const double *D, bool inside, int p,
                                                      double mu, double nu, double s[6])
{
    const int pj[6] = {0, 0, 0, 1, 1, 2}, pk[6] = {0, 1, 2, 1, 2, 2};
    const double lambda = 2.0 * mu * nu / (1.0 - 2.0 * nu);
    const int j0 = pj[p], k0 = pk[p];
    double e[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    e[j0][k0] = 1.0;
    e[k0][j0] = 1.0;
    const double tr = e[0][0] + e[1][1] + e[2][2];
    double m[3][3], mv[6], gu[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) m[i][j] = 2.0 * mu * e[i][j] + (i == j ? lambda * tr : 0.0);
    mv[0] = m[0][0]; mv[1] = m[0][1]; mv[2] = m[0][2];
    mv[3] = m[1][1]; mv[4] = m[1][2]; mv[5] = m[2][2];
    const double ck = 1.0 / (8.0 * M_PI * (1.0 - nu));
    const double ci = 1.0 / (16.0 * M_PI * mu * (1.0 - nu));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    if (e[j][k] != 0.0) acc += e[j][k] * h8_psi(K, i, j, k, l);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (e[i][k] != 0.0) acc -= 4.0 * (1.0 - nu) * e[i][k] * h8_phi(K, k, l);
            acc -= 2.0 * nu * tr * h8_phi(K, i, l);
            gu[i][l] = acc * ck;
        }
    // image part: the diagonal and the symmetric sums of its grad u
    double gi[6] = {0, 0, 0, 0, 0, 0};
    if (IMG) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double g = 0.0;
#pragma unroll
            for (int q = 0; q < 6; ++q) g += D[6 * q + c] * mv[q];
            gi[c] = g * ci;
        }
    }
    const double div = gu[0][0] + gu[1][1] + gu[2][2] + (gi[0] + gi[1] + gi[2]);
    const double in = inside ? 1.0 : 0.0;
    s[0] = 2.0 * mu * (gu[0][0] + gi[0]) + lambda * div - in * m[0][0];
    s[1] = mu * (gu[0][1] + gu[1][0] + gi[3]) - in * m[0][1];
    s[2] = mu * (gu[0][2] + gu[2][0] + gi[4]) - in * m[0][2];
    s[3] = 2.0 * mu * (gu[1][1] + gi[1]) + lambda * div - in * m[1][1];
    s[4] = mu * (gu[1][2] + gu[2][1] + gi[5]) - in * m[1][2];
    s[5] = 2.0 * mu * (gu[2][2] + gi[2]) + lambda * div - in * m[2][2];
}

// Stress s[6] for the unit eigenstrain component p (0..5 = xx xy xz yy yz zz).
__device__ void hex8_stress_unit(const Hex8Pair &P, int p, double mu, double nu, double s[6])
{
    hex8_stress_unit_body<true>(P.K, P.D, P.inside, p, mu, nu, s);
}

// The same without the image part (the caller adds it).
__device__ void hex8_stress_unit_kelvin(const Hex8Kelvin &K, bool inside, int p, double mu, double nu, double s[6])
{
    hex8_stress_unit_body<false>(K, nullptr, inside, p, mu, nu, s);
}

}  // namespace oq
