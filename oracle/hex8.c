/*
 * oracle/hex8.c -- TEST INFRASTRUCTURE (the CPU oracle).
 *
 * Stress in an elastic half-space (z <= 0, traction-free at z = 0) due to a
 * uniform eigenstrain in a cuboid.  The reference computes this with
 * GeoGreensFunctions.jl's `stress_vol_hex8!` (un-vendored, after Barbot,
 * Moore & Lambert 2017, BSSA 107(2)), called at src/BEM/GF.jl:215-221 and
 * GF.jl:277-283 with (T, L, W) <=> (x, y, z) extents, reference point
 * q = (cx, cy - dy/2, cz + dz/2) (mesh.jl:181-183), rotation 0.
 *
 * Formulation (restated from the mathematics, not from that package):
 *   grad u = m : int_V d2 G(x, y) / dx dy dV,   m = C : eps*,
 * with Mindlin's half-space Green's function G split as G = G_K + G_I:
 *   - G_K (Kelvin, full space) is integrated in closed form through
 *     Eshelby's potentials phi = int_V 1/r, psi = int_V r (Eshelby 1957):
 *       8 pi (1-nu) u_i = psi_,jki e*_jk - 4(1-nu) phi_,k e*_ik
 *                         - 2 nu phi_,i e*_kk,
 *     whose 2nd/4th derivatives over a cuboid are Chinnery corner sums of
 *     elementary functions (log(R+X), atan, rational);
 *   - G_I (image + surface correction, regular inside the medium) is
 *     integrated by tensor Gauss-Legendre quadrature (nquad points per axis)
 *     on sub-boxes, bisected along their longest axis until every sub-box
 *     is at least 5 half-diagonals away from the receiver's image point.
 *   sigma = lambda tr(grad u) I + mu (grad u + grad u^T) - chi_V C : eps*.
 * A receiver on a face is treated as the limit from the +x/+y/+z side
 * (the cuboid is the half-open box [x0,x1) x [y0,y1) x [z0,z1)); traction
 * components across a face are continuous, so this choice only affects the
 * discontinuous ones.
 *
 * Parity status: the reference's hex8 values are unpinned (no tests, package
 * absent).  This function is pinned by tests/test_oracle_hex8.py: the
 * laterally infinite slab limit, traction-free free surface, equilibrium,
 * additivity when splitting a cuboid, traction continuity across faces, and
 * quadrature convergence.
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

static inline double oq_ipow(double a, int n)
{
    double r = 1.0;
    int m = n < 0 ? -n : n;
    while (m-- > 0) r *= a;
    return n < 0 ? 1.0 / r : r;
}
#define OQ_IPOW(a, n) oq_ipow((a), (n))

/* Regular (image) part: D[i*18 + l*6 + p] = sum over pair p of
 * d2 G_I_ij / dx_l dy_k, without the 1/(16 pi mu (1-nu)) factor. */
static void image_kernel(double X, double Y, double zr, double ys, double nu,
                         double D[54])
{
    const double R = sqrt(X * X + Y * Y + (zr + ys) * (zr + ys));
#include "hex8_image.inc"
}

/* log(R + X_c) in the cancellation-free form; for an edge-line point
 * (rho2 = 0, X_c < 0) the form -log(R - X_c) is used, which differs by the
 * constant log(rho2) that cancels between the two corners on that line. */
static double logrx(double R, double xc, double rho2)
{
    if (xc >= 0.0) return log(R + xc);
    if (rho2 > 0.0) return log(rho2 / (R - xc));
    return -log(R - xc);
}

/* atan(a b / (c R)) with the c -> +0 limit at c = 0. */
static double atanq(double a, double b, double c, double R)
{
    double num = a * b;
    if (c == 0.0) return num == 0.0 ? 0.0 : copysign(M_PI / 2.0, num);
    return atan(num / (c * R));
}

static double ratio(double num, double den)
{
    return den == 0.0 ? 0.0 : num / den;
}

/* Corner functions of phi_,kl (6, order 11 12 13 22 23 33) and of the
 * needed psi_,ijkl (indexed by the sorted multiset) at offset X. */
typedef struct {
    double phi[3][3];
    double psi[3][3][3][3];
} kcorner_t;

static void kelvin_corner(const double X[3], kcorner_t *o)
{
    const double R = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    double A[3], L[3], rho2[3];
    int a, b, c, i, j, k, l;
    if (R == 0.0) {
        memset(o, 0, sizeof(*o));
        return;
    }
    /* rho2[c] = X_a^2 + X_b^2 for the pair {a,b} excluding c */
    for (c = 0; c < 3; ++c) {
        a = (c + 1) % 3;
        b = (c + 2) % 3;
        rho2[c] = X[a] * X[a] + X[b] * X[b];
        L[c] = logrx(R, X[c], rho2[c]);
        A[c] = atanq(X[a], X[b], X[c], R);
    }
    for (a = 0; a < 3; ++a) {
        for (b = 0; b < 3; ++b) {
            if (a == b) o->phi[a][a] = -A[a];
            else o->phi[a][b] = L[3 - a - b];
        }
    }
    /* psi_,ijkl depends only on the multiset {i,j,k,l} */
    for (i = 0; i < 3; ++i)
        for (j = 0; j < 3; ++j)
            for (k = 0; k < 3; ++k)
                for (l = 0; l < 3; ++l) {
                    int cnt[3] = {0, 0, 0};
                    double v;
                    cnt[i]++; cnt[j]++; cnt[k]++; cnt[l]++;
                    if (cnt[0] == 4 || cnt[1] == 4 || cnt[2] == 4) {
                        int p = cnt[0] == 4 ? 0 : (cnt[1] == 4 ? 1 : 2);
                        int q = (p + 1) % 3, r = (p + 2) % 3;
                        /* psi_pppp = 2 phi_pp - psi_ppqq - psi_pprr */
                        double bpq = -X[p] * X[q] * ratio(X[r], rho2[r] * R);
                        double bpr = -X[p] * X[r] * ratio(X[q], rho2[q] * R);
                        v = -2.0 * A[p] - bpq - bpr;
                    } else if (cnt[0] == 3 || cnt[1] == 3 || cnt[2] == 3) {
                        int p = cnt[0] == 3 ? 0 : (cnt[1] == 3 ? 1 : 2);
                        int q = cnt[0] == 1 ? 0 : (cnt[1] == 1 ? 1 : 2);
                        int r = 3 - p - q;
                        /* psi_pppq = X_q^2 X_r / (rho_pq^2 R) + L_r - X_r/R */
                        v = X[q] * X[q] * ratio(X[r], rho2[r] * R) + L[r] - X[r] / R;
                    } else if (cnt[0] == 2 && cnt[1] == 2) {
                        v = -X[0] * X[1] * ratio(X[2], rho2[2] * R);
                    } else if (cnt[0] == 2 && cnt[2] == 2) {
                        v = -X[0] * X[2] * ratio(X[1], rho2[1] * R);
                    } else if (cnt[1] == 2 && cnt[2] == 2) {
                        v = -X[1] * X[2] * ratio(X[0], rho2[0] * R);
                    } else {
                        int p = cnt[0] == 2 ? 0 : (cnt[1] == 2 ? 1 : 2);
                        v = X[p] / R; /* psi_ppqr */
                    }
                    o->psi[i][j][k][l] = v;
                }
}

/* 3x3 tensor e from the 6-vector [xx xy xz yy yz zz]. */
static void unpack6(const double e6[6], double e[3][3])
{
    e[0][0] = e6[0]; e[0][1] = e[1][0] = e6[1]; e[0][2] = e[2][0] = e6[2];
    e[1][1] = e6[3]; e[1][2] = e[2][1] = e6[4]; e[2][2] = e6[5];
}

static const double GL_X[11][10] = {
    {0},
    {0.0},
    {-0.5773502691896257645, 0.5773502691896257645},
    {-0.7745966692414833770, 0.0, 0.7745966692414833770},
    {-0.8611363115940525752, -0.3399810435848562648, 0.3399810435848562648, 0.8611363115940525752},
    {-0.9061798459386639928, -0.5384693101056830910, 0.0, 0.5384693101056830910, 0.9061798459386639928},
    {-0.9324695142031520278, -0.6612093864662645137, -0.2386191860831969086, 0.2386191860831969086,
     0.6612093864662645137, 0.9324695142031520278},
    {-0.9491079123427585245, -0.7415311855993944399, -0.4058451513773971669, 0.0, 0.4058451513773971669,
     0.7415311855993944399, 0.9491079123427585245},
    {-0.9602898564975362317, -0.7966664774136267396, -0.5255324099163289858, -0.1834346424956498049,
     0.1834346424956498049, 0.5255324099163289858, 0.7966664774136267396, 0.9602898564975362317},
    {-0.9681602395076260898, -0.8360311073266357943, -0.6133714327005903973, -0.3242534234038089290, 0.0,
     0.3242534234038089290, 0.6133714327005903973, 0.8360311073266357943, 0.9681602395076260898},
    {-0.9739065285171717200, -0.8650633666889845107, -0.6794095682990244062, -0.4333953941292471908,
     -0.1488743389816312108, 0.1488743389816312108, 0.4333953941292471908, 0.6794095682990244062,
     0.8650633666889845107, 0.9739065285171717200}};
static const double GL_W[11][10] = {
    {0},
    {2.0},
    {1.0, 1.0},
    {0.5555555555555555556, 0.8888888888888888889, 0.5555555555555555556},
    {0.3478548451374538574, 0.6521451548625461426, 0.6521451548625461426, 0.3478548451374538574},
    {0.2369268850561890875, 0.4786286704993664680, 0.5688888888888888889, 0.4786286704993664680,
     0.2369268850561890875},
    {0.1713244923791703450, 0.3607615730481386076, 0.4679139345726910474, 0.4679139345726910474,
     0.3607615730481386076, 0.1713244923791703450},
    {0.1294849661688696933, 0.2797053914892766679, 0.3818300505051189449, 0.4179591836734693878,
     0.3818300505051189449, 0.2797053914892766679, 0.1294849661688696933},
    {0.1012285362903762591, 0.2223810344533744706, 0.3137066458778872873, 0.3626837833783619830,
     0.3626837833783619830, 0.3137066458778872873, 0.2223810344533744706, 0.1012285362903762591},
    {0.0812743883615744120, 0.1806481606948574041, 0.2606106964029354623, 0.3123470770400028401,
     0.3302393550012597632, 0.3123470770400028401, 0.2606106964029354623, 0.1806481606948574041,
     0.0812743883615744120},
    {0.0666713443086881376, 0.1494513491505805932, 0.2190863625159820440, 0.2692667193099963551,
     0.2955242247147528702, 0.2955242247147528702, 0.2692667193099963551, 0.2190863625159820440,
     0.1494513491505805932, 0.0666713443086881376}};

#define OQ_HEX8_KAPPA 5.0
#define OQ_HEX8_MAXDEPTH 48

/* Adds the quadrature of D over the box lo..hi into acc[54].  The box is
 * bisected along its longest axis until its half-diagonal is at most 1/5
 * of its distance to the receiver's image point (x, y, -z), where the
 * kernel is singular; leaves get a tensor nq^3 Gauss-Legendre rule. */
static void image_box(double x, double y, double z, const double lo0[3],
                      const double hi0[3], double nu, int nq, double acc[54])
{
    double stk[OQ_HEX8_MAXDEPTH + 2][6], D[54];
    int dep[OQ_HEX8_MAXDEPTH + 2];
    int top = 0, a;
    const double xi[3] = {x, y, -z};
    for (a = 0; a < 3; ++a) {
        stk[0][a] = lo0[a];
        stk[0][3 + a] = hi0[a];
    }
    dep[0] = 0;
    top = 1;
    while (top > 0) {
        double lo[3], hi[3], h[3], ctr[3], dist2 = 0.0, hd2 = 0.0;
        int d, i, j, k, m, amax = 0;
        --top;
        for (a = 0; a < 3; ++a) {
            double dd;
            lo[a] = stk[top][a];
            hi[a] = stk[top][3 + a];
            h[a] = 0.5 * (hi[a] - lo[a]);
            ctr[a] = 0.5 * (hi[a] + lo[a]);
            dd = fabs(xi[a] - ctr[a]) - h[a];
            if (dd > 0.0) dist2 += dd * dd;
            hd2 += h[a] * h[a];
            if (h[a] > h[amax]) amax = a;
        }
        d = dep[top];
        if (d < OQ_HEX8_MAXDEPTH && dist2 < OQ_HEX8_KAPPA * OQ_HEX8_KAPPA * hd2) {
            /* children: lower half pushed last so it is processed first */
            for (a = 0; a < 3; ++a) {
                stk[top][a] = lo[a];
                stk[top][3 + a] = hi[a];
                stk[top + 1][a] = lo[a];
                stk[top + 1][3 + a] = hi[a];
            }
            stk[top][amax] = ctr[amax];     /* upper half */
            stk[top + 1][3 + amax] = ctr[amax]; /* lower half */
            dep[top] = dep[top + 1] = d + 1;
            top += 2;
            continue;
        }
        for (i = 0; i < nq; ++i)
            for (j = 0; j < nq; ++j)
                for (k = 0; k < nq; ++k) {
                    double ys1 = ctr[0] + h[0] * GL_X[nq][i];
                    double ys2 = ctr[1] + h[1] * GL_X[nq][j];
                    double ys3 = ctr[2] + h[2] * GL_X[nq][k];
                    double w = GL_W[nq][i] * GL_W[nq][j] * GL_W[nq][k] * h[0] * h[1] * h[2];
                    image_kernel(x - ys1, y - ys2, z, ys3, nu, D);
                    for (m = 0; m < 54; ++m) acc[m] += w * D[m];
                }
    }
}

void oq_stress_vol_hex8(double x, double y, double z, double qx, double qy,
                        double qz, double T, double L, double W,
                        const double e6[6], double mu, double nu, int nquad,
                        double s[6])
{
    const double lo[3] = {qx - T / 2.0, qy, qz - W};
    const double hi[3] = {qx + T / 2.0, qy + L, qz};
    const double lambda = 2.0 * mu * nu / (1.0 - 2.0 * nu);
    double e[3][3], m[3][3], mv[6], gu[3][3], acc[54], tr;
    double phi[3][3], psi[3][3][3][3];
    int sidx, a, i, j, k, l, p, inside;
    kcorner_t kc;

    if (nquad < 1) nquad = 1;
    if (nquad > 10) nquad = 10;
    unpack6(e6, e);
    tr = e[0][0] + e[1][1] + e[2][2];
    for (i = 0; i < 3; ++i)
        for (j = 0; j < 3; ++j) m[i][j] = 2.0 * mu * e[i][j] + (i == j ? lambda * tr : 0.0);

    /* Kelvin part: corner sums of Eshelby potential derivatives */
    memset(phi, 0, sizeof(phi));
    memset(psi, 0, sizeof(psi));
    for (sidx = 0; sidx < 8; ++sidx) {
        double X[3], sg = 1.0;
        for (a = 0; a < 3; ++a) {
            int bit = (sidx >> a) & 1;
            X[a] = (a == 0 ? x : (a == 1 ? y : z)) - (bit ? hi[a] : lo[a]);
            if (bit) sg = -sg;
        }
        kelvin_corner(X, &kc);
        for (i = 0; i < 3; ++i)
            for (j = 0; j < 3; ++j) {
                phi[i][j] += sg * kc.phi[i][j];
                for (k = 0; k < 3; ++k)
                    for (l = 0; l < 3; ++l) psi[i][j][k][l] += sg * kc.psi[i][j][k][l];
            }
    }
    for (i = 0; i < 3; ++i)
        for (l = 0; l < 3; ++l) {
            double acc_k = 0.0;
            for (j = 0; j < 3; ++j)
                for (k = 0; k < 3; ++k) acc_k += e[j][k] * psi[i][j][k][l];
            for (k = 0; k < 3; ++k) acc_k -= 4.0 * (1.0 - nu) * e[i][k] * phi[k][l];
            acc_k -= 2.0 * nu * tr * phi[i][l];
            gu[i][l] = acc_k / (8.0 * M_PI * (1.0 - nu));
        }

    /* image part: quadrature of m : d2 G_I */
    memset(acc, 0, sizeof(acc));
    image_box(x, y, z, lo, hi, nu, nquad, acc);
    mv[0] = m[0][0]; mv[1] = m[0][1]; mv[2] = m[0][2];
    mv[3] = m[1][1]; mv[4] = m[1][2]; mv[5] = m[2][2];
    for (i = 0; i < 3; ++i)
        for (l = 0; l < 3; ++l) {
            double g = 0.0;
            for (p = 0; p < 6; ++p) g += acc[i * 18 + l * 6 + p] * mv[p];
            gu[i][l] += g / (16.0 * M_PI * mu * (1.0 - nu));
        }

    inside = (x >= lo[0] && x < hi[0] && y >= lo[1] && y < hi[1] && z >= lo[2] && z < hi[2]);
    {
        double div = gu[0][0] + gu[1][1] + gu[2][2];
        double sg3[3][3];
        for (i = 0; i < 3; ++i)
            for (j = 0; j < 3; ++j)
                sg3[i][j] = mu * (gu[i][j] + gu[j][i]) + (i == j ? lambda * div : 0.0) -
                            (inside ? m[i][j] : 0.0);
        s[0] = sg3[0][0]; s[1] = sg3[0][1]; s[2] = sg3[0][2];
        s[3] = sg3[1][1]; s[4] = sg3[1][2]; s[5] = sg3[2][2];
    }
}
