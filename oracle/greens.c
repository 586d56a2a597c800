/*
 * oracle/greens.c -- TEST INFRASTRUCTURE (the CPU oracle).
 *
 * Mesh generation (src/BEM/mesh.jl) and the four stress Green's-function
 * builders (src/BEM/GF.jl), FP64, restated in C.  Each function cites the
 * reference lines it follows.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* ---------- degree trigonometry (Julia sind/cosd semantics) ---------- */
static double sin_deg_small(double r) { return sin(r * (M_PI / 180.0)); }
static double cos_deg_small(double r) { return cos(r * (M_PI / 180.0)); }

double oq_sind(double deg)
{
    double rx, arx, s;
    if (!isfinite(deg)) return NAN;
    rx = fmod(deg, 360.0);
    arx = fabs(rx);
    if (rx == 0.0) return copysign(0.0, deg);
    if (arx == 180.0) return copysign(0.0, rx);
    if (arx < 45.0) s = sin_deg_small(arx);
    else if (arx <= 135.0) s = cos_deg_small(90.0 - arx);
    else if (arx < 225.0) s = sin_deg_small(180.0 - arx);
    else if (arx <= 315.0) s = -cos_deg_small(270.0 - arx);
    else s = -sin_deg_small(360.0 - arx);
    return rx < 0.0 ? -s : s;
}

double oq_cosd(double deg)
{
    double arx;
    if (!isfinite(deg)) return NAN;
    arx = fabs(fmod(deg, 360.0));
    if (arx < 45.0) return cos_deg_small(arx);
    if (arx <= 135.0) return sin_deg_small(90.0 - arx);
    if (arx < 225.0) return -cos_deg_small(180.0 - arx);
    if (arx <= 315.0) return sin_deg_small(arx - 270.0);
    return cos_deg_small(360.0 - arx);
}

/* ---------- mesh.jl ---------- */

/* Length of Julia's range(start, stop=stop, step=step): floor((stop-start)
 * /step) + 1, with the quotient snapped to an integer when it is within a
 * few ulps (Julia rationalises float ranges, so 50/2 gives exactly 25). */
int64_t oq_range_len(double start, double stop, double step)
{
    double qt = (stop - start) / step;
    double rq = nearbyint(qt);
    if (fabs(qt - rq) <= 1e-10 * fmax(1.0, fabs(rq))) qt = rq;
    if (qt < 0.0) return 0;
    return (int64_t)floor(qt) + 1;
}

/* _equidist_mesh_strike (mesh.jl:52-56). */
void oq_mesh_strike(double xlen, double dx, int64_t nx, double *x, double *ax)
{
    int64_t i;
    double x0 = -xlen / 2.0 + dx / 2.0;
    for (i = 0; i < nx; ++i) {
        x[i] = x0 + (double)i * dx;
        ax[2 * i] = x[i] - dx / 2.0;
        ax[2 * i + 1] = x[i] + dx / 2.0;
    }
}

/* _equidist_mesh_downdip (mesh.jl:45-50): xi_j = -(j-1) dxi - dxi/2,
 * y = xi cosd(dip), z = xi sind(dip). */
void oq_mesh_downdip(double xilen, double dxi, double dip, int64_t nxi,
                     double *xi, double *axi, double *y, double *z)
{
    int64_t j;
    double cd = oq_cosd(dip), sd = oq_sind(dip);
    (void)xilen;
    for (j = 0; j < nxi; ++j) {
        double w = (0.0 - (double)j * dxi) - dxi / 2.0;
        xi[j] = w;
        axi[2 * j] = w - dxi / 2.0;
        axi[2 * j + 1] = w + dxi / 2.0;
        y[j] = w * cd;
        z[j] = w * sd;
    }
}

/* Structured equivalent of gen_gmsh_mesh (mesh.jl:95-133: uniform
 * transfinite x/y, nz extruded layers with heights normalize(cumsum(rfzh),
 * Inf)) followed by gen_mesh(:BEMHex8Mesh) (mesh.jl:149-186: barycentre c,
 * sizes, q = (cx, cy - dy/2, cz + dz/2)).  Cell order: x fastest, then y,
 * then layer (gmsh's own order is not reproducible without gmsh). */
void oq_mesh_hex8(double llx, double lly, double llz, double dx, double dy,
                  double dz, int64_t nx, int64_t ny, int64_t nz,
                  const double *rfzh, double *cx, double *cy, double *cz,
                  double *qx, double *qy, double *qz, double *ddx, double *ddy,
                  double *ddz)
{
    int64_t i, j, k, e = 0;
    double *zf = (double *)malloc((size_t)(nz + 1) * sizeof(double));
    double acc = 0.0, tot = 0.0;
    for (k = 0; k < nz; ++k) tot += rfzh ? rfzh[k] : 1.0;
    zf[0] = llz;
    for (k = 0; k < nz; ++k) {
        acc += rfzh ? rfzh[k] : 1.0;
        zf[k + 1] = llz + dz * (acc / tot);
    }
    for (k = 0; k < nz; ++k) {
        for (j = 0; j < ny; ++j) {
            for (i = 0; i < nx; ++i, ++e) {
                double x0 = llx + dx * (double)i / (double)nx;
                double x1 = llx + dx * (double)(i + 1) / (double)nx;
                double y0 = lly + dy * (double)j / (double)ny;
                double y1 = lly + dy * (double)(j + 1) / (double)ny;
                double z0 = zf[k], z1 = zf[k + 1];
                cx[e] = 0.5 * (x0 + x1);
                cy[e] = 0.5 * (y0 + y1);
                cz[e] = 0.5 * (z0 + z1);
                ddx[e] = fabs(x1 - x0);
                ddy[e] = fabs(y1 - y0);
                ddz[e] = fabs(z1 - z0);
                qx[e] = cx[e];
                qy[e] = cy[e] - ddy[e] / 2.0;
                qz[e] = cz[e] + ddz[e] / 2.0;
            }
        }
    }
    free(zf);
}

/* ---------- GF.jl ---------- */

/* shear_traction_dc3d (GF.jl:76-87). */
static double traction_dc3d(int ftype, const double *u, double lambda,
                            double mu, double dip)
{
    if (ftype == 0) {
        double sxy = mu * (u[4] + u[6]);
        double sxz = mu * (u[5] + u[9]);
        return -sxy * oq_sind(dip) + sxz * oq_cosd(dip);
    } else {
        double szz = (lambda + 2.0 * mu) * u[11] + lambda * u[3] + lambda * u[7];
        double syy = (lambda + 2.0 * mu) * u[7] + lambda * u[3] + lambda * u[11];
        double syz = mu * (u[10] + u[8]);
        return (szz - syy) / 2.0 * oq_sind(2.0 * dip) + syz * oq_cosd(2.0 * dip);
    }
}

/* shear_traction_vol (GF.jl:89-96). */
static double traction_vol(int ftype, const double *s, double dip)
{
    if (ftype == 0) return -s[1] * oq_sind(dip) + s[2] * oq_cosd(dip);
    return (s[5] - s[3]) / 2.0 * oq_sind(2.0 * dip) + s[4] * oq_cosd(2.0 * dip);
}

/* K_ff Toeplitz kernel (GF.jl:31-58): st[i,j,l] = traction at (x_i, row j)
 * from unit slip on (column 1, row l), summed over periodic images
 * p*lrept, p in -nrept..nrept, lrept = (1+buffer_ratio)*dx*nx. */
void oq_gf_ff(const oq_fault_mesh_t *m, double lambda, double mu, int ftype,
              int64_t nrept, double buffer_ratio, double *st)
{
    const double lrept = (buffer_ratio + 1.0) * (m->dx * (double)m->nx);
    const double alpha = (lambda + mu) / (lambda + 2.0 * mu);
    const double ds = ftype == 0 ? 1.0 : 0.0, dd = ftype == 0 ? 0.0 : 1.0;
    int64_t l;
#pragma omp parallel for schedule(dynamic)
    for (l = 0; l < m->nxi; ++l) {
        int64_t i, j, p;
        double u[12], c[12];
        for (j = 0; j < m->nxi; ++j) {
            for (i = 0; i < m->nx; ++i) {
                int k;
                memset(u, 0, sizeof(u));
                for (p = -nrept; p <= nrept; ++p) {
                    double jump = (double)p * lrept;
                    oq_dc3d(alpha, m->x[i], m->y[j], m->z[j], m->dep, m->dip,
                            m->ax[0] + jump, m->ax[1] + jump, m->axi[2 * l],
                            m->axi[2 * l + 1], ds, dd, 0.0, c);
                    for (k = 0; k < 12; ++k) u[k] += c[k];
                }
                st[i + m->nx * j + m->nx * m->nxi * l] =
                    traction_dc3d(ftype, u, lambda, mu, m->dip);
            }
        }
    }
}

/* K_mf = gf12 (GF.jl:123-174): 6ne x nf column-major; quadrature over the
 * receiver cell at c + l .* delta/2 with weights qw. */
void oq_gf_mf(const oq_fault_mesh_t *mf, const oq_hex8_mesh_t *ma,
              double lambda, double mu, int ftype, const double *qc,
              const double *qw, int64_t nq, int64_t nrept,
              double buffer_ratio, double *st)
{
    const int64_t ne = ma->ne, nf = mf->nx * mf->nxi;
    const double lrept = (buffer_ratio + 1.0) * (mf->dx * (double)mf->nx);
    const double alpha = (lambda + mu) / (lambda + 2.0 * mu);
    const double ds = ftype == 0 ? 1.0 : 0.0, dd = ftype == 0 ? 0.0 : 1.0;
    int64_t j;
    memset(st, 0, sizeof(double) * (size_t)(6 * ne * nf));
#pragma omp parallel for schedule(dynamic)
    for (j = 0; j < nf; ++j) {
        const int64_t q1 = j % mf->nx, q2 = j / mf->nx;
        double *col = st + 6 * ne * j;
        int64_t i, w, p;
        double u[12], c[12];
        for (i = 0; i < ne; ++i) {
            for (w = 0; w < nq; ++w) {
                double rx = ma->cx[i] + qc[3 * w] * ma->dx[i] / 2.0;
                double ry = ma->cy[i] + qc[3 * w + 1] * ma->dy[i] / 2.0;
                double rz = ma->cz[i] + qc[3 * w + 2] * ma->dz[i] / 2.0;
                double lek;
                int k;
                memset(u, 0, sizeof(u));
                for (p = -nrept; p <= nrept; ++p) {
                    double jump = (double)p * lrept;
                    oq_dc3d(alpha, rx, ry, rz, mf->dep, mf->dip,
                            mf->ax[2 * q1] + jump, mf->ax[2 * q1 + 1] + jump,
                            mf->axi[2 * q2], mf->axi[2 * q2 + 1], ds, dd, 0.0,
                            c);
                    for (k = 0; k < 12; ++k) u[k] += c[k];
                }
                lek = lambda * (u[3] + u[7] + u[11]);
                col[i] += qw[w] * (lek + 2.0 * mu * u[3]);
                col[i + ne] += qw[w] * (mu * (u[4] + u[6]));
                col[i + 2 * ne] += qw[w] * (mu * (u[5] + u[9]));
                col[i + 3 * ne] += qw[w] * (lek + 2.0 * mu * u[7]);
                col[i + 4 * ne] += qw[w] * (mu * (u[8] + u[10]));
                col[i + 5 * ne] += qw[w] * (lek + 2.0 * mu * u[11]);
            }
        }
    }
}

/* K_fm = gf21 (GF.jl:194-227): nf x 6ne column-major, column (p-1)ne + j
 * is the traction at every patch centre from unit eigenstrain component p
 * in cell j.  nu = lambda / (2 (lambda + mu)) (GF.jl:203). */
void oq_gf_fm(const oq_hex8_mesh_t *ma, const oq_fault_mesh_t *mf,
              double lambda, double mu, int ftype, int nquad, double *st)
{
    const int64_t ne = ma->ne, nf = mf->nx * mf->nxi;
    const double nu = lambda / 2.0 / (lambda + mu);
    int64_t j;
#pragma omp parallel for schedule(dynamic)
    for (j = 0; j < ne; ++j) {
        int p;
        int64_t i;
        for (p = 0; p < 6; ++p) {
            double e[6] = {0, 0, 0, 0, 0, 0}, s[6];
            double *col = st + nf * (p * ne + j);
            e[p] = 1.0;
            for (i = 0; i < nf; ++i) {
                const int64_t a = i % mf->nx, b = i / mf->nx;
                oq_stress_vol_hex8(mf->x[a], mf->y[b], mf->z[b], ma->qx[j],
                                   ma->qy[j], ma->qz[j], ma->dx[j], ma->dy[j],
                                   ma->dz[j], e, mu, nu, nquad, s);
                col[i] = traction_vol(ftype, s, mf->dip);
            }
        }
    }
}

/* K_mm = gf22 (GF.jl:250-296): 6ne x 6ne column-major; entry
 * [(k-1)ne + j, (p-1)ne + i] += w_q sigma_k at quadrature points of
 * receiver cell j. */
void oq_gf_mm_rows(const oq_hex8_mesh_t *src, const oq_hex8_mesh_t *rcv,
                   double lambda, double mu, const double *qc, const double *qw,
                   int64_t nq, int nquad, double *st)
{
    /* GF.jl:250-296 with the receiver cells (quadrature-averaged rows)
     * taken from `rcv` and the eigenstrain sources from `src`: entry
     * [(k-1)nr + j, (p-1)ns + i] column-major, (6 nr) x (6 ns). */
    const int64_t ns = src->ne, nr = rcv->ne, r6 = 6 * nr;
    const double nu = lambda / 2.0 / (lambda + mu);
    int64_t i;
    memset(st, 0, sizeof(double) * (size_t)(r6 * 6 * ns));
#pragma omp parallel for schedule(dynamic)
    for (i = 0; i < ns; ++i) {
        int p, k;
        int64_t j, q;
        for (p = 0; p < 6; ++p) {
            double e[6] = {0, 0, 0, 0, 0, 0}, s[6];
            double *col = st + r6 * (p * ns + i);
            e[p] = 1.0;
            for (j = 0; j < nr; ++j) {
                for (q = 0; q < nq; ++q) {
                    double rx = rcv->cx[j] + qc[3 * q] * rcv->dx[j] / 2.0;
                    double ry = rcv->cy[j] + qc[3 * q + 1] * rcv->dy[j] / 2.0;
                    double rz = rcv->cz[j] + qc[3 * q + 2] * rcv->dz[j] / 2.0;
                    oq_stress_vol_hex8(rx, ry, rz, src->qx[i], src->qy[i],
                                       src->qz[i], src->dx[i], src->dy[i],
                                       src->dz[i], e, mu, nu, nquad, s);
                    for (k = 0; k < 6; ++k) col[k * nr + j] += s[k] * qw[q];
                }
            }
        }
    }
}

void oq_gf_mm(const oq_hex8_mesh_t *ma, double lambda, double mu,
              const double *qc, const double *qw, int64_t nq, int nquad,
              double *st)
{
    oq_gf_mm_rows(ma, ma, lambda, mu, qc, qw, nq, nquad, st);
}
