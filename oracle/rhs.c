/*
 * oracle/rhs.c -- TEST INFRASTRUCTURE (the CPU oracle).
 *
 * The in-place ODE right-hand side of src/BEM/equation.jl, FP64:
 * rate-and-state friction with the aging law and regularised friction,
 * power-law (Maxwell for n = 0) mantle viscosity, the K_ff product by the
 * length-(2nx-1) real DFT convolution, and the dense mantle products.
 * OpenMP threads mirror the reference's @batch / BLAS threading; with one
 * thread the sums run in a fixed order.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* update_fault! + dθ_dt, aging law (equation.jl:233-246, :279). */
void oq_update_fault(const oq_rs_t *p, int64_t nf, const double *v,
                     const double *theta, const double *dtau, double *dv,
                     double *dtheta, double *ddelta)
{
    int64_t i;
#pragma omp parallel for
    for (i = 0; i < nf; ++i) {
        double th = theta[i] > 0.0 ? theta[i] : 0.0;
        double psi1 = exp((p->f0 + p->b[i] * log(p->v0 * th / p->L[i])) / p->a[i]) /
                      (2.0 * p->v0);
        double psi2 = p->sigma[i] * psi1 / hypot(1.0, v[i] * psi1);
        double dmu_dv = p->a[i] * psi2;
        double dmu_dth = p->b[i] / theta[i] * v[i] * psi2;
        double dth = 1.0 - v[i] * theta[i] / p->L[i];
        dtheta[i] = dth;
        dv[i] = (dtau[i] - dmu_dth * dth) / (dmu_dv + p->eta);
        ddelta[i] = v[i];
    }
}

/* update_fault_with_dilatancy! (equation.jl:248-276) with the aging law
 * (:279) and the pore-pressure law d𝓅_dt (:282): power-law friction
 * mu = f0 vf^(a/f0) thf^(b/f0) on the effective normal stress sigma - p,
 * vf = max(0, v/v0), thf = max(0, theta v0/L), no radiation damping. */
void oq_update_fault_dila(const oq_rs_t *p, const oq_dila_t *d, int64_t nf,
                          const double *v, const double *theta, const double *pp,
                          const double *dtau, double *dv, double *dtheta,
                          double *ddelta, double *dpp)
{
    int64_t i;
#pragma omp parallel for
    for (i = 0; i < nf; ++i) {
        double dth = 1.0 - v[i] * theta[i] / p->L[i];
        double dp = -(pp[i] - d->p0[i]) / d->tp[i] + d->eps[i] / d->beta[i] / theta[i] * dth;
        double af = p->a[i] / p->f0, bf = p->b[i] / p->f0;
        double vf = v[i] / p->v0 > 0.0 ? v[i] / p->v0 : 0.0;
        double thf = theta[i] * p->v0 / p->L[i];
        double sig = p->sigma[i] - pp[i];
        double va, vam1, tb, tbm1;
        thf = thf > 0.0 ? thf : 0.0;
        va = pow(vf, af);
        vam1 = pow(vf, af - 1.0);
        tb = pow(thf, bf);
        tbm1 = pow(thf, bf - 1.0);
        dtheta[i] = dth;
        dpp[i] = dp;
        dv[i] = (dtau[i] + p->f0 * dp * va * tb -
                 p->f0 * sig * va * tbm1 * bf * p->v0 / p->L[i] * dth) /
                (p->f0 * sig * vam1 * tb * af / p->v0);
        ddelta[i] = v[i];
    }
}

/* ode with dilatancy (equation.jl:168-183): Toeplitz-by-DFT dtau, then
 * update_fault_with_dilatancy!. */
void oq_rhs_fault_dila(const oq_rs_t *p, const oq_dila_t *d, const double *st,
                       int64_t nx, int64_t nxi, const double *v, const double *theta,
                       const double *pp, double *dv, double *dtheta, double *ddelta,
                       double *dpp)
{
    const int64_t nf = nx * nxi;
    double *s = (double *)malloc(sizeof(double) * (size_t)nf);
    double *dtau = (double *)malloc(sizeof(double) * (size_t)nf);
    int64_t i;
    for (i = 0; i < nf; ++i) s[i] = v[i] - p->vpl;
    oq_dtau_fft(st, nx, nxi, s, dtau);
    oq_update_fault_dila(p, d, nf, v, theta, pp, dtau, dv, dtheta, ddelta, dpp);
    free(s);
    free(dtau);
}

/* update_strain_rate! with dϵ_dt = γ σ τ^n (equation.jl:207-222, :285). */
void oq_update_strain_rate(const oq_visc_t *p, int64_t ne, const double *sg,
                           double *de)
{
    int64_t i;
#pragma omp parallel for
    for (i = 0; i < ne; ++i) {
        double skk = (sg[i] + sg[i + 3 * ne] + sg[i + 5 * ne]) / 3.0;
        double sxx = sg[i] - skk, syy = sg[i + 3 * ne] - skk,
               szz = sg[i + 5 * ne] - skk;
        double sxy = sg[i + ne], sxz = sg[i + 2 * ne], syz = sg[i + 4 * ne];
        double tn = sqrt(sxx * sxx + syy * syy + szz * szz +
                         2.0 * (sxy * sxy + sxz * sxz + syz * syz));
        double f = p->gamma[i] * pow(tn, p->n[i]);
        de[i] = f * sxx;
        de[i + ne] = f * sxy;
        de[i + 2 * ne] = f * sxz;
        de[i + 3 * ne] = f * syy;
        de[i + 4 * ne] = f * syz;
        de[i + 5 * ne] = f * szz;
    }
}

/* dτ_dt! (equation.jl:44-61): rfft of the zero-padded source along strike
 * (length M = 2nx-1), per-frequency contraction with the DFT of the
 * symmetric circulant [st; reverse(st[2:end])] (GF.jl:60-67), inverse real
 * DFT, first nx rows.  Direct DFT sums with an exact twiddle index. */
void oq_dtau_fft(const double *st, int64_t nx, int64_t nxi, const double *s,
                 double *dtau)
{
    const int64_t M = 2 * nx - 1, nn = nx * nxi;
    double *cs = (double *)malloc(sizeof(double) * (size_t)M);
    double *sn = (double *)malloc(sizeof(double) * (size_t)M);
    double *gre = (double *)malloc(sizeof(double) * (size_t)(nn * nxi));
    double *gim = (double *)malloc(sizeof(double) * (size_t)(nn * nxi));
    double *sre = (double *)malloc(sizeof(double) * (size_t)nn);
    double *sim = (double *)malloc(sizeof(double) * (size_t)nn);
    int64_t jl, l, j, f;
    for (f = 0; f < M; ++f) {
        cs[f] = cos(2.0 * M_PI * (double)f / (double)M);
        sn[f] = sin(2.0 * M_PI * (double)f / (double)M);
    }
    /* kernel DFT: gf[f, j, l] */
#pragma omp parallel for
    for (jl = 0; jl < nxi * nxi; ++jl) {
        const double *col = st + nx * jl;
        int64_t ff, n;
        for (ff = 0; ff < nx; ++ff) {
            double re = 0.0, im = 0.0;
            for (n = 0; n < M; ++n) {
                double x = n < nx ? col[n] : col[M - n];
                int64_t k = (ff * n) % M;
                re += x * cs[k];
                im -= x * sn[k];
            }
            gre[ff + nx * jl] = re;
            gim[ff + nx * jl] = im;
        }
    }
    /* source DFT (zero padded) */
#pragma omp parallel for
    for (l = 0; l < nxi; ++l) {
        int64_t ff, n;
        for (ff = 0; ff < nx; ++ff) {
            double re = 0.0, im = 0.0;
            for (n = 0; n < nx; ++n) {
                int64_t k = (ff * n) % M;
                re += s[n + nx * l] * cs[k];
                im -= s[n + nx * l] * sn[k];
            }
            sre[ff + nx * l] = re;
            sim[ff + nx * l] = im;
        }
    }
    /* per frequency contraction and inverse real DFT */
#pragma omp parallel for
    for (j = 0; j < nxi; ++j) {
        double *tre = (double *)malloc(sizeof(double) * (size_t)nx);
        double *tim = (double *)malloc(sizeof(double) * (size_t)nx);
        int64_t ff, n, ll;
        for (ff = 0; ff < nx; ++ff) {
            double re = 0.0, im = 0.0;
            for (ll = 0; ll < nxi; ++ll) {
                double gr = gre[ff + nx * (j + nxi * ll)];
                double gi = gim[ff + nx * (j + nxi * ll)];
                double xr = sre[ff + nx * ll], xi = sim[ff + nx * ll];
                re += gr * xr - gi * xi;
                im += gr * xi + gi * xr;
            }
            tre[ff] = re;
            tim[ff] = im;
        }
        for (n = 0; n < nx; ++n) {
            double acc = tre[0];
            for (ff = 1; ff < nx; ++ff) {
                int64_t k = (ff * n) % M;
                acc += 2.0 * (tre[ff] * cs[k] - tim[ff] * sn[k]);
            }
            dtau[n + nx * j] = acc / (double)M;
        }
        free(tre);
        free(tim);
    }
    free(cs); free(sn); free(gre); free(gim); free(sre); free(sim);
}

/* Dense Toeplitz contraction E[i,j] = sum_{k,l} st[|i-k|, j, l] s[k,l]
 * (test/BEM/tests.jl:46-58). */
void oq_dtau_toeplitz(const double *st, int64_t nx, int64_t nxi,
                      const double *s, double *dtau)
{
    int64_t r;
#pragma omp parallel for
    for (r = 0; r < nx * nxi; ++r) {
        const int64_t i = r % nx, j = r / nx;
        double acc = 0.0;
        int64_t k, l;
        for (l = 0; l < nxi; ++l)
            for (k = 0; k < nx; ++k) {
                int64_t d = i > k ? i - k : k - i;
                acc += st[d + nx * j + nx * nxi * l] * s[k + nx * l];
            }
        dtau[r] = acc;
    }
}

/* ode(du,u,p,t), fault only (equation.jl:156-166). */
void oq_rhs_fault(const oq_rs_t *p, const double *st, int64_t nx, int64_t nxi,
                  const double *v, const double *theta, double *dv,
                  double *dtheta, double *ddelta)
{
    const int64_t nf = nx * nxi;
    double *s = (double *)malloc(sizeof(double) * (size_t)nf);
    double *dtau = (double *)malloc(sizeof(double) * (size_t)nf);
    int64_t i;
    for (i = 0; i < nf; ++i) s[i] = v[i] - p->vpl;
    oq_dtau_fft(st, nx, nxi, s, dtau);
    oq_update_fault(p, nf, v, theta, dtau, dv, dtheta, ddelta);
    free(s);
    free(dtau);
}

/* y += A x with A (m x n) column-major.  OpenMP over blocks of 512 rows;
 * each block sweeps the columns in order, reading 4 KB contiguous per
 * column (the access pattern of a BLAS column-major dgemv), with one
 * accumulator per row: row i's sum is formed in column order and then
 * added to y[i], as a row-by-row loop would form it. */
static void gemv_colmajor(const double *A, int64_t m, int64_t n,
                          const double *x, double *y)
{
    enum { RB = 512 };
    int64_t b;
#pragma omp parallel for schedule(static)
    for (b = 0; b < m; b += RB) {
        double acc[RB];
        const int64_t ib = m - b < RB ? m - b : RB;
        int64_t i, j;
        for (i = 0; i < ib; ++i) acc[i] = 0.0;
        for (j = 0; j < n; ++j) {
            const double *col = A + b + m * j;
            const double xj = x[j];
            for (i = 0; i < ib; ++i) acc[i] += col[i] * xj;
        }
        for (i = 0; i < ib; ++i) y[b + i] += acc[i];
    }
}

/* ode(du,u,p,t), coupled viscoelastic (equation.jl:185-205). */
void oq_rhs_coupled(const oq_rs_t *pf, const oq_visc_t *pa, const double *st,
                    const double *gf12, const double *gf21, const double *gf22,
                    int64_t nx, int64_t nxi, int64_t ne, const double *v,
                    const double *theta, const double *sigma, double *dv,
                    double *dtheta, double *deps, double *dsigma,
                    double *ddelta)
{
    const int64_t nf = nx * nxi, nm = 6 * ne;
    double *s = (double *)malloc(sizeof(double) * (size_t)nf);
    double *dtau = (double *)malloc(sizeof(double) * (size_t)nf);
    double *rde = (double *)malloc(sizeof(double) * (size_t)nm);
    int64_t i;
    for (i = 0; i < nf; ++i) s[i] = v[i] - pf->vpl;
    oq_update_strain_rate(pa, ne, sigma, deps);
    for (i = 0; i < nm; ++i) rde[i] = deps[i] - pa->deps0[i / ne];
    oq_dtau_fft(st, nx, nxi, s, dtau);
    gemv_colmajor(gf21, nf, nm, rde, dtau);
    memset(dsigma, 0, sizeof(double) * (size_t)nm);
    gemv_colmajor(gf12, nm, nf, s, dsigma);
    gemv_colmajor(gf22, nm, nm, rde, dsigma);
    oq_update_fault(pf, nf, v, theta, dtau, dv, dtheta, ddelta);
    free(s);
    free(dtau);
    free(rde);
}

/* Same RHS with a dense receiver-major operator K (N x N, row-major,
 * N = nf + 6ne): rows [0,nf) fault tractions, rows [nf,N) mantle stresses;
 * columns [0,nf) fault slip-rate deficit, [nf,N) relative strain rate. */
void oq_rhs_dense(const oq_rs_t *pf, const oq_visc_t *pa, const double *K,
                  int64_t nf, int64_t ne, const double *v, const double *theta,
                  const double *sigma, double *dv, double *dtheta,
                  double *deps, double *dsigma, double *ddelta)
{
    const int64_t nm = 6 * ne, N = nf + nm;
    double *s = (double *)malloc(sizeof(double) * (size_t)N);
    double *y = (double *)malloc(sizeof(double) * (size_t)N);
    int64_t i;
    for (i = 0; i < nf; ++i) s[i] = v[i] - pf->vpl;
    if (ne > 0) {
        oq_update_strain_rate(pa, ne, sigma, deps);
        for (i = 0; i < nm; ++i) s[nf + i] = deps[i] - pa->deps0[i / ne];
    }
#pragma omp parallel for
    for (i = 0; i < N; ++i) {
        double acc = 0.0;
        int64_t j;
        const double *row = K + N * i;
        for (j = 0; j < N; ++j) acc += row[j] * s[j];
        y[i] = acc;
    }
    for (i = 0; i < nm; ++i) dsigma[i] = y[nf + i];
    oq_update_fault(pf, nf, v, theta, y, dv, dtheta, ddelta);
    free(s);
    free(y);
}

/* Adapter: the dense RHS on a packed ArrayPartition state. */
typedef struct {
    const oq_rs_t *pf;
    const oq_visc_t *pa;
    const double *K;
    int64_t nf, ne;
} dense_ctx_t;

static void dense_packed(void *vctx, double t, const double *u, double *du)
{
    const dense_ctx_t *c = (const dense_ctx_t *)vctx;
    const int64_t nf = c->nf, nm = 6 * c->ne;
    (void)t;
    if (c->ne > 0)
        oq_rhs_dense(c->pf, c->pa, c->K, nf, c->ne, u, u + nf, u + 2 * nf + nm, du, du + nf,
                     du + 2 * nf, du + 2 * nf + nm, du + 2 * nf + 2 * nm);
    else
        oq_rhs_dense(c->pf, c->pa, c->K, nf, 0, u, u + nf, NULL, du, du + nf, NULL, NULL,
                     du + 2 * nf);
}

void oq_cycle_dense(const oq_rs_t *pf, const oq_visc_t *pa, const double *K,
                    int64_t nf, int64_t ne, double *u, double t0, double t1,
                    const oq_ode_opts_t *o, oq_ode_stats_t *st, double *step_t,
                    double *step_v, int64_t cap, int64_t *nrec)
{
    oq_cycle_dense_alg(pf, pa, K, nf, ne, u, t0, t1, o, st, step_t, step_v, cap, nrec, 0);
}

void oq_cycle_dense_alg(const oq_rs_t *pf, const oq_visc_t *pa, const double *K,
                        int64_t nf, int64_t ne, double *u, double t0, double t1,
                        const oq_ode_opts_t *o, oq_ode_stats_t *st, double *step_t,
                        double *step_v, int64_t cap, int64_t *nrec, int alg)
{
    dense_ctx_t c = {pf, pa, K, nf, ne};
    const int64_t n = ne > 0 ? 3 * nf + 12 * ne : 3 * nf;
    if (alg == 1)
        oq_vcabm5_record(dense_packed, &c, n, u, t0, t1, o, st, nf, step_t, step_v, cap, nrec);
    else
        oq_tsit5_record(dense_packed, &c, n, u, t0, t1, o, st, nf, step_t, step_v, cap, nrec);
}
