/*
 * oracle/tsit5.c -- TEST INFRASTRUCTURE (the CPU oracle).
 *
 * Tsitouras (2011) 5(4) explicit Runge-Kutta pair with FSAL, the Tsit5 the
 * reference hands to OrdinaryDiffEq (test/tests.jl:11, io.jl:125-130),
 * restated with OrdinaryDiffEq's step-size semantics: RMS norm of the
 * embedded error scaled by abstol + reltol*max(|u_old|,|u_new|), the PI
 * controller (beta1 = 7/50, beta2 = 2/25, qmin = 1/5, qmax = 10,
 * gamma = 9/10, qoldinit = 1e-4) and the Hairer initial-step heuristic.
 * OrdinaryDiffEq is not vendored in the reference, so the exact step
 * sequence of the reference is unpinned; the GPU stepper implements this
 * same controller, which makes GPU-vs-oracle comparisons algorithm-exact.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* Butcher tableau (Tsitouras 2011, Table 1), b = last row of a (FSAL),
 * btilde = b - bhat (the embedded 4th-order weights). */
static const double TS_C[7] = {0.0, 0.161, 0.327, 0.9, 0.9800255409045097,
                               1.0, 1.0};
static const double TS_A[7][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365,
     -0.09249506636175525, 0, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159,
     -0.071584973281401, -0.028269050394068383, 0, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742,
     -3.290069515436081, 2.324710524099774, 0}};
static const double TS_BT[7] = {-0.00178001105222577714, -0.0008164344596567469,
                                0.007880878010261995, -0.1447110071732629,
                                0.5823571654525552, -0.45808210592918697,
                                0.015151515151515152};

static double tableau_flat[7 + 49 + 7 + 7];

const double *oq_tsit5_tableau(void)
{
    int i, j;
    for (i = 0; i < 7; ++i) tableau_flat[i] = TS_C[i];
    for (i = 0; i < 7; ++i)
        for (j = 0; j < 7; ++j) tableau_flat[7 + 7 * i + j] = TS_A[i][j];
    for (j = 0; j < 7; ++j) tableau_flat[56 + j] = TS_A[6][j];
    for (j = 0; j < 7; ++j) tableau_flat[63 + j] = TS_BT[j];
    return tableau_flat;
}

static double rms_scaled(const double *x, const double *u, int64_t n,
                         double abstol, double reltol)
{
    double acc = 0.0;
    int64_t i;
    for (i = 0; i < n; ++i) {
        double w = x[i] / (abstol + fabs(u[i]) * reltol);
        acc += w * w;
    }
    return sqrt(acc / (double)n);
}

/* ode_determine_initdt (Hairer, Norsett & Wanner, II.4), order 5. */
static double initial_dt(oq_rhs_fn f, void *ctx, int64_t n, const double *u0,
                         double t0, const double *f0, const oq_ode_opts_t *o,
                         double *tmp_u, double *tmp_f)
{
    double d0 = rms_scaled(u0, u0, n, o->abstol, o->reltol);
    double d1 = rms_scaled(f0, u0, n, o->abstol, o->reltol);
    double dt0, d2, acc = 0.0, m, dt1;
    int64_t i;
    dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
    dt0 = fmin(dt0, o->dtmax);
    for (i = 0; i < n; ++i) tmp_u[i] = u0[i] + dt0 * f0[i];
    f(ctx, t0 + dt0, tmp_u, tmp_f);
    for (i = 0; i < n; ++i) {
        double w = (tmp_f[i] - f0[i]) / (o->abstol + fabs(u0[i]) * o->reltol);
        acc += w * w;
    }
    d2 = sqrt(acc / (double)n) / dt0;
    m = fmax(d1, d2);
    dt1 = (m <= 1e-15) ? fmax(1e-6, dt0 * 1e-3) : pow(0.01 / m, 1.0 / 5.0);
    return fmin(fmin(100.0 * dt0, dt1), o->dtmax);
}

void oq_tsit5(oq_rhs_fn f, void *ctx, int64_t n, double *u, double t0,
              double t1, const oq_ode_opts_t *o, oq_ode_stats_t *st)
{
    oq_tsit5_record(f, ctx, n, u, t0, t1, o, st, 0, NULL, NULL, 0, NULL);
}

/* Tsit5 that also records, after each accepted step, t and max|u[0:nrec_n]|
 * (the largest slip rate when u starts with v) into step_t / step_v (up to
 * cap entries; *nrec gets the number of accepted steps). */
void oq_tsit5_record(oq_rhs_fn f, void *ctx, int64_t n, double *u, double t0,
                     double t1, const oq_ode_opts_t *o, oq_ode_stats_t *st,
                     int64_t nrec_n, double *step_t, double *step_v, int64_t cap,
                     int64_t *nrec)
{
    const double beta1 = 7.0 / 50.0, beta2 = 2.0 / 25.0, qmin = 0.2,
                 qmax = 10.0, gamma = 0.9, qoldinit = 1e-4;
    double *k = (double *)malloc(sizeof(double) * (size_t)(7 * n));
    double *y = (double *)malloc(sizeof(double) * (size_t)n);
    double *err = (double *)malloc(sizeof(double) * (size_t)n);
    double t = t0, dt, qold = qoldinit;
    int64_t iter = 0, i;
    int s, j;

    memset(st, 0, sizeof(*st));
    f(ctx, t, u, k);
    st->nf = 1;
    if (o->dt0 > 0.0) {
        dt = o->dt0;
    } else {
        dt = initial_dt(f, ctx, n, u, t, k, o, y, err);
        st->nf += 1;
    }
    while (t < t1) {
        double eest, q11, q, tnext;
        int accept;
        if (iter >= o->maxiters) { st->retcode = 1; break; }
        if (dt > o->dtmax) dt = o->dtmax;
        if (t + dt > t1) dt = t1 - t;
        if (!(dt >= o->dtmin) || !isfinite(dt)) { st->retcode = isfinite(dt) ? 2 : 3; break; }
        ++iter;
        for (s = 1; s < 7; ++s) {
            for (i = 0; i < n; ++i) {
                double acc = 0.0;
                for (j = 0; j < s; ++j) acc += TS_A[s][j] * k[j * n + i];
                y[i] = u[i] + dt * acc;
            }
            f(ctx, t + TS_C[s] * dt, y, k + s * n);
            st->nf += 1;
        }
        /* y holds the 5th-order solution (stage 7 argument, FSAL) */
        for (i = 0; i < n; ++i) {
            double acc = 0.0;
            for (j = 0; j < 7; ++j) acc += TS_BT[j] * k[j * n + i];
            err[i] = dt * acc / (o->abstol + fmax(fabs(u[i]), fabs(y[i])) * o->reltol);
        }
        {
            double a2 = 0.0;
            for (i = 0; i < n; ++i) a2 += err[i] * err[i];
            eest = sqrt(a2 / (double)n);
        }
        if (eest == 0.0) {
            q11 = 0.0;
            q = 1.0 / qmax;
        } else {
            q11 = pow(eest, beta1);
            q = q11 / pow(qold, beta2);
            q = fmax(1.0 / qmax, fmin(1.0 / qmin, q / gamma));
        }
        accept = (eest <= 1.0);
        if (accept) {
            tnext = (t1 - (t + dt) <= 0.0) ? t1 : t + dt;
            t = tnext;
            memcpy(u, y, sizeof(double) * (size_t)n);
            memcpy(k, k + 6 * n, sizeof(double) * (size_t)n);
            if (q >= 1.0 && q <= 1.0) q = 1.0;
            qold = fmax(eest, qoldinit);
            dt = dt / q;
            if (nrec_n > 0 && st->naccept < cap && step_t && step_v) {
                double vm = 0.0;
                for (i = 0; i < nrec_n; ++i) vm = fmax(vm, fabs(u[i]));
                step_t[st->naccept] = t;
                step_v[st->naccept] = vm;
            }
            st->naccept += 1;
        } else {
            if (!isfinite(eest)) { st->retcode = 3; st->nreject += 1; break; }
            dt = dt / fmin(1.0 / qmin, q11 / gamma);
            st->nreject += 1;
        }
    }
    st->t = t;
    st->dt = dt;
    if (nrec) *nrec = st->naccept;
    free(k);
    free(y);
    free(err);
}
