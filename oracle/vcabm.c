/*
 * oracle/vcabm.c -- TEST INFRASTRUCTURE (the CPU oracle).
 *
 * Variable-coefficient Adams-Bashforth-Moulton of order 5 (the VCABM5 the
 * reference's example hands to OrdinaryDiffEq, examples/otf-with-mantle.jl:
 * 160,167-168), restated from the published method: Hairer, Norsett &
 * Wanner, Solving ODEs I, Sec. III.5 (variable step size Adams methods in
 * divided-difference form, after Krogh and Shampine & Gordon):
 *
 *   predictor  p = y_n + h sum_{j<k} g_j(n) Phi*_j(n),  Phi*_j = beta_j Phi_j
 *   evaluate   fp = f(t_{n+1}, p)
 *   corrector  y_{n+1} = p + h g_k(n) Phi^p_k,  Phi^p_k = fp - sum_{j<k} Phi*_j
 *   error      h (g_k - g_{k-1}) Phi^p_k   (order k vs order k+1 corrector)
 *   evaluate   f_{n+1} = f(t_{n+1}, y_{n+1});  Phi_0 = f_{n+1},
 *              Phi_{j+1}(n+1) = Phi_j(n+1) - Phi*_j(n)
 *
 * with beta_j(n) = prod_{i<j} psi_i(n+1)/psi_i(n), psi_i(m) = t_m - t_{m-i-1},
 * and g_j(n) = c_{j,1} from c_{0,q} = 1/q,
 * c_{j,q} = c_{j-1,q} - c_{j-1,q+1} h / psi_{j-1}(n+1).  k = 5 once five
 * derivatives are known; the first steps use k = 1, 2, ... (the order grows
 * by one per accepted step).  Step control: the scaled RMS norm of the error
 * (abstol + reltol max(|y_n|, |y_{n+1}|)), accepted when <= 1, new step
 * h / clamp(err^(1/(k+1)) / 0.9, 1/2, 5) (growth <= 2, shrink >= 1/5 keeps the
 * variable-step formulas stable).  OrdinaryDiffEq's exact step sequence is
 * not vendored in the reference: parity of this method with it is unpinned;
 * the GPU stepper implements this same method.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define VK 5

/* Coefficients g[0..k] and beta[0..k-1] for a step h from t_n with past
 * steps hist[0] = t_n - t_{n-1}, hist[1] = t_{n-1} - t_{n-2}, ... */
void oq_vcabm_coeffs(int k, double h, const double *hist, double *g, double *beta)
{
    double psi_new[VK + 1], psi_old[VK + 1], c[VK + 2];
    int i, j, q;
    if (k < 1) k = 1;
    if (k > VK) k = VK;
    /* psi_i(n+1) = h + hist[0] + ... + hist[i-1];  psi_i(n) = hist[0] + ... + hist[i] */
    psi_new[0] = h;
    psi_old[0] = hist[0];
    for (i = 1; i <= k; ++i) {
        psi_new[i] = psi_new[i - 1] + hist[i - 1];
        psi_old[i] = psi_old[i - 1] + hist[i];
    }
    beta[0] = 1.0;
    for (j = 1; j < k; ++j) beta[j] = beta[j - 1] * psi_new[j - 1] / psi_old[j - 1];
    for (q = 1; q <= k + 1; ++q) c[q] = 1.0 / q;
    g[0] = c[1];
    for (j = 1; j <= k; ++j) {
        for (q = 1; q <= k + 1 - j; ++q) c[q] = c[q] - c[q + 1] * h / psi_new[j - 1];
        g[j] = c[1];
    }
}

void oq_vcabm5_record(oq_rhs_fn f, void *ctx, int64_t n, double *u, double t0, double t1,
                      const oq_ode_opts_t *o, oq_ode_stats_t *st, int64_t nrec_n,
                      double *step_t, double *step_v, int64_t cap, int64_t *nrec)
{
    double *phi = (double *)malloc(sizeof(double) * (size_t)((VK + 1) * n));
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    double *fp = (double *)malloc(sizeof(double) * (size_t)n);
    double *ynew = (double *)malloc(sizeof(double) * (size_t)n);
    double hist[VK + 1] = {0};
    double g[VK + 1], beta[VK + 1];
    double t = t0, h;
    int k = 1;
    int64_t iter = 0, i;
    int j;

    memset(st, 0, sizeof(*st));
    f(ctx, t, u, phi); /* Phi_0 = f(t0, u0) */
    st->nf = 1;
    if (o->dt0 > 0.0) {
        h = o->dt0;
    } else {
        /* the Hairer-Norsett-Wanner initial step (as Tsit5, order 5) */
        double d0 = 0.0, d1 = 0.0, d2 = 0.0, dt0, m, dt1;
        for (i = 0; i < n; ++i) {
            double sc = o->abstol + fabs(u[i]) * o->reltol;
            d0 += (u[i] / sc) * (u[i] / sc);
            d1 += (phi[i] / sc) * (phi[i] / sc);
        }
        d0 = sqrt(d0 / (double)n);
        d1 = sqrt(d1 / (double)n);
        dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
        dt0 = fmin(dt0, o->dtmax);
        for (i = 0; i < n; ++i) p[i] = u[i] + dt0 * phi[i];
        f(ctx, t + dt0, p, fp);
        st->nf += 1;
        for (i = 0; i < n; ++i) {
            double w = (fp[i] - phi[i]) / (o->abstol + fabs(u[i]) * o->reltol);
            d2 += w * w;
        }
        d2 = sqrt(d2 / (double)n) / dt0;
        m = fmax(d1, d2);
        dt1 = (m <= 1e-15) ? fmax(1e-6, dt0 * 1e-3) : pow(0.01 / m, 1.0 / 5.0);
        h = fmin(fmin(100.0 * dt0, dt1), o->dtmax);
    }
    while (t < t1) {
        double hc = h, eest, acc2 = 0.0, q;
        if (iter >= o->maxiters) { st->retcode = 1; break; }
        if (hc > o->dtmax) hc = o->dtmax;
        if (t + hc > t1) hc = t1 - t;
        if (!(hc >= o->dtmin) || !isfinite(hc)) { st->retcode = isfinite(hc) ? 2 : 3; break; }
        ++iter;
        oq_vcabm_coeffs(k, hc, hist, g, beta);
        for (i = 0; i < n; ++i) {
            double acc = 0.0;
            for (j = 0; j < k; ++j) acc += g[j] * beta[j] * phi[j * n + i];
            p[i] = u[i] + hc * acc;
        }
        f(ctx, t + hc, p, fp);
        st->nf += 1;
        for (i = 0; i < n; ++i) {
            double ph = fp[i], w;
            for (j = 0; j < k; ++j) ph -= beta[j] * phi[j * n + i];
            ynew[i] = p[i] + hc * g[k] * ph;
            w = hc * (g[k] - g[k - 1]) * ph / (o->abstol + fmax(fabs(u[i]), fabs(ynew[i])) * o->reltol);
            acc2 += w * w;
        }
        eest = sqrt(acc2 / (double)n);
        q = (eest == 0.0) ? 0.5 : fmax(0.5, fmin(5.0, pow(eest, 1.0 / (k + 1)) / 0.9));
        if (eest <= 1.0) {
            double tn = (t1 - (t + hc) <= 0.0) ? t1 : t + hc;
            /* second evaluation and the divided-difference update */
            f(ctx, tn, ynew, fp);
            st->nf += 1;
            for (i = 0; i < n; ++i) {
                double tmp = fp[i];
                for (j = 0; j < k; ++j) {
                    double old = phi[j * n + i];
                    phi[j * n + i] = tmp;
                    tmp -= beta[j] * old;
                }
                phi[k * n + i] = tmp;
                u[i] = ynew[i];
            }
            for (j = VK; j > 0; --j) hist[j] = hist[j - 1];
            hist[0] = hc;
            t = tn;
            if (k < VK) ++k;
            h = hc / q;
            if (nrec_n > 0 && st->naccept < cap && step_t && step_v) {
                double vm = 0.0;
                for (i = 0; i < nrec_n; ++i) vm = fmax(vm, fabs(u[i]));
                step_t[st->naccept] = t;
                step_v[st->naccept] = vm;
            }
            st->naccept += 1;
        } else {
            st->nreject += 1;
            if (!isfinite(eest)) { st->retcode = 3; break; }
            h = hc / q;
        }
    }
    st->t = t;
    st->dt = h;
    if (nrec) *nrec = st->naccept;
    free(phi);
    free(p);
    free(fp);
    free(ynew);
}

void oq_vcabm5(oq_rhs_fn f, void *ctx, int64_t n, double *u, double t0, double t1,
               const oq_ode_opts_t *o, oq_ode_stats_t *st)
{
    oq_vcabm5_record(f, ctx, n, u, t0, t1, o, st, 0, NULL, NULL, 0, NULL);
}
