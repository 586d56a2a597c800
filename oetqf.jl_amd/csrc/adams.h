// adams.h -- the coefficients of the variable-step Adams methods in
// divided-difference form (Hairer, Norsett & Wanner I, Sec. III.5), shared by
// the single-state and the ensemble VCABM5 steppers.
#pragma once
#include <hip/hip_runtime.h>

namespace oq {

static constexpr int kAbK = 5;  // the largest order

// g[0..k] and beta[0..k-1] of a step of size h after the past steps hist
// (most recent first)
static __device__ void ab_coeffs(int k, double h, const double *hist, double *g, double *beta)
{
    double pn[kAbK + 1], po[kAbK + 1], c[kAbK + 2];
    pn[0] = h;
    po[0] = hist[0];
    for (int i = 1; i <= k; ++i) {
        pn[i] = pn[i - 1] + hist[i - 1];
        po[i] = po[i - 1] + hist[i];
    }
    beta[0] = 1.0;
    for (int j = 1; j < k; ++j) beta[j] = beta[j - 1] * pn[j - 1] / po[j - 1];
    for (int q = 1; q <= k + 1; ++q) c[q] = 1.0 / q;
    g[0] = c[1];
    for (int j = 1; j <= k; ++j) {
        for (int q = 1; q <= k + 1 - j; ++q) c[q] = c[q] - c[q + 1] * h / pn[j - 1];
        g[j] = c[1];
    }
}

}  // namespace oq
