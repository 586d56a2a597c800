// stepper.hip -- on-device adaptive Tsit5 for the earthquake-cycle ODE.
//
// Replaces the OrdinaryDiffEq time loop the reference runs over `ode`
// (solve(prob, Tsit5(); ...), test/tests.jl:11, io.jl:125-130): Tsitouras'
// 5(4) pair with FSAL, the RMS norm of the embedded error scaled by
// abstol + reltol*max(|u_old|,|u_new|), and OrdinaryDiffEq's PI controller
// (beta1 = 7/50, beta2 = 2/25, qmin = 1/5, qmax = 10, gamma = 9/10,
// qoldinit = 1e-4).  All state -- u, the seven stages, the error sums, t,
// dt and the controller memory -- lives in device memory, so one step is a
// fixed launch sequence that is captured once into a HIP graph and replayed;
// the host only polls a small control block every few steps.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "adams.h"
#include "rhs_impl.h"

namespace oq {

// Tsitouras (2011) tableau, b = a[6][:] (FSAL), btilde = b - bhat.
__constant__ double kTsC[7] = {0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0};
#define OQ_TSIT5_A                                                                                   \
    {{0, 0, 0, 0, 0, 0},                                                                             \
     {0.161, 0, 0, 0, 0, 0},                                                                         \
     {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},                                         \
     {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},                           \
     {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},       \
     {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401,                   \
      -0.028269050394068383, 0},                                                                     \
     {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081,          \
      2.324710524099774}}
__constant__ double kTsA[7][6] = OQ_TSIT5_A;
static const double kTsAHost[7][6] = OQ_TSIT5_A;  // the same literals, for stage arguments
__constant__ double kTsBt[7] = {-0.00178001105222577714, -0.0008164344596567469,
                                0.007880878010261995, -0.1447110071732629,
                                0.5823571654525552, -0.45808210592918697,
                                0.015151515151515152};

struct Ctrl {
    double t, t1, dt, dtcur, qold, eest;
    double reltol, abstol, dtmax, dtmin;
    long long iter, maxiters, naccept, nreject;
    int done, retcode, accept, pad;
};

static constexpr int kErrBlocks = 256;

// Start of a step: termination checks, dt clamp to dtmax and t1.
__global__ void k_step_begin(Ctrl *c)
{
    c->accept = 0;
    if (c->done) return;
    if (!(c->t < c->t1)) {
        c->done = 1;
        c->retcode = OQ_RET_SUCCESS;
        return;
    }
    if (c->iter >= c->maxiters) {
        c->done = 1;
        c->retcode = OQ_RET_MAXITERS;
        return;
    }
    double dt = c->dt;
    if (dt > c->dtmax) dt = c->dtmax;
    if (c->t + dt > c->t1) dt = c->t1 - c->t;
    if (!(dt >= c->dtmin) || !isfinite(dt)) {
        c->done = 1;
        c->retcode = isfinite(dt) ? OQ_RET_DTMIN : OQ_RET_NONFINITE;
        return;
    }
    c->dtcur = dt;
    c->iter += 1;
}

// y = u + dt * sum_{j<s} a[s][j] k_j   (stage s = 1..6)
__global__ void __launch_bounds__(256) k_stage(const Ctrl *__restrict__ c, int s, int64_t n,
                                               const double *__restrict__ u,
                                               const double *__restrict__ k,
                                               double *__restrict__ y)
{
    if (c->done) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int j = 0; j < s; ++j) acc += kTsA[s][j] * k[j * n + i];
    y[i] = u[i] + c->dtcur * acc;
}

// Per-block sums of the squared scaled error (deterministic order).
__global__ void __launch_bounds__(256) k_err(const Ctrl *__restrict__ c, int64_t n,
                                             const double *__restrict__ u,
                                             const double *__restrict__ y,
                                             const double *__restrict__ k,
                                             double *__restrict__ part)
{
    __shared__ double red[256];
    if (c->done) return;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double e = 0.0;
#pragma unroll
        for (int j = 0; j < 7; ++j) e += kTsBt[j] * k[j * n + i];
        const double w = c->dtcur * e / (c->abstol + fmax(fabs(u[i]), fabs(y[i])) * c->reltol);
        acc += w * w;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// PI step-size control (OrdinaryDiffEq's PIController semantics).
__global__ void __launch_bounds__(256) k_controller(Ctrl *c, int64_t n, const double *part)
{
    __shared__ double red[256];
    if (c->done) return;
    red[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double beta1 = 7.0 / 50.0, beta2 = 2.0 / 25.0, qmin = 0.2, qmax = 10.0, gamma = 0.9,
                 qoldinit = 1e-4;
    const double eest = sqrt(red[0] / (double)n);
    double q11, q;
    c->eest = eest;
    if (eest == 0.0) {
        q11 = 0.0;
        q = 1.0 / qmax;
    } else {
        q11 = pow(eest, beta1);
        q = q11 / pow(c->qold, beta2);
        q = fmax(1.0 / qmax, fmin(1.0 / qmin, q / gamma));
    }
    if (eest <= 1.0) {
        c->t = (c->t1 - (c->t + c->dtcur) <= 0.0) ? c->t1 : c->t + c->dtcur;
        c->qold = fmax(eest, qoldinit);
        c->dt = c->dtcur / q;
        c->naccept += 1;
        c->accept = 1;
    } else {
        c->nreject += 1;
        if (!isfinite(eest)) {
            c->done = 1;
            c->retcode = OQ_RET_NONFINITE;
            return;
        }
        c->dt = c->dtcur / fmin(1.0 / qmin, q11 / gamma);
    }
}

// Accepted step: u <- y, k1 <- k7 (FSAL).
__global__ void __launch_bounds__(256) k_commit(const Ctrl *__restrict__ c, int64_t n,
                                                double *__restrict__ u,
                                                const double *__restrict__ y,
                                                double *__restrict__ k)
{
    if (!c->accept) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u[i] = y[i];
    k[i] = k[6 * n + i];
}

// Recorder (the device side of wsolve, io.jl:51-66): after an accepted
// step, the step counter advances, (t, max v) goes to a per-step ring, and
// every `stride`-th accepted step the state goes to a snapshot ring.
struct Recorder {
    long long *nstep, *nsnap;  // totals since the last drain
    double *step_t, *step_vmax;
    double *snap_t, *snap_u;
    double *snap_du;       // derivative ring (nullptr: not recorded)
    const double *du;      // f(u) of the accepted state (graph mode)
    long long step_cap, snap_cap, stride, stride_count;
    int64_t n, nf;
};

__global__ void __launch_bounds__(256) k_record(const Ctrl *__restrict__ c, Recorder r,
                                                const double *__restrict__ u)
{
    __shared__ double red[256];
    __shared__ int take;
    if (!c->accept) return;
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < r.nf; i += blockDim.x) m = fmax(m, fabs(u[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const long long k = *r.nstep;
        if (r.step_cap > 0) {
            r.step_t[k % r.step_cap] = c->t;
            r.step_vmax[k % r.step_cap] = red[0];
        }
        *r.nstep = k + 1;
        take = (r.snap_cap > 0) && (c->naccept % r.stride == 0);
    }
    __syncthreads();
    if (!take) return;
    const long long js = *r.nsnap % r.snap_cap;
    for (int64_t i = threadIdx.x; i < r.n; i += blockDim.x) r.snap_u[js * r.n + i] = u[i];
    if (r.snap_du)
        for (int64_t i = threadIdx.x; i < r.n; i += blockDim.x) r.snap_du[js * r.n + i] = r.du[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        r.snap_t[js] = c->t;
        *r.nsnap = *r.nsnap + 1;
    }
}

// ---------------- VCABM5: variable-coefficient Adams-Bashforth-Moulton ----------------
// The solver of the reference's example (VCABM5, examples/otf-with-mantle.jl:
// 160,167-168), restated from Hairer, Norsett & Wanner I, Sec. III.5 (variable
// step size Adams methods in divided-difference form), PECE:
//   predictor p = y_n + h sum_{j<k} g_j beta_j Phi_j, fp = f(p),
//   corrector y_{n+1} = p + h g_k Phi^p_k, Phi^p_k = fp - sum_{j<k} beta_j Phi_j,
//   error h (g_k - g_{k-1}) Phi^p_k, then f_{n+1} = f(y_{n+1}) and
//   Phi_0 = f_{n+1}, Phi_{j+1} = Phi_j(n+1) - beta_j Phi_j(n).
// Two RHS per step; k = 1 at the start, + 1 per accepted step up to 5; step
// factor h_new/h = 1/clamp(err^(1/(k+1))/0.9, 1/2, 5).
struct AbCtrl {
    Ctrl c;  // t, dt (next step), dtcur, eest, tolerances, counters, flags
    int k, kcur;
    double hist[kAbK + 1];  // past steps, most recent first
    double g[kAbK + 1], beta[kAbK + 1];
};

// ab_coeffs on the lanes of one wavefront: lane q (1 <= q <= k + 1) holds
// c_q through the j recursion (each update reads c_{q+1} before its own
// update, so the lanes of a j step are independent) while lane 0 runs the
// beta chain; the same operations in the same order per value, so the same
// coefficients, with five dependent division steps instead of twenty-five.
__device__ void ab_coeffs_wave(int k, double h, const double *hist, double *g, double *beta)
{
    const int lane = threadIdx.x & 63;
    double pn[kAbK + 1], po[kAbK + 1];
    pn[0] = h;
    po[0] = hist[0];
    for (int i = 1; i <= k; ++i) {
        pn[i] = pn[i - 1] + hist[i - 1];
        po[i] = po[i - 1] + hist[i];
    }
    double cq = (lane >= 1 && lane <= k + 1) ? 1.0 / lane : 0.0, b = 1.0;
    if (lane == 0) beta[0] = 1.0;
    if (lane == 1) g[0] = cq;
    for (int j = 1; j <= k; ++j) {
        const double cn = __shfl_down(cq, 1);  // c_{q+1} before this step's update
        if (lane >= 1 && lane <= k + 1 - j) cq = cq - cn * h / pn[j - 1];
        if (lane == 0 && j < k) {
            b = b * pn[j - 1] / po[j - 1];
            beta[j] = b;
        }
        if (lane == 1) g[j] = cq;
    }
}

__global__ void k_ab_begin(AbCtrl *a)
{
    Ctrl *c = &a->c;
    c->accept = 0;
    if (c->done) return;
    if (!(c->t < c->t1)) {
        c->done = 1;
        c->retcode = OQ_RET_SUCCESS;
        return;
    }
    if (c->iter >= c->maxiters) {
        c->done = 1;
        c->retcode = OQ_RET_MAXITERS;
        return;
    }
    double dt = c->dt;
    if (dt > c->dtmax) dt = c->dtmax;
    if (c->t + dt > c->t1) dt = c->t1 - c->t;
    if (!(dt >= c->dtmin) || !isfinite(dt)) {
        c->done = 1;
        c->retcode = isfinite(dt) ? OQ_RET_DTMIN : OQ_RET_NONFINITE;
        return;
    }
    c->dtcur = dt;
    c->iter += 1;
    a->kcur = a->k;
    ab_coeffs(a->k, dt, a->hist, a->g, a->beta);
}

__global__ void __launch_bounds__(256) k_ab_predict(const AbCtrl *__restrict__ a, int64_t n,
                                                    const double *__restrict__ u,
                                                    const double *__restrict__ phi,
                                                    double *__restrict__ p)
{
    if (a->c.done) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = a->kcur;
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc += a->g[j] * a->beta[j] * phi[j * n + i];
    p[i] = u[i] + a->c.dtcur * acc;
}

// corrector and per-block sums of the squared scaled error
__global__ void __launch_bounds__(256) k_ab_correct(const AbCtrl *__restrict__ a, int64_t n,
                                                    const double *__restrict__ u,
                                                    const double *__restrict__ p,
                                                    const double *__restrict__ fp,
                                                    const double *__restrict__ phi,
                                                    double *__restrict__ ynew,
                                                    double *__restrict__ part)
{
    __shared__ double red[256];
    if (a->c.done) return;
    const int k = a->kcur;
    const double h = a->c.dtcur, gk = a->g[k], dg = a->g[k] - a->g[k - 1];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double ph = fp[i];
        for (int j = 0; j < k; ++j) ph -= a->beta[j] * phi[j * n + i];
        const double yn = p[i] + h * gk * ph;
        ynew[i] = yn;
        const double w = h * dg * ph / (a->c.abstol + fmax(fabs(u[i]), fabs(yn)) * a->c.reltol);
        acc += w * w;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_ab_controller(AbCtrl *a, int64_t n, const double *part)
{
    __shared__ double red[256];
    Ctrl *c = &a->c;
    if (c->done) return;
    red[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double eest = sqrt(red[0] / (double)n);
    const int k = a->kcur;
    const double q = (eest == 0.0) ? 0.5 : fmax(0.5, fmin(5.0, pow(eest, 1.0 / (k + 1)) / 0.9));
    c->eest = eest;
    if (eest <= 1.0) {
        c->t = (c->t1 - (c->t + c->dtcur) <= 0.0) ? c->t1 : c->t + c->dtcur;
        for (int j = kAbK; j > 0; --j) a->hist[j] = a->hist[j - 1];
        a->hist[0] = c->dtcur;
        if (a->k < kAbK) a->k += 1;
        c->dt = c->dtcur / q;
        c->naccept += 1;
        c->accept = 1;
    } else {
        c->nreject += 1;
        if (!isfinite(eest)) {
            c->done = 1;
            c->retcode = OQ_RET_NONFINITE;
            return;
        }
        c->dt = c->dtcur / q;
    }
}

// accepted step: u <- y_{n+1}; divided differences of the new derivative
__global__ void __launch_bounds__(256) k_ab_commit(const AbCtrl *__restrict__ a, int64_t n,
                                                   double *__restrict__ u,
                                                   const double *__restrict__ ynew,
                                                   const double *__restrict__ fnew,
                                                   double *__restrict__ phi)
{
    if (!a->c.accept) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = a->kcur;
    double tmp = fnew[i];
    for (int j = 0; j < k; ++j) {
        const double old = phi[j * n + i];
        phi[j * n + i] = tmp;
        tmp -= a->beta[j] * old;
    }
    phi[k * n + i] = tmp;
    u[i] = ynew[i];
}

// ---------------- persistent stepper for small fault-only problems ----------------
// For a few hundred to a few thousand patches the graph above is bound by
// its ~23 kernel launches per step, not by arithmetic.  Here one cooperative
// launch runs up to max_steps attempted steps: workgroup w owns patches
// [w R, w R + R) -- their rows of K live in its LDS, and so do their entries
// of u, the stage state and the seven stages.  Each RHS needs the whole
// source vector s = v - vpl, published through device memory and one grid
// barrier (two buffers by stage parity, so a workgroup that is ahead never
// overwrites a vector another is still reading); the error norm and max|v|
// partials need one more.  The next step's first stage rides on that last
// barrier (both candidates for its source vector are published with the
// partials), so a step after the first of a launch costs six barriers.
// Every workgroup then runs the same controller on the same sums in the same
// order, so all copies of t, dt and the step counters stay identical without
// further communication.  Barrier waits
// are bounded by a wall-clock limit: a workgroup that times out raises the
// abort flag and every workgroup leaves at its next barrier.
struct PersistArgs {
    RhsDev d;
    const double *K;   // FP64 operator, N x ld
    Ctrl *ctrl;
    double *u, *k1;    // packed state and the FSAL stage
    double *s2;        // 2 x N source vectors, then (Tsit5) 4 x N stage-1 candidates
    double *part;      // 2 x nwg: squared-error sums, max |v|
    unsigned *bar;     // barrier arrivals (zero at launch)
    int *abort;
    Recorder rec;
    int recording, R, nwg;
    long long max_steps, wait_ticks;
};

// tau[r] = K row r . s for the owned rows, K rows in LDS at a pitch of
// ldp = ld + 2 doubles, s in LDS.  Sixteen lanes share two rows (r and
// r + 16 of a 32-row block): lane q sums the column pairs 32k + 2q,
// 32k + 2q + 1 of both rows with one 16-byte LDS read of s and one of each
// row per pair (the sixteen lanes of a row read 256 contiguous bytes), two
// accumulators per row, and four DPP steps sum the sixteen lanes.  Rows
// past Rv repeat row Rv - 1 and are not stored.
__device__ __forceinline__ void lds_rows(const double *Ks, const double *sv, int64_t ld, int64_t ldp, int Rv,
                                         int tid, double *tau)
{
    const int q = tid & 15, rp = tid >> 4;
    for (int r0 = 0; r0 < Rv; r0 += 32) {
        const int ra = r0 + rp, rb = ra + 16;
        const double *ka = Ks + (size_t)(ra < Rv ? ra : Rv - 1) * ldp + 2 * q;
        const double *kb = Ks + (size_t)(rb < Rv ? rb : Rv - 1) * ldp + 2 * q;
        const double *sp = sv + 2 * q;
        double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
        for (int64_t c = 0; c < ld; c += 128) {  // ld is a multiple of 512: four pairs per pass, loads first
            double2 s2[4], x2[4], y2[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                s2[u] = *reinterpret_cast<const double2 *>(sp + c + 32 * u);
                x2[u] = *reinterpret_cast<const double2 *>(ka + c + 32 * u);
                y2[u] = *reinterpret_cast<const double2 *>(kb + c + 32 * u);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a0 += x2[u].x * s2[u].x;
                a1 += x2[u].y * s2[u].y;
                b0 += y2[u].x * s2[u].x;
                b1 += y2[u].y * s2[u].y;
            }
        }
        double xa = a0 + a1, xb = b0 + b1;
        xa += dpp_f64<0xB1>(xa);   // quad_perm [1,0,3,2]
        xb += dpp_f64<0xB1>(xb);
        xa += dpp_f64<0x4E>(xa);   // quad_perm [2,3,0,1]
        xb += dpp_f64<0x4E>(xb);
        xa += dpp_f64<0x141>(xa);  // row_half_mirror: the two quads of each 8 lanes
        xb += dpp_f64<0x141>(xb);
        xa += dpp_f64<0x140>(xa);  // row_mirror: the two 8-lane halves of the 16
        xb += dpp_f64<0x140>(xb);
        if (q == 0 && ra < Rv) tau[ra] = xa;
        if (q == 0 && rb < Rv) tau[rb] = xb;
    }
}

template <class F>
__device__ __forceinline__ bool grid_barrier(const PersistArgs &a, unsigned &epoch, int *sh_bad, F side)
{
    __syncthreads();
    ++epoch;
    if (threadIdx.x == 0) {
        const unsigned target = epoch * (unsigned)a.nwg;
        __hip_atomic_fetch_add(a.bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const long long t0 = (long long)wall_clock64();
        int bad = 0;
        while (__hip_atomic_load(a.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(a.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                bad = 1;
                break;
            }
            if ((long long)wall_clock64() - t0 > a.wait_ticks) {
                __hip_atomic_store(a.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                bad = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *sh_bad = bad;
    } else if (threadIdx.x >= 64) {
        side();  // wavefronts 1..3, while thread 0 waits
    }
    __syncthreads();
    return *sh_bad == 0;
}

__device__ __forceinline__ bool grid_barrier(const PersistArgs &a, unsigned &epoch, int *sh_bad)
{
    return grid_barrier(a, epoch, sh_bad, [] {});
}

// Maximum over the 64 lanes of a wavefront (the steps of wave_sum with fmax).
__device__ __forceinline__ double wave_max(double x)
{
    x = fmax(x, dpp_f64<0xB1>(x));
    x = fmax(x, dpp_f64<0x4E>(x));
    x = fmax(x, dpp_f64<0x141>(x));
    x = fmax(x, dpp_f64<0x140>(x));
    x = fmax(x, swizzle_xor16_f64(x));
    return fmax(readlane_f64(x, 0), readlane_f64(x, 32));
}

__global__ void __launch_bounds__(256) k_persist(PersistArgs a)
{
    extern __shared__ __attribute__((aligned(16))) double lds[];  // 16-byte LDS reads of K and s
    __shared__ Ctrl c;
    __shared__ double red[256], red2[256], tau[64], pdv[64], pdt[64], pth[64];
    __shared__ int bad, take;
    __shared__ long long nstep, nsnap;
    const int R = a.R, tid = threadIdx.x;
    const int64_t N = a.d.N, ld = a.d.ld, n = a.d.lay.n;
    const int64_t p0 = (int64_t)blockIdx.x * R;
    const int Rv = (int)((N - p0) < R ? (N - p0) : R);  // owned patches (>= 1)
    const int E = 3 * R;                                 // owned entries: [v | th | del]
    const int64_t ldp = ld + 2;                          // K row pitch in LDS
    double *Ks = lds;                                    // R x ldp
    double *sv = Ks + (size_t)R * ldp;                   // ld
    double *uo = sv + ld, *yo = uo + E, *ko = yo + E;    // E, E, 7 x E
    const int64_t comp_off[3] = {a.d.lay.v, a.d.lay.th, a.d.lay.del};

    for (int64_t x = tid; x < (int64_t)Rv * ld; x += 256) Ks[(x / ld) * ldp + x % ld] = a.K[p0 * ld + x];
    for (int64_t j = tid; j < ld; j += 256) sv[j] = 0.0;
    for (int e = tid; e < E; e += 256) {
        const int cpt = e / R, i = e % R;
        uo[e] = i < Rv ? a.u[comp_off[cpt] + p0 + i] : 0.0;
        ko[e] = i < Rv ? a.k1[comp_off[cpt] + p0 + i] : 0.0;
    }
    if (tid == 0) {
        c = *a.ctrl;
        nstep = a.recording ? *a.rec.nstep : 0;
        nsnap = a.recording ? *a.rec.nsnap : 0;
    }
    // thread 64 + r (r < Rv <= 64) keeps patch p0 + r's friction parameters
    // in registers and evaluates its dtau-free friction terms (fault_pre)
    // while thread 0 waits at the grid barrier; after the LDS sweep only
    // fault_dv remains
    double fa = 1.0, fb = 0.0, fL = 1.0, fs = 0.0;
    const int pr = tid - 64;
    if (pr >= 0 && pr < Rv) {
        fa = a.d.a[p0 + pr];
        fb = a.d.b[p0 + pr];
        fL = a.d.L[p0 + pr];
        fs = a.d.sigma[p0 + pr];
    }
    auto pre = [&]() {
        if (pr >= 0 && pr < Rv) fault_pre(a.d, fa, fb, fL, fs, yo[pr], yo[R + pr], pdv[pr], pdt[pr], pth[pr]);
    };
    __syncthreads();
    unsigned epoch = 0;
    bool s1ready = false;  // stage 1's source vector crossed the last barrier
    int s1acc = 0;
    for (long long it = 0; it < a.max_steps; ++it) {
        if (tid == 0) {  // k_step_begin
            c.accept = 0;
            if (!c.done) {
                if (!(c.t < c.t1)) {
                    c.done = 1;
                    c.retcode = OQ_RET_SUCCESS;
                } else if (c.iter >= c.maxiters) {
                    c.done = 1;
                    c.retcode = OQ_RET_MAXITERS;
                } else {
                    double dt = c.dt;
                    if (dt > c.dtmax) dt = c.dtmax;
                    if (c.t + dt > c.t1) dt = c.t1 - c.t;
                    if (!(dt >= c.dtmin) || !isfinite(dt)) {
                        c.done = 1;
                        c.retcode = isfinite(dt) ? OQ_RET_DTMIN : OQ_RET_NONFINITE;
                    } else {
                        c.dtcur = dt;
                        c.iter += 1;
                    }
                }
            }
        }
        __syncthreads();
        if (c.done) break;
        const double dt = c.dtcur;
        for (int st = 1; st < 7; ++st) {
            for (int e = tid; e < E; e += 256) {
                double acc = 0.0;
                for (int j = 0; j < st; ++j) acc += kTsA[st][j] * ko[j * E + e];
                yo[e] = uo[e] + dt * acc;
            }
            __syncthreads();
            if (st == 1 && s1ready) {
                // the previous step published both candidates for this stage's
                // source vector with its error partials: s = (U + dt B) - vpl
                // with the same operations as yo above
                const double *U = a.s2 + (s1acc ? 2 : 4) * N, *B = U + N;
                for (int64_t j = tid; j < N; j += 256) sv[j] = (U[j] + dt * B[j]) - a.d.vpl;
                pre();
            } else {
                double *sg = a.s2 + (st & 1) * N;
                for (int i = tid; i < Rv; i += 256) sg[p0 + i] = yo[i] - a.d.vpl;
                if (!grid_barrier(a, epoch, &bad, pre)) return;
                for (int64_t j = tid; j < N; j += 256) sv[j] = sg[j];
            }
            __syncthreads();
            lds_rows(Ks, sv, ld, ldp, Rv, tid, tau);
            __syncthreads();
            if (tid < Rv) {  // one patch per thread
                const int r = tid;
                ko[st * E + r] = fault_dv(a.d, tau[r], pdv[r], pdt[r], pth[r]);
                ko[st * E + R + r] = pth[r];
                ko[st * E + 2 * R + r] = yo[r];
            }
            __syncthreads();
        }
        // squared scaled error and max |v| of the owned entries
        double acc = 0.0, vm = 0.0;
        for (int e = tid; e < E; e += 256) {
            if (e % R >= Rv) continue;
            double er = 0.0;
#pragma unroll
            for (int j = 0; j < 7; ++j) er += kTsBt[j] * ko[j * E + e];
            const double w = dt * er / (c.abstol + fmax(fabs(uo[e]), fabs(yo[e])) * c.reltol);
            acc += w * w;
            if (e < R) vm = fmax(vm, fabs(yo[e]));
        }
        // wavefront sums and maxima, then the four wavefronts' in fixed order
        acc = wave_sum(acc);
        vm = wave_max(vm);
        if ((tid & 63) == 0) {
            red[tid >> 6] = acc;
            red2[tid >> 6] = vm;
        }
        __syncthreads();
        if (tid == 0) {
            a.part[2 * blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
            a.part[2 * blockIdx.x + 1] = fmax(fmax(red2[0], red2[1]), fmax(red2[2], red2[3]));
        }
        // Stage 1 of the next attempt needs y = u + dt a21 k1 with (u, k1) =
        // (y_{n+1}, k7) after an acceptance or (u_n, k1) after a rejection,
        // and dt from the controller.  Both (u, a21 k1) candidates of the
        // owned v entries cross this barrier with the partials, so the next
        // step's first stage needs no barrier of its own.
        {
            double *C = a.s2 + 2 * N;  // [U_acc | B_acc | U_rej | B_rej]
            for (int i = tid; i < Rv; i += 256) {
                double b6 = 0.0, b1 = 0.0;
                b6 += kTsA[1][0] * ko[6 * E + i];
                b1 += kTsA[1][0] * ko[i];
                C[p0 + i] = yo[i];
                C[N + p0 + i] = b6;
                C[2 * N + p0 + i] = uo[i];
                C[3 * N + p0 + i] = b1;
            }
        }
        if (!grid_barrier(a, epoch, &bad)) return;
        for (int w = tid; w < 2 * a.nwg; w += 256) red[w] = a.part[w];  // all loads in flight at once
        __syncthreads();
        if (tid < 64) {  // k_controller, on the same sums in every workgroup
            double sum = 0.0, vmax = 0.0;
            for (int w = 0; w < a.nwg; ++w) {
                sum += red[2 * w];
                vmax = fmax(vmax, red[2 * w + 1]);
            }
            const double beta1 = 7.0 / 50.0, beta2 = 2.0 / 25.0, qmin = 0.2, qmax = 10.0,
                         gamma = 0.9, qoldinit = 1e-4;
            const double eest = sqrt(sum / (double)n);
            // eest^beta1 on lane 0 and qold^beta2 on lane 1, side by side
            const double pw = pow(tid == 0 ? eest : c.qold, tid == 0 ? beta1 : beta2);
            const double pw0 = readlane_f64(pw, 0), pw1 = readlane_f64(pw, 1);
            if (tid == 0) {
                double q11, q;
                c.eest = eest;
                if (eest == 0.0) {
                    q11 = 0.0;
                    q = 1.0 / qmax;
                } else {
                    q11 = pw0;
                    q = q11 / pw1;
                    q = fmax(1.0 / qmax, fmin(1.0 / qmin, q / gamma));
                }
                take = 0;
                if (eest <= 1.0) {
                    c.t = (c.t1 - (c.t + c.dtcur) <= 0.0) ? c.t1 : c.t + c.dtcur;
                    c.qold = fmax(eest, qoldinit);
                    c.dt = c.dtcur / q;
                    c.naccept += 1;
                    c.accept = 1;
                    if (a.recording) {  // k_record
                        if (blockIdx.x == 0 && a.rec.step_cap > 0) {
                            a.rec.step_t[nstep % a.rec.step_cap] = c.t;
                            a.rec.step_vmax[nstep % a.rec.step_cap] = vmax;
                        }
                        nstep += 1;
                        take = (a.rec.snap_cap > 0) && (c.naccept % a.rec.stride == 0);
                    }
                } else {
                    c.nreject += 1;
                    if (!isfinite(eest)) {
                        c.done = 1;
                        c.retcode = OQ_RET_NONFINITE;
                    } else {
                        c.dt = c.dtcur / fmin(1.0 / qmin, q11 / gamma);
                    }
                }
            }
        }
        __syncthreads();
        s1acc = c.accept;  // which candidate the next stage 1 takes
        s1ready = true;
        if (c.accept) {  // k_commit (FSAL: k1 <- k7)
            for (int e = tid; e < E; e += 256) {
                uo[e] = yo[e];
                ko[e] = ko[6 * E + e];
            }
            if (take) {
                const long long js = nsnap % a.rec.snap_cap;
                for (int e = tid; e < E; e += 256)
                    if (e % R < Rv) {
                        const int64_t at = js * n + comp_off[e / R] + p0 + e % R;
                        a.rec.snap_u[at] = yo[e];
                        if (a.rec.snap_du) a.rec.snap_du[at] = ko[e];  // k1 = f(u) after FSAL
                    }
                if (blockIdx.x == 0 && tid == 0) a.rec.snap_t[js] = c.t;
            }
            __syncthreads();
            if (tid == 0 && take) nsnap += 1;
        }
        __syncthreads();
    }
    // hand the state back for the next launch (or the graph stepper)
    for (int e = tid; e < E; e += 256) {
        const int cpt = e / R, i = e % R;
        if (i < Rv) {
            a.u[comp_off[cpt] + p0 + i] = uo[e];
            a.k1[comp_off[cpt] + p0 + i] = ko[e];
        }
    }
    if (blockIdx.x == 0 && tid == 0) {
        *a.ctrl = c;
        if (a.recording) {
            *a.rec.nstep = nstep;
            *a.rec.nsnap = nsnap;
        }
    }
}

// Persistent VCABM5: as k_persist, with the method of enqueue_step_ab.  A
// step needs the predicted source vector (one barrier), then the error
// partials and the corrected source vector, published together before one
// more barrier (the corrected vector is used only when the step is
// accepted): two grid barriers instead of Tsit5's seven.  Buffer 0 (predicted
// s), buffer 1 (corrected s) and the partials are each rewritten only after
// a barrier that every workgroup crosses once it has read them.
// Phi_0..Phi_5 of the owned entries stay in LDS across the steps of a launch.
__global__ void __launch_bounds__(256) k_persist_ab(PersistArgs a, AbCtrl *actrl, double *phig)
{
    extern __shared__ __attribute__((aligned(16))) double lds[];  // 16-byte LDS reads of K and s
    __shared__ AbCtrl ac;
    __shared__ double red[256], red2[256], tau[64], pdv[64], pdt[64], pth[64];
    __shared__ int bad, take;
    __shared__ long long nstep, nsnap;
    const int R = a.R, tid = threadIdx.x;
    const int64_t N = a.d.N, ld = a.d.ld, n = a.d.lay.n;
    const int64_t p0 = (int64_t)blockIdx.x * R;
    const int Rv = (int)((N - p0) < R ? (N - p0) : R);
    const int E = 3 * R;
    const int64_t ldp = ld + 2;  // K row pitch in LDS
    double *Ks = lds, *sv = Ks + (size_t)R * ldp;
    double *uo = sv + ld, *po = uo + E, *fpo = po + E, *yo = fpo + E, *fo = yo + E, *ph = fo + E;  // ph: 6E
    const int64_t comp_off[3] = {a.d.lay.v, a.d.lay.th, a.d.lay.del};

    for (int64_t x = tid; x < (int64_t)Rv * ld; x += 256) Ks[(x / ld) * ldp + x % ld] = a.K[p0 * ld + x];
    for (int64_t j = tid; j < ld; j += 256) sv[j] = 0.0;
    for (int e = tid; e < E; e += 256) {
        const int cpt = e / R, i = e % R;
        const int64_t gi = comp_off[cpt] + p0 + i;
        uo[e] = i < Rv ? a.u[gi] : 0.0;
        po[e] = fpo[e] = yo[e] = fo[e] = 0.0;  // padding entries stay zero
        for (int j = 0; j <= kAbK; ++j) ph[j * E + e] = i < Rv ? phig[j * n + gi] : 0.0;
    }
    if (tid == 0) {
        ac = *actrl;
        nstep = a.recording ? *a.rec.nstep : 0;
        nsnap = a.recording ? *a.rec.nsnap : 0;
    }
    double fa = 1.0, fb = 0.0, fL = 1.0, fs = 0.0;  // patch p0 + tid - 64's parameters (as in k_persist)
    const int pr = tid - 64;
    if (pr >= 0 && pr < Rv) {
        fa = a.d.a[p0 + pr];
        fb = a.d.b[p0 + pr];
        fL = a.d.L[p0 + pr];
        fs = a.d.sigma[p0 + pr];
    }
    __syncthreads();
    unsigned epoch = 0;
    // s <- x[v] - vpl (owned), barrier, s -> LDS, owned rows of K s, rates -> out
    // dtau-free friction terms of state x, on wavefronts 1..3 during a barrier
    auto pre_of = [&](const double *x) {
        return [=]() {
            if (pr >= 0 && pr < Rv) fault_pre(a.d, fa, fb, fL, fs, x[pr], x[R + pr], pdv[pr], pdt[pr], pth[pr]);
        };
    };
    // s <- x[v] - vpl for the owned patches, into source buffer buf
    auto publish = [&](const double *x, int buf) {
        double *sg = a.s2 + buf * N;
        for (int i = tid; i < Rv; i += 256) sg[p0 + i] = x[i] - a.d.vpl;
    };
    // after a barrier that follows every workgroup's publish: s -> LDS, the
    // owned rows of K s, rates -> out
    auto consume = [&](const double *x, double *out, int buf) {
        const double *sg = a.s2 + buf * N;
        for (int64_t j = tid; j < N; j += 256) sv[j] = sg[j];
        __syncthreads();
        lds_rows(Ks, sv, ld, ldp, Rv, tid, tau);
        __syncthreads();
        if (tid < Rv) {  // the dtau-free terms of x came with the preceding barrier
            const int r = tid;
            out[r] = fault_dv(a.d, tau[r], pdv[r], pdt[r], pth[r]);
            out[R + r] = pth[r];
            out[2 * R + r] = x[r];
        }
        __syncthreads();
    };
    for (long long it = 0; it < a.max_steps; ++it) {
        if (tid == 0) {  // k_ab_begin
            Ctrl &c = ac.c;
            c.accept = 0;
            if (!c.done) {
                if (!(c.t < c.t1)) {
                    c.done = 1;
                    c.retcode = OQ_RET_SUCCESS;
                } else if (c.iter >= c.maxiters) {
                    c.done = 1;
                    c.retcode = OQ_RET_MAXITERS;
                } else {
                    double dt = c.dt;
                    if (dt > c.dtmax) dt = c.dtmax;
                    if (c.t + dt > c.t1) dt = c.t1 - c.t;
                    if (!(dt >= c.dtmin) || !isfinite(dt)) {
                        c.done = 1;
                        c.retcode = isfinite(dt) ? OQ_RET_DTMIN : OQ_RET_NONFINITE;
                    } else {
                        c.dtcur = dt;
                        c.iter += 1;
                        ac.kcur = ac.k;
                    }
                }
            }
        }
        __syncthreads();
        if (ac.c.done) break;
        if (tid < 64) ab_coeffs_wave(ac.kcur, ac.c.dtcur, ac.hist, ac.g, ac.beta);
        __syncthreads();
        const int k = ac.kcur;
        const double h = ac.c.dtcur;
        for (int e = tid; e < E; e += 256) {  // predictor
            double acc = 0.0;
            for (int j = 0; j < k; ++j) acc += ac.g[j] * ac.beta[j] * ph[j * E + e];
            po[e] = uo[e] + h * acc;
        }
        __syncthreads();
        publish(po, 0);
        if (!grid_barrier(a, epoch, &bad, pre_of(po))) return;
        consume(po, fpo, 0);
        double acc2 = 0.0, vm = 0.0;  // corrector, error and max |v| partials
        for (int e = tid; e < E; e += 256) {
            if (e % R >= Rv) continue;  // padding: yo stays zero
            double phk = fpo[e];
            for (int j = 0; j < k; ++j) phk -= ac.beta[j] * ph[j * E + e];
            const double yn = po[e] + h * ac.g[k] * phk;
            yo[e] = yn;
            const double w = h * (ac.g[k] - ac.g[k - 1]) * phk /
                             (ac.c.abstol + fmax(fabs(uo[e]), fabs(yn)) * ac.c.reltol);
            acc2 += w * w;
            if (e < R) vm = fmax(vm, fabs(yn));
        }
        acc2 = wave_sum(acc2);  // wavefront sums and maxima, as in k_persist
        vm = wave_max(vm);
        if ((tid & 63) == 0) {
            red[tid >> 6] = acc2;
            red2[tid >> 6] = vm;
        }
        __syncthreads();
        if (tid == 0) {
            a.part[2 * blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
            a.part[2 * blockIdx.x + 1] = fmax(fmax(red2[0], red2[1]), fmax(red2[2], red2[3]));
        }
        // the corrected source vector crosses the same barrier as the error
        // partials: an accepted step needs no third barrier (a rejected one
        // leaves buffer 1 unread)
        publish(yo, 1);
        if (!grid_barrier(a, epoch, &bad, pre_of(yo))) return;
        for (int w = tid; w < 2 * a.nwg; w += 256) red[w] = a.part[w];
        __syncthreads();
        if (tid == 0) {  // k_ab_controller, on the same sums in every workgroup
            Ctrl &c = ac.c;
            double sum = 0.0, vmax = 0.0;
            for (int w = 0; w < a.nwg; ++w) {
                sum += red[2 * w];
                vmax = fmax(vmax, red[2 * w + 1]);
            }
            const double eest = sqrt(sum / (double)n);
            const double q = (eest == 0.0) ? 0.5 : fmax(0.5, fmin(5.0, pow(eest, 1.0 / (k + 1)) / 0.9));
            c.eest = eest;
            take = 0;
            if (eest <= 1.0) {
                c.t = (c.t1 - (c.t + c.dtcur) <= 0.0) ? c.t1 : c.t + c.dtcur;
                for (int j = kAbK; j > 0; --j) ac.hist[j] = ac.hist[j - 1];
                ac.hist[0] = c.dtcur;
                if (ac.k < kAbK) ac.k += 1;
                c.dt = c.dtcur / q;
                c.naccept += 1;
                c.accept = 1;
                if (a.recording) {
                    if (blockIdx.x == 0 && a.rec.step_cap > 0) {
                        a.rec.step_t[nstep % a.rec.step_cap] = c.t;
                        a.rec.step_vmax[nstep % a.rec.step_cap] = vmax;
                    }
                    nstep += 1;
                    take = (a.rec.snap_cap > 0) && (c.naccept % a.rec.stride == 0);
                }
            } else {
                c.nreject += 1;
                if (!isfinite(eest)) {
                    c.done = 1;
                    c.retcode = OQ_RET_NONFINITE;
                } else {
                    c.dt = c.dtcur / q;
                }
            }
        }
        __syncthreads();
        if (ac.c.accept) {
            consume(yo, fo, 1);
            for (int e = tid; e < E; e += 256) {  // k_ab_commit
                double tmp = fo[e];
                for (int j = 0; j < k; ++j) {
                    const double old = ph[j * E + e];
                    ph[j * E + e] = tmp;
                    tmp -= ac.beta[j] * old;
                }
                ph[k * E + e] = tmp;
                uo[e] = yo[e];
            }
            if (take) {
                const long long js = nsnap % a.rec.snap_cap;
                for (int e = tid; e < E; e += 256)
                    if (e % R < Rv) {
                        const int64_t at = js * n + comp_off[e / R] + p0 + e % R;
                        a.rec.snap_u[at] = yo[e];
                        if (a.rec.snap_du) a.rec.snap_du[at] = fo[e];  // f(y_{n+1})
                    }
                if (blockIdx.x == 0 && tid == 0) a.rec.snap_t[js] = ac.c.t;
            }
            __syncthreads();
            if (tid == 0 && take) nsnap += 1;
        }
        __syncthreads();
    }
    for (int e = tid; e < E; e += 256) {
        const int cpt = e / R, i = e % R;
        if (i < Rv) {
            const int64_t gi = comp_off[cpt] + p0 + i;
            a.u[gi] = uo[e];
            for (int j = 0; j <= kAbK; ++j) phig[j * n + gi] = ph[j * E + e];
        }
    }
    if (blockIdx.x == 0 && tid == 0) {
        *actrl = ac;  // a.ctrl points at actrl->c
        if (a.recording) {
            *a.rec.nstep = nstep;
            *a.rec.nsnap = nsnap;
        }
    }
}

}  // namespace oq

struct oq_stepper {
    oq_rhs *h = nullptr;
    int64_t n = 0;
    oq::DevBuf u, y, k, part, ctrl;
    oq::Ctrl *ctrl_host = nullptr;  // pinned
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // recorder (disabled until oq_stepper_record)
    oq::DevBuf rcnt, rstep_t, rstep_v, rsnap_t, rsnap_u, rsnap_du;
    oq::Recorder rec{};
    bool recording = false;
    // algorithm: 0 Tsit5, 1 VCABM5 (oq_stepper_create_vcabm5)
    int alg = 0;
    oq::DevBuf ab;  // VCABM5: y_{n+1} and f(y_{n+1}), 2n
    // persistent mode (oq_stepper_set_persistent)
    bool persistent = false;
    int p_R = 0, p_nwg = 0;
    size_t p_lds = 0;
    long long p_wait = 0;
    oq::DevBuf p_s2, p_part, p_bar, p_abort;
};

namespace oq {

static inline unsigned nb(int64_t n) { return (unsigned)((n + 255) / 256); }

// One attempted VCABM5 step: two RHS evaluations (the second also runs on a
// rejected step; its result is not used then).
static int enqueue_step_ab(oq_stepper *st)
{
    oq_rhs *h = st->h;
    const int64_t n = st->n;
    AbCtrl *a = st->ctrl.as<AbCtrl>();
    double *u = st->u.as<double>(), *p = st->y.as<double>(), *phi = st->k.as<double>();
    double *fp = phi + 6 * n, *ynew = st->ab.as<double>(), *fnew = ynew + n;
    int rc;
    hipLaunchKernelGGL(k_ab_begin, dim3(1), dim3(1), 0, h->stream, a);
    hipLaunchKernelGGL(k_ab_predict, dim3(nb(n)), dim3(256), 0, h->stream, a, n, u, phi, p);
    if ((rc = rhs_enqueue(h, p, fp))) return rc;
    hipLaunchKernelGGL(k_ab_correct, dim3(kErrBlocks), dim3(256), 0, h->stream, a, n, u, p, fp, phi,
                       ynew, st->part.as<double>());
    hipLaunchKernelGGL(k_ab_controller, dim3(1), dim3(256), 0, h->stream, a, n, st->part.as<double>());
    if ((rc = rhs_enqueue(h, ynew, fnew))) return rc;
    hipLaunchKernelGGL(k_ab_commit, dim3(nb(n)), dim3(256), 0, h->stream, a, n, u, ynew, fnew, phi);
    if (st->recording)
        hipLaunchKernelGGL(k_record, dim3(1), dim3(256), 0, h->stream, &a->c, st->rec, u);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

// The launch sequence of one attempted step (captured into a graph).
static int enqueue_step(oq_stepper *st)
{
    if (st->alg == 1) return enqueue_step_ab(st);
    oq_rhs *h = st->h;
    const int64_t n = st->n;
    Ctrl *c = st->ctrl.as<Ctrl>();
    double *u = st->u.as<double>(), *y = st->y.as<double>(), *k = st->k.as<double>();
    int rc;
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, h->stream, c);
    // the DFT operator (fault only, no dilatancy) forms each stage inside
    // its forward-FFT kernel: one launch less per stage
    const bool fused = h->kdtype == OQ_K_FFT && !h->dev.dila && h->ne == 0;
    for (int s = 1; s < 7; ++s) {
        if (fused) {
            StageCoef sc{};
            sc.s = s;
            for (int j = 0; j < 6; ++j) sc.a[j] = kTsAHost[s][j];
            if ((rc = rhs_enqueue_fft_stage(h, sc, &c->dtcur, &c->done, u, k, n, y, k + s * n))) return rc;
            continue;
        }
        hipLaunchKernelGGL(k_stage, dim3(nb(n)), dim3(256), 0, h->stream, c, s, n, u, k, y);
        // the RHS kernels have no done-check of their own: evaluating one
        // more RHS on a finished state is harmless and keeps them graph-safe
        if ((rc = rhs_enqueue(h, y, k + s * n))) return rc;
    }
    hipLaunchKernelGGL(k_err, dim3(kErrBlocks), dim3(256), 0, h->stream, c, n, u, y, k,
                       st->part.as<double>());
    hipLaunchKernelGGL(k_controller, dim3(1), dim3(256), 0, h->stream, c, n, st->part.as<double>());
    hipLaunchKernelGGL(k_commit, dim3(nb(n)), dim3(256), 0, h->stream, c, n, u, y, k);
    if (st->recording)
        hipLaunchKernelGGL(k_record, dim3(1), dim3(256), 0, h->stream, c, st->rec, u);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

static int capture_step(oq_stepper *st)
{
    oq_rhs *h = st->h;
    if (st->exec) { (void)hipGraphExecDestroy(st->exec); st->exec = nullptr; }
    if (st->graph) { (void)hipGraphDestroy(st->graph); st->graph = nullptr; }
    OQ_HIP(hipStreamSynchronize(h->stream));
    OQ_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(st);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(h->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    OQ_HIP(ce);
    st->graph = g;
    OQ_HIP(hipGraphInstantiate(&st->exec, g, nullptr, nullptr, 0));
    return OQ_OK;
}

static double rms_scaled(const std::vector<double> &x, const std::vector<double> &u, double at,
                         double rt)
{
    double acc = 0.0;
    for (size_t i = 0; i < x.size(); ++i) {
        const double w = x[i] / (at + fabs(u[i]) * rt);
        acc += w * w;
    }
    return sqrt(acc / (double)x.size());
}

static void destroy_stepper(oq_stepper *s)
{
    if (!s) return;
    if (teardown_sync(s->h ? s->h->stream : nullptr)) {
        if (s->exec) (void)hipGraphExecDestroy(s->exec);
        if (s->graph) (void)hipGraphDestroy(s->graph);
        if (s->ctrl_host) (void)hipHostFree(s->ctrl_host);
    }
    delete s;
}

}  // namespace oq

using namespace oq;

extern "C" {

static int stepper_create(oq_rhs *h, const double *u0, double t0, double t1,
                          const oq_solver_opts *o, int alg, oq_stepper **out)
{
    OQ_REQUIRE(h && u0 && o && out, "NULL argument");
    OQ_REQUIRE(t1 > t0, "tspan must satisfy t1 > t0");
    OQ_REQUIRE(o->reltol > 0.0 && o->abstol >= 0.0, "tolerances must be positive");
    *out = nullptr;
    OQ_HIP(hipSetDevice(h->device));
    oq_stepper *st = new oq_stepper();
    st->h = h;
    st->n = h->lay.n;
    st->alg = alg;
    const int64_t n = st->n;
    int rc = st->u.alloc((size_t)n * 8);
    if (!rc) rc = st->y.alloc((size_t)n * 8);
    if (!rc) rc = st->k.alloc((size_t)(7 * n) * 8);
    if (!rc) rc = st->part.alloc(kErrBlocks * 8);
    if (!rc) rc = st->ctrl.alloc(alg == 1 ? sizeof(AbCtrl) : sizeof(Ctrl));
    if (!rc && alg == 1) rc = st->ab.alloc((size_t)(2 * n) * 8);
    if (!rc && hipHostMalloc((void **)&st->ctrl_host, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess) {
        set_error("hipHostMalloc failed");
        rc = OQ_ERR_NOMEM;
    }
    if (rc) {
        destroy_stepper(st);
        return rc;
    }
    // f0 and the initial step (Hairer-Norsett-Wanner II.4, order 5), as
    // OrdinaryDiffEq's ode_determine_initdt, evaluated with device RHS calls.
    const double dtmax = o->dtmax > 0.0 ? o->dtmax : INFINITY;
    std::vector<double> hu(u0, u0 + n), f0(n), u1(n), f1(n);
    double *du = st->k.as<double>();
    if (hipMemcpyAsync(st->u.p, u0, (size_t)n * 8, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
        rhs_enqueue(h, st->u.as<double>(), du) != OQ_OK ||
        hipMemcpyAsync(f0.data(), du, (size_t)n * 8, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
        set_error("initial RHS evaluation failed");
        destroy_stepper(st);
        return OQ_ERR_HIP;
    }
    double dt;
    if (o->dt0 > 0.0) {
        dt = o->dt0;
    } else {
        const double d0 = rms_scaled(hu, hu, o->abstol, o->reltol);
        const double d1 = rms_scaled(f0, hu, o->abstol, o->reltol);
        double dt0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
        dt0 = fmin(dt0, dtmax);
        for (int64_t i = 0; i < n; ++i) u1[i] = hu[i] + dt0 * f0[i];
        double *yd = st->y.as<double>(), *fd = du + n;
        if (hipMemcpyAsync(yd, u1.data(), (size_t)n * 8, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
            rhs_enqueue(h, yd, fd) != OQ_OK ||
            hipMemcpyAsync(f1.data(), fd, (size_t)n * 8, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            set_error("initial step-size RHS evaluation failed");
            destroy_stepper(st);
            return OQ_ERR_HIP;
        }
        double acc = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            const double w = (f1[i] - f0[i]) / (o->abstol + fabs(hu[i]) * o->reltol);
            acc += w * w;
        }
        const double d2 = sqrt(acc / (double)n) / dt0;
        const double m = fmax(d1, d2);
        const double dt1 = (m <= 1e-15) ? fmax(1e-6, dt0 * 1e-3) : pow(0.01 / m, 1.0 / 5.0);
        dt = fmin(fmin(100.0 * dt0, dt1), dtmax);
    }
    Ctrl c;
    memset(&c, 0, sizeof(c));
    c.t = t0;
    c.t1 = t1;
    c.dt = dt;
    c.qold = 1e-4;
    c.reltol = o->reltol;
    c.abstol = o->abstol;
    c.dtmax = dtmax;
    c.dtmin = o->dtmin;
    c.maxiters = o->maxiters > 0 ? o->maxiters : (long long)1e18;
    *st->ctrl_host = c;
    AbCtrl ab;
    memset(&ab, 0, sizeof(ab));
    ab.c = c;
    ab.k = 1;
    const void *cblk = alg == 1 ? (const void *)&ab : (const void *)st->ctrl_host;
    if (hipMemcpyAsync(st->ctrl.p, cblk, alg == 1 ? sizeof(AbCtrl) : sizeof(Ctrl), hipMemcpyHostToDevice,
                       h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
        set_error("control block upload failed");
        destroy_stepper(st);
        return OQ_ERR_HIP;
    }
    // Capture one attempted step into a graph.
    if ((rc = capture_step(st))) {
        destroy_stepper(st);
        return rc;
    }
    *out = st;
    return OQ_OK;
}

int oq_stepper_create(oq_rhs *h, const double *u0, double t0, double t1,
                      const oq_solver_opts *o, oq_stepper **out)
{
    return stepper_create(h, u0, t0, t1, o, 0, out);
}

int oq_stepper_create_vcabm5(oq_rhs *h, const double *u0, double t0, double t1,
                             const oq_solver_opts *o, oq_stepper **out)
{
    return stepper_create(h, u0, t0, t1, o, 1, out);
}

int oq_stepper_run(oq_stepper *st, int64_t max_steps, int *retcode)
{
    OQ_REQUIRE(st != nullptr && max_steps > 0, "invalid argument");
    oq_rhs *h = st->h;
    OQ_HIP(hipSetDevice(h->device));
    if (st->persistent) {
        PersistArgs a{};
        a.d = h->dev;
        a.K = h->K.as<double>();
        a.ctrl = st->ctrl.as<Ctrl>();
        a.u = st->u.as<double>();
        a.k1 = st->k.as<double>();
        a.s2 = st->p_s2.as<double>();
        a.part = st->p_part.as<double>();
        a.bar = st->p_bar.as<unsigned>();
        a.abort = st->p_abort.as<int>();
        a.rec = st->rec;
        a.recording = st->recording ? 1 : 0;
        a.R = st->p_R;
        a.nwg = st->p_nwg;
        a.max_steps = max_steps;
        a.wait_ticks = st->p_wait;
        OQ_HIP(hipMemsetAsync(st->p_bar.p, 0, sizeof(unsigned), h->stream));
        if (st->alg == 1) {
            AbCtrl *actrl = st->ctrl.as<AbCtrl>();
            double *phig = st->k.as<double>();
            void *args[] = {&a, &actrl, &phig};
            OQ_HIP(hipLaunchCooperativeKernel((const void *)k_persist_ab, dim3(st->p_nwg), dim3(256),
                                              args, st->p_lds, h->stream));
        } else {
            void *args[] = {&a};
            OQ_HIP(hipLaunchCooperativeKernel((const void *)k_persist, dim3(st->p_nwg), dim3(256),
                                              args, st->p_lds, h->stream));
        }
        int ab = 0;
        OQ_HIP(hipMemcpyAsync(&ab, st->p_abort.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        OQ_HIP(hipMemcpyAsync(st->ctrl_host, st->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost,
                              h->stream));
        OQ_HIP(hipStreamSynchronize(h->stream));
        if (ab) {
            st->persistent = false;
            set_error("persistent stepper: a grid barrier timed out; the state is not valid");
            return OQ_ERR_STATE;
        }
        if (retcode) *retcode = st->ctrl_host->done ? st->ctrl_host->retcode : OQ_RET_RUNNING;
        return OQ_OK;
    }
    const int64_t chunk = 16;
    int64_t done_steps = 0;
    while (done_steps < max_steps) {
        const int64_t m = (max_steps - done_steps) < chunk ? (max_steps - done_steps) : chunk;
        for (int64_t i = 0; i < m; ++i) OQ_HIP(hipGraphLaunch(st->exec, h->stream));
        done_steps += m;
        OQ_HIP(hipMemcpyAsync(st->ctrl_host, st->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost,
                              h->stream));
        OQ_HIP(hipStreamSynchronize(h->stream));
        if (st->ctrl_host->done) break;
    }
    if (retcode) *retcode = st->ctrl_host->done ? st->ctrl_host->retcode : OQ_RET_RUNNING;
    return OQ_OK;
}

int oq_stepper_set_persistent(oq_stepper *st, int on)
{
    OQ_REQUIRE(st != nullptr, "stepper is NULL");
    oq_rhs *h = st->h;
    OQ_HIP(hipSetDevice(h->device));
    if (!on) {
        st->persistent = false;
        return OQ_OK;
    }
    OQ_REQUIRE(h->ne == 0 && !h->dev.dila,
               "the persistent stepper takes fault-only problems without dilatancy");
    OQ_REQUIRE(h->kdtype == OQ_K_FP64, "the persistent stepper takes an FP64 operator");
    int coop = 0, cus = 0, maxlds = 0, khz = 0;
    OQ_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, h->device));
    OQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    OQ_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
    OQ_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, h->device));
    OQ_REQUIRE(coop != 0, "the device has no cooperative launch");
    // rows of K per workgroup: as many as fit in 128 KB of LDS
    const int64_t rows = (int64_t)(131072 / (h->ld * 8));
    OQ_REQUIRE(rows >= 1, "the operator rows are too long for the persistent stepper (N > 16384)");
    const int R = (int)(rows < 64 ? rows : 64);
    const int nwg = (int)((h->N + R - 1) / R);
    // per owned entry: Tsit5 keeps u, the stage state and seven stages; VCABM5
    // keeps u, p, f(p), y_{n+1}, f(y_{n+1}) and Phi_0..Phi_5
    const size_t per_entry = st->alg == 1 ? 11 : 9;
    const size_t lds = ((size_t)R * (h->ld + 2) + h->ld + per_entry * 3 * R) * 8;  // K rows at pitch ld + 2
    OQ_REQUIRE((int64_t)lds <= (int64_t)maxlds, "the persistent stepper does not fit in LDS");
    const void *kfn = st->alg == 1 ? (const void *)k_persist_ab : (const void *)k_persist;
    OQ_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    OQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 256, lds));
    OQ_REQUIRE(per_cu >= 1 && (int64_t)per_cu * cus >= nwg,
               "the persistent stepper needs more co-resident workgroups than the device holds");
    OQ_REQUIRE(nwg <= 128, "the persistent stepper takes at most 128 workgroups (N <= 1024)");
    // two source buffers and, for Tsit5, the four stage-1 candidate vectors
    int rc = st->p_s2.alloc((size_t)(6 * h->N) * 8);
    if (!rc) rc = st->p_part.alloc((size_t)(2 * nwg) * 8);
    if (!rc) rc = st->p_bar.alloc(sizeof(unsigned));
    if (!rc) rc = st->p_abort.alloc(sizeof(int));
    if (rc) return rc;
    OQ_HIP(hipMemsetAsync(st->p_abort.p, 0, sizeof(int), h->stream));
    OQ_HIP(hipStreamSynchronize(h->stream));
    st->p_R = R;
    st->p_nwg = nwg;
    st->p_lds = lds;
    // a barrier wait longer than 5 s means a workgroup is gone: abort
    st->p_wait = (long long)(khz > 0 ? khz : 100000) * 5000;
    st->persistent = true;
    return OQ_OK;
}

int oq_stepper_state(oq_stepper *st, double *t, double *dt, double *u, int64_t *naccept,
                     int64_t *nreject)
{
    OQ_REQUIRE(st != nullptr, "stepper is NULL");
    oq_rhs *h = st->h;
    OQ_HIP(hipSetDevice(h->device));
    OQ_HIP(hipMemcpyAsync(st->ctrl_host, st->ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost,
                          h->stream));
    if (u)
        OQ_HIP(hipMemcpyAsync(u, st->u.p, (size_t)st->n * 8, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipStreamSynchronize(h->stream));
    if (t) *t = st->ctrl_host->t;
    if (dt) *dt = st->ctrl_host->dt;
    if (naccept) *naccept = st->ctrl_host->naccept;
    if (nreject) *nreject = st->ctrl_host->nreject;
    return OQ_OK;
}

int oq_stepper_record_du(oq_stepper *st, int64_t stride, int64_t snap_capacity,
                         int64_t step_capacity, int du)
{
    OQ_REQUIRE(st != nullptr, "stepper is NULL");
    OQ_REQUIRE(stride >= 1 && snap_capacity >= 0 && step_capacity >= 0,
               "stride >= 1 and non-negative capacities required");
    oq_rhs *h = st->h;
    OQ_HIP(hipSetDevice(h->device));
    int rc;
    if ((rc = st->rcnt.alloc(2 * sizeof(long long)))) return rc;
    if ((rc = st->rstep_t.alloc((size_t)step_capacity * 8))) return rc;
    if ((rc = st->rstep_v.alloc((size_t)step_capacity * 8))) return rc;
    if ((rc = st->rsnap_t.alloc((size_t)snap_capacity * 8))) return rc;
    if ((rc = st->rsnap_u.alloc((size_t)(snap_capacity * st->n) * 8))) return rc;
    if (du) {
        if ((rc = st->rsnap_du.alloc((size_t)(snap_capacity * st->n) * 8))) return rc;
    } else {
        st->rsnap_du.release();
    }
    const long long zero[2] = {0, 0};
    OQ_HIP(hipMemcpyAsync(st->rcnt.p, zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
    Recorder &r = st->rec;
    r.nstep = st->rcnt.as<long long>();
    r.nsnap = st->rcnt.as<long long>() + 1;
    r.step_t = st->rstep_t.as<double>();
    r.step_vmax = st->rstep_v.as<double>();
    r.snap_t = st->rsnap_t.as<double>();
    r.snap_u = st->rsnap_u.as<double>();
    r.snap_du = du ? st->rsnap_du.as<double>() : nullptr;
    // f(u) of the accepted state: Tsit5 keeps it in k1 (FSAL), VCABM5 in the
    // second half of its corrector buffer
    r.du = st->alg == 1 ? st->ab.as<double>() + st->n : st->k.as<double>();
    r.step_cap = step_capacity;
    r.snap_cap = snap_capacity;
    r.stride = stride;
    r.n = st->n;
    r.nf = h->nf;
    st->recording = true;
    return capture_step(st);
}

int oq_stepper_record(oq_stepper *st, int64_t stride, int64_t snap_capacity,
                      int64_t step_capacity)
{
    return oq_stepper_record_du(st, stride, snap_capacity, step_capacity, 0);
}

// Copies the `keep` newest of `cnt` ring entries (each `w` doubles, ring of
// `cap` slots written at index count % cap) to `out` in the order written.
static int drain_ring(const double *ring, long long cnt, long long cap, int64_t w, double *out,
                      hipStream_t s)
{
    const long long keep = cnt < cap ? cnt : cap;
    if (keep <= 0 || out == nullptr) return OQ_OK;
    const long long first = (cnt - keep) % cap;  // slot of the oldest kept entry
    const long long n1 = (first + keep <= cap) ? keep : cap - first;
    OQ_HIP(hipMemcpyAsync(out, ring + first * w, (size_t)(n1 * w) * 8, hipMemcpyDeviceToHost, s));
    if (keep > n1)
        OQ_HIP(hipMemcpyAsync(out + n1 * w, ring, (size_t)((keep - n1) * w) * 8, hipMemcpyDeviceToHost,
                              s));
    return OQ_OK;
}

int oq_stepper_drain_du(oq_stepper *st, int64_t *nsteps, double *step_t, double *step_vmax,
                        int64_t *nsnaps, double *snap_t, double *snap_u, double *snap_du)
{
    OQ_REQUIRE(st != nullptr, "stepper is NULL");
    OQ_REQUIRE(st->recording, "recording is not enabled (oq_stepper_record)");
    oq_rhs *h = st->h;
    OQ_HIP(hipSetDevice(h->device));
    long long cnt[2];
    OQ_HIP(hipMemcpyAsync(cnt, st->rcnt.p, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipStreamSynchronize(h->stream));
    const Recorder &r = st->rec;
    const long long ks = r.step_cap > 0 ? (cnt[0] < r.step_cap ? cnt[0] : r.step_cap) : 0;
    const long long kn = r.snap_cap > 0 ? (cnt[1] < r.snap_cap ? cnt[1] : r.snap_cap) : 0;
    const long long lost_s = r.step_cap > 0 ? cnt[0] - ks : 0, lost_n = r.snap_cap > 0 ? cnt[1] - kn : 0;
    if (nsteps) *nsteps = r.step_cap > 0 ? ks : cnt[0];
    if (nsnaps) *nsnaps = kn;
    int rc;
    if (r.step_cap > 0) {
        if ((rc = drain_ring(r.step_t, cnt[0], r.step_cap, 1, step_t, h->stream))) return rc;
        if ((rc = drain_ring(r.step_vmax, cnt[0], r.step_cap, 1, step_vmax, h->stream))) return rc;
    }
    if (r.snap_cap > 0) {
        if ((rc = drain_ring(r.snap_t, cnt[1], r.snap_cap, 1, snap_t, h->stream))) return rc;
        if ((rc = drain_ring(r.snap_u, cnt[1], r.snap_cap, st->n, snap_u, h->stream))) return rc;
        if (r.snap_du && (rc = drain_ring(r.snap_du, cnt[1], r.snap_cap, st->n, snap_du, h->stream)))
            return rc;
    }
    const long long zero[2] = {0, 0};
    OQ_HIP(hipMemcpyAsync(st->rcnt.p, zero, sizeof(zero), hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipStreamSynchronize(h->stream));
    if (lost_s > 0 || lost_n > 0) {
        set_error("recorder overflow: " + std::to_string(lost_s) + " step records and " +
                  std::to_string(lost_n) +
                  " snapshots were overwritten before this drain; the newest ones were handed out "
                  "and the recorder restarted (drain more often)");
        return OQ_ERR_STATE;
    }
    return OQ_OK;
}

int oq_stepper_drain(oq_stepper *st, int64_t *nsteps, double *step_t, double *step_vmax,
                     int64_t *nsnaps, double *snap_t, double *snap_u)
{
    return oq_stepper_drain_du(st, nsteps, step_t, step_vmax, nsnaps, snap_t, snap_u, nullptr);
}

int oq_stepper_destroy(oq_stepper *st)
{
    if (st) {
        if (runtime_usable()) (void)hipSetDevice(st->h->device);
        destroy_stepper(st);
    }
    return OQ_OK;
}

}  // extern "C"
