// rhs_impl.h -- the device RHS handle shared by rhs.hip and stepper.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "oq_internal.h"

namespace oq {

// Packed ArrayPartition offsets (equation.jl:159-165, :193-195).
struct StateLayout {
    int64_t nf = 0, ne = 0;
    int64_t v = 0, th = 0, eps = 0, sig = 0, del = 0, n = 0;
    int64_t pp = 0;  // pore pressure (dilatancy: [v th del p], equation.jl:172-173)
    // The dilatancy model (fault only): u = (v, theta, delta, p).
    void set_dilatancy(int64_t nf_)
    {
        set(nf_, 0);
        pp = 3 * nf_;
        n = 4 * nf_;
    }
    void set(int64_t nf_, int64_t ne_)
    {
        nf = nf_;
        ne = ne_;
        v = 0;
        th = nf;
        if (ne > 0) {
            eps = 2 * nf;
            sig = 2 * nf + 6 * ne;
            del = 2 * nf + 12 * ne;
            n = 3 * nf + 12 * ne;
        } else {
            eps = sig = 0;
            del = 2 * nf;
            n = 3 * nf;
        }
    }
};

// Arguments of the fused RHS kernels (all device pointers).
struct RhsDev {
    const double *a, *b, *L, *sigma;  // nf each
    double eta, vpl, f0, v0;
    const double *gamma, *nexp;       // nlaws*ne
    int nlaws;
    double deps0[6];
    StateLayout lay;
    int64_t N, ld;
    // DilatancyProperty (property.jl:27-32), per patch; dila = 0: off
    const double *tp, *deps_d, *beta, *p0;
    int dila;
};

// Cross-lane moves of a double as two 32-bit halves.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ double swizzle_xor16_f64(double x)
{
    // ds_swizzle bit-mask mode: and 0x1F, or 0, xor 0x10 (lane ^ 16 within 32)
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_ds_swizzle((int)b, 0x401F);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), 0x401F);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ double readlane_f64(double x, int lane)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the 64 lanes of a wavefront, the same value in every lane, in a
// fixed order: DPP quad permutations (lane ^ 1, lane ^ 2), then the
// half-row and row mirrors (lane <-> 7 - lane, 15 - lane: after the quad
// steps all lanes of a quad hold the same value, so a mirror pairs whole
// groups as an xor would), ds_swizzle across the rows of each 32-lane half,
// and the two halves read out as scalars.
__device__ __forceinline__ double wave_sum(double x)
{
    x += dpp_f64<0xB1>(x);   // quad_perm [1,0,3,2]
    x += dpp_f64<0x4E>(x);   // quad_perm [2,3,0,1]
    x += dpp_f64<0x141>(x);  // row_half_mirror
    x += dpp_f64<0x140>(x);  // row_mirror
    x += swizzle_xor16_f64(x);
    return readlane_f64(x, 0) + readlane_f64(x, 32);
}

// Rate-and-state rates of a fault patch with parameters a, b, L, sigma
// (update_fault! with the aging law, equation.jl:237-245 and :279): dv/dt and
// dtheta/dt from v, theta and the traction rate dtau.  fault_pre holds every
// term that does not involve dtau (the persistent steppers evaluate it while
// the traction rate is still being gathered), fault_dv the rest.
__device__ __forceinline__ void fault_pre(const RhsDev &d, double a, double b, double L, double sigma, double v,
                                          double th, double &dmu_dv, double &dmu_dth, double &dth)
{
    const double thp = th > 0.0 ? th : 0.0;
    const double psi1 = exp((d.f0 + b * log(d.v0 * thp / L)) / a) / (2.0 * d.v0);
    const double psi2 = sigma * psi1 / hypot(1.0, v * psi1);
    dmu_dv = a * psi2;
    dmu_dth = b / th * v * psi2;
    dth = 1.0 - v * th / L;
}

__device__ __forceinline__ double fault_dv(const RhsDev &d, double dtau, double dmu_dv, double dmu_dth, double dth)
{
    return (dtau - dmu_dth * dth) / (dmu_dv + d.eta);
}

__device__ __forceinline__ void fault_rates_p(const RhsDev &d, double a, double b, double L, double sigma,
                                              double v, double th, double dtau, double &dv, double &dth)
{
    double dmu_dv, dmu_dth;
    fault_pre(d, a, b, L, sigma, v, th, dmu_dv, dmu_dth, dth);
    dv = fault_dv(d, dtau, dmu_dv, dmu_dth, dth);
}

// The same for patch i, its parameters read from d (the persistent steppers
// hold them in registers across the RHS evaluations of a launch instead).
__device__ __forceinline__ void fault_rates(const RhsDev &d, int64_t i, double v, double th,
                                            double dtau, double &dv, double &dth)
{
    fault_rates_p(d, d.a[i], d.b[i], d.L[i], d.sigma[i], v, th, dtau, dv, dth);
}

// Rates with dilatancy (update_fault_with_dilatancy!, equation.jl:248-276;
// aging law :279, pore-pressure law :282): power-law friction
// f0 (sigma - p) (v/v0)^(a/f0) (theta v0/L)^(b/f0), no radiation damping.
__device__ __forceinline__ void fault_rates_dila(const RhsDev &d, int64_t i, double v, double th,
                                                 double pp, double dtau, double &dv, double &dth,
                                                 double &dp)
{
    const double L = d.L[i], af = d.a[i] / d.f0, bf = d.b[i] / d.f0;
    dth = 1.0 - v * th / L;
    dp = -(pp - d.p0[i]) / d.tp[i] + d.deps_d[i] / d.beta[i] / th * dth;
    const double vf = v / d.v0 > 0.0 ? v / d.v0 : 0.0;
    const double tf0 = th * d.v0 / L, thf = tf0 > 0.0 ? tf0 : 0.0;
    const double va = pow(vf, af), vam1 = pow(vf, af - 1.0);
    const double tb = pow(thf, bf), tbm1 = pow(thf, bf - 1.0);
    const double sg = d.sigma[i] - pp;
    dv = (dtau + d.f0 * dp * va * tb - d.f0 * sg * va * tbm1 * bf * d.v0 / L * dth) /
         (d.f0 * sg * vam1 * tb * af / d.v0);
}

// The same on the packed state vectors: du gets (dv, dtheta, ddelta = v)
// and, with dilatancy, dp.
__device__ __forceinline__ void fault_update(const RhsDev &d, int64_t i, double dtau,
                                            const double *__restrict__ u, double *__restrict__ du)
{
    const double v = u[d.lay.v + i], th = u[d.lay.th + i];
    double dv, dth;
    if (d.dila) {
        double dp;
        fault_rates_dila(d, i, v, th, u[d.lay.pp + i], dtau, dv, dth, dp);
        du[d.lay.pp + i] = dp;
    } else {
        fault_rates(d, i, v, th, dtau, dv, dth);
    }
    du[d.lay.th + i] = dth;
    du[d.lay.v + i] = dv;
    du[d.lay.del + i] = v;
}

// K_mm of a mantle whose cells form a uniform horizontal grid (cell
// c = ix + nx (iy + ny iz)): entry [(k, jz, jy, jx), (p, iz, iy, ix)] depends
// on (k, jz, p, iz) and the offsets (jy - iy, jx - ix) only, so it is kept as
// 36 nz^2 tiles of (2ny - 1) x (2nx - 1) and applied as 2D convolutions.
struct MmGrid {
    int64_t nx = 0, ny = 0, nz = 0;
    int r = 0;        // outputs per thread along x (register window)
    int nxp = 0;      // nx rounded up to r (tile rows hold 2 nxp entries)
    int tp = 0;       // source planes (p, iz) per workgroup
    int nchunk = 0;   // workgroups per output plane (k, jz) = partial sums per entry
    int threads = 0;  // workgroup size (256)
    size_t lds = 0;   // dynamic LDS bytes
    int fs = 0;       // column chunks of the fault rows' K_fm pass
    int64_t fcw = 0;  // columns per chunk
    // Along strike: when the mantle's x spacing is m times the fault's strike
    // spacing, K_fm and K_mf depend on a patch-cell pair's x offset only
    // through a - m jx and are kept as tables of D = nx_f + m (nx - 1)
    // offsets per line (strike = true): F[b][q][a - m jx + m (nx - 1)] over
    // fault rows b and mantle source lines q = (p nz + jz) ny + jy, and
    // G[q'][b][m jx - a + nx_f - 1] over mantle receiver lines q' (stress
    // component k in place of p) and fault rows b.
    bool strike = false;
    int64_t m = 0, D = 0;
    int fqs = 0, fqb = 0, nfc = 0;  // F blocks: lines per block, per LDS batch; blocks per fault row
    int gbs = 0, gbb = 0, ngc = 0;  // G blocks: fault rows per block, per LDS batch; blocks per line
    int nr = 0, fseg = 0, gseg = 0;  // ceil(nx_f / m); register-window segments of F and G outputs
    int Fp = 0, Gp = 0;              // LDS row widths of the F and G tiles
    int flpp = 0;                    // F lines in flight per block
    int sw = 0;                      // G: LDS pitch of the s_f rows
    int nblocks = 0;                // blocks of the fused launch (role table entries)
    int nconv = 0;                  // convolution blocks (the first launch when split)
    size_t lds_fused = 0, lds_other = 0;  // dynamic LDS of the fused launch, of the second split launch
    bool split = false;             // convolutions and the other roles as two launches (measured faster)
};

}  // namespace oq

struct oq_rhs {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t nx = 0, nxi = 0, ne = 0, nf = 0, N = 0, ld = 0;
    int kdtype = OQ_K_FP64;
    oq::StateLayout lay;
    oq::DevBuf K;                       // N x ld, kdtype
    oq::DevBuf a, b, L, sigma, gamma, nexp;
    oq::DevBuf tp, deps_d, beta, p0;    // dilatancy (oq_rhs_set_dilatancy)
    oq::DevBuf s64;                     // source vector, ld doubles (zero padded)
    oq::DevBuf s32;                     // FP32 copy of s (BF16 operator)
    oq::DevBuf u, du;                   // packed state scratch for the host API
    oq::DevBuf tpart;                   // Toeplitz operator: per-chunk partial sums
    oq::DevBuf ftw, fg, fsh, fth;       // DFT operator: twiddles, kernel DFT, S^, T^
    // coupled model with the DFT K_ff (OQ_K_FFT, ne > 0): K holds the mantle
    // rows [K_mf K_mm] (6ne x ld); Kf the fault rows' K_fm (nf x ldm), which
    // act on sm, the mantle part of s (ldm doubles, zero padded); stk the
    // Toeplitz kernel; tauff the K_ff traction rates
    oq::DevBuf Kf, sm, stk, tauff;
    int64_t ldm = 0;
    // OQ_K_FFT_GRID (the coupled DFT model with a grid mantle): K holds only
    // the mantle rows' K_mf (6ne x ldf); K_mm is the grid kernel table Tmm,
    // applied by the convolution blocks of k_grid_fused into nchunk partial
    // sums per mantle row (mpart, nchunk x 6ne) that k_grid_finish adds
    oq::DevBuf Tmm, mpart, fpart, mvout;  // fpart: K_ff rates + K_fm chunk partials; mvout: K_mf s_f
                                          // partials (1 or ngc x 6ne)
    oq::DevBuf Ftab, Gtab, roles, roles2;  // along-strike tables; block roles (fused, interleaved; split:
                                           // convolutions first)
    bool mmgrid = false;
    int64_t ldf = 0;
    oq::MmGrid grid{};
    oq::RhsDev dev{};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace oq {
// Enqueues one RHS evaluation du = f(u) on h->stream (device pointers).
int rhs_enqueue(oq_rhs *h, const double *u, double *du);
// Enqueues only the K-streaming kernel (requires a prior prologue).
int rhs_enqueue_matvec(oq_rhs *h, const double *u, double *du);
// Stage coefficients of an explicit Runge-Kutta stage: y = u + dt sum_{j<s} a[j] k_j.
struct StageCoef {
    int s;
    double a[6];
};
// DFT operator (OQ_K_FFT), fault-only without dilatancy: a Tsit5 stage and
// the RHS at it in three launches (the stage fused into the forward FFT);
// dtp and done point into the stepper's control block.
int rhs_enqueue_fft_stage(oq_rhs *h, const StageCoef &sc, const double *dtp, const int *done, const double *u,
                          const double *k, int64_t n, double *y, double *du);
}  // namespace oq
