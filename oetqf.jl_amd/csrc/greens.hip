// greens.hip -- stress Green's-function builders on gfx950 (FP64), one
// thread per source-receiver pair, and their C-ABI entry points.
//
// Reference: src/BEM/GF.jl (K_ff :31-71, K_mf :123-174, K_fm :194-227,
// K_mm :250-296), threaded there over source columns (P1); here the whole
// (receiver x source) grid is one launch.  Output element (r, c) of a block
// goes to out[r*rs + c*cs], so the same kernels fill the caller's Julia
// column-major arrays and the library's receiver-major dense operator.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <complex>
#include <atomic>
#include <vector>

#include "hex8_dev.h"
#include "okada_dev.h"
#include "oq_internal.h"

namespace oq {

// ---------------- small host utilities ----------------
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// Cleared once a teardown call finds the runtime unusable (deinitialised at
// process exit, or a sticky device error): later releases then skip their
// HIP calls and only drop the host objects.
static std::atomic<bool> g_runtime_ok{true};
bool runtime_usable() { return g_runtime_ok.load(std::memory_order_relaxed); }
bool teardown_sync(hipStream_t s)
{
    if (!runtime_usable()) return false;
    const hipError_t e = s ? hipStreamSynchronize(s) : hipDeviceSynchronize();
    if (e != hipSuccess) {
        g_runtime_ok.store(false, std::memory_order_relaxed);
        return false;
    }
    return true;
}

static double sin_deg_small(double r) { return sin(r * (M_PI / 180.0)); }
static double cos_deg_small(double r) { return cos(r * (M_PI / 180.0)); }

double sind(double deg)
{
    if (!std::isfinite(deg)) return NAN;
    const double rx = fmod(deg, 360.0), arx = fabs(rx);
    if (rx == 0.0) return copysign(0.0, deg);
    if (arx == 180.0) return copysign(0.0, rx);
    double s;
    if (arx < 45.0) s = sin_deg_small(arx);
    else if (arx <= 135.0) s = cos_deg_small(90.0 - arx);
    else if (arx < 225.0) s = sin_deg_small(180.0 - arx);
    else if (arx <= 315.0) s = -cos_deg_small(270.0 - arx);
    else s = -sin_deg_small(360.0 - arx);
    return rx < 0.0 ? -s : s;
}

double cosd(double deg)
{
    if (!std::isfinite(deg)) return NAN;
    const double arx = fabs(fmod(deg, 360.0));
    if (arx < 45.0) return cos_deg_small(arx);
    if (arx <= 135.0) return sin_deg_small(90.0 - arx);
    if (arx < 225.0) return -cos_deg_small(180.0 - arx);
    if (arx <= 315.0) return sin_deg_small(arx - 270.0);
    return cos_deg_small(360.0 - arx);
}

int DevBuf::alloc(size_t n)
{
    release();
    if (n == 0) n = 8;
    if (hipMalloc(&p, n) != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc of " + std::to_string(n) + " bytes failed");
        return OQ_ERR_NOMEM;
    }
    bytes = n;
    return OQ_OK;
}

int DevBuf::upload(const void *h, size_t n, hipStream_t s)
{
    int rc = alloc(n);
    if (rc) return rc;
    OQ_HIP(hipMemcpyAsync(p, h, n, hipMemcpyHostToDevice, s));
    OQ_HIP(hipStreamSynchronize(s));
    return OQ_OK;
}

void DevBuf::release()
{
    if (p && runtime_usable()) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int check_fault(const oq_fault_mesh *m)
{
    OQ_REQUIRE(m != nullptr, "fault mesh is NULL");
    OQ_REQUIRE(m->nx > 0 && m->nxi > 0, "fault mesh needs nx > 0 and nxi > 0");
    OQ_REQUIRE(m->x && m->ax && m->y && m->z && m->axi, "fault mesh array is NULL");
    OQ_REQUIRE(std::isfinite(m->dip) && std::isfinite(m->dx) && std::isfinite(m->dep),
               "fault mesh scalars must be finite");
    return OQ_OK;
}

int check_hex8(const oq_hex8_mesh *m)
{
    OQ_REQUIRE(m != nullptr, "hex8 mesh is NULL");
    OQ_REQUIRE(m->ne > 0, "hex8 mesh needs ne > 0");
    OQ_REQUIRE(m->cx && m->cy && m->cz && m->qx && m->qy && m->qz && m->dx && m->dy && m->dz,
               "hex8 mesh array is NULL");
    return OQ_OK;
}

int check_quad(const oq_quadrature *q)
{
    OQ_REQUIRE(q != nullptr && q->coords && q->weights && q->nq > 0,
               "Wrong format of quadrature!");
    return OQ_OK;
}

int FaultUpload::load(const oq_fault_mesh *m)
{
    int rc;
    if ((rc = check_fault(m))) return rc;
    const size_t nx = (size_t)m->nx, nxi = (size_t)m->nxi;
    if ((rc = x.upload(m->x, nx * 8))) return rc;
    if ((rc = ax.upload(m->ax, 2 * nx * 8))) return rc;
    if ((rc = y.upload(m->y, nxi * 8))) return rc;
    if ((rc = z.upload(m->z, nxi * 8))) return rc;
    if ((rc = axi.upload(m->axi, 2 * nxi * 8))) return rc;
    dev.x = x.as<double>();
    dev.ax = ax.as<double>();
    dev.y = y.as<double>();
    dev.z = z.as<double>();
    dev.axi = axi.as<double>();
    dev.nx = m->nx;
    dev.nxi = m->nxi;
    dev.dep = m->dep;
    dev.sd = sind(m->dip);
    dev.cd = cosd(m->dip);
    dev.s2d = sind(2.0 * m->dip);
    dev.c2d = cosd(2.0 * m->dip);
    dev.lrept = 0.0;
    return OQ_OK;
}

int Hex8Upload::load(const oq_hex8_mesh *m)
{
    int rc;
    if ((rc = check_hex8(m))) return rc;
    const size_t n = (size_t)m->ne * 8;
    DevBuf *bufs[9] = {&cx, &cy, &cz, &qx, &qy, &qz, &dx, &dy, &dz};
    const double *src[9] = {m->cx, m->cy, m->cz, m->qx, m->qy, m->qz, m->dx, m->dy, m->dz};
    for (int i = 0; i < 9; ++i)
        if ((rc = bufs[i]->upload(src[i], n))) return rc;
    dev.cx = cx.as<double>(); dev.cy = cy.as<double>(); dev.cz = cz.as<double>();
    dev.qx = qx.as<double>(); dev.qy = qy.as<double>(); dev.qz = qz.as<double>();
    dev.dx = dx.as<double>(); dev.dy = dy.as<double>(); dev.dz = dz.as<double>();
    dev.ne = m->ne;
    return OQ_OK;
}

int QuadUpload::load(const oq_quadrature *q)
{
    int rc;
    if ((rc = check_quad(q))) return rc;
    if ((rc = c.upload(q->coords, 3 * (size_t)q->nq * 8))) return rc;
    if ((rc = w.upload(q->weights, (size_t)q->nq * 8))) return rc;
    nq = (int)q->nq;
    return OQ_OK;
}

// ---------------- kernels ----------------

// Shear traction on the fault from the 9 gradients (shear_traction_dc3d,
// GF.jl:76-87); g = (uxx uyx uzx uxy uyy uzy uxz uyz uzz).
__device__ __forceinline__ double traction_grad(int ftype, const double g[9], double lambda,
                                                double mu, const DevFault &f)
{
    if (ftype == 0) {
        const double sxy = mu * (g[1] + g[3]);
        const double sxz = mu * (g[2] + g[6]);
        return -sxy * f.sd + sxz * f.cd;
    }
    const double szz = (lambda + 2.0 * mu) * g[8] + lambda * g[0] + lambda * g[4];
    const double syy = (lambda + 2.0 * mu) * g[4] + lambda * g[0] + lambda * g[8];
    const double syz = mu * (g[7] + g[5]);
    return (szz - syy) / 2.0 * f.s2d + syz * f.c2d;
}

// shear_traction_vol (GF.jl:89-96).
__device__ __forceinline__ double traction_vol(int ftype, const double s[6], const DevFault &f)
{
    if (ftype == 0) return -s[1] * f.sd + s[2] * f.cd;
    return (s[5] - s[3]) / 2.0 * f.s2d + s[4] * f.c2d;
}

// K_ff Toeplitz kernel, thread (i, j, l): st[i + nx j + nx nxi l]
// (GF.jl:42-56), images p*lrept for p in -nrept..nrept.
template <int SLIP>
__global__ void __launch_bounds__(256) k_gf_ff(DevFault f, OkConst c, double lambda, double mu,
                                               int nrept, double *st)
{
    const int64_t n = f.nx * f.nxi * f.nxi;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t i = t % f.nx, j = (t / f.nx) % f.nxi, l = t / (f.nx * f.nxi);
    double u[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, g[9];
    for (int p = -nrept; p <= nrept; ++p) {
        const double jump = (double)p * f.lrept;
        okada_grad<SLIP>(c, f.x[i], f.y[j], f.z[j], f.dep, f.ax[0] + jump, f.ax[1] + jump,
                         f.axi[2 * l], f.axi[2 * l + 1], g);
#pragma unroll
        for (int k = 0; k < 9; ++k) u[k] += g[k];
    }
    st[t] = traction_grad(SLIP, u, lambda, mu, f);
}

// K_mf = gf12 (GF.jl:141-170): thread (source patch j fastest, receiver
// cell i); rows k*ne + i, column j.  Held to two wavefronts per SIMD (256
// registers, 84 B/lane of scratch for strike slip, against 282 registers at
// one): 6.8 -> 6.1 ms for 12288 receivers x 4096 patches x 5 images.
template <int SLIP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_gf_mf(DevFault f, DevHex8 h, const double *qc, const double *qw, int nq, OkConst c, double lambda, double mu,
        int nrept, double *out, int64_t rs, int64_t cs)
{
    const int64_t nf = f.nx * f.nxi;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * h.ne) return;
    const int64_t j = t % nf, i = t / nf;
    const int64_t q1 = j % f.nx, q2 = j / f.nx;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int w = 0; w < nq; ++w) {
        const double rx = h.cx[i] + qc[3 * w] * h.dx[i] / 2.0;
        const double ry = h.cy[i] + qc[3 * w + 1] * h.dy[i] / 2.0;
        const double rz = h.cz[i] + qc[3 * w + 2] * h.dz[i] / 2.0;
        double u[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, g[9];
        for (int p = -nrept; p <= nrept; ++p) {
            const double jump = (double)p * f.lrept;
            okada_grad<SLIP>(c, rx, ry, rz, f.dep, f.ax[2 * q1] + jump, f.ax[2 * q1 + 1] + jump,
                             f.axi[2 * q2], f.axi[2 * q2 + 1], g);
#pragma unroll
            for (int k = 0; k < 9; ++k) u[k] += g[k];
        }
        const double lek = lambda * (u[0] + u[4] + u[8]);
        const double ww = qw[w];
        acc[0] += ww * (lek + 2.0 * mu * u[0]);
        acc[1] += ww * (mu * (u[1] + u[3]));
        acc[2] += ww * (mu * (u[2] + u[6]));
        acc[3] += ww * (lek + 2.0 * mu * u[4]);
        acc[4] += ww * (mu * (u[5] + u[7]));
        acc[5] += ww * (lek + 2.0 * mu * u[8]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) out[(k * h.ne + i) * rs + j * cs] = acc[k];
}

// K_fm = gf21 (GF.jl:206-224): thread (source cell j fastest, receiver
// patch i); all 6 unit eigenstrains from one geometric evaluation.  Only
// the shear traction on the fault is needed, so the image quadrature
// accumulates its 6-term contraction (h8_image_traction) instead of the 54
// integrals; the full-space part and the eigenstress follow as before.
// Two wavefronts per SIMD (at most 256 registers): dip slip needs 266
// unconstrained and spills 44 B/lane under the bound, 67 -> 56 ms at
// 4096 x 2048 pairs.
template <int FT>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2)))
k_gf_fm(DevHex8 h, DevFault f, double mu, double nu, double *out, int64_t rs, int64_t cs)
{
    const int64_t nf = f.nx * f.nxi;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * h.ne) return;
    const int64_t j = t % h.ne, i = t / h.ne;
    const int64_t a = i % f.nx, b = i / f.nx;
    const double x = f.x[a], y = f.y[b], z = f.z[b];
    const double lo[3] = {h.qx[j] - h.dx[j] / 2.0, h.qy[j], h.qz[j] - h.dz[j]};
    const double hi[3] = {h.qx[j] + h.dx[j] / 2.0, h.qy[j] + h.dy[j], h.qz[j]};
    const double ca = FT == 0 ? -mu * f.sd : mu * f.s2d, cb = FT == 0 ? mu * f.cd : mu * f.c2d;
    double E[6] = {0, 0, 0, 0, 0, 0};
    h8_image_traction<FT>(x, y, z, lo, hi, nu, ca, cb, E);
    Hex8Kelvin K;
    const bool inside = hex8_kelvin(x, y, z, lo, hi, K);
    const double ci = 1.0 / (16.0 * M_PI * mu * (1.0 - nu));
    for (int p = 0; p < 6; ++p) {
        double s[6], mv[6];
        hex8_stress_unit_kelvin(K, inside, p, mu, nu, s);
        hex8_moment(p, mu, nu, mv);
        double ti = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) ti += E[q] * mv[q];
        out[i * rs + (p * h.ne + j) * cs] = traction_vol(FT, s, f) + ti * ci;
    }
}

// K_mm = gf22 (GF.jl:262-289): thread (source cell i fastest, receiver
// cell j); entry [(k)nrec + j, (p)nsrc + i] += w_q sigma_k.  Receivers come
// from rm, sources from sm: the same mesh for the dense block, a set of
// shifted receiver cells and one source column for the grid kernel table.
__global__ void __launch_bounds__(128) k_gf_mm(DevHex8 rm, DevHex8 sm, const double *qc,
                                               const double *qw, int nq, double mu, double nu,
                                               double *out, int64_t rs, int64_t cs)
{
    const int64_t nrec = rm.ne, nsrc = sm.ne;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrec * nsrc) return;
    const int64_t i = t % nsrc, j = t / nsrc;
    // The 36 sums live in LDS (a column per thread): in registers they would
    // be live through the image quadrature and push it into scratch spills.
    __shared__ double acc[36][128];
    const int c = threadIdx.x;
#pragma unroll
    for (int m = 0; m < 36; ++m) acc[m][c] = 0.0;
    for (int q = 0; q < nq; ++q) {
        const double rx = rm.cx[j] + qc[3 * q] * rm.dx[j] / 2.0;
        const double ry = rm.cy[j] + qc[3 * q + 1] * rm.dy[j] / 2.0;
        const double rz = rm.cz[j] + qc[3 * q + 2] * rm.dz[j] / 2.0;
        Hex8Pair P;
        hex8_pair(rx, ry, rz, sm.qx[i], sm.qy[i], sm.qz[i], sm.dx[i], sm.dy[i], sm.dz[i], nu, P);
        for (int p = 0; p < 6; ++p) {
            double s[6];
            hex8_stress_unit(P, p, mu, nu, s);
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[p * 6 + k][c] += s[k] * qw[q];
        }
    }
    for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int k = 0; k < 6; ++k) out[(k * nrec + j) * rs + (p * nsrc + i) * cs] = acc[p * 6 + k][c];
}

// Dense fault-fault block: K[(i,j),(k,l)] = st[|i-k|, j, l] (the expansion
// test/BEM/tests.jl:46-49 checks the FFT form against).
__global__ void __launch_bounds__(256) k_expand_toeplitz(const double *st, int64_t nx,
                                                         int64_t nxi, double *k, int64_t ld)
{
    const int64_t nf = nx * nxi;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nf * nf; t += step) {
        const int64_t c = t % nf, r = t / nf;
        const int64_t i = r % nx, j = r / nx, kk = c % nx, l = c / nx;
        const int64_t d = i > kk ? i - kk : kk - i;
        k[r * ld + c] = st[d + nx * j + nx * nxi * l];
    }
}

static inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }
// one thread per pair: a dispatch counts at most 2^32 - 1 work-items
static constexpr int64_t kMaxPairs = (int64_t)0xFFFFFFFFu - 256;
#define OQ_PAIRS_FIT(n) OQ_REQUIRE((n) <= kMaxPairs, "too many source-receiver pairs for one launch")

int launch_gf_ff(const DevFault &f, double lambda, double mu, int ftype, int nrept, double *st,
                 hipStream_t s)
{
    const OkConst c = ok_const((lambda + mu) / (lambda + 2.0 * mu), f.sd, f.cd);
    const int64_t n = f.nx * f.nxi * f.nxi;
    OQ_PAIRS_FIT(n);
    if (ftype == 0)
        hipLaunchKernelGGL(k_gf_ff<0>, dim3(nblocks(n, 256)), dim3(256), 0, s, f, c, lambda, mu,
                           nrept, st);
    else
        hipLaunchKernelGGL(k_gf_ff<1>, dim3(nblocks(n, 256)), dim3(256), 0, s, f, c, lambda, mu,
                           nrept, st);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

int launch_gf_mf(const DevFault &f, const DevHex8 &h, const double *qc, const double *qw, int nq,
                 double lambda, double mu, int ftype, int nrept, double *out, int64_t rs,
                 int64_t cs, hipStream_t s)
{
    const OkConst c = ok_const((lambda + mu) / (lambda + 2.0 * mu), f.sd, f.cd);
    const int64_t n = f.nx * f.nxi * h.ne;
    OQ_PAIRS_FIT(n);
    if (ftype == 0)
        hipLaunchKernelGGL(k_gf_mf<0>, dim3(nblocks(n, 256)), dim3(256), 0, s, f, h, qc, qw, nq, c,
                           lambda, mu, nrept, out, rs, cs);
    else
        hipLaunchKernelGGL(k_gf_mf<1>, dim3(nblocks(n, 256)), dim3(256), 0, s, f, h, qc, qw, nq, c,
                           lambda, mu, nrept, out, rs, cs);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

int launch_gf_fm(const DevHex8 &h, const DevFault &f, double lambda, double mu, int ftype,
                 double *out, int64_t rs, int64_t cs, hipStream_t s)
{
    const double nu = lambda / 2.0 / (lambda + mu);  // GF.jl:203
    const int64_t n = f.nx * f.nxi * h.ne;
    OQ_PAIRS_FIT(n);
    if (ftype == 0)
        hipLaunchKernelGGL(k_gf_fm<0>, dim3(nblocks(n, 128)), dim3(128), 0, s, h, f, mu, nu, out, rs, cs);
    else
        hipLaunchKernelGGL(k_gf_fm<1>, dim3(nblocks(n, 128)), dim3(128), 0, s, h, f, mu, nu, out, rs, cs);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

int launch_gf_mm(const DevHex8 &h, const double *qc, const double *qw, int nq, double lambda,
                 double mu, double *out, int64_t rs, int64_t cs, hipStream_t s)
{
    return launch_gf_mm_pairs(h, h, qc, qw, nq, lambda, mu, out, rs, cs, s);
}

int launch_gf_mm_pairs(const DevHex8 &rm, const DevHex8 &sm, const double *qc, const double *qw,
                       int nq, double lambda, double mu, double *out, int64_t rs, int64_t cs,
                       hipStream_t s)
{
    const double nu = lambda / 2.0 / (lambda + mu);  // GF.jl:259
    const int64_t n = rm.ne * sm.ne;
    OQ_PAIRS_FIT(n);
    hipLaunchKernelGGL(k_gf_mm, dim3(nblocks(n, 128)), dim3(128), 0, s, rm, sm, qc, qw, nq, mu, nu,
                       out, rs, cs);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

int launch_expand_toeplitz(const double *st, int64_t nx, int64_t nxi, double *k, int64_t ld,
                           hipStream_t s)
{
    const int64_t nf = nx * nxi;
    hipLaunchKernelGGL(k_expand_toeplitz, dim3(nf * nf < (256 << 16) ? nblocks(nf * nf, 256) : 65536u), dim3(256), 0, s, st, nx,
                       nxi, k, ld);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

// ---------------- DFT of the Toeplitz kernel (GF.jl:60-67) ----------------
// The reference stores the real DFT, length M = 2nx-1, of the symmetric
// sequence x = [st[:,j,l]; reverse(st[2:end,j,l])] (x_n = x_{M-n}) for every
// (j, l).  Its transform is real:
//   X_f = x_0 + 2 sum_{n=1}^{nx-1} x_n cos(2 pi f n / M),   f < nx,
// and the imaginary parts are exactly zero here (FFTW returns rounding-level
// values).  The inverse recovers st from nx complex coefficients c_f:
//   st_n = (Re c_0 + 2 sum_{f>=1} (Re c_f cos - Im c_f sin)(2 pi f n / M)) / M.
// One workgroup per (j, l) column; the column and the quarter-symmetric
// twiddles cos/sin(2 pi k / M), k < nx, sit in LDS (16 or 32 bytes per
// strike cell).  Phases k = f n mod M are advanced by addition.
static constexpr int kDftThreads = 256;

__device__ inline void dft_twiddles(double *cs, double *sn, int64_t nx, int64_t M)
{
    for (int64_t k = threadIdx.x; k < nx; k += blockDim.x) {
        double sv, cv;
        sincospi(2.0 * (double)k / (double)M, &sv, &cv);
        cs[k] = cv;
        if (sn) sn[k] = sv;
    }
}

__global__ void __launch_bounds__(kDftThreads) k_toeplitz_dft(const double *__restrict__ st, int64_t nx,
                                                             double *__restrict__ out)
{
    extern __shared__ double lds[];
    double *cs = lds, *x = lds + nx;
    const int64_t M = 2 * nx - 1, jl = blockIdx.x;
    dft_twiddles(cs, nullptr, nx, M);
    for (int64_t n = threadIdx.x; n < nx; n += blockDim.x) x[n] = st[jl * nx + n];
    __syncthreads();
    for (int64_t f = threadIdx.x; f < nx; f += blockDim.x) {
        double acc = 0.0;
        int64_t k = 0;
        for (int64_t n = 1; n < nx; ++n) {
            k += f;
            if (k >= M) k -= M;
            acc += x[n] * cs[k < nx ? k : M - k];
        }
        out[2 * (jl * nx + f)] = x[0] + 2.0 * acc;
        out[2 * (jl * nx + f) + 1] = 0.0;
    }
}

__global__ void __launch_bounds__(kDftThreads) k_toeplitz_idft(const double *__restrict__ dft, int64_t nx,
                                                              double *__restrict__ st)
{
    extern __shared__ double lds[];
    double *cs = lds, *sn = lds + nx, *re = lds + 2 * nx, *im = lds + 3 * nx;
    const int64_t M = 2 * nx - 1, jl = blockIdx.x;
    dft_twiddles(cs, sn, nx, M);
    for (int64_t f = threadIdx.x; f < nx; f += blockDim.x) {
        re[f] = dft[2 * (jl * nx + f)];
        im[f] = dft[2 * (jl * nx + f) + 1];
    }
    __syncthreads();
    for (int64_t n = threadIdx.x; n < nx; n += blockDim.x) {
        double acc = 0.0;
        int64_t k = 0;
        for (int64_t f = 1; f < nx; ++f) {
            k += n;
            if (k >= M) k -= M;
            const double c = cs[k < nx ? k : M - k], sv = k < nx ? sn[k] : -sn[M - k];
            acc += re[f] * c - im[f] * sv;
        }
        st[jl * nx + n] = (re[0] + 2.0 * acc) / (double)M;
    }
}

int launch_toeplitz_dft(const double *st, int64_t nx, int64_t nxi, double *out, hipStream_t s)
{
    OQ_REQUIRE(nx <= kMaxDftNx, "the device DFT of the Toeplitz kernel takes nx <= 5120");
    const size_t lds = (size_t)(2 * nx) * 8;
    if (lds > 64 * 1024)
        OQ_HIP(hipFuncSetAttribute((const void *)k_toeplitz_dft,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_toeplitz_dft, dim3((unsigned)(nxi * nxi)), dim3(kDftThreads), lds, s, st, nx, out);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

int launch_toeplitz_idft(const double *dft, int64_t nx, int64_t nxi, double *st, hipStream_t s)
{
    OQ_REQUIRE(nx <= kMaxDftNx, "the device DFT of the Toeplitz kernel takes nx <= 5120");
    const size_t lds = (size_t)(4 * nx) * 8;
    if (lds > 64 * 1024)
        OQ_HIP(hipFuncSetAttribute((const void *)k_toeplitz_idft,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_toeplitz_idft, dim3((unsigned)(nxi * nxi)), dim3(kDftThreads), lds, s, dft, nx, st);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

static int lrept_of(const oq_fault_mesh *m, double buffer_ratio, double *lrept)
{
    OQ_REQUIRE(buffer_ratio >= 0.0, "Argument `buffer_ratio` must be ≥ 0.");
    *lrept = (buffer_ratio + 1.0) * (m->dx * (double)m->nx);
    return OQ_OK;
}

}  // namespace oq

using namespace oq;

extern "C" {

const char *oq_last_error(void) { return g_err.c_str(); }

int oq_version(void) { return 1; }

int oq_device_count(int *count)
{
    OQ_REQUIRE(count != nullptr, "count is NULL");
    OQ_HIP(hipGetDeviceCount(count));
    return OQ_OK;
}

int oq_gf_okada_ff(const oq_fault_mesh *mesh, double lambda, double mu, int ftype, int64_t nrept,
                   double buffer_ratio, double *st)
{
    int rc;
    FaultUpload fu;
    OQ_REQUIRE(st != nullptr, "output is NULL");
    OQ_REQUIRE(ftype == 0 || ftype == 1, "ftype must be StrikeSlip (0) or DipSlip (1)");
    OQ_REQUIRE(nrept >= 0, "nrept must be >= 0");
    if ((rc = check_fault(mesh))) return rc;
    double lrept = 0.0;
    if ((rc = lrept_of(mesh, buffer_ratio, &lrept))) return rc;
    if ((rc = fu.load(mesh))) return rc;
    fu.dev.lrept = lrept;
    const size_t n = (size_t)(mesh->nx * mesh->nxi * mesh->nxi);
    DevBuf out;
    if ((rc = out.alloc(n * 8))) return rc;
    if ((rc = launch_gf_ff(fu.dev, lambda, mu, ftype, (int)nrept, out.as<double>(), 0))) return rc;
    OQ_HIP(hipMemcpy(st, out.p, n * 8, hipMemcpyDeviceToHost));
    return OQ_OK;
}

int oq_gf_okada_ff_dft(const oq_fault_mesh *mesh, double lambda, double mu, int ftype,
                       int64_t nrept, double buffer_ratio, double *st_dft)
{
    int rc;
    FaultUpload fu;
    OQ_REQUIRE(st_dft != nullptr, "output is NULL");
    OQ_REQUIRE(ftype == 0 || ftype == 1, "ftype must be StrikeSlip (0) or DipSlip (1)");
    OQ_REQUIRE(nrept >= 0, "nrept must be >= 0");
    if ((rc = check_fault(mesh))) return rc;
    OQ_REQUIRE(mesh->nx <= kMaxDftNx, "the device DFT of the Toeplitz kernel takes nx <= 5120");
    double lrept = 0.0;
    if ((rc = lrept_of(mesh, buffer_ratio, &lrept))) return rc;
    if ((rc = fu.load(mesh))) return rc;
    fu.dev.lrept = lrept;
    const size_t n = (size_t)(mesh->nx * mesh->nxi * mesh->nxi);
    DevBuf st, out;  // st: the real kernel, scratch between the two launches
    if ((rc = st.alloc(n * 8)) || (rc = out.alloc(2 * n * 8))) return rc;
    if ((rc = launch_gf_ff(fu.dev, lambda, mu, ftype, (int)nrept, st.as<double>(), 0))) return rc;
    if ((rc = launch_toeplitz_dft(st.as<double>(), mesh->nx, mesh->nxi, out.as<double>(), 0))) return rc;
    OQ_HIP(hipMemcpy(st_dft, out.p, 2 * n * 8, hipMemcpyDeviceToHost));
    return OQ_OK;
}

int oq_gf_okada_mf(const oq_fault_mesh *mf, const oq_hex8_mesh *ma, double lambda, double mu,
                   int ftype, const oq_quadrature *quad, int64_t nrept, double buffer_ratio,
                   double *st)
{
    int rc;
    FaultUpload fu;
    Hex8Upload hu;
    QuadUpload qu;
    OQ_REQUIRE(st != nullptr, "output is NULL");
    OQ_REQUIRE(ftype == 0 || ftype == 1, "ftype must be StrikeSlip (0) or DipSlip (1)");
    OQ_REQUIRE(nrept >= 0, "nrept must be >= 0");
    if ((rc = check_fault(mf)) || (rc = check_hex8(ma)) || (rc = check_quad(quad))) return rc;
    double lrept = 0.0;
    if ((rc = lrept_of(mf, buffer_ratio, &lrept))) return rc;
    if ((rc = fu.load(mf))) return rc;
    fu.dev.lrept = lrept;
    if ((rc = hu.load(ma))) return rc;
    if ((rc = qu.load(quad))) return rc;
    const int64_t nf = mf->nx * mf->nxi, m6 = 6 * ma->ne;
    DevBuf out;
    if ((rc = out.alloc((size_t)(nf * m6) * 8))) return rc;
    if ((rc = launch_gf_mf(fu.dev, hu.dev, qu.c.as<double>(), qu.w.as<double>(), qu.nq, lambda,
                           mu, ftype, (int)nrept, out.as<double>(), 1, m6, 0)))
        return rc;
    OQ_HIP(hipMemcpy(st, out.p, (size_t)(nf * m6) * 8, hipMemcpyDeviceToHost));
    return OQ_OK;
}

int oq_gf_hex8_fm(const oq_hex8_mesh *ma, const oq_fault_mesh *mf, double lambda, double mu,
                  int ftype, double *st)
{
    int rc;
    FaultUpload fu;
    Hex8Upload hu;
    OQ_REQUIRE(st != nullptr, "output is NULL");
    OQ_REQUIRE(ftype == 0 || ftype == 1, "ftype must be StrikeSlip (0) or DipSlip (1)");
    if ((rc = check_fault(mf)) || (rc = check_hex8(ma))) return rc;
    if ((rc = fu.load(mf))) return rc;
    if ((rc = hu.load(ma))) return rc;
    const int64_t nf = mf->nx * mf->nxi, m6 = 6 * ma->ne;
    DevBuf out;
    if ((rc = out.alloc((size_t)(nf * m6) * 8))) return rc;
    if ((rc = launch_gf_fm(hu.dev, fu.dev, lambda, mu, ftype, out.as<double>(), 1, nf, 0)))
        return rc;
    OQ_HIP(hipMemcpy(st, out.p, (size_t)(nf * m6) * 8, hipMemcpyDeviceToHost));
    return OQ_OK;
}

int oq_gf_hex8_mm(const oq_hex8_mesh *ma, double lambda, double mu, const oq_quadrature *quad,
                  double *st)
{
    int rc;
    Hex8Upload hu;
    QuadUpload qu;
    OQ_REQUIRE(st != nullptr, "output is NULL");
    if ((rc = check_hex8(ma)) || (rc = check_quad(quad))) return rc;
    if ((rc = hu.load(ma))) return rc;
    if ((rc = qu.load(quad))) return rc;
    const int64_t m6 = 6 * ma->ne;
    DevBuf out;
    if ((rc = out.alloc((size_t)(m6 * m6) * 8))) return rc;
    if ((rc = launch_gf_mm(hu.dev, qu.c.as<double>(), qu.w.as<double>(), qu.nq, lambda, mu,
                           out.as<double>(), 1, m6, 0)))
        return rc;
    OQ_HIP(hipMemcpy(st, out.p, (size_t)(m6 * m6) * 8, hipMemcpyDeviceToHost));
    return OQ_OK;
}

}  // extern "C"
