// rhs.hip -- the quasi-dynamic earthquake-cycle RHS on gfx950.
//
// Replaces `ode(du,u,p,t)` of src/BEM/equation.jl (:156-166 fault only,
// :185-205 coupled) for the reference's `assemble` (equation.jl:81-89,
// :141-154).  The four Green's blocks are held as ONE dense receiver-major
// operator K (N = nf + 6ne rows; row r = receiver, column c = source):
//
//     [tau_dot  ]   [K_ff  K_fm] [v - vpl       ]
//     [sigma_dot] = [K_mf  K_mm] [eps_dot - e0  ]
//
// and each RHS is two launches:
//   1. prologue  (O(N)):   s = [v - vpl ; eps_dot(sigma) - deps0], also
//                          writes du.eps = eps_dot (power-law viscosity,
//                          equation.jl:207-230);
//   2. matvec    (O(N^2)): every 64-wide wavefront streams R rows of K from
//                          HBM with 16-byte non-temporal loads, FP64
//                          accumulation, a cross-lane reduction, then the
//                          fused epilogue: rate-and-state update for fault
//                          rows (update_fault!, equation.jl:233-246) or
//                          du.sigma for mantle rows.
// K may be stored as FP64, FP32 or BF16; row sums are accumulated in FP64
// (the FP64 FMA rate is far above what an 8 TB/s stream needs); BF16 forms
// its products in FP32 per 16-byte chunk.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "rhs_impl.h"

namespace oq {


static constexpr int64_t kPad = 512;  // row pitch multiple (elements)

// ---------------- prologue ----------------
__global__ void __launch_bounds__(256) k_rhs_prologue(RhsDev d, const double *__restrict__ u,
                                                      double *__restrict__ du,
                                                      double *__restrict__ s,
                                                      float *__restrict__ s32,
                                                      double *__restrict__ sm)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nf = d.lay.nf, ne = d.lay.ne;
    if (t < nf) {
        const double x = u[d.lay.v + t] - d.vpl;  // relative_velocity! (equation.jl:35-42)
        s[t] = x;
        if (s32) s32[t] = (float)x;
        return;
    }
    const int64_t i = t - nf;
    if (i >= ne) return;
    // update_strain_rate! + relative_strain_rate! (equation.jl:207-230)
    const double *sg = u + d.lay.sig;
    const double sxx0 = sg[i], sxy = sg[i + ne], sxz = sg[i + 2 * ne];
    const double syy0 = sg[i + 3 * ne], syz = sg[i + 4 * ne], szz0 = sg[i + 5 * ne];
    const double skk = (sxx0 + syy0 + szz0) / 3.0;
    const double sxx = sxx0 - skk, syy = syy0 - skk, szz = szz0 - skk;
    const double tn = sqrt(sxx * sxx + syy * syy + szz * szz +
                           2.0 * (sxy * sxy + sxz * sxz + syz * syz));
    double f = 0.0;
    for (int l = 0; l < d.nlaws; ++l) f += d.gamma[l * ne + i] * pow(tn, d.nexp[l * ne + i]);
    const double e[6] = {f * sxx, f * sxy, f * sxz, f * syy, f * syz, f * szz};
    double *de = du + d.lay.eps;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double x = e[k] - d.deps0[k];
        de[i + k * ne] = e[k];
        s[nf + i + k * ne] = x;
        if (s32) s32[nf + i + k * ne] = (float)x;
        if (sm) sm[i + k * ne] = x;
    }
}

// ---------------- K streaming + fused epilogue ----------------
// Per storage type: elements per 16-byte load, rows per wavefront and
// wavefronts per workgroup (the shapes that streamed fastest on MI355X at
// N = 16384: FP64 6.15 TB/s, FP32 6.87 TB/s, BF16 6.46 TB/s).
template <int KT> struct KTraits;
template <> struct KTraits<OQ_K_FP64> {
    using T = double;
    static constexpr int kPerLane = 2;
    static constexpr int kRows = 8;
    static constexpr int kWaves = 4;
};
template <> struct KTraits<OQ_K_FP32> {
    using T = float;
    static constexpr int kPerLane = 4;
    static constexpr int kRows = 8;
    static constexpr int kWaves = 2;
};
template <> struct KTraits<OQ_K_BF16> {
    using T = unsigned short;
    static constexpr int kPerLane = 8;
    static constexpr int kRows = 8;
    static constexpr int kWaves = 4;
};

// 16-byte vector types of one lane's load, per storage type.
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <int KT> struct KVec;
template <> struct KVec<OQ_K_FP64> { using V = f64x2; };
template <> struct KVec<OQ_K_FP32> { using V = f32x4; };
template <> struct KVec<OQ_K_BF16> { using V = u32x4; };

// Streaming (non-temporal) 128-bit load: K is read once per RHS, so it
// should not evict the source vector from the caches.
template <class V> __device__ __forceinline__ V ld_nt(const V *p) { return __builtin_nontemporal_load(p); }

// The slice of the source vector matching one lane's 16 bytes of K, and the
// dot product of the two.
template <int KT> struct SVec;
template <> struct SVec<OQ_K_FP64> { double2 a; };
template <> struct SVec<OQ_K_FP32> { double4 a; };
template <> struct SVec<OQ_K_BF16> { float4 a, b; };

template <int KT> __device__ __forceinline__ SVec<KT> s_load(const double *__restrict__ s, int64_t c);
template <>
__device__ __forceinline__ SVec<OQ_K_FP64> s_load<OQ_K_FP64>(const double *__restrict__ s, int64_t c)
{
    return {*reinterpret_cast<const double2 *>(s + c)};
}
template <>
__device__ __forceinline__ SVec<OQ_K_FP32> s_load<OQ_K_FP32>(const double *__restrict__ s, int64_t c)
{
    return {*reinterpret_cast<const double4 *>(s + c)};
}
template <>
__device__ __forceinline__ SVec<OQ_K_BF16> s_load<OQ_K_BF16>(const double *__restrict__ sd, int64_t c)
{
    const float *s = reinterpret_cast<const float *>(sd);
    return {*reinterpret_cast<const float4 *>(s + c), *reinterpret_cast<const float4 *>(s + c + 4)};
}

__device__ __forceinline__ double dot16(const f64x2 kv, const SVec<OQ_K_FP64> &sv)
{
    return fma(kv.x, sv.a.x, kv.y * sv.a.y);
}

__device__ __forceinline__ double dot16(const f32x4 kv, const SVec<OQ_K_FP32> &sv)
{
    double r = (double)kv.x * sv.a.x;
    r = fma((double)kv.y, sv.a.y, r);
    r = fma((double)kv.z, sv.a.z, r);
    return fma((double)kv.w, sv.a.w, r);
}

// BF16: the eight products of a 16-byte chunk are formed in FP32 against
// the FP32 copy of s (product rounding 2^-24, far below BF16's 2^-9), then
// added to the FP64 accumulator.
__device__ __forceinline__ double dot16(const u32x4 kv, const SVec<OQ_K_BF16> &sv)
{
    const float4 a = sv.a, b = sv.b;
    float r = __uint_as_float(kv.x << 16) * a.x;
    r = fmaf(__uint_as_float(kv.x & 0xffff0000u), a.y, r);
    r = fmaf(__uint_as_float(kv.y << 16), a.z, r);
    r = fmaf(__uint_as_float(kv.y & 0xffff0000u), a.w, r);
    float q = __uint_as_float(kv.z << 16) * b.x;
    q = fmaf(__uint_as_float(kv.z & 0xffff0000u), b.y, q);
    q = fmaf(__uint_as_float(kv.w << 16), b.z, q);
    q = fmaf(__uint_as_float(kv.w & 0xffff0000u), b.w, q);
    return (double)r + (double)q;
}

// Row sums of R rows over npass passes of 64 16-byte chunks (rp[r] points at
// row r's chunk of this lane in pass 0; pass q covers columns q * STEP ..),
// starting at pass p0 and wrapping around, reduced over the wavefront.
// Software pipeline: the R loads of the next pass are issued before the
// FMAs of the current one, so 2R 16-byte loads per lane are in flight; the
// scheduling barriers keep the compiler from interleaving each load with
// its use (which it does by default to save registers).
template <int KT, int R>
__device__ __forceinline__ void stream_rows(const typename KVec<KT>::V *const (&rp)[R],
                                            const double *__restrict__ s, int64_t npass, int64_t p0,
                                            int lane, double (&acc)[R])
{
    using V = typename KVec<KT>::V;
    constexpr int E = KTraits<KT>::kPerLane;
    constexpr int64_t STEP = 64 * E;
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    int64_t p = p0;
    V cur[R];
#pragma unroll
    for (int r = 0; r < R; ++r) cur[r] = ld_nt(rp[r] + p * 64);
    for (int64_t it = 1; it < npass; ++it) {
        const int64_t pn = (p + 1 == npass) ? 0 : p + 1;
        const SVec<KT> sv = s_load<KT>(s, p * STEP + lane * E);
        V nxt[R];
#pragma unroll
        for (int r = 0; r < R; ++r) nxt[r] = ld_nt(rp[r] + pn * 64);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] += dot16(cur[r], sv);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < R; ++r) cur[r] = nxt[r];
        p = pn;
    }
    {
        const SVec<KT> sv = s_load<KT>(s, p * STEP + lane * E);
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] += dot16(cur[r], sv);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = wave_sum(acc[r]);
}

// One wavefront streams KTraits::kRows consecutive rows.  Each wavefront
// starts its sweep over the columns at a different pass (rotation by a
// multiple of its index): with all wavefronts reading the same column
// offset of rows whose pitch is a large power of two, the loads would
// concentrate on a few HBM channels (measured 4.9 vs 6.1 TB/s, FP64).
template <int KT>
__global__ void __launch_bounds__(KTraits<KT>::kWaves * 64)
    k_rhs_matvec(RhsDev d, const void *__restrict__ Kp, int64_t ld, int64_t nrows, int64_t rowbase,
                 const double *__restrict__ s, const double *__restrict__ tadd, int tcount,
                 int64_t tstride, const double *__restrict__ u, double *__restrict__ du)
{
    using T = typename KTraits<KT>::T;
    constexpr int E = KTraits<KT>::kPerLane;  // elements per 16-byte load
    constexpr int R = KTraits<KT>::kRows;
    constexpr int W = KTraits<KT>::kWaves;
    constexpr int64_t STEP = 64 * E;          // columns per wavefront pass
    const int lane = threadIdx.x & 63;
    const int64_t gw = (int64_t)blockIdx.x * W + (threadIdx.x >> 6);
    // rows [0, nrows) of the operator at Kp (pitch ld) are receivers
    // rowbase + r; row r adds tadd[c * tstride + r] for c < tcount, parts of
    // the operator applied elsewhere (the DFT K_ff of the coupled model, the
    // grid convolution partials of K_mm), in that order
    const int64_t row0 = gw * R;
    if (row0 >= nrows) return;
    using V = typename KVec<KT>::V;
    const T *K = static_cast<const T *>(Kp);
    const V *rp[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t row = row0 + r < nrows ? row0 + r : nrows - 1;
        rp[r] = reinterpret_cast<const V *>(K + row * ld) + lane;
    }
    double acc[R];
    stream_rows<KT, R>(rp, s, ld / STEP, (gw * 13) % (ld / STEP), lane, acc);
    if (lane < R) {
        double y = acc[0];
#pragma unroll
        for (int r = 1; r < R; ++r)
            if (lane == r) y = acc[r];
        if (row0 + lane < nrows) {
            const int64_t loc = row0 + lane, row = rowbase + loc;
            for (int c = 0; c < tcount; ++c) y += tadd[c * tstride + loc];
            if (row < d.lay.nf) fault_update(d, row, y, u, du);
            else du[d.lay.sig + (row - d.lay.nf)] = y;
        }
    }
}

static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <int KT>
static void launch_matvec(oq_rhs *h, const void *K, int64_t ld, int64_t nrows, int64_t rowbase,
                          const double *s, const double *tadd, int tcount, int64_t tstride,
                          const double *u, double *du)
{
    constexpr int rows = KTraits<KT>::kRows * KTraits<KT>::kWaves;
    hipLaunchKernelGGL(k_rhs_matvec<KT>, dim3(nblk(nrows, rows)), dim3(KTraits<KT>::kWaves * 64), 0,
                       h->stream, h->dev, K, ld, nrows, rowbase, s, tadd, tcount, tstride, u, du);
}

template <int KT>
static void launch_matvec(oq_rhs *h, const void *K, int64_t ld, int64_t nrows, int64_t rowbase,
                          const double *s, const double *tadd, const double *u, double *du)
{
    launch_matvec<KT>(h, K, ld, nrows, rowbase, s, tadd, tadd ? 1 : 0, 0, u, du);
}

// ---------------- Toeplitz K_ff (fault-only problems, OQ_K_TOEPLITZ) ----------------
// The reference keeps K_ff as its Toeplitz kernel st[|i-k|, j, l] (GF.jl:41-58)
// and applies it by FFT convolution along strike (equation.jl:44-61).  Here
// the same kernel (nx*nxi^2 doubles instead of nf^2) is applied directly:
// block (j, c) stages, for the dip rows l of chunk c, the reflected kernel
// T2[m] = st[|m|, j, l] (m in [-(nx-1), nx-1], zero-padded above) and s[:, l]
// in LDS; thread (i-block, l) forms kTR consecutive outputs
// out[i + r] = sum_k T2[i + r - k] s[k, l] with a register window over T2
// that slides by one element per column (one LDS load and one broadcast
// load per kTR multiply-adds); the l partials of the block are added in a
// fixed order, and k_toep_finish adds the chunks and applies the
// rate-and-state update.
static constexpr int kTR = 8;

__global__ void __launch_bounds__(256) k_toep_partial(const double *__restrict__ st,
                                                      const double *__restrict__ s, int64_t nx,
                                                      int64_t nxi, int lchunk,
                                                      double *__restrict__ part)
{
    extern __shared__ double tsh[];
    const int64_t j = blockIdx.x, c = blockIdx.y, nf = nx * nxi;
    const int64_t l0 = c * lchunk, l1 = (l0 + lchunk < nxi) ? l0 + lchunk : nxi;
    const int64_t nl = l1 - l0, W = 2 * nx - 1 + kTR;
    double *T2 = tsh;                       // lchunk x W
    double *S = T2 + (size_t)lchunk * W;    // lchunk x nx
    double *Rp = S + (size_t)lchunk * nx;   // lchunk x nx partial outputs
    for (int64_t x = threadIdx.x; x < nl * W; x += blockDim.x) {
        const int64_t l = x / W, m = x % W - (nx - 1);
        const int64_t am = m < 0 ? -m : m;
        T2[x] = am < nx ? st[am + nx * j + nf * (l0 + l)] : 0.0;
    }
    for (int64_t x = threadIdx.x; x < nl * nx; x += blockDim.x) S[x] = s[x % nx + nx * (l0 + x / nx)];
    __syncthreads();
    const int64_t nib = (nx + kTR - 1) / kTR;
    for (int64_t t = threadIdx.x; t < nib * nl; t += blockDim.x) {
        const int64_t l = t / nib, i = (t % nib) * kTR;
        const double *T = T2 + l * W + (nx - 1);  // T[m], m in [-(nx-1), nx-1+kTR]
        const double *Sl = S + l * nx;
        double w[kTR], acc[kTR];
#pragma unroll
        for (int r = 0; r < kTR; ++r) {
            w[r] = T[i + r];  // window for column k = 0: slot r holds T[i + r]
            acc[r] = 0.0;
        }
        for (int64_t k = 0; k < nx; k += kTR) {
#pragma unroll
            for (int kk = 0; kk < kTR; ++kk) {
                // column k + kk: slot (r - kk) mod kTR holds T[i + r - k - kk]
                if (k + kk < nx) {
                    const double sv = Sl[k + kk];
#pragma unroll
                    for (int r = 0; r < kTR; ++r) acc[r] += w[(r - kk + kTR) % kTR] * sv;
                    if (k + kk + 1 < nx) w[(kTR - 1 - kk) % kTR] = T[i - (k + kk) - 1];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kTR; ++r)
            if (i + r < nx) Rp[l * nx + i + r] = acc[r];
    }
    __syncthreads();
    for (int64_t i = threadIdx.x; i < nx; i += blockDim.x) {
        double acc = 0.0;
        for (int64_t l = 0; l < nl; ++l) acc += Rp[l * nx + i];
        part[c * nf + i + nx * j] = acc;
    }
}

__global__ void __launch_bounds__(256) k_toep_finish(RhsDev d, const double *__restrict__ part,
                                                     int nchunk, const double *__restrict__ u,
                                                     double *__restrict__ du)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nf = d.lay.nf;
    if (q >= nf) return;
    double tau = 0.0;
    for (int c = 0; c < nchunk; ++c) tau += part[c * nf + q];
    fault_update(d, q, tau, u, du);
}

// ---------------- DFT K_ff (fault-only problems, OQ_K_FFT) ----------------
// The reference's own algorithm (equation.jl:44-61): the source s(:, l) is
// zero-padded and transformed along strike, each frequency is contracted
// over the dip rows with the DFT of the circulant of st (GF.jl:60-67), and
// the inverse transform gives the first nx rows.  The reference pads to
// 2nx - 1; here the length is the power of two L >= 2nx - 1, where the
// circular convolution of the first nx rows is the same (the padding adds
// zeros between st[nx-1] and its mirror, so no product wraps around), and
// the transforms are radix-2 FFTs in LDS.  The padded circulant is still
// symmetric (x_m = x_{L-m}), so its DFT is real:
//   G[f, j, l] = sum_{m<L} x_m cos(2 pi f m / L),  f = 0..L/2.
// Layouts: G[f][l][j] (j fastest), S^[f][l] and T^[j][f] complex
// (interleaved re, im), f = 0..L/2; the twiddle table holds cos(2 pi k / L)
// then sin(2 pi k / L), k < L/2.  Three launches per RHS: the forward FFT
// (which forms s = v - vpl from the state itself, so no prologue runs), the
// per-frequency contraction, and the inverse FFT with the rate-and-state
// update.

static inline int64_t fft_len(int64_t nx)
{
    int64_t L = 2;  // at least one butterfly, so the twiddle table is never empty
    while (L < 2 * nx - 1) L <<= 1;
    return L;
}

// one-time: G from st (grid-stride over (L/2 + 1) * nxi^2 entries)
__global__ void __launch_bounds__(256) k_fft_kernel(const double *__restrict__ st, int64_t nx,
                                                    int64_t nxi, int64_t L, const double *__restrict__ tw,
                                                    double *__restrict__ G)
{
    const int64_t nf = nx * nxi, h = L / 2, tot = (h + 1) * nxi * nxi;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t % nxi, l = (t / nxi) % nxi, f = t / (nxi * nxi);
        const double *col = st + nx * j + nf * l;
        // x_0 + sum_{0<m<nx} x_m (cos(2 pi f m / L) + cos(2 pi f (L - m) / L)) = x_0 + 2 sum x_m cos
        double acc = 0.0;
        int64_t k = 0;
        for (int64_t m = 1; m < nx; ++m) {
            k += f;
            if (k >= L) k -= L;
            const int64_t kk = k < h ? k : L - k;  // cos(2 pi k / L) = cos(2 pi (L - k) / L)
            acc += col[m] * (k < h ? tw[kk] : (kk == h ? -1.0 : tw[kk]));
        }
        G[t] = col[0] + 2.0 * acc;
    }
}

// x1 <- x0 - w x1, x0 <- x0 + w x1 on registers
__device__ __forceinline__ void bfly(double &r0, double &i0, double &r1, double &i1, double wr, double wi)
{
    const double xr = r1 * wr - i1 * wi, xi = r1 * wi + i1 * wr;
    r1 = r0 - xr;
    i1 = i0 - xi;
    r0 += xr;
    i0 += xi;
}

// In-place radix-2 FFT of re/im[0..L) (bit-reversed input order), sign -1
// (forward) or +1 (inverse); the block's threads share the butterflies of
// each stage.  Stages len, 2len and 4len run together: a thread that owns
// the eight entries p + k len/2 (k < 8) of a 4len-block holds every pair of
// the three stages, so three stages cost one barrier (3 for L = 512 instead
// of 9); one or two stages left over run the same way with four or two
// entries per thread.
__device__ __forceinline__ void lds_fft(double *re, double *im, int64_t L, const double *cs,
                                        const double *sn, double sign)
{
    int64_t len = 2;
    for (; 4 * len <= L; len <<= 3) {
        const int64_t half = len >> 1, s1 = L / len, s2 = s1 >> 1, s4 = s1 >> 2;
        for (int64_t b = threadIdx.x; b < L / 8; b += blockDim.x) {
            const int64_t p = b & (half - 1), a0 = (b - p) * 8 + p;
            double r[8], i[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                r[k] = re[a0 + k * half];
                i[k] = im[a0 + k * half];
            }
            const double w1r = cs[p * s1], w1i = sign * sn[p * s1];
#pragma unroll
            for (int k = 0; k < 8; k += 2) bfly(r[k], i[k], r[k + 1], i[k + 1], w1r, w1i);  // stage len
#pragma unroll
            for (int k = 0; k < 2; ++k) {  // stage 2len: positions p, p + half
                const int64_t x = (p + k * half) * s2;
                bfly(r[k], i[k], r[k + 2], i[k + 2], cs[x], sign * sn[x]);
                bfly(r[k + 4], i[k + 4], r[k + 6], i[k + 6], cs[x], sign * sn[x]);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {  // stage 4len: positions p + k half
                const int64_t x = (p + k * half) * s4;
                bfly(r[k], i[k], r[k + 4], i[k + 4], cs[x], sign * sn[x]);
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                re[a0 + k * half] = r[k];
                im[a0 + k * half] = i[k];
            }
        }
        __syncthreads();
    }
    if (2 * len <= L) {  // two stages left: len and 2len
        const int64_t half = len >> 1, s1 = L / len, s2 = s1 >> 1;
        for (int64_t b = threadIdx.x; b < L / 4; b += blockDim.x) {
            const int64_t p = b & (half - 1), a0 = (b - p) * 4 + p, a1 = a0 + half, a2 = a0 + len,
                          a3 = a2 + half;
            double r0 = re[a0], i0 = im[a0], r1 = re[a1], i1 = im[a1];
            double r2 = re[a2], i2 = im[a2], r3 = re[a3], i3 = im[a3];
            const double w1r = cs[p * s1], w1i = sign * sn[p * s1];
            bfly(r0, i0, r1, i1, w1r, w1i);
            bfly(r2, i2, r3, i3, w1r, w1i);
            bfly(r0, i0, r2, i2, cs[p * s2], sign * sn[p * s2]);
            bfly(r1, i1, r3, i3, cs[(p + half) * s2], sign * sn[(p + half) * s2]);
            re[a0] = r0;
            im[a0] = i0;
            re[a1] = r1;
            im[a1] = i1;
            re[a2] = r2;
            im[a2] = i2;
            re[a3] = r3;
            im[a3] = i3;
        }
        __syncthreads();
    } else if (len <= L) {  // one stage left (len == L)
        const int64_t half = len >> 1, step = L / len;
        for (int64_t b = threadIdx.x; b < L / 2; b += blockDim.x) {
            const int64_t p = b & (half - 1), a0 = (b - p) * 2 + p, a1 = a0 + half;
            double r0 = re[a0], i0 = im[a0], r1 = re[a1], i1 = im[a1];
            bfly(r0, i0, r1, i1, cs[p * step], sign * sn[p * step]);
            re[a0] = r0;
            im[a0] = i0;
            re[a1] = r1;
            im[a1] = i1;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int64_t bitrev(int64_t x, int bits)
{
    return (int64_t)(__brev((unsigned)x) >> (32 - bits));  // bits >= 1 since L >= 2
}

// relative_velocity! (equation.jl:35-42) and the forward FFT of the padded
// s(:, l): block l; bins 0..L/2 to S^[f][l]
__global__ void __launch_bounds__(256) k_fft_fwd(RhsDev d, const double *__restrict__ u, int64_t nx,
                                                 int64_t nxi, int64_t L, int bits,
                                                 const double *__restrict__ tw, double *__restrict__ Sh)
{
    extern __shared__ double fsm[];
    const int64_t l = blockIdx.x, h = L / 2;
    double *re = fsm, *im = re + L, *cs = im + L, *sn = cs + h;
    for (int64_t x = threadIdx.x; x < h; x += blockDim.x) {
        cs[x] = tw[x];
        sn[x] = tw[h + x];
    }
    for (int64_t n = threadIdx.x; n < L; n += blockDim.x) {
        const int64_t r = bitrev(n, bits);
        re[r] = n < nx ? u[d.lay.v + n + nx * l] - d.vpl : 0.0;
        im[r] = 0.0;
    }
    __syncthreads();
    lds_fft(re, im, L, cs, sn, -1.0);
    for (int64_t f = threadIdx.x; f <= h; f += blockDim.x) {
        Sh[2 * (l + nxi * f)] = re[f];
        Sh[2 * (l + nxi * f) + 1] = im[f];
    }
}

// A Tsit5 stage fused with k_fft_fwd (fault-only state [v th del], graph
// stepper): block l forms y = u + dt sum_{j<s} a_j k_j for the 3 nx entries
// of dip row l, with the operations of the stepper's stage kernel, stores y
// for the inverse pass and the rest of the step, and transforms s = y_v -
// vpl.  *done set: y is left as it is (the stage kernel's skip) and s is
// formed from it.
__global__ void __launch_bounds__(256) k_fft_fwd_stage(RhsDev d, StageCoef sc, const double *__restrict__ dtp,
                                                       const int *__restrict__ done, const double *__restrict__ u,
                                                       const double *__restrict__ k, int64_t n,
                                                       double *__restrict__ y, int64_t nx, int64_t nxi, int64_t L,
                                                       int bits, const double *__restrict__ tw,
                                                       double *__restrict__ Sh)
{
    extern __shared__ double fsm[];
    const int64_t l = blockIdx.x, h = L / 2;
    double *re = fsm, *im = re + L, *cs = im + L, *sn = cs + h, *yv = sn + h;  // yv: nx
    for (int64_t x = threadIdx.x; x < h; x += blockDim.x) {
        cs[x] = tw[x];
        sn[x] = tw[h + x];
    }
    const bool skip = *done != 0;
    const int64_t off[3] = {d.lay.v, d.lay.th, d.lay.del};
    const double dt = *dtp;
    for (int64_t x = threadIdx.x; x < nx; x += blockDim.x) {
        // the three components of patch x: every load issued before the sums
        // (k_j past the stage index s re-reads k_{s-1} and is not added)
        double uv[3], kv[3][6];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int64_t i = off[c] + x + nx * l;
            uv[c] = skip ? y[i] : u[i];
#pragma unroll
            for (int j = 0; j < 6; ++j) kv[c][j] = k[(j < sc.s ? j : sc.s - 1) * n + i];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double yi = uv[c];
            if (!skip) {
                double acc = 0.0;
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    if (j < sc.s) acc += sc.a[j] * kv[c][j];
                yi = uv[c] + dt * acc;
                y[off[c] + x + nx * l] = yi;
            }
            if (c == 0) yv[x] = yi;
        }
    }
    __syncthreads();
    for (int64_t x = threadIdx.x; x < L; x += blockDim.x) {
        const int64_t r = bitrev(x, bits);
        re[r] = x < nx ? yv[x] - d.vpl : 0.0;
        im[r] = 0.0;
    }
    __syncthreads();
    lds_fft(re, im, L, cs, sn, -1.0);
    for (int64_t f = threadIdx.x; f <= h; f += blockDim.x) {
        Sh[2 * (l + nxi * f)] = re[f];
        Sh[2 * (l + nxi * f) + 1] = im[f];
    }
}

// per-frequency contraction T^[j][f] = sum_l G[f][l][j] S^[f][l]: block f,
// eight l-partials per row so that a thread's loads go out together
__global__ void __launch_bounds__(512) k_fft_mid(const double *__restrict__ G,
                                                 const double *__restrict__ Sh, int64_t nbin,
                                                 int64_t nxi, double *__restrict__ Th)
{
    __shared__ double red[2][512];
    const int64_t f = blockIdx.x;
    const int tid = threadIdx.x, ji = tid & 63, q = tid >> 6;
    const double *Gf = G + nxi * nxi * f;
    const double *Sf = Sh + 2 * nxi * f;
    for (int64_t jb = 0; jb < nxi; jb += 64) {
        const int64_t j = jb + ji;
        double re = 0.0, im = 0.0;
        if (j < nxi) {
#pragma unroll 8
            for (int64_t l = q; l < nxi; l += 8) {
                const double g = Gf[j + nxi * l];
                re += g * Sf[2 * l];
                im += g * Sf[2 * l + 1];
            }
        }
        red[0][tid] = re;
        red[1][tid] = im;
        __syncthreads();
        if (q < 2 && j < nxi) {  // wave 0 the real parts, wave 1 the imaginary parts
            const double *r = red[q] + ji;
            Th[2 * (f + nbin * j) + q] = ((r[0] + r[64]) + (r[128] + r[192])) + ((r[256] + r[320]) + (r[384] + r[448]));
        }
        __syncthreads();
    }
}

// inverse FFT of the Hermitian spectrum of row j (bins 0..L/2 given), the
// first nx rows, and the rate-and-state update: block j
__global__ void __launch_bounds__(256) k_fft_inv(RhsDev d, const double *__restrict__ Th, int64_t nx,
                                                 int64_t L, int bits, const double *__restrict__ tw,
                                                 const double *__restrict__ u, double *__restrict__ du)
{
    extern __shared__ double fsm[];
    const int64_t j = blockIdx.x, h = L / 2;
    double *re = fsm, *im = re + L, *cs = im + L, *sn = cs + h;
    for (int64_t x = threadIdx.x; x < h; x += blockDim.x) {
        cs[x] = tw[x];
        sn[x] = tw[h + x];
    }
    const double *T = Th + 2 * (h + 1) * j;
    for (int64_t f = threadIdx.x; f < L; f += blockDim.x) {
        const int64_t r = bitrev(f, bits);
        if (f <= h) {
            re[r] = T[2 * f];
            im[r] = T[2 * f + 1];
        } else {  // X[L - f] = conj(X[f])
            re[r] = T[2 * (L - f)];
            im[r] = -T[2 * (L - f) + 1];
        }
    }
    __syncthreads();
    lds_fft(re, im, L, cs, sn, 1.0);
    // fault only: the rate-and-state update; coupled (ne > 0): the K_ff
    // traction rates go to du = tauff for the fault rows' K_fm pass
    for (int64_t n = threadIdx.x; n < nx; n += blockDim.x) {
        if (d.lay.ne > 0) du[n + nx * j] = re[n] / (double)L;
        else fault_update(d, n + nx * j, re[n] / (double)L, u, du);
    }
}

// ---------------- grid K_mm (coupled DFT model, OQ_K_FFT_GRID) ----------------
// Over a mantle whose cells form a uniform horizontal grid (nx x ny cells per
// layer, nz layers, cell ix + nx (iy + ny iz)) the half-space stress at
// receiver cell (jx, jy, jz) from source cell (ix, iy, iz) depends on
// (jx - ix, jy - iy, jz, iz) only (the free surface is horizontal).  K_mm is
// kept as tiles T[PO][PI] over output planes PO = k nz + jz and source
// planes PI = p nz + iz (stress component k, eigenstrain component p), row
// yy = jy - iy + ny - 1 in [0, 2ny - 1), and applied as 2D convolutions:
//   out[PO][jy][jx] = sum_PI sum_{iy, ix} T[PO][PI][jy - iy][jx - ix] s[PI][iy][ix].
// A tile row holds the offsets m = jx - ix at m + nxp for m in [-nxp, nxp]
// (Wp = 2 nxp + 1 entries, nxp = nx rounded up to R), zero where |m| >= nx,
// so that a block stages it with straight copies and the sliding window
// below runs over nxp source columns (s zero-padded the same way) without
// bounds tests; the odd pitch puts the rows of the 16 lanes of one LDS
// access phase (4 rows x 4 windows) on distinct banks.
__host__ __device__ inline int grid_nxp(int64_t nx, int r) { return (int)((nx + r - 1) / r * r); }

// Block (PO, chunk c) stages the tiles of its tp source planes and those
// planes of s in LDS; thread (plane q, row jy, R consecutive jx) forms its
// outputs with a register window over the tile row that slides by one entry
// per source column (one LDS load and one broadcast load per R
// multiply-adds, as in k_toep_partial); the tp plane sums are added in a
// fixed order into part[c] (c-th partial of every mantle row).
template <int R>
__device__ __forceinline__ void mm_grid_conv_block(const double *__restrict__ T, const double *__restrict__ s,
                                                   int nx, int ny, int nz, int tp, double *__restrict__ part,
                                                   int PO, int c, double *gsh)
{
    const int npl = 6 * nz, nxy = nx * ny;
    const int pi0 = c * tp, np = (pi0 + tp < npl ? pi0 + tp : npl) - pi0;
    const int nxp = grid_nxp(nx, R), Wp = 2 * nxp + 1, Wy = 2 * ny - 1, tile = Wy * Wp;
    double *Ts = gsh;                 // tp x Wy x Wp
    double *Ss = Ts + tp * tile;      // tp x ny x nxp
    double *Os = Ss + tp * ny * nxp;  // tp x ny x nxp
    {
        const double *src = T + (size_t)(PO * npl + pi0) * tile;
        for (int x = threadIdx.x; x < np * tile; x += blockDim.x) Ts[x] = src[x];
    }
    for (int x = threadIdx.x; x < np * ny * nxp; x += blockDim.x) {
        const int row = x / nxp, ix = x - row * nxp;  // row = q ny + iy
        Ss[x] = ix < nx ? s[(size_t)pi0 * nxy + row * nx + ix] : 0.0;
    }
    __syncthreads();
    const int nxg = nxp / R, U = ny * nxg;
    for (int t = threadIdx.x; t < np * U; t += blockDim.x) {
        const int q = t / U, g = t - q * U, jy = g / nxg, jx0 = (g - jy * nxg) * R;
        double acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = 0.0;
        for (int iy = 0; iy < ny; ++iy) {
            const double *Tr = Ts + (q * Wy + (jy - iy + ny - 1)) * Wp + nxp;  // Tr[m], m in [-nxp, nxp)
            const double *Sr = Ss + (q * ny + iy) * nxp;
            double w[R];
#pragma unroll
            for (int r = 0; r < R; ++r) w[r] = Tr[jx0 + r];  // column 0: slot r holds Tr[jx0 + r]
            for (int ix = 0; ix < nxp; ix += R) {
#pragma unroll
                for (int kk = 0; kk < R; ++kk) {
                    // column ix + kk: slot (r - kk) mod R holds Tr[jx0 + r - ix - kk]
                    const double sv = Sr[ix + kk];
#pragma unroll
                    for (int r = 0; r < R; ++r) acc[r] += w[(r - kk + R) % R] * sv;
                    w[(R - 1 - kk) % R] = Tr[jx0 - (ix + kk) - 1];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) Os[(q * ny + jy) * nxp + jx0 + r] = acc[r];
    }
    __syncthreads();
    for (int o = threadIdx.x; o < nxy; o += blockDim.x) {
        const int jy = o / nx, jx = o - jy * nx;
        double a = 0.0;
        for (int q = 0; q < np; ++q) a += Os[(q * ny + jy) * nxp + jx];
        part[((size_t)c * npl + PO) * nxy + o] = a;
    }
}

// The padded tiles from the pair kernel's block over the shifted receivers
// (receiver xx + Wx (yy + Wy jz), Wx = 2nx - 1, offset m = xx - (nx - 1);
// row k nrec + receiver) and the source column (cells (0, 0, iz), column
// p nz + iz, row pitch 6 nz).
__global__ void __launch_bounds__(256) k_mm_grid_table(const double *__restrict__ raw, int64_t nx,
                                                       int64_t ny, int64_t nz, int64_t nxp,
                                                       double *__restrict__ T)
{
    const int64_t Wx = 2 * nx - 1, Wy = 2 * ny - 1, Wp = 2 * nxp + 1, npl = 6 * nz, nrec = Wx * Wy * nz;
    const int64_t n = npl * npl * Wy * Wp;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += step) {
        const int64_t xp = t % Wp, r1 = t / Wp, yy = r1 % Wy, r2 = r1 / Wy;
        const int64_t PI = r2 % npl, PO = r2 / npl, k = PO / nz, jz = PO % nz;
        const int64_t m = xp - nxp;
        T[t] = (m > -nx && m < nx) ? raw[(k * nrec + (m + nx - 1) + Wx * (yy + Wy * jz)) * npl + PI] : 0.0;
    }
}

// The dense K_mm block (row pitch ld) from the tiles (operator expansion).
__global__ void __launch_bounds__(256) k_expand_mm_grid(const double *__restrict__ T, int64_t nx,
                                                        int64_t ny, int64_t nz, int64_t nxp,
                                                        double *__restrict__ k, int64_t ld)
{
    const int64_t nxy = nx * ny, m6 = 6 * nz * nxy, npl = 6 * nz, Wp = 2 * nxp + 1, Wy = 2 * ny - 1;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m6 * m6; t += step) {
        const int64_t row = t / m6, col = t % m6;
        const int64_t PO = row / nxy, o = row % nxy, PI = col / nxy, o2 = col % nxy;
        const int64_t jy = o / nx, jx = o % nx, iy = o2 / nx, ix = o2 % nx;
        k[row * ld + col] = T[((PO * npl + PI) * Wy + (jy - iy + ny - 1)) * Wp + (jx - ix + nxp)];
    }
}

// The grid model's dense blocks: wavefronts [0, nfw) stream the fault rows'
// K_fm (pitch ldm) over s_m in S column chunks of cw columns into partial
// sums fpart[(1 + c) nf + row] (fpart[row] holds the K_ff traction rates),
// the rest the mantle rows' K_mf (pitch ldf) over s_f into mvout.  cw is
// chosen near ldf, so every wavefront streams about the same bytes.
struct GridMv {
    const double *Kf, *K, *sf, *sm;
    int64_t ldm, ldf, cw;
    int S;
    double *fpart, *mvout;
};

__device__ __forceinline__ void grid_matvec_wave(const RhsDev &d, const GridMv &a, int64_t gw, int lane)
{
    constexpr int KT = OQ_K_FP64;
    constexpr int R = KTraits<KT>::kRows;
    constexpr int64_t STEP = 64 * KTraits<KT>::kPerLane;
    using V = typename KVec<KT>::V;
    const int64_t nf = d.lay.nf, m6 = 6 * d.lay.ne, nfw = (nf + R - 1) / R * a.S;
    const V *rp[R];
    const double *s;
    int64_t row0, nrows, npass, c = -1;
    if (gw < nfw) {
        c = gw % a.S;
        row0 = (gw / a.S) * R;
        nrows = nf;
        const int64_t c0 = c * a.cw, ncol = (c0 + a.cw < a.ldm ? c0 + a.cw : a.ldm) - c0;
        npass = ncol / STEP;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t row = row0 + r < nf ? row0 + r : nf - 1;
            rp[r] = reinterpret_cast<const V *>(a.Kf + row * a.ldm + c0) + lane;
        }
        s = a.sm + c0;
    } else {
        row0 = (gw - nfw) * R;
        if (row0 >= m6) return;
        nrows = m6;
        npass = a.ldf / STEP;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t row = row0 + r < m6 ? row0 + r : m6 - 1;
            rp[r] = reinterpret_cast<const V *>(a.K + row * a.ldf) + lane;
        }
        s = a.sf;
    }
    double acc[R];
    stream_rows<KT, R>(rp, s, npass, (gw * 13) % npass, lane, acc);
    if (lane < R) {
        double y = acc[0];
#pragma unroll
        for (int r = 1; r < R; ++r)
            if (lane == r) y = acc[r];
        const int64_t loc = row0 + lane;
        if (loc < nrows) {
            if (c >= 0) a.fpart[(1 + c) * nf + loc] = y;
            else a.mvout[loc] = y;
        }
    }
}

// Along-strike tables of the grid model (see MmGrid).  F blocks (fault row
// b, lines [c fqs, (c + 1) fqs)) stage fqb lines of F (rows padded to Fp
// entries with zeros) and of s_m at a time; thread (line slot li, residue
// a0 = a mod m, segment sg) accumulates the kRW outputs a = a0 + m (sg kRW +
// r) with a register window over F[.][a - m jx + m (nxc - 1)] that slides
// by one slot (m entries) per source column jx, one LDS load and one
// broadcast load per kRW multiply-adds; the line slots are added in a
// fixed order into fpart[(1 + c) nf + a + nx b].  G blocks (mantle line q',
// fault rows [c gbs, (c + 1) gbs)) stage gbb rows of G (m leading zeros,
// rows padded to Gp) and of s_f (rows padded to m nr) at a time; thread
// (segment sg, row slot bi, residue a0) accumulates the kRW outputs jx = sg
// kRW + r over a = a0 + m k with the window over G[.][m jx - a + nx - 1]
// sliding by one slot per k; wavefront sums, then the four wavefronts'
// sums in a fixed order, go to gpart[c 6ne + q' nxc + jx].
static constexpr int kRW = 8;
static constexpr size_t kGridOtherLds = 40 * 1024;  // LDS of an F or G block (4 per CU)

struct GridSt {
    const double *F, *G, *sf, *sm;
    int nx, nxi, nxc, Q, m, D;
    int fqs, fqb, gbs, gbb;
    int nr, fseg, gseg, Fp, Gp;  // nr = ceil(nx / m); segments of kRW outputs; padded row widths
    int flpp;                    // F: lines in flight (<= 256 / (m fseg))
    int sw;                      // G: LDS pitch of the s_f rows (>= m nr)
    int64_t nf, m6;
    double *fpart, *gpart;
};

__device__ __forceinline__ void grid_f_block(const GridSt &g, int b, int c, double *sh)
{
    const int nx = g.nx, D = g.D, nxc = g.nxc, m = g.m, Q = g.Q, Fp = g.Fp;
    const int q0 = c * g.fqs, q1 = (q0 + g.fqs < Q ? q0 + g.fqs : Q);
    const int tpl = m * g.fseg, lpp = g.flpp;  // line slots
    const int t = threadIdx.x, li = t / tpl, rr = t - li * tpl, sg = rr / m, a0 = rr - sg * m;
    const bool active = li < lpp;
    double acc[kRW];
#pragma unroll
    for (int r = 0; r < kRW; ++r) acc[r] = 0.0;
    double *Fs = sh, *Ss = sh + g.fqb * Fp;
    for (int qb = q0; qb < q1; qb += g.fqb) {
        const int nq = (g.fqb < q1 - qb ? g.fqb : q1 - qb);
        __syncthreads();
        for (int x = t; x < nq * Fp; x += 256) {
            const int q = x / Fp, dd = x - q * Fp;
            Fs[x] = dd < D ? g.F[((size_t)b * Q + qb + q) * D + dd] : 0.0;
        }
        for (int x = t; x < nq * nxc; x += 256) Ss[x] = g.sm[(size_t)qb * nxc + x];
        __syncthreads();
        if (active)
            for (int q = li; q < nq; q += lpp) {
                // Fq[m x] = F[q][a0 + m sg kRW + m (nxc - 1) + m x]; output r, column jx: x = r - jx
                const double *Fq = Fs + q * Fp + a0 + m * (sg * kRW) + m * (nxc - 1);
                const double *Sq = Ss + q * nxc;
                double w[kRW];
#pragma unroll
                for (int r = 0; r < kRW; ++r) w[r] = Fq[m * r];
                for (int j0 = 0; j0 < nxc; j0 += kRW) {
#pragma unroll
                    for (int kk = 0; kk < kRW; ++kk) {
                        // column jx = j0 + kk: slot (r - kk) mod kRW holds Fq[m (r - jx)]
                        const int jx = j0 + kk;
                        if (jx < nxc) {
                            const double sv = Sq[jx];
#pragma unroll
                            for (int r = 0; r < kRW; ++r) acc[r] += w[(r - kk + kRW) % kRW] * sv;
                            if (jx + 1 < nxc) w[(kRW - 1 - kk) % kRW] = Fq[-m * (jx + 1)];
                        }
                    }
                }
            }
    }
    __syncthreads();
    double *Rs = sh;  // lpp x nx partial sums
    if (active)
#pragma unroll
        for (int r = 0; r < kRW; ++r) {
            const int a = a0 + m * (sg * kRW + r);
            if (a < nx) Rs[li * nx + a] = acc[r];
        }
    __syncthreads();
    for (int a = t; a < nx; a += 256) {
        double y = 0.0;
        for (int l = 0; l < lpp; ++l) y += Rs[l * nx + a];
        g.fpart[(size_t)(1 + c) * g.nf + a + (size_t)nx * b] = y;
    }
}

__device__ __forceinline__ void grid_g_block(const GridSt &g, int qq, int c, double *sh)
{
    const int nx = g.nx, D = g.D, nxc = g.nxc, m = g.m, nxi = g.nxi, Gp = g.Gp, nr = g.nr, sw = g.sw;
    const int b0 = c * g.gbs, b1 = (b0 + g.gbs < nxi ? b0 + g.gbs : nxi);
    // a segment takes whole wavefronts (wps of them); in it, thread (row slot bi, residue a0)
    const int wps = 4 / g.gseg, tps = 64 * wps, bp = tps / m;
    const int t = threadIdx.x, sg = t / tps, x0 = t - sg * tps, bi = x0 / m, a0 = x0 - bi * m;
    const int w = t >> 6, lane = t & 63;
    const bool active = sg < g.gseg && bi < bp;
    double acc[kRW];
#pragma unroll
    for (int r = 0; r < kRW; ++r) acc[r] = 0.0;
    double *Gs = sh, *Sf = sh + g.gbb * Gp;
    for (int bb = b0; bb < b1; bb += g.gbb) {
        const int nb = (g.gbb < b1 - bb ? g.gbb : b1 - bb);
        __syncthreads();
        for (int x = t; x < nb * Gp; x += 256) {
            const int b = x / Gp, e = x - b * Gp - m;
            Gs[x] = (e >= 0 && e < D) ? g.G[((size_t)qq * nxi + bb + b) * D + e] : 0.0;
        }
        for (int x = t; x < nb * sw; x += 256) {
            const int b = x / sw, a = x - b * sw;
            Sf[x] = a < nx ? g.sf[(size_t)(bb + b) * nx + a] : 0.0;  // a < m nr <= sw read below
        }
        __syncthreads();
        for (int b = bi; active && b < nb; b += bp) {
            // Gq[m x] = G[b][m (sg kRW + x) - a0 + nx - 1] (m leading zeros); output r, step k: x = r - k
            const double *Gq = Gs + b * Gp + m + (nx - 1 - a0) + m * (sg * kRW);
            const double *Sq = Sf + b * sw + a0;
            double wv[kRW];
#pragma unroll
            for (int r = 0; r < kRW; ++r) wv[r] = Gq[m * r];
            for (int k0 = 0; k0 < nr; k0 += kRW) {
#pragma unroll
                for (int kk = 0; kk < kRW; ++kk) {
                    const int k = k0 + kk;
                    if (k < nr) {
                        const double sv = Sq[m * k];
#pragma unroll
                        for (int r = 0; r < kRW; ++r) acc[r] += wv[(r - kk + kRW) % kRW] * sv;
                        if (k + 1 < nr) wv[(kRW - 1 - kk) % kRW] = Gq[-m * (k + 1)];
                    }
                }
            }
        }
    }
    // wavefront sums (a wavefront holds one segment), then the wavefronts of a segment in order
    __syncthreads();
    double *Rs = sh;  // 4 x kRW
#pragma unroll
    for (int r = 0; r < kRW; ++r) {
        const double y = wave_sum(acc[r]);  // every thread takes part; idle ones hold zeros
        if (lane == 0) Rs[w * kRW + r] = y;
    }
    __syncthreads();
    for (int o = t; o < g.gseg * kRW; o += 256) {
        const int s2 = o / kRW, r = o - s2 * kRW, jx = s2 * kRW + r;
        if (jx < nxc) {
            double y = 0.0;
            for (int v = 0; v < wps; ++v) y += Rs[(s2 * wps + v) * kRW + r];
            g.gpart[(size_t)c * g.m6 + (size_t)qq * nxc + jx] = y;
        }
    }
}

// One launch for the grid model's O(N^2) work.  roles[block] = role << 28 |
// index: 0 a K_mm convolution block (PO = index mod 6nz, chunk index / 6nz),
// 1 four wavefronts of the dense K_fm / K_mf stream, 2 an F block (fault row
// index mod nxi, chunk index / nxi), 3 a G block (line index mod Q, chunk
// index / Q).  The host interleaves the roles evenly, so the LDS and FP64
// work of the convolutions runs beside the HBM streams.
// MASK selects the roles compiled into an instance (bit r for role r), so
// that a launch of one kind of block gets that role's register budget.
template <int R, int MASK>
__global__ void __launch_bounds__(256) k_grid_fused(RhsDev d, GridMv a, GridSt g2, const double *__restrict__ T,
                                                    int nx, int ny, int nz, int tp, double *__restrict__ mpart,
                                                    const unsigned *__restrict__ roles)
{
    extern __shared__ double gsh[];
    const unsigned rv = roles[blockIdx.x];
    const int role = (int)(rv >> 28), idx = (int)(rv & 0x0fffffffu);
    if constexpr ((MASK & 1) != 0) {
        if (role == 0) {
            const int npl = 6 * nz;
            mm_grid_conv_block<R>(T, a.sm, nx, ny, nz, tp, mpart, idx % npl, idx / npl, gsh);
            return;
        }
    }
    if constexpr ((MASK & 2) != 0) {
        if (role == 1) {
            grid_matvec_wave(d, a, (int64_t)idx * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
            return;
        }
    }
    if constexpr ((MASK & 4) != 0) {
        if (role == 2) {
            grid_f_block(g2, idx % g2.nxi, idx / g2.nxi, gsh);
            return;
        }
    }
    if constexpr ((MASK & 8) != 0) {
        if (role == 3) grid_g_block(g2, idx % g2.Q, idx / g2.Q, gsh);
    }
}


// Fault rows: the K_ff rates and K_fm chunk partials, then the friction
// update; mantle rows: du.sigma = K_mf s_f + the convolution partials.
__global__ void __launch_bounds__(256) k_grid_finish(RhsDev d, const double *__restrict__ fpart, int nfp,
                                                     const double *__restrict__ mvout, int nmv,
                                                     const double *__restrict__ mpart, int nchunk,
                                                     const double *__restrict__ u, double *__restrict__ du)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nf = d.lay.nf, m6 = 6 * d.lay.ne;
    if (q < nf) {
        double tau = 0.0;
        for (int c = 0; c < nfp; ++c) tau += fpart[c * nf + q];
        fault_update(d, q, tau, u, du);
    } else if (q < nf + m6) {
        const int64_t r = q - nf;
        double y = 0.0;
        for (int c = 0; c < nmv; ++c) y += mvout[c * m6 + r];
        for (int c = 0; c < nchunk; ++c) y += mpart[c * m6 + r];
        du[d.lay.sig + r] = y;
    }
}

// F[b][q][d] from the fm pair kernel's block over the shifted virtual
// patches (receiver d + D b, row pitch Q, column q).
__global__ void __launch_bounds__(256) k_grid_ftable(const double *__restrict__ raw, int64_t nxi, int64_t Q,
                                                     int64_t D, double *__restrict__ F)
{
    const int64_t n = nxi * Q * D, step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += step) {
        const int64_t dd = t % D, r = t / D, q = r % Q, b = r / Q;
        F[t] = raw[(dd + D * b) * Q + q];
    }
}

// G[q'][b][e] from the mf pair kernel's block over the shifted virtual cells
// (receiver e + D (jy + ny jz), row (k ne' + receiver), row pitch nxi,
// column b).
__global__ void __launch_bounds__(256) k_grid_gtable(const double *__restrict__ raw, int64_t nxi, int64_t ny,
                                                     int64_t nz, int64_t D, double *__restrict__ G)
{
    const int64_t Q = 6 * ny * nz, nev = D * ny * nz, n = Q * nxi * D;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += step) {
        const int64_t e = t % D, r = t / D, b = r % nxi, qq = r / nxi;
        const int64_t k = qq / (ny * nz), rem = qq % (ny * nz), jz = rem / ny, jy = rem % ny;
        G[t] = raw[(k * nev + e + D * (jy + ny * jz)) * nxi + b];
    }
}

// Dense K_fm (fault rows, mantle columns from nf) and K_mf (mantle rows,
// fault columns) of a receiver-major operator (pitch ld) from the tables.
__global__ void __launch_bounds__(256) k_expand_strike(const double *__restrict__ F, const double *__restrict__ G,
                                                       int64_t nx, int64_t nxi, int64_t nxc, int64_t Q, int64_t m,
                                                       int64_t D, double *__restrict__ k, int64_t ld)
{
    const int64_t nf = nx * nxi, m6 = Q * nxc, step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 2 * nf * m6; t += step) {
        if (t < nf * m6) {  // K_fm: row f = a + nx b, column nf + q nxc + jx
            const int64_t f = t / m6, col = t % m6, a = f % nx, b = f / nx, q = col / nxc, jx = col % nxc;
            k[f * ld + nf + col] = F[(b * Q + q) * D + a - m * jx + m * (nxc - 1)];
        } else {            // K_mf: row nf + q' nxc + jx, column a + nx b
            const int64_t u = t - nf * m6, row = u / nf, f = u % nf, a = f % nx, b = f / nx;
            const int64_t qq = row / nxc, jx = row % nxc;
            k[(nf + row) * ld + f] = G[(qq * nxi + b) * D + m * jx - a + nx - 1];
        }
    }
}

static inline size_t fft_lds(int64_t nx) { return (size_t)(3 * fft_len(nx)) * 8; }

// LDS bytes of a block with lchunk dip rows
static inline int64_t toep_lds(const oq_rhs *h, int64_t lc)
{
    return lc * (2 * h->nx - 1 + kTR + 2 * h->nx) * 8;
}

// chunking of the dip rows: enough blocks to fill the device, LDS <= 64 KB
static void toep_shape(const oq_rhs *h, int *lchunk, int *nchunk, int *threads)
{
    int nc = (int)((1024 + h->nxi - 1) / h->nxi);
    if (nc > h->nxi) nc = (int)h->nxi;
    int lc = (int)((h->nxi + nc - 1) / nc);
    while (lc > 1 && toep_lds(h, lc) > 65536) --lc;
    *lchunk = lc;
    *nchunk = (int)((h->nxi + lc - 1) / lc);
    const int64_t t = ((h->nx + kTR - 1) / kTR) * lc;
    const int64_t t64 = (t + 63) / 64 * 64;
    *threads = (int)(t64 < 64 ? 64 : (t64 > 256 ? 256 : t64));
}

template <int R, int MASK>
static void grid_launch1(const oq_rhs *h, int nblocks, size_t lds, const unsigned *roles, const GridMv &a,
                         const GridSt &g2)
{
    const MmGrid &g = h->grid;
    hipLaunchKernelGGL((k_grid_fused<R, MASK>), dim3((unsigned)nblocks), dim3(256), lds, h->stream, h->dev, a, g2,
                       h->Tmm.as<double>(), (int)g.nx, (int)g.ny, (int)g.nz, g.tp, h->mpart.as<double>(), roles);
}

template <int R>
static void grid_launch_r(const oq_rhs *h, const GridMv &a, const GridSt &g2)
{
    const MmGrid &g = h->grid;
    const unsigned *r1 = h->roles.as<unsigned>(), *r2 = h->roles2.as<unsigned>();
    if (!g.split) {
        if (g.strike) grid_launch1<R, 1 | 4 | 8>(h, g.nblocks, g.lds_fused, r1, a, g2);
        else grid_launch1<R, 1 | 2>(h, g.nblocks, g.lds_fused, r1, a, g2);
        return;
    }
    grid_launch1<R, 1>(h, g.nconv, g.lds, r2, a, g2);
    if (g.strike) grid_launch1<R, 4 | 8>(h, g.nblocks - g.nconv, g.lds_other, r2 + g.nconv, a, g2);
    else grid_launch1<R, 2>(h, g.nblocks - g.nconv, g.lds_other, r2 + g.nconv, a, g2);
}

template <int R>
static int grid_lds_attrs(size_t lds)
{
    const void *f[5] = {(const void *)k_grid_fused<R, 1 | 4 | 8>, (const void *)k_grid_fused<R, 1 | 2>,
                        (const void *)k_grid_fused<R, 1>, (const void *)k_grid_fused<R, 4 | 8>,
                        (const void *)k_grid_fused<R, 2>};
    for (const void *p : f) OQ_HIP(hipFuncSetAttribute(p, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    return OQ_OK;
}

int rhs_enqueue_matvec(oq_rhs *h, const double *u, double *du)
{
    if (h->kdtype == OQ_K_FFT && h->ne > 0) {
        // coupled model (equation.jl:185-205 with dτ_dt! by DFT, :200): K_ff s_f
        // by the DFT convolution into tauff, then the fault rows' K_fm s_m plus
        // tauff with the rate-and-state update, and the mantle rows [K_mf K_mm] s
        const int64_t L = fft_len(h->nx);
        int bits = 0;
        while (((int64_t)1 << bits) < L) ++bits;
        hipLaunchKernelGGL(k_fft_fwd, dim3((unsigned)h->nxi), dim3(256), fft_lds(h->nx), h->stream, h->dev, u,
                           h->nx, h->nxi, L, bits, h->ftw.as<double>(), h->fsh.as<double>());
        hipLaunchKernelGGL(k_fft_mid, dim3((unsigned)(L / 2 + 1)), dim3(512), 0, h->stream, h->fg.as<double>(),
                           h->fsh.as<double>(), L / 2 + 1, h->nxi, h->fth.as<double>());
        if (h->mmgrid) {
            // K_ff rates into fpart[0, nf); the grid convolution of K_mm with K_fm and K_mf
            // (column-chunked dense streams, or along strike the F and G tables) in one launch;
            // then the friction update and du.sigma
            const MmGrid &g = h->grid;
            hipLaunchKernelGGL(k_fft_inv, dim3((unsigned)h->nxi), dim3(256), fft_lds(h->nx), h->stream, h->dev,
                               h->fth.as<double>(), h->nx, L, bits, h->ftw.as<double>(), u, h->fpart.as<double>());
            static_assert(KTraits<OQ_K_FP64>::kWaves == 4, "k_grid_fused runs four matvec wavefronts per block");
            const GridMv a = {h->Kf.as<double>(), h->K.as<double>(), h->s64.as<double>(), h->sm.as<double>(),
                              h->ldm, h->ldf, g.fcw, g.fs, h->fpart.as<double>(), h->mvout.as<double>()};
            const GridSt g2 = {h->Ftab.as<double>(), h->Gtab.as<double>(), h->s64.as<double>(), h->sm.as<double>(),
                               (int)h->nx, (int)h->nxi, (int)g.nx, (int)(6 * g.ny * g.nz), (int)g.m, (int)g.D,
                               g.fqs, g.fqb, g.gbs, g.gbb, g.nr, g.fseg, g.gseg, g.Fp, g.Gp, g.flpp, g.sw, h->nf, 6 * h->ne,
                               h->fpart.as<double>(), h->mvout.as<double>()};
            if (g.r == 8) grid_launch_r<8>(h, a, g2);
            else grid_launch_r<4>(h, a, g2);
            hipLaunchKernelGGL(k_grid_finish, dim3(nblk(h->nf + 6 * h->ne, 256)), dim3(256), 0, h->stream, h->dev,
                               h->fpart.as<double>(), 1 + (g.strike ? g.nfc : g.fs), h->mvout.as<double>(),
                               g.strike ? g.ngc : 1, h->mpart.as<double>(), g.nchunk, u, du);
            OQ_HIP(hipGetLastError());
            return OQ_OK;
        }
        hipLaunchKernelGGL(k_fft_inv, dim3((unsigned)h->nxi), dim3(256), fft_lds(h->nx), h->stream, h->dev,
                           h->fth.as<double>(), h->nx, L, bits, h->ftw.as<double>(), u, h->tauff.as<double>());
        launch_matvec<OQ_K_FP64>(h, h->Kf.p, h->ldm, h->nf, 0, h->sm.as<double>(), h->tauff.as<double>(), u, du);
        launch_matvec<OQ_K_FP64>(h, h->K.p, h->ld, 6 * h->ne, h->nf, h->s64.as<double>(), nullptr, u, du);
        OQ_HIP(hipGetLastError());
        return OQ_OK;
    }
    if (h->kdtype == OQ_K_FFT) {  // no separate prologue: k_fft_fwd forms s from u
        const int64_t L = fft_len(h->nx);
        int bits = 0;
        while (((int64_t)1 << bits) < L) ++bits;
        hipLaunchKernelGGL(k_fft_fwd, dim3((unsigned)h->nxi), dim3(256), fft_lds(h->nx), h->stream, h->dev, u,
                           h->nx, h->nxi, L, bits, h->ftw.as<double>(), h->fsh.as<double>());
        hipLaunchKernelGGL(k_fft_mid, dim3((unsigned)(L / 2 + 1)), dim3(512), 0, h->stream, h->fg.as<double>(),
                           h->fsh.as<double>(), L / 2 + 1, h->nxi, h->fth.as<double>());
        hipLaunchKernelGGL(k_fft_inv, dim3((unsigned)h->nxi), dim3(256), fft_lds(h->nx), h->stream, h->dev,
                           h->fth.as<double>(), h->nx, L, bits, h->ftw.as<double>(), u, du);
        OQ_HIP(hipGetLastError());
        return OQ_OK;
    }
    if (h->kdtype == OQ_K_TOEPLITZ) {
        int lc, nc, nt;
        toep_shape(h, &lc, &nc, &nt);
        hipLaunchKernelGGL(k_toep_partial, dim3((unsigned)h->nxi, (unsigned)nc), dim3(nt),
                           (size_t)toep_lds(h, lc), h->stream, h->K.as<double>(),
                           h->s64.as<double>(), h->nx, h->nxi, lc, h->tpart.as<double>());
        hipLaunchKernelGGL(k_toep_finish, dim3(nblk(h->nf, 256)), dim3(256), 0, h->stream, h->dev,
                           h->tpart.as<double>(), nc, u, du);
        OQ_HIP(hipGetLastError());
        return OQ_OK;
    }
    switch (h->kdtype) {
    case OQ_K_FP64: launch_matvec<OQ_K_FP64>(h, h->K.p, h->ld, h->N, 0, h->s64.as<double>(), nullptr, u, du); break;
    case OQ_K_FP32: launch_matvec<OQ_K_FP32>(h, h->K.p, h->ld, h->N, 0, h->s64.as<double>(), nullptr, u, du); break;
    default:
        launch_matvec<OQ_K_BF16>(h, h->K.p, h->ld, h->N, 0, reinterpret_cast<const double *>(h->s32.p), nullptr,
                                 u, du);
        break;
    }
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

int rhs_enqueue_fft_stage(oq_rhs *h, const StageCoef &sc, const double *dtp, const int *done, const double *u,
                          const double *k, int64_t n, double *y, double *du)
{
    const int64_t L = fft_len(h->nx);
    int bits = 0;
    while (((int64_t)1 << bits) < L) ++bits;
    hipLaunchKernelGGL(k_fft_fwd_stage, dim3((unsigned)h->nxi), dim3(256), fft_lds(h->nx) + (size_t)h->nx * 8,
                       h->stream, h->dev, sc, dtp, done, u, k, n, y, h->nx, h->nxi, L, bits, h->ftw.as<double>(),
                       h->fsh.as<double>());
    hipLaunchKernelGGL(k_fft_mid, dim3((unsigned)(L / 2 + 1)), dim3(512), 0, h->stream, h->fg.as<double>(),
                       h->fsh.as<double>(), L / 2 + 1, h->nxi, h->fth.as<double>());
    hipLaunchKernelGGL(k_fft_inv, dim3((unsigned)h->nxi), dim3(256), fft_lds(h->nx), h->stream, h->dev,
                       h->fth.as<double>(), h->nx, L, bits, h->ftw.as<double>(), y, du);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

int rhs_enqueue(oq_rhs *h, const double *u, double *du)
{
    if (h->kdtype == OQ_K_FFT && h->ne == 0) return rhs_enqueue_matvec(h, u, du);  // s is formed there
    const int64_t n = h->nf + h->ne;
    hipLaunchKernelGGL(k_rhs_prologue, dim3(nblk(n, 256)), dim3(256), 0, h->stream, h->dev, u, du,
                       h->s64.as<double>(), h->kdtype == OQ_K_BF16 ? h->s32.as<float>() : nullptr,
                       h->sm.as<double>());
    OQ_HIP(hipGetLastError());
    return rhs_enqueue_matvec(h, u, du);
}

// ---------------- operator assembly helpers ----------------

// Element-wise kernels over operator-sized arrays (up to N*ld entries, more
// than the 2^32 work-items one dispatch can count) use grid-stride loops over
// a capped grid.
static constexpr int64_t kStrideBlocks = 1 << 16;
static inline unsigned grid_capped(int64_t n, int bs)
{
    const int64_t b = (n + bs - 1) / bs;
    return (unsigned)(b < kStrideBlocks ? (b > 0 ? b : 1) : kStrideBlocks);
}

// dst[r*ld + c] = src[r*srs + c*scs] for an m x n block.
__global__ void __launch_bounds__(256) k_copy_block(const double *__restrict__ src, int64_t m,
                                                    int64_t n, int64_t srs, int64_t scs,
                                                    double *__restrict__ dst, int64_t ld)
{
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m * n; t += step) {
        const int64_t c = t % n, r = t / n;
        dst[r * ld + c] = src[r * srs + c * scs];
    }
}

__global__ void __launch_bounds__(256) k_to_fp32(const double *__restrict__ src, float *dst, int64_t n)
{
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += step)
        dst[t] = (float)src[t];
}

// FP64 -> BF16, round to nearest even.  The FP64 value is first rounded to
// FP32 precision by round-to-odd (truncate, then set the last bit when
// inexact); rounding that 24-bit value to nearest-even at 8 bits gives the
// correctly rounded result (round-to-odd to p >= q + 2 bits composes with
// any rounding to q bits).
__device__ inline unsigned short bf16_rne(double x)
{
    float f = (float)x;
    if (fabs((double)f) > fabs(x)) f = __uint_as_float(__float_as_uint(f) - 1u);  // toward zero
    unsigned int w = __float_as_uint(f);
    if ((double)f != x) w |= 1u;
    const unsigned int r = w + 0x7fffu + ((w >> 16) & 1u);
    return (unsigned short)(r >> 16);
}

__global__ void __launch_bounds__(256) k_to_bf16(const double *__restrict__ src,
                                                 unsigned short *dst, int64_t n)
{
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += step)
        dst[t] = bf16_rne(src[t]);
}

__global__ void __launch_bounds__(256) k_from_lowp(const void *src, int kt, double *dst, int64_t n)
{
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += step) {
        if (kt == OQ_K_FP32) dst[t] = (double)static_cast<const float *>(src)[t];
        else dst[t] = (double)__uint_as_float(((unsigned int)static_cast<const unsigned short *>(src)[t]) << 16);
    }
}

static int alloc_common(oq_rhs *h, int device, int64_t nx, int64_t nxi, int64_t ne, int kdtype)
{
    OQ_REQUIRE(nx > 0 && nxi > 0 && ne >= 0, "sizes must be positive");
    OQ_REQUIRE(kdtype != OQ_K_FFT_GRID, "OQ_K_FFT_GRID needs the mantle mesh: use oq_rhs_create_from_mesh");
    OQ_REQUIRE(kdtype == OQ_K_FP64 || kdtype == OQ_K_FP32 || kdtype == OQ_K_BF16 ||
                   kdtype == OQ_K_TOEPLITZ || kdtype == OQ_K_FFT,
               "kdtype must be OQ_K_FP64, OQ_K_FP32, OQ_K_BF16, OQ_K_TOEPLITZ or OQ_K_FFT");
    OQ_REQUIRE(kdtype != OQ_K_TOEPLITZ || ne == 0, "the Toeplitz operator is for fault-only problems");
    OQ_REQUIRE((kdtype != OQ_K_TOEPLITZ && kdtype != OQ_K_FFT) || nx <= 2040,
               "the Toeplitz and DFT operators take nx <= 2040 (a dip row of the kernel or the twiddle "
               "table with one row of s in LDS)");
    int ndev = 0;
    OQ_HIP(hipGetDeviceCount(&ndev));
    OQ_REQUIRE(device >= 0 && device < ndev, "invalid device index");
    OQ_HIP(hipSetDevice(device));
    h->device = device;
    OQ_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    OQ_HIP(hipEventCreate(&h->ev0));
    OQ_HIP(hipEventCreate(&h->ev1));
    h->nx = nx;
    h->nxi = nxi;
    h->ne = ne;
    h->nf = nx * nxi;
    h->N = h->nf + 6 * ne;
    h->ld = (h->N + kPad - 1) / kPad * kPad;
    h->kdtype = kdtype;
    h->lay.set(h->nf, ne);
    int rc;
    if ((rc = h->s64.alloc((size_t)h->ld * 8))) return rc;
    OQ_HIP(hipMemsetAsync(h->s64.p, 0, (size_t)h->ld * 8, h->stream));
    if ((rc = h->s32.alloc((size_t)h->ld * 4))) return rc;
    OQ_HIP(hipMemsetAsync(h->s32.p, 0, (size_t)h->ld * 4, h->stream));
    if ((rc = h->u.alloc((size_t)h->lay.n * 8))) return rc;
    if ((rc = h->du.alloc((size_t)h->lay.n * 8))) return rc;
    OQ_HIP(hipMemsetAsync(h->u.p, 0, (size_t)h->lay.n * 8, h->stream));
    OQ_HIP(hipMemsetAsync(h->du.p, 0, (size_t)h->lay.n * 8, h->stream));
    return OQ_OK;
}

static int upload_props(oq_rhs *h, const oq_rs_props *f, const oq_visc_props *v)
{
    OQ_REQUIRE(f && f->a && f->b && f->L && f->sigma, "fault property array is NULL");
    // RateStateQuasiDynamicProperty invariants (property.jl:20-24)
    OQ_REQUIRE(f->f0 > 0.0, "f₀ > 0 required");
    OQ_REQUIRE(f->v0 > 0.0, "v₀ > 0 required");
    OQ_REQUIRE(f->eta > 0.0, "η > 0 required");
    OQ_REQUIRE(f->vpl > 0.0, "vpl > 0 required");
    const size_t nf8 = (size_t)h->nf * 8;
    int rc;
    if ((rc = h->a.upload(f->a, nf8, h->stream))) return rc;
    if ((rc = h->b.upload(f->b, nf8, h->stream))) return rc;
    if ((rc = h->L.upload(f->L, nf8, h->stream))) return rc;
    if ((rc = h->sigma.upload(f->sigma, nf8, h->stream))) return rc;
    RhsDev &d = h->dev;
    d.a = h->a.as<double>();
    d.b = h->b.as<double>();
    d.L = h->L.as<double>();
    d.sigma = h->sigma.as<double>();
    d.eta = f->eta;
    d.vpl = f->vpl;
    d.f0 = f->f0;
    d.v0 = f->v0;
    d.nlaws = 0;
    d.gamma = d.nexp = nullptr;
    if (h->ne > 0) {
        OQ_REQUIRE(v && v->gamma && v->n && v->nlaws >= 1, "viscosity property is NULL");
        const size_t n8 = (size_t)(v->nlaws * h->ne) * 8;
        if ((rc = h->gamma.upload(v->gamma, n8, h->stream))) return rc;
        if ((rc = h->nexp.upload(v->n, n8, h->stream))) return rc;
        d.gamma = h->gamma.as<double>();
        d.nexp = h->nexp.as<double>();
        d.nlaws = (int)v->nlaws;
        for (int k = 0; k < 6; ++k) d.deps0[k] = v->deps0[k];
    }
    d.lay = h->lay;
    d.N = h->N;
    d.ld = h->ld;
    d.tp = d.deps_d = d.beta = d.p0 = nullptr;
    d.dila = 0;
    return OQ_OK;
}

// The Toeplitz and DFT operators: h->K holds st (nx*nxi*nxi doubles, st[i +
// nx j + nf l]) already on the device in `st`.  Toeplitz: allocate the chunk
// partials.  DFT: the twiddle table, G (from st, on the device) and the
// transform buffers.
// DFT operator set-up from the Toeplitz kernel st (device): twiddles, the
// kernel's DFT G and the spectra buffers.
static int fft_setup(oq_rhs *h, const double *st)
{
    const int64_t nx = h->nx, nxi = h->nxi, L = fft_len(nx), hl = L / 2;
    std::vector<double> tw((size_t)(2 * hl));
    for (int64_t k = 0; k < hl; ++k) {  // cos, sin of 2 pi k / L
        tw[(size_t)k] = cos(2.0 * M_PI * (double)k / (double)L);
        tw[(size_t)(hl + k)] = sin(2.0 * M_PI * (double)k / (double)L);
    }
    int rc = h->ftw.upload(tw.data(), tw.size() * 8, h->stream);
    if (!rc) rc = h->fg.alloc((size_t)((hl + 1) * nxi * nxi) * 8);
    if (!rc) rc = h->fsh.alloc((size_t)(2 * (hl + 1) * nxi) * 8);
    if (!rc) rc = h->fth.alloc((size_t)(2 * (hl + 1) * nxi) * 8);
    if (rc) return rc;
    hipLaunchKernelGGL(k_fft_kernel, dim3(grid_capped((hl + 1) * nxi * nxi, 256)), dim3(256), 0, h->stream,
                       st, nx, nxi, L, h->ftw.as<double>(), h->fg.as<double>());
    OQ_HIP(hipGetLastError());
    OQ_HIP(hipFuncSetAttribute((const void *)k_fft_fwd, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)fft_lds(nx)));
    OQ_HIP(hipFuncSetAttribute((const void *)k_fft_inv, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)fft_lds(nx)));
    OQ_HIP(hipFuncSetAttribute((const void *)k_fft_fwd_stage, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(fft_lds(nx) + (size_t)nx * 8)));
    OQ_HIP(hipStreamSynchronize(h->stream));
    return OQ_OK;
}

static int finalize_toeplitz(oq_rhs *h, DevBuf &st)
{
    h->K.p = st.p;
    h->K.bytes = st.bytes;
    st.p = nullptr;
    st.bytes = 0;
    if (h->kdtype == OQ_K_FFT) return fft_setup(h, h->K.as<double>());
    int lc, nc, nt;
    toep_shape(h, &lc, &nc, &nt);
    return h->tpart.alloc((size_t)(nc * h->nf) * 8);
}

// Grid mantle (OQ_K_FFT_GRID): the cells must be ix + nx (iy + ny iz) with x
// and y centres on uniform spacings hx, hy, the same extents within a layer
// and the same reference-point offsets (to 1e-11 of the largest cell
// extent; a generated box is exact to rounding).
static int mm_grid_detect(const oq_hex8_mesh *m, MmGrid &g, double &hx, double &hy)
{
    const int64_t ne = m->ne;
    const double ext = fmax(fabs(m->dx[0]), fmax(fabs(m->dy[0]), fabs(m->dz[0])));
    const double tol = 1e-11 * ext;
    auto near = [tol](double a, double b) { return fabs(a - b) <= tol; };
    int64_t nx = 1;
    while (nx < ne && near(m->cy[nx], m->cy[0]) && near(m->cz[nx], m->cz[0])) ++nx;
    int64_t nxy = nx;
    while (nxy < ne && near(m->cz[nxy], m->cz[0])) ++nxy;
    OQ_REQUIRE(nxy % nx == 0 && ne % nxy == 0,
               "OQ_K_FFT_GRID: the mantle cells do not form a grid (x fastest, then y, then layer)");
    const int64_t ny = nxy / nx, nz = ne / nxy;
    hx = nx > 1 ? (m->cx[nx - 1] - m->cx[0]) / (double)(nx - 1) : 0.0;
    hy = ny > 1 ? (m->cy[(ny - 1) * nx] - m->cy[0]) / (double)(ny - 1) : 0.0;
    for (int64_t c = 0; c < ne; ++c) {
        const int64_t ix = c % nx, iy = (c / nx) % ny, b = (c / nxy) * nxy;
        const bool ok = near(m->cx[c], m->cx[0] + (double)ix * hx) && near(m->cy[c], m->cy[0] + (double)iy * hy) &&
                        near(m->cz[c], m->cz[b]) && near(m->dx[c], m->dx[0]) && near(m->dy[c], m->dy[0]) &&
                        near(m->dz[c], m->dz[b]) && near(m->qx[c] - m->cx[c], m->qx[0] - m->cx[0]) &&
                        near(m->qy[c] - m->cy[c], m->qy[0] - m->cy[0]) &&
                        near(m->qz[c] - m->cz[c], m->qz[b] - m->cz[b]);
        OQ_REQUIRE(ok, "OQ_K_FFT_GRID: the mantle cells are not a uniform horizontal grid (equal spacing "
                       "and extents along x and y, the same layers under every column)");
    }
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    return OQ_OK;
}

// Shape of the convolution blocks: 256 threads over (source planes x outputs
// / R), at most 64 KB of LDS unless one plane needs more.
static int mm_grid_shape(MmGrid &g)
{
    // 8 outputs per thread only where that still gives 256 of them per plane
    g.r = g.ny * ((g.nx + 7) / 8) >= 256 ? 8 : 4;
    g.nxp = grid_nxp(g.nx, g.r);
    const int64_t Wy = 2 * g.ny - 1, Wp = 2 * g.nxp + 1, nxg = g.nxp / g.r;
    const int64_t U = g.ny * nxg, per = Wy * Wp + 2 * g.ny * g.nxp;
    OQ_REQUIRE(per * 8 <= 160 * 1024,
               "OQ_K_FFT_GRID: one source plane of the grid kernel exceeds the LDS of a workgroup "
               "(at most about 60 x 60 cells per layer)");
    int64_t tp = 256 / U;
    if (tp < 1) tp = 1;
    if (tp > 6 * g.nz) tp = 6 * g.nz;
    while (tp > 1 && tp * per * 8 > 64 * 1024) --tp;
    g.tp = (int)tp;
    g.nchunk = (int)((6 * g.nz + tp - 1) / tp);
    g.threads = 256;
    g.lds = (size_t)(tp * per * 8);
    const size_t lmax = g.lds > kGridOtherLds ? g.lds : kGridOtherLds;
    return g.r == 8 ? grid_lds_attrs<8>(lmax) : grid_lds_attrs<4>(lmax);
}

// Along strike: the fault's centres x[a] and extents ax[a] on one spacing
// dxf (a RectOkada mesh is), and the mantle's x spacing hx = m dxf for an
// integer m >= 1 (to 1e-11 of dxf and 1e-9 of hx).  Otherwise K_fm and K_mf
// stay dense (no error).
static bool mm_strike_detect(const oq_fault_mesh *mf, MmGrid &g, double hx)
{
    const int64_t nx = mf->nx;
    const double dxf = nx > 1 ? (mf->x[nx - 1] - mf->x[0]) / (double)(nx - 1) : mf->dx;
    if (!(dxf > 0.0) || g.nx > 64) return false;
    const double tol = 1e-11 * dxf;
    for (int64_t a = 0; a < nx; ++a) {
        const double sh = (double)a * dxf;
        if (fabs(mf->x[a] - (mf->x[0] + sh)) > tol || fabs(mf->ax[2 * a] - (mf->ax[0] + sh)) > tol ||
            fabs(mf->ax[2 * a + 1] - (mf->ax[1] + sh)) > tol)
            return false;
    }
    int64_t m = 1;
    if (g.nx > 1) {
        const double r = hx / dxf;
        m = (int64_t)llround(r);
        if (m < 1 || fabs(r - (double)m) > 1e-9 * r) return false;
    }
    g.m = m;
    g.D = nx + m * (g.nx - 1);
    return true;
}

// LDS batches and blocks of the F and G roles (within the convolution
// blocks' LDS); false if a role does not fit (then K_fm and K_mf stay dense).
static bool mm_strike_shape(MmGrid &g, int64_t nx, int64_t nxi)
{
    const int64_t cap = (int64_t)(kGridOtherLds / 8), Q = 6 * g.ny * g.nz, m = g.m;
    const int64_t nr = (nx + m - 1) / m, fseg = (nr + kRW - 1) / kRW, gseg = (g.nx + kRW - 1) / kRW;
    // G's row pitches are 8 mod 32 doubles, so that the eight row slots of a
    // wavefront fall on different LDS banks
    const auto pitch8 = [](int64_t w) { return w + ((8 - w % 32) + 32) % 32; };
    const int64_t Fp = m * (fseg * kRW + g.nx - 1), Gp = pitch8(nx + m * gseg * kRW), sw = pitch8(m * nr);
    if (m * fseg > 256 || gseg > 4 || m > 64 * (4 / gseg) || nx > 2048) return false;
    int64_t lpp = 256 / (m * fseg);
    if (lpp * nx > cap) lpp = cap / nx;  // the line slots' partial sums share the LDS
    const int64_t fqb = cap / (Fp + g.nx), gbb = cap / (Gp + sw);
    if (fqb < 1 || gbb < 1 || lpp < 1 || 4 * kRW > cap) return false;
    g.flpp = (int)lpp;
    g.nr = (int)nr;
    g.fseg = (int)fseg;
    g.gseg = (int)gseg;
    g.Fp = (int)Fp;
    g.Gp = (int)Gp;
    g.sw = (int)sw;
    g.fqb = (int)fqb;
    g.gbb = (int)gbb;
    const int64_t nfc = Q < 16 ? Q : 16, fqs = (Q + nfc - 1) / nfc;
    g.fqs = (int)fqs;
    g.nfc = (int)((Q + fqs - 1) / fqs);
    g.gbs = (int)nxi;  // one G block per mantle line takes every fault row
    g.ngc = 1;
    return true;
}

// F and G from the fm and mf pair kernels (GF.jl:194-227, :123-174) on
// virtual meshes: for F, D patches per fault row at x[0] + dxf (d - m (nx -
// 1)) as receivers of the cells (0, jy, jz); for G, D cells per (jy, jz) at
// X0 + dxf (e - (nx_f - 1)) as receivers of the patches (0, b), with the
// real fault's image spacing.
static int mm_strike_build(const oq_fault_mesh *mf, const oq_hex8_mesh *ma, double lrept, const QuadUpload &qu,
                           double lambda, double mu, int ftype, int nrept, oq_rhs *h)
{
    const MmGrid &g = h->grid;
    const int64_t nx = mf->nx, nxi = mf->nxi, D = g.D, nyz = g.ny * g.nz, Q = 6 * nyz, nxy = g.nx * g.ny;
    const double dxf = nx > 1 ? (mf->x[nx - 1] - mf->x[0]) / (double)(nx - 1) : mf->dx;
    // virtual receiver patches of F
    std::vector<double> vx((size_t)D), vax((size_t)(2 * D));
    for (int64_t d = 0; d < D; ++d) {
        const double sh = (double)(d - g.m * (g.nx - 1)) * dxf;
        vx[d] = mf->x[0] + sh;
        vax[2 * d] = mf->ax[0] + sh;
        vax[2 * d + 1] = mf->ax[1] + sh;
    }
    oq_fault_mesh vf = *mf;
    vf.x = vx.data();
    vf.ax = vax.data();
    vf.nx = D;
    // the source column of F: cells (0, jy, jz), index jy + ny jz
    std::vector<double> col[9];
    const double *src[9] = {ma->cx, ma->cy, ma->cz, ma->qx, ma->qy, ma->qz, ma->dx, ma->dy, ma->dz};
    for (int a = 0; a < 9; ++a)
        for (int64_t jz = 0; jz < g.nz; ++jz)
            for (int64_t jy = 0; jy < g.ny; ++jy) col[a].push_back(src[a][jz * nxy + jy * g.nx]);
    const oq_hex8_mesh cm = {col[0].data(), col[1].data(), col[2].data(), col[3].data(), col[4].data(),
                             col[5].data(), col[6].data(), col[7].data(), col[8].data(), nyz};
    // virtual receiver cells of G: index e + D (jy + ny jz)
    std::vector<double> vc[9];
    for (auto &v : vc) v.resize((size_t)(D * nyz));
    for (int64_t j = 0; j < nyz; ++j)
        for (int64_t e = 0; e < D; ++e) {
            const size_t i = (size_t)(e + D * j);
            const double sh = (double)(e - (nx - 1)) * dxf;
            vc[0][i] = col[0][j] + sh;
            vc[1][i] = col[1][j];
            vc[2][i] = col[2][j];
            vc[3][i] = vc[0][i] + (col[3][j] - col[0][j]);
            vc[4][i] = col[4][j];
            vc[5][i] = col[5][j];
            vc[6][i] = col[6][j];
            vc[7][i] = col[7][j];
            vc[8][i] = col[8][j];
        }
    const oq_hex8_mesh vm = {vc[0].data(), vc[1].data(), vc[2].data(), vc[3].data(), vc[4].data(),
                             vc[5].data(), vc[6].data(), vc[7].data(), vc[8].data(), D * nyz};
    // the source column of G: patches (0, b)
    oq_fault_mesh sf = *mf;
    sf.nx = 1;
    FaultUpload fvu, fsu;
    Hex8Upload cu, vu;
    DevBuf raw;
    const size_t nt = (size_t)(nxi * Q * D);
    int rc = fvu.load(&vf);
    if (!rc) rc = fsu.load(&sf);
    if (!rc) rc = cu.load(&cm);
    if (!rc) rc = vu.load(&vm);
    if (rc) return rc;
    fsu.dev.lrept = lrept;
    if ((rc = raw.alloc(nt * 8)) || (rc = h->Ftab.alloc(nt * 8)) || (rc = h->Gtab.alloc(nt * 8))) return rc;
    if ((rc = launch_gf_fm(cu.dev, fvu.dev, lambda, mu, ftype, raw.as<double>(), Q, 1, h->stream))) return rc;
    hipLaunchKernelGGL(k_grid_ftable, dim3(grid_capped((int64_t)nt, 256)), dim3(256), 0, h->stream,
                       raw.as<double>(), nxi, Q, D, h->Ftab.as<double>());
    OQ_HIP(hipGetLastError());
    if ((rc = launch_gf_mf(fsu.dev, vu.dev, qu.c.as<double>(), qu.w.as<double>(), qu.nq, lambda, mu, ftype, nrept,
                           raw.as<double>(), nxi, 1, h->stream)))
        return rc;
    hipLaunchKernelGGL(k_grid_gtable, dim3(grid_capped((int64_t)nt, 256)), dim3(256), 0, h->stream,
                       raw.as<double>(), nxi, g.ny, g.nz, D, h->Gtab.as<double>());
    OQ_HIP(hipGetLastError());
    OQ_HIP(hipStreamSynchronize(h->stream));
    return OQ_OK;
}

// The role table of k_grid_fused: every role's blocks spread evenly over the
// launch (block b of role r sorts at (b + 1/2) / count_r).
static int mm_roles_build(oq_rhs *h)
{
    MmGrid &g = h->grid;
    constexpr int R = KTraits<OQ_K_FP64>::kRows, W = KTraits<OQ_K_FP64>::kWaves;
    int64_t cnt[4] = {(int64_t)(6 * g.nz) * g.nchunk, 0, 0, 0};
    if (g.strike) {
        cnt[2] = h->nxi * g.nfc;
        cnt[3] = 6 * g.ny * g.nz * g.ngc;
    } else {
        const int64_t waves = (h->nf + R - 1) / R * g.fs + (6 * h->ne + R - 1) / R;
        cnt[1] = (waves + W - 1) / W;
    }
    std::vector<std::pair<double, unsigned>> e;
    for (int r = 0; r < 4; ++r) {
        OQ_REQUIRE(cnt[r] < (1 << 28), "OQ_K_FFT_GRID: too many blocks for the role table");
        for (int64_t b = 0; b < cnt[r]; ++b)
            e.push_back({((double)b + 0.5) / (double)cnt[r], ((unsigned)r << 28) | (unsigned)b});
    }
    std::vector<unsigned> t2;  // split: role order (convolutions first)
    for (auto &x : e) t2.push_back(x.second);
    std::stable_sort(e.begin(), e.end(), [](const std::pair<double, unsigned> &x,
                                            const std::pair<double, unsigned> &y) { return x.first < y.first; });
    std::vector<unsigned> t;
    for (auto &x : e) t.push_back(x.second);
    g.nblocks = (int)t.size();
    g.nconv = (int)cnt[0];
    g.lds_other = g.strike ? kGridOtherLds : 0;
    g.lds_fused = g.lds > g.lds_other ? g.lds : g.lds_other;
    int rc = h->roles.upload(t.data(), t.size() * sizeof(unsigned), h->stream);
    if (!rc) rc = h->roles2.upload(t2.data(), t2.size() * sizeof(unsigned), h->stream);
    return rc;
}

// Column chunks of the fault rows' K_fm pass (pitch ldm) near the mantle
// rows' length ldf, in whole passes of 128 columns.
static void mm_grid_chunks(MmGrid &g, int64_t ldm, int64_t ldf)
{
    int64_t S = (ldm + ldf / 2) / ldf;
    if (S < 1) S = 1;
    const int64_t cw = ((ldm + S - 1) / S + 127) / 128 * 128;
    g.fcw = cw;
    g.fs = (int)((ldm + cw - 1) / cw);
}

// h->Tmm: the pair kernel (GF.jl:262-289) over receivers shifted by every
// horizontal offset from the source column (cells (0, 0, iz)), rearranged
// into the convolution tiles; h->mpart for the partial sums.
static int mm_grid_build(oq_hex8_mesh const *m, double hx, double hy, const QuadUpload &qu, double lambda,
                         double mu, oq_rhs *h)
{
    const MmGrid &g = h->grid;
    const int64_t Wx = 2 * g.nx - 1, Wy = 2 * g.ny - 1, nxy = g.nx * g.ny, nrec = Wx * Wy * g.nz;
    std::vector<double> r[9];
    for (auto &v : r) v.resize((size_t)nrec);
    for (int64_t jz = 0; jz < g.nz; ++jz)
        for (int64_t yy = 0; yy < Wy; ++yy)
            for (int64_t xx = 0; xx < Wx; ++xx) {
                const size_t i = (size_t)(xx + Wx * (yy + Wy * jz)), b = (size_t)(jz * nxy);
                const double cx = m->cx[0] + (double)(xx - (g.nx - 1)) * hx;
                const double cy = m->cy[0] + (double)(yy - (g.ny - 1)) * hy;
                r[0][i] = cx;
                r[1][i] = cy;
                r[2][i] = m->cz[b];
                r[3][i] = cx + (m->qx[0] - m->cx[0]);
                r[4][i] = cy + (m->qy[0] - m->cy[0]);
                r[5][i] = m->qz[b];
                r[6][i] = m->dx[0];
                r[7][i] = m->dy[0];
                r[8][i] = m->dz[b];
            }
    std::vector<double> c[9];
    const double *src[9] = {m->cx, m->cy, m->cz, m->qx, m->qy, m->qz, m->dx, m->dy, m->dz};
    for (int a = 0; a < 9; ++a)
        for (int64_t iz = 0; iz < g.nz; ++iz) c[a].push_back(src[a][iz * nxy]);
    const oq_hex8_mesh rm = {r[0].data(), r[1].data(), r[2].data(), r[3].data(), r[4].data(),
                             r[5].data(), r[6].data(), r[7].data(), r[8].data(), nrec};
    const oq_hex8_mesh sm = {c[0].data(), c[1].data(), c[2].data(), c[3].data(), c[4].data(),
                             c[5].data(), c[6].data(), c[7].data(), c[8].data(), g.nz};
    Hex8Upload ru, su;
    DevBuf raw;
    const int64_t npl = 6 * g.nz, nt = npl * npl * Wy * (2 * g.nxp + 1);
    int rc = ru.load(&rm);
    if (!rc) rc = su.load(&sm);
    if (!rc) rc = raw.alloc((size_t)(6 * nrec * npl) * 8);
    if (!rc) rc = h->Tmm.alloc((size_t)nt * 8);
    if (!rc) rc = h->mpart.alloc((size_t)(h->grid.nchunk * 6 * h->ne) * 8);
    const int nfp = 1 + (h->grid.strike ? h->grid.nfc : h->grid.fs), nmv = h->grid.strike ? h->grid.ngc : 1;
    if (!rc) rc = h->fpart.alloc((size_t)(nfp * h->nf) * 8);
    if (!rc) rc = h->mvout.alloc((size_t)(nmv * 6 * h->ne) * 8);
    if (!rc)
        rc = launch_gf_mm_pairs(ru.dev, su.dev, qu.c.as<double>(), qu.w.as<double>(), qu.nq, lambda, mu,
                                raw.as<double>(), npl, 1, h->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mm_grid_table, dim3(grid_capped(nt, 256)), dim3(256), 0, h->stream, raw.as<double>(),
                       g.nx, g.ny, g.nz, (int64_t)g.nxp, h->Tmm.as<double>());
    OQ_HIP(hipGetLastError());
    OQ_HIP(hipStreamSynchronize(h->stream));
    return OQ_OK;
}

// Fused or split launches for the grid model: 20 operator applications of
// each on zeroed state after 3 of warm-up, keep the faster (ties: fused).
static int mm_grid_pick(oq_rhs *h)
{
    OQ_HIP(hipMemsetAsync(h->u.p, 0, h->u.bytes, h->stream));
    OQ_HIP(hipMemsetAsync(h->s64.p, 0, h->s64.bytes, h->stream));
    OQ_HIP(hipMemsetAsync(h->sm.p, 0, h->sm.bytes, h->stream));
    double *u = h->u.as<double>(), *du = h->du.as<double>();
    float ms[2] = {0.f, 0.f};
    for (int k = 0; k < 2; ++k) {
        h->grid.split = k == 1;
        int rc;
        for (int i = 0; i < 3; ++i)
            if ((rc = rhs_enqueue_matvec(h, u, du))) return rc;
        OQ_HIP(hipEventRecord(h->ev0, h->stream));
        for (int i = 0; i < 20; ++i)
            if ((rc = rhs_enqueue_matvec(h, u, du))) return rc;
        OQ_HIP(hipEventRecord(h->ev1, h->stream));
        OQ_HIP(hipEventSynchronize(h->ev1));
        OQ_HIP(hipEventElapsedTime(&ms[k], h->ev0, h->ev1));
    }
    h->grid.split = ms[1] < ms[0];
    return OQ_OK;
}

// Coupled model with the DFT K_ff (OQ_K_FFT, ne > 0): keeps st, sets up the
// DFT, and allocates (zeroed) the fault rows' K_fm (nf x ldm), the mantle
// rows [K_mf K_mm] (6ne x ld, in h->K), the mantle source copy and the K_ff
// traction rates.  The blocks are filled by the caller.
static int fft_coupled_alloc(oq_rhs *h, DevBuf &st)
{
    h->stk.p = st.p;
    h->stk.bytes = st.bytes;
    st.p = nullptr;
    st.bytes = 0;
    int rc = fft_setup(h, h->stk.as<double>());
    const int64_t m6 = 6 * h->ne;
    h->ldm = (m6 + kPad - 1) / kPad * kPad;
    h->ldf = (h->nf + kPad - 1) / kPad * kPad;
    if (h->mmgrid) mm_grid_chunks(h->grid, h->ldm, h->ldf);
    // grid mantle: the mantle rows hold only K_mf (6ne x ldf); along strike
    // neither dense block exists
    const bool dense = !(h->mmgrid && h->grid.strike);
    const size_t bf = (size_t)(h->nf * h->ldm) * 8, bm = (size_t)(m6 * (h->mmgrid ? h->ldf : h->ld)) * 8;
    if (!rc && dense) rc = h->Kf.alloc(bf);
    if (!rc && dense) rc = h->K.alloc(bm);
    if (!rc) rc = h->sm.alloc((size_t)h->ldm * 8);
    if (!rc && !h->mmgrid) rc = h->tauff.alloc((size_t)h->nf * 8);  // the grid model uses fpart
    if (rc) return rc;
    if (dense) OQ_HIP(hipMemsetAsync(h->Kf.p, 0, bf, h->stream));
    if (dense) OQ_HIP(hipMemsetAsync(h->K.p, 0, bm, h->stream));
    OQ_HIP(hipMemsetAsync(h->sm.p, 0, (size_t)h->ldm * 8, h->stream));
    return OQ_OK;
}

// Converts the FP64 operator in k64 (N x ld) to h->K in h->kdtype.
static int finalize_k(oq_rhs *h, DevBuf &k64)
{
    const int64_t n = h->N * h->ld;
    if (h->kdtype == OQ_K_FP64) {
        h->K.p = k64.p;
        h->K.bytes = k64.bytes;
        k64.p = nullptr;
        k64.bytes = 0;
        return OQ_OK;
    }
    int rc;
    if (h->kdtype == OQ_K_FP32) {
        if ((rc = h->K.alloc((size_t)n * 4))) return rc;
        hipLaunchKernelGGL(k_to_fp32, dim3(grid_capped(n, 256)), dim3(256), 0, h->stream, k64.as<double>(),
                           h->K.as<float>(), n);
    } else {
        if ((rc = h->K.alloc((size_t)n * 2))) return rc;
        hipLaunchKernelGGL(k_to_bf16, dim3(grid_capped(n, 256)), dim3(256), 0, h->stream, k64.as<double>(),
                           h->K.as<unsigned short>(), n);
    }
    OQ_HIP(hipGetLastError());
    OQ_HIP(hipStreamSynchronize(h->stream));
    k64.release();
    return OQ_OK;
}

// gf11 from the host into `st` on the device: the real Toeplitz kernel as
// given, or recovered from its DFT (the reference builder's default output)
// by the device inverse DFT.
static int upload_toeplitz(const double *gf11, int is_dft, int64_t nx, int64_t nxi, DevBuf &st,
                           hipStream_t s)
{
    const size_t n = (size_t)(nx * nxi * nxi);
    if (!is_dft) return st.upload(gf11, n * 8, s);
    DevBuf dft;
    int rc = dft.upload(gf11, 2 * n * 8, s);
    if (!rc) rc = st.alloc(n * 8);
    if (!rc) rc = launch_toeplitz_idft(dft.as<double>(), nx, nxi, st.as<double>(), s);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) {
        set_error("hipStreamSynchronize failed after the inverse DFT");
        rc = OQ_ERR_HIP;
    }
    return rc;
}

static void destroy(oq_rhs *h)
{
    if (!h) return;
    if (teardown_sync(h->stream)) {
        if (h->ev0) (void)hipEventDestroy(h->ev0);
        if (h->ev1) (void)hipEventDestroy(h->ev1);
        if (h->stream) (void)hipStreamDestroy(h->stream);
    }
    delete h;
}

}  // namespace oq

using namespace oq;

extern "C" {

int oq_rhs_create(int device, int64_t nx, int64_t nxi, int64_t ne, const oq_rs_props *fault,
                  const oq_visc_props *visc, const double *gf11, int gf11_is_dft,
                  const double *gf12, const double *gf21, const double *gf22, int kdtype,
                  oq_rhs **out)
{
    OQ_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    OQ_REQUIRE(gf11 != nullptr, "gf11 is NULL");
    OQ_REQUIRE(ne == 0 || (gf12 && gf21 && gf22), "coupled model needs gf12, gf21 and gf22");
    OQ_REQUIRE(nx > 0 && nxi > 0 && ne >= 0, "sizes must be positive");
    OQ_REQUIRE(fault && fault->a && fault->b && fault->L && fault->sigma,
               "fault property array is NULL");
    OQ_REQUIRE(fault->f0 > 0.0 && fault->v0 > 0.0 && fault->eta > 0.0 && fault->vpl > 0.0,
               "RateStateQuasiDynamicProperty requires f₀ > 0, v₀ > 0, η > 0, vpl > 0");
    OQ_REQUIRE(ne == 0 || (visc && visc->gamma && visc->n && visc->nlaws >= 1),
               "viscosity property is NULL");
    oq_rhs *h = new oq_rhs();
    int rc = alloc_common(h, device, nx, nxi, ne, kdtype);
    if (!rc) rc = upload_props(h, fault, visc);
    if (rc) {
        destroy(h);
        return rc;
    }
    const int64_t nf = h->nf, m6 = 6 * ne, ld = h->ld;
    DevBuf k64, tmp;
    if (kdtype == OQ_K_FFT && ne > 0) {
        rc = upload_toeplitz(gf11, gf11_is_dft, nx, nxi, tmp, h->stream);
        if (!rc) rc = fft_coupled_alloc(h, tmp);
        struct Blk { const double *src; int64_t m, n, srs, scs; double *dst; int64_t ldd; };
        const Blk blks[3] = {{gf21, nf, m6, 1, nf, h->Kf.as<double>(), h->ldm},  // K_fm: nf x 6ne
                             {gf12, m6, nf, 1, m6, h->K.as<double>(), ld},       // K_mf: 6ne x nf
                             {gf22, m6, m6, 1, m6, h->K.as<double>() + nf, ld}}; // K_mm: 6ne x 6ne
        for (int bi = 0; !rc && bi < 3; ++bi) {
            const Blk &b = blks[bi];
            if (hipStreamSynchronize(h->stream) != hipSuccess) { rc = OQ_ERR_HIP; break; }
            rc = tmp.upload(b.src, (size_t)(b.m * b.n) * 8, h->stream);
            if (rc) break;
            hipLaunchKernelGGL(k_copy_block, dim3(grid_capped(b.m * b.n, 256)), dim3(256), 0, h->stream,
                               tmp.as<double>(), b.m, b.n, b.srs, b.scs, b.dst, b.ldd);
            if (hipGetLastError() != hipSuccess) {
                set_error("k_copy_block launch failed");
                rc = OQ_ERR_HIP;
            }
        }
        if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = OQ_ERR_HIP;
        if (rc) {
            destroy(h);
            return rc;
        }
        *out = h;
        return OQ_OK;
    }
    if (kdtype == OQ_K_TOEPLITZ || kdtype == OQ_K_FFT) {
        rc = upload_toeplitz(gf11, gf11_is_dft, nx, nxi, tmp, h->stream);
        if (!rc) rc = finalize_toeplitz(h, tmp);
        if (rc) {
            destroy(h);
            return rc;
        }
        *out = h;
        return OQ_OK;
    }
    rc = k64.alloc((size_t)(h->N * ld) * 8);
    if (!rc && hipMemsetAsync(k64.p, 0, (size_t)(h->N * ld) * 8, h->stream) != hipSuccess) {
        set_error("hipMemsetAsync failed");
        rc = OQ_ERR_HIP;
    }
    if (!rc) {
        rc = upload_toeplitz(gf11, gf11_is_dft, nx, nxi, tmp, h->stream);
        if (!rc) rc = launch_expand_toeplitz(tmp.as<double>(), nx, nxi, k64.as<double>(), ld, h->stream);
    }
    struct Blk { const double *src; int64_t m, n, srs, scs, r0, c0; };
    const Blk blks[3] = {{gf21, nf, m6, 1, nf, 0, nf},     // K_fm: nf x 6ne col-major
                         {gf12, m6, nf, 1, m6, nf, 0},     // K_mf: 6ne x nf
                         {gf22, m6, m6, 1, m6, nf, nf}};   // K_mm: 6ne x 6ne
    for (int bi = 0; !rc && ne > 0 && bi < 3; ++bi) {
        const Blk &b = blks[bi];
        if (hipStreamSynchronize(h->stream) != hipSuccess) { rc = OQ_ERR_HIP; break; }
        rc = tmp.upload(b.src, (size_t)(b.m * b.n) * 8, h->stream);
        if (rc) break;
        hipLaunchKernelGGL(k_copy_block, dim3(grid_capped(b.m * b.n, 256)), dim3(256), 0, h->stream,
                           tmp.as<double>(), b.m, b.n, b.srs, b.scs,
                           k64.as<double>() + b.r0 * ld + b.c0, ld);
        if (hipGetLastError() != hipSuccess) {
            set_error("k_copy_block launch failed");
            rc = OQ_ERR_HIP;
        }
    }
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) {
        set_error("assembly failed");
        rc = OQ_ERR_HIP;
    }
    tmp.release();
    if (!rc) rc = finalize_k(h, k64);
    if (rc) {
        destroy(h);
        return rc;
    }
    *out = h;
    return OQ_OK;
}

int oq_rhs_create_from_mesh(int device, const oq_fault_mesh *mf, const oq_hex8_mesh *ma,
                            double lambda, double mu, int ftype, int64_t nrept,
                            double buffer_ratio, const oq_quadrature *quad,
                            const oq_rs_props *fault, const oq_visc_props *visc, int kdtype,
                            oq_rhs **out)
{
    OQ_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    int rc;
    if ((rc = check_fault(mf))) return rc;
    OQ_REQUIRE(ftype == 0 || ftype == 1, "ftype must be StrikeSlip (0) or DipSlip (1)");
    OQ_REQUIRE(buffer_ratio >= 0.0, "Argument `buffer_ratio` must be ≥ 0.");
    OQ_REQUIRE(nrept >= 0, "nrept must be >= 0");
    const int64_t ne = ma ? ma->ne : 0;
    if (ma && (rc = check_hex8(ma))) return rc;
    if (ma && (rc = check_quad(quad))) return rc;
    // OQ_K_FFT_GRID = the coupled OQ_K_FFT model with K_mm on the mantle grid
    const bool grid = kdtype == OQ_K_FFT_GRID;
    MmGrid g;
    double hx = 0.0, hy = 0.0;
    if (grid) {
        OQ_REQUIRE(ne > 0, "OQ_K_FFT_GRID is for the coupled model (a hex8 mantle)");
        if ((rc = mm_grid_detect(ma, g, hx, hy))) return rc;
        kdtype = OQ_K_FFT;
    }
    oq_rhs *h = new oq_rhs();
    rc = alloc_common(h, device, mf->nx, mf->nxi, ne, kdtype);
    if (!rc && grid) {
        h->mmgrid = true;
        h->grid = g;
        rc = mm_grid_shape(h->grid);
        if (!rc)
            h->grid.strike = mm_strike_detect(mf, h->grid, hx) && mm_strike_shape(h->grid, mf->nx, mf->nxi);
    }
    if (!rc) rc = upload_props(h, fault, visc);
    FaultUpload fu;
    Hex8Upload hu;
    QuadUpload qu;
    DevBuf k64, st;
    if (!rc) rc = fu.load(mf);
    if (!rc) fu.dev.lrept = (buffer_ratio + 1.0) * (mf->dx * (double)mf->nx);
    if (!rc && ma) rc = hu.load(ma);
    if (!rc && ma) rc = qu.load(quad);
    const int64_t nf = h->nf, ld = h->ld;
    if (!rc && kdtype == OQ_K_FFT && ne > 0) {
        rc = st.alloc((size_t)(nf * mf->nxi) * 8);
        if (!rc) rc = launch_gf_ff(fu.dev, lambda, mu, ftype, (int)nrept, st.as<double>(), h->stream);
        if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = OQ_ERR_HIP;
        if (!rc) rc = fft_coupled_alloc(h, st);
        if (h->mmgrid && h->grid.strike) {
            if (!rc) rc = mm_strike_build(mf, ma, fu.dev.lrept, qu, lambda, mu, ftype, (int)nrept, h);
        } else {
            if (!rc) rc = launch_gf_fm(hu.dev, fu.dev, lambda, mu, ftype, h->Kf.as<double>(), h->ldm, 1, h->stream);
            if (!rc)
                rc = launch_gf_mf(fu.dev, hu.dev, qu.c.as<double>(), qu.w.as<double>(), qu.nq, lambda, mu, ftype,
                                  (int)nrept, h->K.as<double>(), h->mmgrid ? h->ldf : ld, 1, h->stream);
        }
        if (!rc && h->mmgrid) rc = mm_grid_build(ma, hx, hy, qu, lambda, mu, h);
        if (!rc && h->mmgrid) rc = mm_roles_build(h);
        if (!rc && h->mmgrid) rc = mm_grid_pick(h);
        else if (!rc)
            rc = launch_gf_mm(hu.dev, qu.c.as<double>(), qu.w.as<double>(), qu.nq, lambda, mu,
                              h->K.as<double>() + nf, ld, 1, h->stream);
        if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) {
            set_error("device assembly failed");
            rc = OQ_ERR_HIP;
        }
        if (rc) {
            destroy(h);
            return rc;
        }
        *out = h;
        return OQ_OK;
    }
    if (!rc && (kdtype == OQ_K_TOEPLITZ || kdtype == OQ_K_FFT)) {
        rc = st.alloc((size_t)(nf * mf->nxi) * 8);
        if (!rc) rc = launch_gf_ff(fu.dev, lambda, mu, ftype, (int)nrept, st.as<double>(), h->stream);
        if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) {
            set_error("device assembly failed");
            rc = OQ_ERR_HIP;
        }
        if (!rc) rc = finalize_toeplitz(h, st);
        if (rc) {
            destroy(h);
            return rc;
        }
        *out = h;
        return OQ_OK;
    }
    if (!rc) rc = k64.alloc((size_t)(h->N * ld) * 8);
    if (!rc && hipMemsetAsync(k64.p, 0, (size_t)(h->N * ld) * 8, h->stream) != hipSuccess) {
        set_error("hipMemsetAsync failed");
        rc = OQ_ERR_HIP;
    }
    if (!rc) rc = st.alloc((size_t)(nf * mf->nxi) * 8);
    if (!rc) rc = launch_gf_ff(fu.dev, lambda, mu, ftype, (int)nrept, st.as<double>(), h->stream);
    if (!rc) rc = launch_expand_toeplitz(st.as<double>(), mf->nx, mf->nxi, k64.as<double>(), ld, h->stream);
    if (!rc && ne > 0) {
        double *K = k64.as<double>();
        rc = launch_gf_fm(hu.dev, fu.dev, lambda, mu, ftype, K + nf, ld, 1, h->stream);
        if (!rc)
            rc = launch_gf_mf(fu.dev, hu.dev, qu.c.as<double>(), qu.w.as<double>(), qu.nq, lambda,
                              mu, ftype, (int)nrept, K + nf * ld, ld, 1, h->stream);
        if (!rc)
            rc = launch_gf_mm(hu.dev, qu.c.as<double>(), qu.w.as<double>(), qu.nq, lambda, mu,
                              K + nf * ld + nf, ld, 1, h->stream);
    }
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) {
        set_error("device assembly failed");
        rc = OQ_ERR_HIP;
    }
    if (!rc) rc = finalize_k(h, k64);
    if (rc) {
        destroy(h);
        return rc;
    }
    *out = h;
    return OQ_OK;
}

int oq_rhs_eval(oq_rhs *h, double t, const double *v, const double *theta, const double *eps,
                const double *sigma, const double *delta, double *dv, double *dtheta,
                double *deps, double *dsigma, double *ddelta)
{
    (void)t;
    OQ_REQUIRE(h != nullptr, "handle is NULL");
    OQ_REQUIRE(v && theta && delta && dv && dtheta && ddelta, "state pointer is NULL");
    OQ_REQUIRE(h->ne == 0 || (eps && sigma && deps && dsigma), "mantle state pointer is NULL");
    OQ_REQUIRE(!h->dev.dila, "a dilatancy handle takes oq_rhs_eval_dila");
    OQ_HIP(hipSetDevice(h->device));
    const StateLayout &L = h->lay;
    double *u = h->u.as<double>(), *du = h->du.as<double>();
    const size_t f8 = (size_t)L.nf * 8, m8 = (size_t)(6 * L.ne) * 8;
    OQ_HIP(hipMemcpyAsync(u + L.v, v, f8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(u + L.th, theta, f8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(u + L.del, delta, f8, hipMemcpyHostToDevice, h->stream));
    if (L.ne > 0) {
        OQ_HIP(hipMemcpyAsync(u + L.eps, eps, m8, hipMemcpyHostToDevice, h->stream));
        OQ_HIP(hipMemcpyAsync(u + L.sig, sigma, m8, hipMemcpyHostToDevice, h->stream));
    }
    int rc = rhs_enqueue(h, u, du);
    if (rc) return rc;
    OQ_HIP(hipMemcpyAsync(dv, du + L.v, f8, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipMemcpyAsync(dtheta, du + L.th, f8, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipMemcpyAsync(ddelta, du + L.del, f8, hipMemcpyDeviceToHost, h->stream));
    if (L.ne > 0) {
        OQ_HIP(hipMemcpyAsync(deps, du + L.eps, m8, hipMemcpyDeviceToHost, h->stream));
        OQ_HIP(hipMemcpyAsync(dsigma, du + L.sig, m8, hipMemcpyDeviceToHost, h->stream));
    }
    OQ_HIP(hipStreamSynchronize(h->stream));
    return OQ_OK;
}

int oq_rhs_set_dilatancy(oq_rhs *h, const oq_dila_props *dila)
{
    OQ_REQUIRE(h != nullptr && dila != nullptr, "NULL argument");
    OQ_REQUIRE(dila->tp && dila->eps && dila->beta && dila->p0, "dilatancy property array is NULL");
    OQ_REQUIRE(h->ne == 0, "dilatancy is a fault-only model (equation.jl:108-117)");
    OQ_HIP(hipSetDevice(h->device));
    const size_t nf8 = (size_t)h->nf * 8;
    int rc;
    if ((rc = h->tp.upload(dila->tp, nf8, h->stream))) return rc;
    if ((rc = h->deps_d.upload(dila->eps, nf8, h->stream))) return rc;
    if ((rc = h->beta.upload(dila->beta, nf8, h->stream))) return rc;
    if ((rc = h->p0.upload(dila->p0, nf8, h->stream))) return rc;
    h->lay.set_dilatancy(h->nf);
    if ((rc = h->u.alloc((size_t)h->lay.n * 8))) return rc;
    if ((rc = h->du.alloc((size_t)h->lay.n * 8))) return rc;
    OQ_HIP(hipMemsetAsync(h->u.p, 0, (size_t)h->lay.n * 8, h->stream));
    OQ_HIP(hipMemsetAsync(h->du.p, 0, (size_t)h->lay.n * 8, h->stream));
    OQ_HIP(hipStreamSynchronize(h->stream));
    RhsDev &d = h->dev;
    d.lay = h->lay;
    d.tp = h->tp.as<double>();
    d.deps_d = h->deps_d.as<double>();
    d.beta = h->beta.as<double>();
    d.p0 = h->p0.as<double>();
    d.dila = 1;
    return OQ_OK;
}

int oq_rhs_eval_dila(oq_rhs *h, double t, const double *v, const double *theta,
                     const double *delta, const double *p, double *dv, double *dtheta,
                     double *ddelta, double *dp)
{
    (void)t;
    OQ_REQUIRE(h != nullptr, "handle is NULL");
    OQ_REQUIRE(h->dev.dila, "the handle has no dilatancy (oq_rhs_set_dilatancy)");
    OQ_REQUIRE(v && theta && delta && p && dv && dtheta && ddelta && dp, "state pointer is NULL");
    OQ_HIP(hipSetDevice(h->device));
    const StateLayout &L = h->lay;
    double *u = h->u.as<double>(), *du = h->du.as<double>();
    const size_t f8 = (size_t)L.nf * 8;
    OQ_HIP(hipMemcpyAsync(u + L.v, v, f8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(u + L.th, theta, f8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(u + L.del, delta, f8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(u + L.pp, p, f8, hipMemcpyHostToDevice, h->stream));
    int rc = rhs_enqueue(h, u, du);
    if (rc) return rc;
    OQ_HIP(hipMemcpyAsync(dv, du + L.v, f8, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipMemcpyAsync(dtheta, du + L.th, f8, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipMemcpyAsync(ddelta, du + L.del, f8, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipMemcpyAsync(dp, du + L.pp, f8, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipStreamSynchronize(h->stream));
    return OQ_OK;
}

int oq_rhs_eval_device(oq_rhs *h, double t, const double *u, double *du)
{
    (void)t;
    OQ_REQUIRE(h != nullptr && u != nullptr && du != nullptr, "NULL argument");
    OQ_HIP(hipSetDevice(h->device));
    return rhs_enqueue(h, u, du);
}

int oq_rhs_info(oq_rhs *h, int64_t *nstate, int64_t *n, int64_t *kbytes)
{
    OQ_REQUIRE(h != nullptr, "handle is NULL");
    if (nstate) *nstate = h->lay.n;
    if (n) *n = h->N;
    if (kbytes)
        *kbytes = (int64_t)(h->K.bytes + h->Kf.bytes + h->stk.bytes + h->Tmm.bytes + h->Ftab.bytes + h->Gtab.bytes);
    return OQ_OK;
}

int oq_rhs_get_k(oq_rhs *h, double *k)
{
    OQ_REQUIRE(h != nullptr && k != nullptr, "NULL argument");
    OQ_HIP(hipSetDevice(h->device));
    const int64_t n = h->N * h->ld;
    DevBuf tmp;
    const double *src = h->K.as<double>();
    if (h->kdtype == OQ_K_FFT && h->ne > 0) {  // [expanded K_ff  K_fm ; K_mf K_mm]
        const int64_t nf = h->nf, m6 = 6 * h->ne;
        int rc = tmp.alloc((size_t)n * 8);
        if (rc) return rc;
        OQ_HIP(hipMemsetAsync(tmp.p, 0, (size_t)n * 8, h->stream));
        if ((rc = launch_expand_toeplitz(h->stk.as<double>(), h->nx, h->nxi, tmp.as<double>(), h->ld,
                                         h->stream)))
            return rc;
        if (!(h->mmgrid && h->grid.strike))
            OQ_HIP(hipMemcpy2DAsync(tmp.as<double>() + nf, (size_t)h->ld * 8, h->Kf.p, (size_t)h->ldm * 8,
                                    (size_t)m6 * 8, (size_t)nf, hipMemcpyDeviceToDevice, h->stream));
        if (h->mmgrid && h->grid.strike) {
            const MmGrid &g = h->grid;
            hipLaunchKernelGGL(k_expand_strike, dim3(grid_capped(2 * nf * m6, 256)), dim3(256), 0, h->stream,
                               h->Ftab.as<double>(), h->Gtab.as<double>(), h->nx, h->nxi, g.nx, 6 * g.ny * g.nz, g.m,
                               g.D, tmp.as<double>(), h->ld);
            OQ_HIP(hipGetLastError());
        }
        if (h->mmgrid) {
            if (!h->grid.strike)
                OQ_HIP(hipMemcpy2DAsync(tmp.as<double>() + nf * h->ld, (size_t)h->ld * 8, h->K.p,
                                        (size_t)h->ldf * 8, (size_t)nf * 8, (size_t)m6, hipMemcpyDeviceToDevice,
                                        h->stream));
            const MmGrid &g = h->grid;
            hipLaunchKernelGGL(k_expand_mm_grid, dim3(grid_capped(m6 * m6, 256)), dim3(256), 0, h->stream,
                               h->Tmm.as<double>(), g.nx, g.ny, g.nz, (int64_t)g.nxp,
                               tmp.as<double>() + nf * h->ld + nf, h->ld);
            OQ_HIP(hipGetLastError());
        } else {
            OQ_HIP(hipMemcpy2DAsync(tmp.as<double>() + nf * h->ld, (size_t)h->ld * 8, h->K.p, (size_t)h->ld * 8,
                                    (size_t)h->N * 8, (size_t)m6, hipMemcpyDeviceToDevice, h->stream));
        }
        src = tmp.as<double>();
    } else if (h->kdtype == OQ_K_TOEPLITZ || h->kdtype == OQ_K_FFT) {
        int rc = tmp.alloc((size_t)n * 8);
        if (rc) return rc;
        OQ_HIP(hipMemsetAsync(tmp.p, 0, (size_t)n * 8, h->stream));
        if ((rc = launch_expand_toeplitz(h->K.as<double>(), h->nx, h->nxi, tmp.as<double>(), h->ld,
                                         h->stream)))
            return rc;
        src = tmp.as<double>();
    } else if (h->kdtype != OQ_K_FP64) {
        int rc = tmp.alloc((size_t)n * 8);
        if (rc) return rc;
        hipLaunchKernelGGL(k_from_lowp, dim3(grid_capped(n, 256)), dim3(256), 0, h->stream, h->K.p,
                           h->kdtype, tmp.as<double>(), n);
        OQ_HIP(hipGetLastError());
        src = tmp.as<double>();
    }
    OQ_HIP(hipMemcpy2DAsync(k, (size_t)h->N * 8, src, (size_t)h->ld * 8, (size_t)h->N * 8,
                            (size_t)h->N, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipStreamSynchronize(h->stream));
    return OQ_OK;
}

int oq_rhs_time(oq_rhs *h, int iters, double *ms_eval, double *ms_matvec)
{
    OQ_REQUIRE(h != nullptr && iters > 0, "invalid argument");
    OQ_HIP(hipSetDevice(h->device));
    double *u = h->u.as<double>(), *du = h->du.as<double>();
    int rc;
    OQ_HIP(hipEventRecord(h->ev0, h->stream));
    for (int i = 0; i < iters; ++i)
        if ((rc = rhs_enqueue(h, u, du))) return rc;
    OQ_HIP(hipEventRecord(h->ev1, h->stream));
    OQ_HIP(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    OQ_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    if (ms_eval) *ms_eval = (double)ms / iters;
    OQ_HIP(hipEventRecord(h->ev0, h->stream));
    for (int i = 0; i < iters; ++i)
        if ((rc = rhs_enqueue_matvec(h, u, du))) return rc;
    OQ_HIP(hipEventRecord(h->ev1, h->stream));
    OQ_HIP(hipEventSynchronize(h->ev1));
    OQ_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    if (ms_matvec) *ms_matvec = (double)ms / iters;
    return OQ_OK;
}

int oq_rhs_destroy(oq_rhs *h)
{
    if (h) {
        if (runtime_usable()) (void)hipSetDevice(h->device);
        destroy(h);
    }
    return OQ_OK;
}

}  // extern "C"
