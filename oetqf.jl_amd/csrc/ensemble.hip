// ensemble.hip -- many frictional-parameter realizations sharing one geometry.
//
// The reference integrates one `assemble`d problem at a time; a parameter
// study runs it once per realization (typically as an OrdinaryDiffEq
// EnsembleProblem over remade problems, examples/otf-with-mantle.jl:154-161).
// Here M members that share the Green's operator K advance together:
//
//   - every RHS is ONE GEMM  Y = K S  (N x N by N x M), K streamed once in
//     FP32 (or BF16) through v_mfma_f32_32x32x2f32 tiles, FP32 products
//     accumulated in the MFMA registers for 256-column chunks and folded into
//     an FP32 (hi, lo) two-sum pair between chunks, then the fused epilogue
//     (rate-and-state update / viscous stress rate) per (row, member);
//   - a lockstep Tsit5: all members evaluate stage s together, each with its
//     own dt, error norm, PI controller state and accept/reject decision.
//
// Layout: every per-member field is entry-major, x[i * M + m], so that the
// members of one row are contiguous (coalesced prologue/epilogue/stage
// loads, and the GEMM's B operand S is stored [source][member]).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "adams.h"
#include "rhs_impl.h"

typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4x __attribute__((ext_vector_type(4)));
typedef unsigned int u4x __attribute__((ext_vector_type(4)));

namespace oq {

struct EnsDev {
    int64_t nf, ne, N, ld, M, Mp, n;  // Mp: padded member count (multiple of 64; 256 for FP32 and split FP16)
    StateLayout lay;                   // offsets in entries (rows)
    const double *a, *b, *L, *sigma;   // nf x M
    const double *eta, *vpl, *f0, *v0; // M
    const double *gamma, *nexp;        // nlaws*ne (shared)
    int nlaws;
    double deps0[6];
    // 0 once every member of a lockstep run is done: the step's remaining
    // kernels then return at once, so the host can queue many steps
    // without watching for the end (1 outside lockstep runs)
    const int *live;
};

// ---------------- prologue: S = [v - vpl ; eps_dot - deps0] ----------------
__global__ void __launch_bounds__(256) k_ens_prologue(EnsDev d, const double *__restrict__ u,
                                                      double *__restrict__ du,
                                                      float *__restrict__ S)
{
    if (!*d.live) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t M = d.M;
    if (t < d.nf * M) {
        const int64_t i = t / M, m = t % M;
        S[i * d.Mp + m] = (float)(u[(d.lay.v + i) * M + m] - d.vpl[m]);
        return;
    }
    const int64_t t2 = t - d.nf * M;
    if (t2 >= d.ne * M) return;
    const int64_t i = t2 / M, m = t2 % M, ne = d.ne;
    const double *sg = u + d.lay.sig * M;
    const double sxx0 = sg[i * M + m], sxy = sg[(i + ne) * M + m], sxz = sg[(i + 2 * ne) * M + m];
    const double syy0 = sg[(i + 3 * ne) * M + m], syz = sg[(i + 4 * ne) * M + m];
    const double szz0 = sg[(i + 5 * ne) * M + m];
    const double skk = (sxx0 + syy0 + szz0) / 3.0;
    const double sxx = sxx0 - skk, syy = syy0 - skk, szz = szz0 - skk;
    const double tn = sqrt(sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz));
    double f = 0.0;
    for (int l = 0; l < d.nlaws; ++l) f += d.gamma[l * ne + i] * pow(tn, d.nexp[l * ne + i]);
    const double e[6] = {f * sxx, f * sxy, f * sxz, f * syy, f * syz, f * szz};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        du[(d.lay.eps + i + k * ne) * M + m] = e[k];
        S[(d.nf + i + k * ne) * d.Mp + m] = (float)(e[k] - d.deps0[k]);
    }
}

__device__ __forceinline__ void two_sum(float &hi, float &lo, float x)
{
    const float s = hi + x;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (x - bp);
    hi = s;
    lo += err;
}

// Epilogue of one (row, member) of Y = K S: update_fault! (equation.jl:
// 237-245) with member m's parameters on fault rows, the stress rate on
// mantle rows.
__device__ __forceinline__ void ens_epilogue(const EnsDev &d, int64_t row, int64_t m, double y,
                                             const double *__restrict__ u, double *__restrict__ du)
{
    const int64_t M = d.M;
    if (row < d.nf) {
        const int64_t ix = row * M + m;
        const double v = u[(d.lay.v + row) * M + m], th = u[(d.lay.th + row) * M + m];
        const double a = d.a[ix], b = d.b[ix], L = d.L[ix];
        const double thp = th > 0.0 ? th : 0.0;
        const double psi1 = exp((d.f0[m] + b * log(d.v0[m] * thp / L)) / a) / (2.0 * d.v0[m]);
        const double psi2 = d.sigma[ix] * psi1 / hypot(1.0, v * psi1);
        const double dth = 1.0 - v * th / L;
        du[(d.lay.th + row) * M + m] = dth;
        du[(d.lay.v + row) * M + m] = (y - b / th * v * psi2 * dth) / (a * psi2 + d.eta[m]);
        du[(d.lay.del + row) * M + m] = v;
    } else {
        du[(d.lay.sig + row - d.nf) * M + m] = y;
    }
}

// ---------------- GEMM + fused epilogue ----------------
// Block: 256 threads = 4 wavefronts; tile 128 rows x 64 members; each
// wavefront owns 32 rows x 64 members = two 32x32 MFMA tiles.  K-slices of
// BK = 16 columns are staged through LDS.
static constexpr int EBM = 128, EBN = 64, EBK = 16, ECHUNK = 256;

template <int KT>
__global__ void __launch_bounds__(256) k_ens_gemm(EnsDev d, const void *__restrict__ Kp,
                                                  const float *__restrict__ S,
                                                  const double *__restrict__ u,
                                                  double *__restrict__ du)
{
    if (!*d.live) return;
    __shared__ float As[EBM][EBK + 1];
    __shared__ float Bs[EBK][EBN + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t rb = (int64_t)blockIdx.x * EBM, cb = (int64_t)blockIdx.y * EBN;
    // global -> LDS assignments
    const int ar = tid >> 1, ak = (tid & 1) * 8;      // A: row ar, 8 columns from ak
    const int bk = tid >> 4, bc = (tid & 15) * 4;     // B: k row bk, 4 members from bc
    const int64_t arow = rb + ar < d.N ? rb + ar : d.N - 1;
    f16v acc0 = {0}, acc1 = {0};
    float hi0[16], lo0[16], hi1[16], lo1[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) hi0[i] = lo0[i] = hi1[i] = lo1[i] = 0.f;
    const int64_t nk = d.ld;  // zero-padded columns
    for (int64_t k0 = 0; k0 < nk; k0 += EBK) {
        // stage A (K rows) and B (S slice)
        if (KT == OQ_K_FP32) {
            const float *Krow = static_cast<const float *>(Kp) + arow * d.ld + k0 + ak;
            const f4x x0 = __builtin_nontemporal_load(reinterpret_cast<const f4x *>(Krow));
            const f4x x1 = __builtin_nontemporal_load(reinterpret_cast<const f4x *>(Krow) + 1);
            As[ar][ak + 0] = x0.x; As[ar][ak + 1] = x0.y; As[ar][ak + 2] = x0.z; As[ar][ak + 3] = x0.w;
            As[ar][ak + 4] = x1.x; As[ar][ak + 5] = x1.y; As[ar][ak + 6] = x1.z; As[ar][ak + 7] = x1.w;
        } else {
            const unsigned short *Krow = static_cast<const unsigned short *>(Kp) + arow * d.ld + k0 + ak;
            const u4x w = __builtin_nontemporal_load(reinterpret_cast<const u4x *>(Krow));
            As[ar][ak + 0] = __uint_as_float(w.x << 16); As[ar][ak + 1] = __uint_as_float(w.x & 0xffff0000u);
            As[ar][ak + 2] = __uint_as_float(w.y << 16); As[ar][ak + 3] = __uint_as_float(w.y & 0xffff0000u);
            As[ar][ak + 4] = __uint_as_float(w.z << 16); As[ar][ak + 5] = __uint_as_float(w.z & 0xffff0000u);
            As[ar][ak + 6] = __uint_as_float(w.w << 16); As[ar][ak + 7] = __uint_as_float(w.w & 0xffff0000u);
        }
        {
            const float4 s4 = *reinterpret_cast<const float4 *>(S + (k0 + bk) * d.Mp + cb + bc);
            Bs[bk][bc + 0] = s4.x; Bs[bk][bc + 1] = s4.y; Bs[bk][bc + 2] = s4.z; Bs[bk][bc + 3] = s4.w;
        }
        __syncthreads();
        const int r = wave * 32 + (lane & 31), kk = lane >> 5;
#pragma unroll
        for (int k = 0; k < EBK; k += 2) {
            const float a = As[r][k + kk];
            const float b0 = Bs[k + kk][lane & 31];
            const float b1 = Bs[k + kk][32 + (lane & 31)];
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b1, acc1, 0, 0, 0);
        }
        __syncthreads();
        if (((k0 + EBK) % ECHUNK) == 0 || k0 + EBK >= nk) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                two_sum(hi0[i], lo0[i], acc0[i]);
                two_sum(hi1[i], lo1[i], acc1[i]);
                acc0[i] = 0.f;
                acc1[i] = 0.f;
            }
        }
    }
    // fused epilogue: element i of tile t sits at row (i/4)*8 + (lane/32)*4 + i%4
    // and member 32 t + lane%32 of this wavefront's 32 x 64 block
    const int64_t M = d.M;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int64_t m = cb + 32 * t + (lane & 31);
        if (m >= M) continue;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int64_t row = rb + wave * 32 + (i / 4) * 8 + (lane >> 5) * 4 + (i % 4);
            if (row >= d.N) continue;
            const double y = t == 0 ? (double)hi0[i] + (double)lo0[i] : (double)hi1[i] + (double)lo1[i];
            ens_epilogue(d, row, m, y, u, du);
        }
    }
}

// ---------------- split-FP16 GEMM (OQ_K_FP16X2) ----------------
// FP32-class products on the FP16 matrix cores.  Each row r of K is kept as
// K[r, j] = (Kh + 2^-11 Kl) * rsc[r] with Kh = fp16(x), Kl = fp16(2^11 (x -
// Kh)), x = K sigma_r and sigma_r the power of two that brings max_j |K[r, j]|
// into [2^14, 2^15): 22 significant bits per entry.  Kl carries the 2^11 so
// that it stays a normal FP16 number wherever Kh is (a bare remainder would
// fall into the subnormal range for entries below 2^-18 of the row maximum).
// Each member's source vector is split the same way per chunk of kHCh
// columns, with its own power-of-two scale per (chunk, member), so that
// fault slip rates and mantle strain rates (many decades apart) each keep 22
// bits.  Per chunk, Kh Sh accumulates in one FP32 accumulator and Kh Sl + Kl
// Sh in another (the dropped Kl Sl is 2^-22 relative), on
// v_mfma_f32_32x32x16_f16; the chunk sum (acc_h + 2^-11 acc_x) is scaled and
// added into FP64.  The column space puts the mantle columns at the next
// chunk boundary after the fault columns, so no chunk mixes the two.
//
// Operand layout (global and LDS): per row of K (per member of S), every 8
// consecutive columns are stored as 8 hi halves then 8 lo halves (32 B).  In
// one MFMA k-step of 16 columns, lane l supplies columns 8 (l / 32) + 0..7 of
// row (member) l % 32 for both A and B, so a lane reads one 16-byte hi and
// one 16-byte lo vector.
typedef _Float16 h8x __attribute__((ext_vector_type(8)));

static constexpr int kHCh = 512;               // columns per chunk
static constexpr int kHBM = 64, kHBN = 256;    // block tile: rows x members
static constexpr int kHKS = 32;                // columns per LDS stage
static constexpr int kHRow = 144;              // LDS bytes per row/member (128 + 16 pad)
static constexpr int kHStage = (kHBM + kHBN) * kHRow;
static constexpr size_t kHLds = 2 * (size_t)kHStage;  // double-buffered

struct EnsH3 {
    int64_t nfp, ldc, nchunk;  // padded fault columns, columns, chunks
    const _Float16 *K;         // N x 2 ldc halves
    const double *rsc;         // N: K = (hi + lo) rsc
    _Float16 *S;               // Mp x 2 ldc halves (member-major)
    double *ssc;               // nchunk x Mp: S = (hi + lo) ssc
    double *Y;                 // N x Mp
};

// column of the split layout -> column of K (or -1 for padding)
__device__ __forceinline__ int64_t h3_col(const EnsDev &d, const EnsH3 &h, int64_t jc)
{
    if (jc < d.nf) return jc;
    if (jc >= h.nfp && jc - h.nfp < d.N - d.nf) return d.nf + (jc - h.nfp);
    return -1;
}

// power of two s with mx s in [2^14, 2^15) (1 for mx = 0)
__device__ __forceinline__ double h3_scale(double mx)
{
    if (!(mx > 0.0)) return 1.0;
    int e;
    (void)frexp(mx, &e);  // mx in [2^(e-1), 2^e)
    return ldexp(1.0, 15 - e);
}

__device__ __forceinline__ void h3_split8(const double *x, double sc, h8x &hi, h8x &lo)
{
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const double v = x[e] * sc;
        const _Float16 a = (_Float16)v;
        hi[e] = a;
        lo[e] = (_Float16)((v - (double)a) * 2048.0);
    }
}

// one-time: FP64 operator (N x ld) -> split rows and row scales
__global__ void __launch_bounds__(256) k_h3_split_k(EnsDev d, EnsH3 h, const double *__restrict__ K64,
                                                    _Float16 *__restrict__ Kh, double *__restrict__ rsc)
{
    __shared__ double red[256];
    for (int64_t row = blockIdx.x; row < d.N; row += gridDim.x) {
        const double *kr = K64 + row * d.ld;
        double mx = 0.0;
        for (int64_t j = threadIdx.x; j < d.N; j += 256) mx = fmax(mx, fabs(kr[j]));
        red[threadIdx.x] = mx;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
            __syncthreads();
        }
        const double sc = h3_scale(red[0]);
        if (threadIdx.x == 0) rsc[row] = 1.0 / sc;
        _Float16 *out = Kh + row * 2 * h.ldc;
        for (int64_t g = threadIdx.x; g < h.ldc / 8; g += 256) {
            double x[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int64_t j = h3_col(d, h, 8 * g + e);
                x[e] = j >= 0 ? kr[j] : 0.0;
            }
            h8x hi, lo;
            h3_split8(x, sc, hi, lo);
            *reinterpret_cast<h8x *>(out + 16 * g) = hi;
            *reinterpret_cast<h8x *>(out + 16 * g + 8) = lo;
        }
        __syncthreads();  // red is reused by the next row
    }
}

// Per RHS: the member-major split of S with per-(chunk, member) scales.
// Block (chunk of kHCh columns, 16 members) stages its S values in LDS (eight
// loads in flight per thread), takes each member's maximum over the chunk,
// writes the chunk scale and splits, each member's row written as 512
// contiguous bytes per 16 lanes.  (Measured: a separate max pass over
// 64-column blocks and an LDS-transpose split took 8.6 + 18.5 us against
// 17.9 us here.)
__global__ void __launch_bounds__(256) k_h3_prep_s(EnsDev d, EnsH3 h, const float *__restrict__ S32)
{
    if (!*d.live) return;
    __shared__ float tile[kHCh][17];
    __shared__ float red[16][17];
    const int tid = threadIdx.x, ml = tid & 15, cg = tid >> 4;
    const int64_t chunk = blockIdx.x, mb = (int64_t)blockIdx.y * 16, c0 = chunk * kHCh;
    float mx = 0.f;
    for (int j0 = cg; j0 < kHCh; j0 += 16 * 8) {  // eight loads in flight per pass
        float xs[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int64_t j = h3_col(d, h, c0 + j0 + 16 * u);
            xs[u] = S32[(j >= 0 ? j : 0) * d.Mp + mb + ml];
            if (j < 0) xs[u] = 0.f;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            tile[j0 + 16 * u][ml] = xs[u];
            mx = fmaxf(mx, fabsf(xs[u]));
        }
    }
    red[cg][ml] = mx;
    __syncthreads();
    const int mm = tid >> 4, g = tid & 15;  // output member and group lane
    float m2 = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) m2 = fmaxf(m2, red[k][mm]);
    const double sc = h3_scale((double)m2);
    if (g == 0) h.ssc[chunk * d.Mp + mb + mm] = 1.0 / sc;
    _Float16 *out = h.S + (mb + mm) * 2 * h.ldc + 2 * c0;
    for (int gg = g; gg < kHCh / 8; gg += 16) {
        double x[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) x[e] = (double)tile[8 * gg + e][mm];
        h8x hi, lo;
        h3_split8(x, sc, hi, lo);
        *reinterpret_cast<h8x *>(out + 16 * gg) = hi;
        *reinterpret_cast<h8x *>(out + 16 * gg + 8) = lo;
    }
}

// Y = K S for a 64-row x 256-member block tile, 8 wavefronts.  Wave w owns
// rows [32 (w % 2), +32) x members [64 (w / 2), +64): two 32 x 32 MFMA
// tiles.  Stages of 32 columns of K (64 rows) and S (256 members) go through
// double-buffered LDS; global loads run two stages ahead in registers.
// (Measured alternatives: 4 waves with 2 x 2 tiles spilled into AGPR
// copies; 128-member blocks two per CU streamed K twice and ran 20 %
// slower; three stages loaded straight into LDS with global_load_lds and
// XOR-swizzled rows ran 12 % slower.)
__global__ void __launch_bounds__(512) k_ens_gemm_h3(EnsDev d, EnsH3 h)
{
    if (!*d.live) return;
    extern __shared__ __attribute__((aligned(16))) unsigned char hlds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wm = wave >> 1;
    const int64_t rb = (int64_t)blockIdx.x * kHBM, m0 = (int64_t)blockIdx.y * kHBN;
    const int64_t rowb = 4 * h.ldc;  // bytes per split row
    // global -> LDS pieces of 16 B: K 64 rows x 8 (one per thread), S 256
    // members x 8 (four per thread)
    const unsigned char *kg, *sg[4];
    int kl, sl[4];
    {
        const int r = tid >> 3, pc = tid & 7;
        const int64_t row = rb + r < d.N ? rb + r : d.N - 1;
        kg = reinterpret_cast<const unsigned char *>(h.K) + row * rowb + 16 * pc;
        kl = r * kHRow + 16 * pc;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int p = tid + 512 * i, mm = p >> 3, qc = p & 7;
            sg[i] = reinterpret_cast<const unsigned char *>(h.S) + (m0 + mm) * rowb + 16 * qc;
            sl[i] = (kHBM + mm) * kHRow + 16 * qc;
        }
    }
    u4x pa[5], pb[5];
    auto gload = [&](u4x *pr, int64_t k0) {
        pr[0] = __builtin_nontemporal_load(reinterpret_cast<const u4x *>(kg + 4 * k0));
#pragma unroll
        for (int i = 0; i < 4; ++i) pr[1 + i] = *reinterpret_cast<const u4x *>(sg[i] + 4 * k0);
    };
    auto lstore = [&](const u4x *pr, unsigned char *st) {
        *reinterpret_cast<u4x *>(st + kl) = pr[0];
#pragma unroll
        for (int i = 0; i < 4; ++i) *reinterpret_cast<u4x *>(st + sl[i]) = pr[1 + i];
    };
    f16v acc[2], acx[2];
    double y[2][16];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        acc[b] = f16v{0};
        acx[b] = f16v{0};
#pragma unroll
        for (int i = 0; i < 16; ++i) y[b][i] = 0.0;
    }
    const int lr = lane & 31, lg = lane >> 5;
    auto compute = [&](const unsigned char *cur) {
        // both k-steps' fragments first: the second step's LDS latency hides
        // behind the first step's MFMAs
        h8x ah[2], al[2], bh[2][2], bl[2][2];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int go = (2 * ks + lg) * 32;
            const unsigned char *pA = cur + (wr * 32 + lr) * kHRow + go;
            ah[ks] = *reinterpret_cast<const h8x *>(pA);
            al[ks] = *reinterpret_cast<const h8x *>(pA + 16);
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const unsigned char *pB = cur + (kHBM + wm * 64 + b * 32 + lr) * kHRow + go;
                bh[ks][b] = *reinterpret_cast<const h8x *>(pB);
                bl[ks][b] = *reinterpret_cast<const h8x *>(pB + 16);
            }
        }
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[ks], bh[ks][b], acc[b], 0, 0, 0);
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acx[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[ks], bl[ks][b], acx[b], 0, 0, 0);
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acx[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[ks], bh[ks][b], acx[b], 0, 0, 0);
        }
    };
    // chunk scales of this lane's two members, loaded one chunk ahead so the
    // fold does not wait on the stage prefetches
    const double *scp = h.ssc + m0 + wm * 64 + lr;
    double inv[2] = {scp[0], scp[32]};
    auto fold = [&](int64_t st) {  // chunk end: scale by the member's chunk scale, add into FP64
        if (((st + 1) * kHKS) % kHCh != 0) return;
        const int64_t cn = (st + 1) * kHKS / kHCh;  // next chunk
        const int64_t cl = cn < h.nchunk ? cn : h.nchunk - 1;
        const double nx0 = scp[cl * d.Mp], nx1 = scp[cl * d.Mp + 32];
#pragma unroll
        for (int b = 0; b < 2; ++b) {
#pragma unroll
            for (int i = 0; i < 16; ++i)
                y[b][i] += ((double)acc[b][i] + (double)acx[b][i] * (1.0 / 2048.0)) * inv[b];
            acc[b] = f16v{0};
            acx[b] = f16v{0};
        }
        inv[0] = nx0;
        inv[1] = nx1;
    };
    // Pipeline: at the top of stage st, registers hold stage st + 1 (set
    // (st + 1) % 2) and stage st + 2 is issued into the other set; stage
    // st + 1 goes to LDS after stage st's MFMAs.  nst is even (ldc is a
    // multiple of kHCh), so the loop runs in pairs with static register sets.
    // The body has no data-dependent branches (past the end, the last stage
    // is loaded again and stored into the buffer nobody reads), so the
    // compiler's wait counts keep both prefetched stages in flight.
    const int64_t nst = h.ldc / kHKS;
    auto kof = [&](int64_t st) { return (st < nst ? st : nst - 1) * kHKS; };
    gload(pa, 0);
    gload(pb, kof(1));
    lstore(pa, hlds);
    __syncthreads();
    for (int64_t st = 0; st < nst; st += 2) {
        // even stage st: pb holds st + 1, pa is free
        gload(pa, kof(st + 2));
        compute(hlds);
        lstore(pb, hlds + kHStage);
        fold(st);
        __syncthreads();
        // odd stage st + 1: pa holds st + 2, pb is free
        gload(pb, kof(st + 3));
        compute(hlds + kHStage);
        lstore(pa, hlds);
        fold(st + 1);
        __syncthreads();
    }
    // element i of tile b: row 32 wr + (i / 4) 8 + (lane / 32) 4 + i % 4,
    // member 64 wm + 32 b + lane % 32
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int64_t row = rb + wr * 32 + (i / 4) * 8 + lg * 4 + (i % 4);
        if (row >= d.N) continue;
        const double rs = h.rsc[row];
#pragma unroll
        for (int b = 0; b < 2; ++b) h.Y[row * d.Mp + m0 + wm * 64 + b * 32 + lr] = y[b][i] * rs;
    }
}

// ---------------- FP32 GEMM streaming K once (OQ_K_FP32) ----------------
// The same block shape as k_ens_gemm_h3: a 64-row x 256-member tile per
// workgroup of 8 wavefronts, so each row of K is read from HBM once per RHS
// (the 128 x 64 tiles of k_ens_gemm read it once per 64 members).  Wave w
// owns rows [32 (w % 2), +32) x members [64 (w / 2), +64): two 32 x 32 tiles
// of v_mfma_f32_32x32x2f32.  Stages of 32 columns (K: 64 rows x 128 B, S:
// 256 members x 128 B, member-major) go through double-buffered LDS with the
// global loads two stages ahead in registers.  Within a stage the 16 MFMA
// k-steps take columns 16 (lane / 32) + j, so a lane's A and B fragments
// for all 16 steps are 64 contiguous bytes each.  The FP32 accumulators are
// added into FP64 every 256 columns.
static constexpr int kFBM = 64, kFBN = 256, kFKS = 32, kFRow = 144, kFChunk = 256;
static constexpr int kFStage = (kFBM + kFBN) * kFRow;
static constexpr size_t kFLds = 2 * (size_t)kFStage;

// S (entry-major [column][member], pitch Mp) -> member-major [member][column]
// (pitch ld), 64 x 64 tiles through LDS.
__global__ void __launch_bounds__(256) k_ens_s_member_major(EnsDev d, const float *__restrict__ S,
                                                            float *__restrict__ Sm)
{
    if (!*d.live) return;
    __shared__ float tile[64][65];
    const int64_t c0 = (int64_t)blockIdx.x * 64, m0 = (int64_t)blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) tile[r][tx] = S[(c0 + r) * d.Mp + m0 + tx];  // r: column
    __syncthreads();
    for (int r = ty; r < 64; r += 4) Sm[(m0 + r) * d.ld + c0 + tx] = tile[tx][r];  // r: member
}

__global__ void __launch_bounds__(512) k_ens_gemm_f32(EnsDev d, const float *__restrict__ K,
                                                      const float *__restrict__ Sm,
                                                      double *__restrict__ Y)
{
    if (!*d.live) return;
    extern __shared__ __attribute__((aligned(16))) unsigned char flds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wm = wave >> 1;
    const int64_t rb = (int64_t)blockIdx.x * kFBM, m0 = (int64_t)blockIdx.y * kFBN;
    const int64_t rowb = 4 * d.ld;  // bytes per row of K and of Sm
    const unsigned char *kg, *sg[4];
    int kl, sl[4];
    {
        const int r = tid >> 3, pc = tid & 7;
        const int64_t row = rb + r < d.N ? rb + r : d.N - 1;
        kg = reinterpret_cast<const unsigned char *>(K) + row * rowb + 16 * pc;
        kl = r * kFRow + 16 * pc;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int p = tid + 512 * i, mm = p >> 3, qc = p & 7;
            sg[i] = reinterpret_cast<const unsigned char *>(Sm) + (m0 + mm) * rowb + 16 * qc;
            sl[i] = (kFBM + mm) * kFRow + 16 * qc;
        }
    }
    u4x pa[5], pb[5];
    auto gload = [&](u4x *pr, int64_t k0) {  // k0: first column of the stage
        pr[0] = __builtin_nontemporal_load(reinterpret_cast<const u4x *>(kg + 4 * k0));
#pragma unroll
        for (int i = 0; i < 4; ++i) pr[1 + i] = *reinterpret_cast<const u4x *>(sg[i] + 4 * k0);
    };
    auto lstore = [&](const u4x *pr, unsigned char *st) {
        *reinterpret_cast<u4x *>(st + kl) = pr[0];
#pragma unroll
        for (int i = 0; i < 4; ++i) *reinterpret_cast<u4x *>(st + sl[i]) = pr[1 + i];
    };
    f16v acc[2];
    double y[2][16];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        acc[b] = f16v{0};
#pragma unroll
        for (int i = 0; i < 16; ++i) y[b][i] = 0.0;
    }
    const int lr = lane & 31, lg = lane >> 5;
    auto compute = [&](const unsigned char *cur) {
        const unsigned char *pA = cur + (wr * 32 + lr) * kFRow + 64 * lg;
        f4x a[4], bv[2][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] = *reinterpret_cast<const f4x *>(pA + 16 * q);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const unsigned char *pB = cur + (kFBM + wm * 64 + b * 32 + lr) * kFRow + 64 * lg;
#pragma unroll
            for (int q = 0; q < 4; ++q) bv[b][q] = *reinterpret_cast<const f4x *>(pB + 16 * q);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[j / 4][j % 4], bv[b][j / 4][j % 4], acc[b],
                                                              0, 0, 0);
    };
    auto fold = [&](int64_t st) {
        if (((st + 1) * kFKS) % kFChunk != 0) return;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
#pragma unroll
            for (int i = 0; i < 16; ++i) y[b][i] += (double)acc[b][i];
            acc[b] = f16v{0};
        }
    };
    // pipeline as k_ens_gemm_h3 (nst even: ld is a multiple of 512)
    const int64_t nst = d.ld / kFKS;
    auto kof = [&](int64_t st) { return (st < nst ? st : nst - 1) * kFKS; };
    gload(pa, 0);
    gload(pb, kof(1));
    lstore(pa, flds);
    __syncthreads();
    for (int64_t st = 0; st < nst; st += 2) {
        gload(pa, kof(st + 2));
        compute(flds);
        lstore(pb, flds + kFStage);
        fold(st);
        __syncthreads();
        gload(pb, kof(st + 3));
        compute(flds + kFStage);
        lstore(pa, flds);
        fold(st + 1);
        __syncthreads();
    }
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int i = 0; i < 16; ++i) y[b][i] += (double)acc[b][i];  // the last partial chunk
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int64_t row = rb + wr * 32 + (i / 4) * 8 + lg * 4 + (i % 4);
        if (row >= d.N) continue;
#pragma unroll
        for (int b = 0; b < 2; ++b) Y[row * d.Mp + m0 + wm * 64 + b * 32 + lr] = y[b][i];
    }
}

// epilogue of the split-FP16 GEMM: one thread per (row, member)
__global__ void __launch_bounds__(256) k_ens_finish(EnsDev d, const double *__restrict__ Y,
                                                    const double *__restrict__ u, double *__restrict__ du)
{
    if (!*d.live) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.N * d.M) return;
    const int64_t row = t / d.M, m = t % d.M;
    ens_epilogue(d, row, m, Y[row * d.Mp + m], u, du);
}

// ---------------- batched Tsit5 ----------------
__constant__ double kEA[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401,
     -0.028269050394068383, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081,
     2.324710524099774}};
__constant__ double kEBt[7] = {-0.00178001105222577714, -0.0008164344596567469,
                               0.007880878010261995, -0.1447110071732629, 0.5823571654525552,
                               -0.45808210592918697, 0.015151515151515152};

struct EnsCtrl {  // arrays of M
    double *t, *dt, *dtcur, *qold;
    long long *naccept, *nreject, *iter;
    int *done, *accept, *retcode;
    double t1, reltol, abstol, dtmax, dtmin;
    long long maxiters;
    int M;
    // accepted-step recorder (rec_cap == 0: off): per-chunk max|v| partials
    // and per-member rings rec_t/rec_v [M x rec_cap], slot (naccept-1) % cap
    long long nf, rec_cap;
    double *vpart, *rec_t, *rec_v;
    int *live;  // the flag EnsDev::live reads (see k_ens_live)
};

// After k_ens_begin: live = 1 while any member is still running.  One block.
__global__ void __launch_bounds__(1024) k_ens_live(EnsCtrl c)
{
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    int mine = 0;
    for (int m = threadIdx.x; m < c.M; m += blockDim.x) mine |= !c.done[m];
    if (mine) any = 1;  // benign race: every writer writes 1
    __syncthreads();
    if (threadIdx.x == 0) *c.live = any;
}

__global__ void k_ens_begin(EnsCtrl c)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= c.M) return;
    c.accept[m] = 0;
    c.dtcur[m] = 0.0;
    if (c.done[m]) return;
    if (!(c.t[m] < c.t1)) { c.done[m] = 1; c.retcode[m] = OQ_RET_SUCCESS; return; }
    if (c.iter[m] >= c.maxiters) { c.done[m] = 1; c.retcode[m] = OQ_RET_MAXITERS; return; }
    double dt = c.dt[m];
    if (dt > c.dtmax) dt = c.dtmax;
    if (c.t[m] + dt > c.t1) dt = c.t1 - c.t[m];
    if (!(dt >= c.dtmin) || !isfinite(dt)) {
        c.done[m] = 1;
        c.retcode[m] = isfinite(dt) ? OQ_RET_DTMIN : OQ_RET_NONFINITE;
        return;
    }
    c.dtcur[m] = dt;
    c.iter[m] += 1;
}

__global__ void __launch_bounds__(256) k_ens_stage(EnsCtrl c, int s, int64_t nM, int64_t M,
                                                   const double *__restrict__ u,
                                                   const double *__restrict__ k,
                                                   double *__restrict__ y)
{
    if (!*c.live) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nM) return;
    const double dt = c.dtcur[i % M];
    double acc = 0.0;
    for (int j = 0; j < s; ++j) acc += kEA[s][j] * k[j * nM + i];
    y[i] = u[i] + dt * acc;
}

// per-(row chunk, member) partial sums of the squared scaled error
static constexpr int kEnsChunks = 512;
__global__ void __launch_bounds__(256) k_ens_err(EnsCtrl c, int64_t n, int64_t M,
                                                 const double *__restrict__ u,
                                                 const double *__restrict__ y,
                                                 const double *__restrict__ k,
                                                 double *__restrict__ part)
{
    if (!*c.live) return;
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    if (m >= M) return;
    const int64_t nM = n * M;
    const int64_t r0 = n * chunk / kEnsChunks, r1 = n * (chunk + 1) / kEnsChunks;
    const double dt = c.dtcur[m];
    double acc = 0.0, vm = 0.0;
    for (int64_t r = r0; r < r1; ++r) {
        const int64_t i = r * M + m;
        double e = 0.0;
#pragma unroll
        for (int j = 0; j < 7; ++j) e += kEBt[j] * k[j * nM + i];
        const double ay = fabs(y[i]);
        const double w = dt * e / (c.abstol + fmax(fabs(u[i]), ay) * c.reltol);
        acc += w * w;
        if (r < c.nf) vm = fmax(vm, ay);  // rows [0, nf) of the state are v
    }
    part[chunk * M + m] = acc;
    if (c.rec_cap > 0) c.vpart[chunk * M + m] = vm;
}

__global__ void k_ens_controller(EnsCtrl c, int64_t n, const double *__restrict__ part)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= c.M || c.done[m]) return;
    double acc = 0.0;
    for (int ch = 0; ch < kEnsChunks; ++ch) acc += part[ch * c.M + m];
    const double beta1 = 7.0 / 50.0, beta2 = 2.0 / 25.0, qmin = 0.2, qmax = 10.0, gamma = 0.9,
                 qoldinit = 1e-4;
    const double eest = sqrt(acc / (double)n);
    double q11, q;
    if (eest == 0.0) {
        q11 = 0.0;
        q = 1.0 / qmax;
    } else {
        q11 = pow(eest, beta1);
        q = q11 / pow(c.qold[m], beta2);
        q = fmax(1.0 / qmax, fmin(1.0 / qmin, q / gamma));
    }
    const double dtc = c.dtcur[m];
    if (eest <= 1.0) {
        c.t[m] = (c.t1 - (c.t[m] + dtc) <= 0.0) ? c.t1 : c.t[m] + dtc;
        c.qold[m] = fmax(eest, qoldinit);
        c.dt[m] = dtc / q;
        c.naccept[m] += 1;
        c.accept[m] = 1;
        if (c.rec_cap > 0) {
            double vm = 0.0;
            for (int ch = 0; ch < kEnsChunks; ++ch) vm = fmax(vm, c.vpart[ch * c.M + m]);
            const long long slot = (long long)m * c.rec_cap + (c.naccept[m] - 1) % c.rec_cap;
            c.rec_t[slot] = c.t[m];
            c.rec_v[slot] = vm;
        }
    } else {
        c.nreject[m] += 1;
        if (!isfinite(eest)) {
            c.done[m] = 1;
            c.retcode[m] = OQ_RET_NONFINITE;
            return;
        }
        c.dt[m] = dtc / fmin(1.0 / qmin, q11 / gamma);
    }
}

__global__ void __launch_bounds__(256) k_ens_commit(EnsCtrl c, int64_t nM, int64_t M,
                                                    double *__restrict__ u,
                                                    const double *__restrict__ y,
                                                    double *__restrict__ k)
{
    if (!*c.live) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nM || !c.accept[i % M]) return;
    u[i] = y[i];
    k[i] = k[6 * nM + i];
}

// ---------------- lockstep VCABM5 ----------------
// The single-state VCABM5 (PECE in divided-difference form, order 1 rising
// to 5, two RHS per attempted step) with per-member order, step history,
// coefficients, error norm and acceptance, in the ensemble's entry-major
// layout.  Buffers: phi = k[0, 6 nM) (divided differences of f), f(p) =
// k[6 nM, 7 nM), p = y, and yn, fn (the corrector and its f) in ab.
struct EnsAb {
    int *k, *kcur;             // M: order of the next step / of the current one
    double *hist, *g, *beta;   // M x (kAbK + 1), member-major
    double *yn, *fn;           // nM each
};

__global__ void k_eab_coeffs(EnsCtrl c, EnsAb a)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= c.M || c.done[m] || !(c.dtcur[m] > 0.0)) return;
    const int k = a.k[m];
    a.kcur[m] = k;
    ab_coeffs(k, c.dtcur[m], a.hist + m * (kAbK + 1), a.g + m * (kAbK + 1), a.beta + m * (kAbK + 1));
}

// p = u + h sum_{j<k} g_j beta_j Phi_j (members not stepping: p = u)
__global__ void __launch_bounds__(256) k_eab_predict(EnsCtrl c, EnsAb a, int64_t nM, int64_t M,
                                                     const double *__restrict__ u,
                                                     const double *__restrict__ phi, double *__restrict__ p)
{
    if (!*c.live) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nM) return;
    const int64_t m = i % M;
    const double h = c.dtcur[m];
    if (!(h > 0.0)) {
        p[i] = u[i];
        return;
    }
    const int k = a.kcur[m];
    const double *g = a.g + m * (kAbK + 1), *be = a.beta + m * (kAbK + 1);
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc += g[j] * be[j] * phi[j * nM + i];
    p[i] = u[i] + h * acc;
}

// corrector y_{n+1} = p + h g_k Phi^p_k, Phi^p_k = f(p) - sum_{j<k} beta_j
// Phi_j, and per-(row chunk, member) sums of the squared scaled error
// h (g_k - g_{k-1}) Phi^p_k (and the max |v| of y_{n+1} for the recorder)
__global__ void __launch_bounds__(256) k_eab_correct(EnsCtrl c, EnsAb a, int64_t n, int64_t M,
                                                     const double *__restrict__ u,
                                                     const double *__restrict__ p,
                                                     const double *__restrict__ fp,
                                                     const double *__restrict__ phi,
                                                     double *__restrict__ part)
{
    if (!*c.live) return;
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    if (m >= M) return;
    const int64_t nM = n * M;
    const int64_t r0 = n * chunk / kEnsChunks, r1 = n * (chunk + 1) / kEnsChunks;
    const double h = c.dtcur[m];
    const int k = a.kcur[m];
    const double *g = a.g + m * (kAbK + 1), *be = a.beta + m * (kAbK + 1);
    const double gk = h > 0.0 ? g[k] : 0.0, dg = h > 0.0 ? g[k] - g[k - 1] : 0.0;
    double acc = 0.0, vm = 0.0;
    for (int64_t r = r0; r < r1; ++r) {
        const int64_t i = r * M + m;
        double ph = fp[i];
        if (h > 0.0)
            for (int j = 0; j < k; ++j) ph -= be[j] * phi[j * nM + i];
        const double yn = p[i] + h * gk * ph;
        a.yn[i] = yn;
        const double w = h * dg * ph / (c.abstol + fmax(fabs(u[i]), fabs(yn)) * c.reltol);
        acc += w * w;
        if (r < c.nf) vm = fmax(vm, fabs(yn));  // rows [0, nf) of the state are v
    }
    part[chunk * M + m] = acc;
    if (c.rec_cap > 0) c.vpart[chunk * M + m] = vm;
}

// per member: step factor 1/clamp(err^(1/(k+1))/0.9, 1/2, 5), acceptance,
// step history and order
__global__ void k_eab_controller(EnsCtrl c, EnsAb a, int64_t n, const double *__restrict__ part)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= c.M || c.done[m] || !(c.dtcur[m] > 0.0)) return;
    double acc = 0.0;
    for (int ch = 0; ch < kEnsChunks; ++ch) acc += part[ch * c.M + m];
    const double eest = sqrt(acc / (double)n);
    const int k = a.kcur[m];
    const double q = (eest == 0.0) ? 0.5 : fmax(0.5, fmin(5.0, pow(eest, 1.0 / (k + 1)) / 0.9));
    const double dtc = c.dtcur[m];
    if (eest <= 1.0) {
        c.t[m] = (c.t1 - (c.t[m] + dtc) <= 0.0) ? c.t1 : c.t[m] + dtc;
        double *hist = a.hist + m * (kAbK + 1);
        for (int j = kAbK; j > 0; --j) hist[j] = hist[j - 1];
        hist[0] = dtc;
        if (a.k[m] < kAbK) a.k[m] += 1;
        c.dt[m] = dtc / q;
        c.naccept[m] += 1;
        c.accept[m] = 1;
        if (c.rec_cap > 0) {
            double vm = 0.0;
            for (int ch = 0; ch < kEnsChunks; ++ch) vm = fmax(vm, c.vpart[ch * c.M + m]);
            const long long slot = (long long)m * c.rec_cap + (c.naccept[m] - 1) % c.rec_cap;
            c.rec_t[slot] = c.t[m];
            c.rec_v[slot] = vm;
        }
    } else {
        c.nreject[m] += 1;
        if (!isfinite(eest)) {
            c.done[m] = 1;
            c.retcode[m] = OQ_RET_NONFINITE;
            return;
        }
        c.dt[m] = dtc / q;
    }
}

// accepted members: u <- y_{n+1}; divided differences of the new f
__global__ void __launch_bounds__(256) k_eab_commit(EnsCtrl c, EnsAb a, int64_t nM, int64_t M,
                                                    double *__restrict__ u, double *__restrict__ phi)
{
    if (!*c.live) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nM) return;
    const int64_t m = i % M;
    if (!c.accept[m]) return;
    const int k = a.kcur[m];
    const double *be = a.beta + m * (kAbK + 1);
    double tmp = a.fn[i];
    for (int j = 0; j < k; ++j) {
        const double old = phi[j * nM + i];
        phi[j * nM + i] = tmp;
        tmp -= be[j] * old;
    }
    phi[k * nM + i] = tmp;
    u[i] = a.yn[i];
}

}  // namespace oq

struct oq_ensemble {
    oq_rhs *h = nullptr;  // owns K (FP32/BF16), stream, sizes, shared viscosity
    int kdtype = OQ_K_FP32;
    int64_t M = 0, Mp = 0, n = 0;
    oq::DevBuf hK, hrsc, hS, hssc, hY;  // OQ_K_FP16X2: split operator, split S, scales, Y
    oq::DevBuf Sm, fY;                  // OQ_K_FP32: member-major S, Y
    oq::EnsH3 h3{};
    oq::DevBuf a, b, L, sigma, eta, vpl, f0, v0, S;
    oq::DevBuf u, y, k, part;
    oq::DevBuf ct, cdt, cdtcur, cqold, cna, cnr, citer, cdone, cacc, cret;
    oq::DevBuf vpart, rec_t, rec_v;  // recorder (oq_ensemble_record)
    oq::DevBuf live;                 // EnsDev::live
    int64_t rec_cap = 0;
    std::vector<long long> rec_drained;  // records already handed out, per member
    oq::EnsDev dev{};
    oq::EnsCtrl ctrl{};
    bool stepping = false;
    bool started = false;  // the controller arrays hold a run (get_ctrl)
    int alg = 0;           // 0 Tsit5, 1 VCABM5 (oq_ensemble_start_vcabm5)
    oq::DevBuf abk, abkcur, abhist, abg, abbeta, ab;  // VCABM5 state (ab: yn, fn)
    oq::EnsAb eab{};
    // the captured lockstep step: [0] one step, [1] kGraphSteps steps back
    // to back (fewer launches; steps past the end of the run are no-ops)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
};

namespace oq {

static constexpr int kGraphSteps = 8;  // steps per launch of the multi-step graph

static void drop_graphs(oq_ensemble *e)
{
    for (int q = 0; q < 2; ++q) {
        if (e->exec[q]) { (void)hipGraphExecDestroy(e->exec[q]); e->exec[q] = nullptr; }
        if (e->graph[q]) { (void)hipGraphDestroy(e->graph[q]); e->graph[q] = nullptr; }
    }
}

static inline unsigned nbk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

static int ens_rhs(oq_ensemble *e, const double *u, double *du)
{
    oq_rhs *h = e->h;
    const int64_t rows = (h->nf + h->ne) * e->M;
    hipLaunchKernelGGL(k_ens_prologue, dim3(nbk(rows, 256)), dim3(256), 0, h->stream, e->dev, u, du,
                       e->S.as<float>());
    if (e->kdtype == OQ_K_FP16X2) {
        hipLaunchKernelGGL(k_h3_prep_s, dim3((unsigned)e->h3.nchunk, (unsigned)(e->Mp / 16)), dim3(256), 0,
                           h->stream, e->dev, e->h3, e->S.as<float>());
        hipLaunchKernelGGL(k_ens_gemm_h3, dim3(nbk(h->N, kHBM), (unsigned)(e->Mp / kHBN)), dim3(512), kHLds,
                           h->stream, e->dev, e->h3);
        hipLaunchKernelGGL(k_ens_finish, dim3(nbk(h->N * e->M, 256)), dim3(256), 0, h->stream, e->dev,
                           e->h3.Y, u, du);
        OQ_HIP(hipGetLastError());
        return OQ_OK;
    }
    if (h->kdtype == OQ_K_FP32) {
        hipLaunchKernelGGL(k_ens_s_member_major, dim3((unsigned)(h->ld / 64), (unsigned)(e->Mp / 64)),
                           dim3(256), 0, h->stream, e->dev, e->S.as<float>(), e->Sm.as<float>());
        hipLaunchKernelGGL(k_ens_gemm_f32, dim3(nbk(h->N, kFBM), (unsigned)(e->Mp / kFBN)), dim3(512), kFLds,
                           h->stream, e->dev, h->K.as<float>(), e->Sm.as<float>(), e->fY.as<double>());
        hipLaunchKernelGGL(k_ens_finish, dim3(nbk(h->N * e->M, 256)), dim3(256), 0, h->stream, e->dev,
                           e->fY.as<double>(), u, du);
        OQ_HIP(hipGetLastError());
        return OQ_OK;
    }
    const dim3 grid(nbk(h->N, EBM), (unsigned)(e->Mp / EBN));
        hipLaunchKernelGGL(k_ens_gemm<OQ_K_BF16>, grid, dim3(256), 0, h->stream, e->dev, h->K.p,
                           e->S.as<float>(), u, du);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

static int ens_step_ab(oq_ensemble *e);

static int ens_step(oq_ensemble *e)
{
    if (e->alg == 1) return ens_step_ab(e);
    oq_rhs *h = e->h;
    const int64_t nM = e->n * e->M;
    double *u = e->u.as<double>(), *y = e->y.as<double>(), *k = e->k.as<double>();
    int rc;
    hipLaunchKernelGGL(k_ens_begin, dim3(nbk(e->M, 256)), dim3(256), 0, h->stream, e->ctrl);
    hipLaunchKernelGGL(k_ens_live, dim3(1), dim3(1024), 0, h->stream, e->ctrl);
    for (int s = 1; s < 7; ++s) {
        hipLaunchKernelGGL(k_ens_stage, dim3(nbk(nM, 256)), dim3(256), 0, h->stream, e->ctrl, s, nM,
                           e->M, u, k, y);
        if ((rc = ens_rhs(e, y, k + s * nM))) return rc;
    }
    hipLaunchKernelGGL(k_ens_err, dim3(nbk(e->M, 256), kEnsChunks), dim3(256), 0, h->stream, e->ctrl,
                       e->n, e->M, u, y, k, e->part.as<double>());
    hipLaunchKernelGGL(k_ens_controller, dim3(nbk(e->M, 256)), dim3(256), 0, h->stream, e->ctrl, e->n,
                       e->part.as<double>());
    hipLaunchKernelGGL(k_ens_commit, dim3(nbk(nM, 256)), dim3(256), 0, h->stream, e->ctrl, nM, e->M,
                       u, y, k);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

// One attempted lockstep VCABM5 step: two ensemble RHS (the second also runs
// for members that rejected; its result is not used for them).
static int ens_step_ab(oq_ensemble *e)
{
    oq_rhs *h = e->h;
    const int64_t nM = e->n * e->M;
    double *u = e->u.as<double>(), *p = e->y.as<double>(), *phi = e->k.as<double>(), *fp = phi + 6 * nM;
    const dim3 gm(nbk(e->M, 256)), gn(nbk(nM, 256)), blk(256);
    int rc;
    hipLaunchKernelGGL(k_ens_begin, gm, blk, 0, h->stream, e->ctrl);
    hipLaunchKernelGGL(k_eab_coeffs, gm, blk, 0, h->stream, e->ctrl, e->eab);
    hipLaunchKernelGGL(k_ens_live, dim3(1), dim3(1024), 0, h->stream, e->ctrl);
    hipLaunchKernelGGL(k_eab_predict, gn, blk, 0, h->stream, e->ctrl, e->eab, nM, e->M, u, phi, p);
    if ((rc = ens_rhs(e, p, fp))) return rc;
    hipLaunchKernelGGL(k_eab_correct, dim3(nbk(e->M, 256), kEnsChunks), blk, 0, h->stream, e->ctrl, e->eab,
                       e->n, e->M, u, p, fp, phi, e->part.as<double>());
    hipLaunchKernelGGL(k_eab_controller, gm, blk, 0, h->stream, e->ctrl, e->eab, e->n, e->part.as<double>());
    if ((rc = ens_rhs(e, e->eab.yn, e->eab.fn))) return rc;
    hipLaunchKernelGGL(k_eab_commit, gn, blk, 0, h->stream, e->ctrl, e->eab, nM, e->M, u, phi);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

// Outside lockstep runs the RHS kernels must run: live = 1 (a copy on the
// stream, ordered before the RHS that follows).
static int set_live(oq_ensemble *e)
{
    static const int one = 1;
    OQ_HIP(hipMemcpyAsync(e->live.p, &one, sizeof(int), hipMemcpyHostToDevice, e->h->stream));
    return OQ_OK;
}

static void ens_destroy(oq_ensemble *e)
{
    if (!e) return;
    if (teardown_sync(e->h ? e->h->stream : nullptr)) {
        for (int q = 0; q < 2; ++q) {
            if (e->exec[q]) (void)hipGraphExecDestroy(e->exec[q]);
            if (e->graph[q]) (void)hipGraphDestroy(e->graph[q]);
        }
    }
    if (e->h) oq_rhs_destroy(e->h);
    delete e;
}

// OQ_K_FP16X2: split the FP64 operator, allocate the per-RHS buffers and
// drop the FP64 copy.
static int build_h3(oq_ensemble *e)
{
    oq_rhs *h = e->h;
    EnsH3 &s = e->h3;
    const int64_t nm = 6 * h->ne;
    s.nfp = h->ne > 0 ? (h->nf + kHCh - 1) / kHCh * kHCh : h->nf;
    s.ldc = (s.nfp + nm + kHCh - 1) / kHCh * kHCh;
    s.nchunk = s.ldc / kHCh;
    int rc = e->hK.alloc((size_t)(h->N * s.ldc) * 4);
    if (!rc) rc = e->hrsc.alloc((size_t)h->N * 8);
    if (!rc) rc = e->hS.alloc((size_t)(e->Mp * s.ldc) * 4);
    if (!rc) rc = e->hssc.alloc((size_t)(s.nchunk * e->Mp) * 8);
    if (!rc) rc = e->hY.alloc((size_t)(h->N * e->Mp) * 8);
    if (rc) return rc;
    s.K = e->hK.as<_Float16>();
    s.rsc = e->hrsc.as<double>();
    s.S = e->hS.as<_Float16>();
    s.ssc = e->hssc.as<double>();
    s.Y = e->hY.as<double>();
    const unsigned g = (unsigned)(h->N < 65536 ? h->N : 65536);
    hipLaunchKernelGGL(k_h3_split_k, dim3(g), dim3(256), 0, h->stream, e->dev, s, h->K.as<double>(),
                       e->hK.as<_Float16>(), e->hrsc.as<double>());
    OQ_HIP(hipGetLastError());
    OQ_HIP(hipStreamSynchronize(h->stream));
    h->K.release();
    OQ_HIP(hipFuncSetAttribute((const void *)k_ens_gemm_h3, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kHLds));
    return OQ_OK;
}

// entry-major copy of M member arrays (each length n): out[i*M + m]
static std::vector<double> interleave(const double *const *src, int64_t n, int64_t M)
{
    std::vector<double> out((size_t)(n * M));
    for (int64_t m = 0; m < M; ++m)
        for (int64_t i = 0; i < n; ++i) out[(size_t)(i * M + m)] = src[m][i];
    return out;
}

}  // namespace oq

using namespace oq;

extern "C" {

int oq_ensemble_create(int device, const oq_fault_mesh *mf, const oq_hex8_mesh *ma, double lambda,
                       double mu, int ftype, int64_t nrept, double buffer_ratio,
                       const oq_quadrature *quad, int64_t nmembers, const oq_rs_props *members,
                       const oq_visc_props *visc, int kdtype, oq_ensemble **out)
{
    OQ_REQUIRE(out != nullptr && members != nullptr, "NULL argument");
    *out = nullptr;
    OQ_REQUIRE(nmembers >= 1, "nmembers must be >= 1");
    OQ_REQUIRE(kdtype == OQ_K_FP32 || kdtype == OQ_K_BF16 || kdtype == OQ_K_FP16X2,
               "ensemble operator is FP32, BF16 or FP16X2");
    for (int64_t m = 0; m < nmembers; ++m) {
        const oq_rs_props &p = members[m];
        OQ_REQUIRE(p.a && p.b && p.L && p.sigma, "member property array is NULL");
        OQ_REQUIRE(p.f0 > 0.0 && p.v0 > 0.0 && p.eta > 0.0 && p.vpl > 0.0,
                   "RateStateQuasiDynamicProperty requires f₀ > 0, v₀ > 0, η > 0, vpl > 0");
    }
    oq_ensemble *e = new oq_ensemble();
    e->kdtype = kdtype;
    // the split operator is made from the FP64 assembly
    int rc = oq_rhs_create_from_mesh(device, mf, ma, lambda, mu, ftype, nrept, buffer_ratio, quad,
                                     &members[0], visc, kdtype == OQ_K_FP16X2 ? OQ_K_FP64 : kdtype,
                                     &e->h);
    if (rc) {
        ens_destroy(e);
        return rc;
    }
    oq_rhs *h = e->h;
    const int64_t M = nmembers, nf = h->nf;
    e->M = M;
    const int64_t mb = kdtype == OQ_K_FP16X2 ? kHBN : (kdtype == OQ_K_FP32 ? kFBN : EBN);
    e->Mp = (M + mb - 1) / mb * mb;
    e->n = h->lay.n;
    std::vector<const double *> pa(M), pb(M), pL(M), ps(M);
    std::vector<double> eta(M), vpl(M), f0(M), v0(M);
    for (int64_t m = 0; m < M; ++m) {
        pa[m] = members[m].a; pb[m] = members[m].b; pL[m] = members[m].L; ps[m] = members[m].sigma;
        eta[m] = members[m].eta; vpl[m] = members[m].vpl; f0[m] = members[m].f0; v0[m] = members[m].v0;
    }
    const size_t fM = (size_t)(nf * M) * 8, M8 = (size_t)M * 8;
    if (!rc) rc = e->a.upload(interleave(pa.data(), nf, M).data(), fM, h->stream);
    if (!rc) rc = e->b.upload(interleave(pb.data(), nf, M).data(), fM, h->stream);
    if (!rc) rc = e->L.upload(interleave(pL.data(), nf, M).data(), fM, h->stream);
    if (!rc) rc = e->sigma.upload(interleave(ps.data(), nf, M).data(), fM, h->stream);
    if (!rc) rc = e->eta.upload(eta.data(), M8, h->stream);
    if (!rc) rc = e->vpl.upload(vpl.data(), M8, h->stream);
    if (!rc) rc = e->f0.upload(f0.data(), M8, h->stream);
    if (!rc) rc = e->v0.upload(v0.data(), M8, h->stream);
    if (!rc) rc = e->S.alloc((size_t)(h->ld * e->Mp) * 4);
    if (!rc && hipMemsetAsync(e->S.p, 0, (size_t)(h->ld * e->Mp) * 4, h->stream) != hipSuccess) rc = OQ_ERR_HIP;
    const int64_t nM = e->n * M;
    if (!rc) rc = e->u.alloc((size_t)nM * 8);
    if (!rc) rc = e->y.alloc((size_t)nM * 8);
    if (!rc) rc = e->k.alloc((size_t)(7 * nM) * 8);
    if (!rc) rc = e->part.alloc((size_t)(kEnsChunks * M) * 8);
    {
        const int one = 1;
        if (!rc) rc = e->live.upload(&one, sizeof(int), h->stream);
    }
    DevBuf *cb8[] = {&e->ct, &e->cdt, &e->cdtcur, &e->cqold, &e->cna, &e->cnr, &e->citer};
    for (DevBuf *bp : cb8)
        if (!rc) rc = bp->alloc(M8);
    DevBuf *cb4[] = {&e->cdone, &e->cacc, &e->cret};
    for (DevBuf *bp : cb4)
        if (!rc) rc = bp->alloc((size_t)M * 4);
    if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = OQ_ERR_HIP;
    if (rc) {
        ens_destroy(e);
        return rc;
    }
    EnsDev &d = e->dev;
    d.nf = nf; d.ne = h->ne; d.N = h->N; d.ld = h->ld; d.M = M; d.Mp = e->Mp; d.n = e->n;
    d.lay = h->lay;
    d.a = e->a.as<double>(); d.b = e->b.as<double>(); d.L = e->L.as<double>(); d.sigma = e->sigma.as<double>();
    d.eta = e->eta.as<double>(); d.vpl = e->vpl.as<double>(); d.f0 = e->f0.as<double>(); d.v0 = e->v0.as<double>();
    d.gamma = h->dev.gamma; d.nexp = h->dev.nexp; d.nlaws = h->dev.nlaws;
    d.live = e->live.as<int>();
    for (int k = 0; k < 6; ++k) d.deps0[k] = h->dev.deps0[k];
    EnsCtrl &c = e->ctrl;
    c.t = e->ct.as<double>(); c.dt = e->cdt.as<double>(); c.dtcur = e->cdtcur.as<double>();
    c.qold = e->cqold.as<double>(); c.naccept = e->cna.as<long long>(); c.nreject = e->cnr.as<long long>();
    c.iter = e->citer.as<long long>(); c.done = e->cdone.as<int>(); c.accept = e->cacc.as<int>();
    c.retcode = e->cret.as<int>(); c.M = (int)M;
    c.nf = nf; c.rec_cap = 0; c.vpart = c.rec_t = c.rec_v = nullptr;
    c.live = e->live.as<int>();
    if (kdtype == OQ_K_FP16X2 && (rc = build_h3(e))) {
        ens_destroy(e);
        return rc;
    }
    if (kdtype == OQ_K_FP32) {
        const size_t sb = (size_t)(e->Mp * h->ld) * 4;
        rc = e->Sm.alloc(sb);
        if (!rc && hipMemsetAsync(e->Sm.p, 0, sb, h->stream) != hipSuccess) rc = OQ_ERR_HIP;
        if (!rc) rc = e->fY.alloc((size_t)(h->N * e->Mp) * 8);
        if (!rc && hipFuncSetAttribute((const void *)k_ens_gemm_f32,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFLds) != hipSuccess)
            rc = OQ_ERR_HIP;
        if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = OQ_ERR_HIP;
        if (rc) {
            set_error("FP32 ensemble buffers: allocation or kernel set-up failed");
            ens_destroy(e);
            return rc;
        }
    }
    *out = e;
    return OQ_OK;
}

int oq_ensemble_eval(oq_ensemble *e, const double *u, double *du)
{
    OQ_REQUIRE(e && u && du, "NULL argument");
    oq_rhs *h = e->h;
    OQ_HIP(hipSetDevice(h->device));
    const int64_t nM = e->n * e->M;
    std::vector<const double *> src(e->M);
    for (int64_t m = 0; m < e->M; ++m) src[m] = u + m * e->n;
    const std::vector<double> ui = interleave(src.data(), e->n, e->M);
    OQ_HIP(hipMemcpyAsync(e->y.p, ui.data(), (size_t)nM * 8, hipMemcpyHostToDevice, h->stream));
    int rc = set_live(e);
    if (!rc) rc = ens_rhs(e, e->y.as<double>(), e->k.as<double>());
    if (rc) return rc;
    std::vector<double> o((size_t)nM);
    OQ_HIP(hipMemcpyAsync(o.data(), e->k.p, (size_t)nM * 8, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipStreamSynchronize(h->stream));
    for (int64_t m = 0; m < e->M; ++m)
        for (int64_t i = 0; i < e->n; ++i) du[m * e->n + i] = o[(size_t)(i * e->M + m)];
    return OQ_OK;
}

// Uploads the members' states (member-major host layout) and controller
// arrays, seeds k1 = f(u) and captures the step graph.
static int ens_begin(oq_ensemble *e, const double *u0, const double *tv, const double *dtv, const double *qv,
                     const long long *na, const long long *nr, const long long *it, const int *done,
                     const int *ret, double t1, const oq_solver_opts *o, int alg)
{
    oq_rhs *h = e->h;
    OQ_HIP(hipSetDevice(h->device));
    const int64_t M = e->M, nM = e->n * M;
    if (alg == 1) {  // VCABM5: order 1, no step history; phi_0 = f(u) is the seed below
        int rc = OQ_OK;
        if (!e->ab.p) {
            const size_t h6 = (size_t)(M * (kAbK + 1)) * 8;
            rc = e->abk.alloc((size_t)M * 4);
            if (!rc) rc = e->abkcur.alloc((size_t)M * 4);
            if (!rc) rc = e->abhist.alloc(h6);
            if (!rc) rc = e->abg.alloc(h6);
            if (!rc) rc = e->abbeta.alloc(h6);
            if (!rc) rc = e->ab.alloc((size_t)(2 * nM) * 8);
            if (rc) return rc;
        }
        const std::vector<int> one(M, 1);
        const std::vector<double> zh((size_t)(M * (kAbK + 1)), 0.0);
        OQ_HIP(hipMemcpyAsync(e->abk.p, one.data(), M * 4, hipMemcpyHostToDevice, h->stream));
        OQ_HIP(hipMemcpyAsync(e->abkcur.p, one.data(), M * 4, hipMemcpyHostToDevice, h->stream));
        OQ_HIP(hipMemcpyAsync(e->abhist.p, zh.data(), zh.size() * 8, hipMemcpyHostToDevice, h->stream));
        OQ_HIP(hipMemcpyAsync(e->abg.p, zh.data(), zh.size() * 8, hipMemcpyHostToDevice, h->stream));
        OQ_HIP(hipMemcpyAsync(e->abbeta.p, zh.data(), zh.size() * 8, hipMemcpyHostToDevice, h->stream));
        e->eab = EnsAb{e->abk.as<int>(), e->abkcur.as<int>(), e->abhist.as<double>(), e->abg.as<double>(),
                       e->abbeta.as<double>(), e->ab.as<double>(), e->ab.as<double>() + nM};
        // the host vectors above are read by the copies before they go out of scope
        OQ_HIP(hipStreamSynchronize(h->stream));
    }
    e->alg = alg;
    std::vector<const double *> src(M);
    for (int64_t m = 0; m < M; ++m) src[m] = u0 + m * e->n;
    const std::vector<double> ui = interleave(src.data(), e->n, M);
    const std::vector<int> zi(M, 0);
    OQ_HIP(hipMemcpyAsync(e->u.p, ui.data(), (size_t)nM * 8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(e->ct.p, tv, M * 8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(e->cdt.p, dtv, M * 8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(e->cqold.p, qv, M * 8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(e->cna.p, na, M * 8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(e->cnr.p, nr, M * 8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(e->citer.p, it, M * 8, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(e->cdone.p, done, M * 4, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(e->cacc.p, zi.data(), M * 4, hipMemcpyHostToDevice, h->stream));
    OQ_HIP(hipMemcpyAsync(e->cret.p, ret, M * 4, hipMemcpyHostToDevice, h->stream));
    EnsCtrl &c = e->ctrl;
    c.t1 = t1;
    c.reltol = o->reltol;
    c.abstol = o->abstol;
    c.dtmax = o->dtmax > 0.0 ? o->dtmax : INFINITY;
    c.dtmin = o->dtmin;
    c.maxiters = o->maxiters > 0 ? o->maxiters : (long long)1e18;
    c.rec_cap = e->rec_cap;
    c.vpart = e->rec_cap ? e->vpart.as<double>() : nullptr;
    c.rec_t = e->rec_cap ? e->rec_t.as<double>() : nullptr;
    c.rec_v = e->rec_cap ? e->rec_v.as<double>() : nullptr;
    e->rec_drained.assign(na, na + M);  // records from before a resume are not in the rings
    // FSAL seed: k1 = f(u) (on a resume the same bits as the k7 of the last
    // accepted step: the RHS kernels are deterministic); for VCABM5 phi_0 = f(u)
    int rc = set_live(e);
    if (!rc) rc = ens_rhs(e, e->u.as<double>(), e->k.as<double>());
    if (rc) return rc;
    drop_graphs(e);
    OQ_HIP(hipStreamSynchronize(h->stream));
    for (int q = 0; q < 2 && !rc; ++q) {
        OQ_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < (q ? kGraphSteps : 1) && !rc; ++i) rc = ens_step(e);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(h->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        OQ_HIP(ce);
        e->graph[q] = g;
        OQ_HIP(hipGraphInstantiate(&e->exec[q], g, nullptr, nullptr, 0));
    }
    e->stepping = true;
    e->started = true;
    return OQ_OK;
}

int oq_ensemble_start(oq_ensemble *e, const double *u0, double t0, double t1,
                      const oq_solver_opts *o)
{
    OQ_REQUIRE(e && u0 && o, "NULL argument");
    OQ_REQUIRE(t1 > t0, "tspan must satisfy t1 > t0");
    OQ_REQUIRE(o->dt0 > 0.0, "the ensemble stepper needs an initial step dt0 > 0");
    const int64_t M = e->M;
    const std::vector<double> tv(M, t0), dtv(M, o->dt0), qv(M, 1e-4);
    const std::vector<long long> zl(M, 0);
    const std::vector<int> zi(M, 0);
    return ens_begin(e, u0, tv.data(), dtv.data(), qv.data(), zl.data(), zl.data(), zl.data(), zi.data(),
                     zi.data(), t1, o, 0);
}

int oq_ensemble_start_vcabm5(oq_ensemble *e, const double *u0, double t0, double t1, const oq_solver_opts *o)
{
    OQ_REQUIRE(e && u0 && o, "NULL argument");
    OQ_REQUIRE(t1 > t0, "tspan must satisfy t1 > t0");
    OQ_REQUIRE(o->dt0 > 0.0, "the ensemble stepper needs an initial step dt0 > 0");
    const int64_t M = e->M;
    const std::vector<double> tv(M, t0), dtv(M, o->dt0), qv(M, 1e-4);
    const std::vector<long long> zl(M, 0);
    const std::vector<int> zi(M, 0);
    return ens_begin(e, u0, tv.data(), dtv.data(), qv.data(), zl.data(), zl.data(), zl.data(), zi.data(),
                     zi.data(), t1, o, 1);
}

int oq_ensemble_get_ctrl(oq_ensemble *e, oq_lockstep_ctrl *c)
{
    OQ_REQUIRE(e && c, "NULL argument");
    OQ_REQUIRE(e->started, "the ensemble stepper has not been started");
    oq_rhs *h = e->h;
    OQ_HIP(hipSetDevice(h->device));
    const int64_t M = e->M;
    const std::pair<void *, const DevBuf *> d8[] = {{c->t, &e->ct}, {c->dt, &e->cdt}, {c->qold, &e->cqold},
                                                     {c->naccept, &e->cna}, {c->nreject, &e->cnr},
                                                     {c->iter, &e->citer}};
    for (const auto &x : d8)
        if (x.first) OQ_HIP(hipMemcpyAsync(x.first, x.second->p, M * 8, hipMemcpyDeviceToHost, h->stream));
    if (c->done) OQ_HIP(hipMemcpyAsync(c->done, e->cdone.p, M * 4, hipMemcpyDeviceToHost, h->stream));
    if (c->retcode) OQ_HIP(hipMemcpyAsync(c->retcode, e->cret.p, M * 4, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipStreamSynchronize(h->stream));
    return OQ_OK;
}

int oq_ensemble_resume(oq_ensemble *e, const double *u, const oq_lockstep_ctrl *c, double t1,
                       const oq_solver_opts *o)
{
    OQ_REQUIRE(e && u && c && o, "NULL argument");
    OQ_REQUIRE(c->t && c->dt && c->qold && c->naccept && c->nreject && c->iter && c->done && c->retcode,
               "every controller array is needed to resume");
    for (int64_t m = 0; m < e->M; ++m)
        OQ_REQUIRE(c->done[m] || (c->t[m] < t1 && c->dt[m] > 0.0), "a running member needs t < t1 and dt > 0");
    static_assert(sizeof(long long) == sizeof(int64_t), "int64_t counters");
    return ens_begin(e, u, c->t, c->dt, c->qold, reinterpret_cast<const long long *>(c->naccept),
                     reinterpret_cast<const long long *>(c->nreject), reinterpret_cast<const long long *>(c->iter),
                     c->done, c->retcode, t1, o, 0);
}

int oq_ensemble_run(oq_ensemble *e, int64_t max_steps, int64_t *nrunning)
{
    OQ_REQUIRE(e && max_steps > 0, "invalid argument");
    OQ_REQUIRE(e->stepping, "call oq_ensemble_start first");
    OQ_REQUIRE(e->rec_cap == 0 || max_steps <= e->rec_cap,
               "with the recorder on, max_steps must be <= step_capacity (one drain per run)");
    oq_rhs *h = e->h;
    OQ_HIP(hipSetDevice(h->device));
    // Steps are queued back to back: once every member is done, the
    // remaining ones return at their first kernel (EnsDev::live), so the
    // host looks at the done flags only every kRunSync steps.
    constexpr int64_t kRunSync = 256;
    std::vector<int> done(e->M);
    int64_t left = e->M;
    for (int64_t s = 0; s < max_steps; s += kRunSync) {
        const int64_t m = (max_steps - s) < kRunSync ? (max_steps - s) : kRunSync;
        for (int64_t i = 0; i + kGraphSteps <= m; i += kGraphSteps) OQ_HIP(hipGraphLaunch(e->exec[1], h->stream));
        for (int64_t i = m / kGraphSteps * kGraphSteps; i < m; ++i) OQ_HIP(hipGraphLaunch(e->exec[0], h->stream));
        OQ_HIP(hipMemcpyAsync(done.data(), e->cdone.p, e->M * 4, hipMemcpyDeviceToHost, h->stream));
        OQ_HIP(hipStreamSynchronize(h->stream));
        left = 0;
        for (int v : done) left += v ? 0 : 1;
        if (left == 0) break;
    }
    if (nrunning) *nrunning = left;
    return OQ_OK;
}

int oq_ensemble_record(oq_ensemble *e, int64_t step_capacity)
{
    OQ_REQUIRE(e != nullptr, "NULL argument");
    OQ_REQUIRE(step_capacity >= 0, "step_capacity must be >= 0");
    OQ_HIP(hipSetDevice(e->h->device));
    if (e->rec_cap == step_capacity) return OQ_OK;
    // the captured step holds the recorder's buffers as kernel arguments:
    // drop it before they are released (oq_ensemble_start captures anew)
    OQ_HIP(hipStreamSynchronize(e->h->stream));
    drop_graphs(e);
    e->stepping = false;
    e->vpart.release();
    e->rec_t.release();
    e->rec_v.release();
    e->rec_cap = 0;
    if (step_capacity > 0) {
        const size_t rb = (size_t)(e->M * step_capacity) * 8;
        int rc = e->vpart.alloc((size_t)(kEnsChunks * e->M) * 8);
        if (!rc) rc = e->rec_t.alloc(rb);
        if (!rc) rc = e->rec_v.alloc(rb);
        if (rc) {
            e->vpart.release();
            e->rec_t.release();
            e->rec_v.release();
            return rc;
        }
        e->rec_cap = step_capacity;
    }
    e->stepping = false;  // the captured step must be rebuilt with the recorder
    return OQ_OK;
}

int oq_ensemble_drain(oq_ensemble *e, int64_t *counts, double *step_t, double *step_vmax)
{
    OQ_REQUIRE(e && counts && step_t && step_vmax, "NULL argument");
    OQ_REQUIRE(e->rec_cap > 0, "the recorder is off (oq_ensemble_record)");
    OQ_REQUIRE(e->stepping, "call oq_ensemble_start first");
    oq_rhs *h = e->h;
    OQ_HIP(hipSetDevice(h->device));
    const int64_t M = e->M, cap = e->rec_cap;
    std::vector<long long> na((size_t)M);
    std::vector<double> rt((size_t)(M * cap)), rv((size_t)(M * cap));
    OQ_HIP(hipMemcpyAsync(na.data(), e->cna.p, M * 8, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipMemcpyAsync(rt.data(), e->rec_t.p, rt.size() * 8, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipMemcpyAsync(rv.data(), e->rec_v.p, rv.size() * 8, hipMemcpyDeviceToHost, h->stream));
    OQ_HIP(hipStreamSynchronize(h->stream));
    int64_t lost = 0;
    for (int64_t m = 0; m < M; ++m) {
        const long long d0 = e->rec_drained[m], fresh = na[m] - d0;
        const long long keep = fresh > cap ? cap : fresh;
        if (fresh > cap) lost += fresh - cap;
        for (long long j = 0; j < keep; ++j) {
            const long long slot = m * cap + (na[m] - keep + j) % cap;
            step_t[m * cap + j] = rt[(size_t)slot];
            step_vmax[m * cap + j] = rv[(size_t)slot];
        }
        counts[m] = keep;
        e->rec_drained[m] = na[m];
    }
    if (lost) {
        set_error("step record overflow: " + std::to_string(lost) +
                  " accepted steps were overwritten before this drain (run at most "
                  "step_capacity steps between drains)");
        return OQ_ERR_STATE;
    }
    return OQ_OK;
}

int oq_ensemble_state(oq_ensemble *e, double *t, double *u, int64_t *naccept, int64_t *nreject,
                      int *retcode)
{
    OQ_REQUIRE(e != nullptr, "NULL argument");
    oq_rhs *h = e->h;
    OQ_HIP(hipSetDevice(h->device));
    const int64_t M = e->M, nM = e->n * M;
    if (t) OQ_HIP(hipMemcpyAsync(t, e->ct.p, M * 8, hipMemcpyDeviceToHost, h->stream));
    if (naccept) OQ_HIP(hipMemcpyAsync(naccept, e->cna.p, M * 8, hipMemcpyDeviceToHost, h->stream));
    if (nreject) OQ_HIP(hipMemcpyAsync(nreject, e->cnr.p, M * 8, hipMemcpyDeviceToHost, h->stream));
    if (retcode) OQ_HIP(hipMemcpyAsync(retcode, e->cret.p, M * 4, hipMemcpyDeviceToHost, h->stream));
    std::vector<double> o;
    if (u) {
        o.resize((size_t)nM);
        OQ_HIP(hipMemcpyAsync(o.data(), e->u.p, (size_t)nM * 8, hipMemcpyDeviceToHost, h->stream));
    }
    OQ_HIP(hipStreamSynchronize(h->stream));
    if (u)
        for (int64_t m = 0; m < M; ++m)
            for (int64_t i = 0; i < e->n; ++i) u[m * e->n + i] = o[(size_t)(i * M + m)];
    return OQ_OK;
}

int oq_ensemble_time(oq_ensemble *e, int iters, double *ms_eval)
{
    OQ_REQUIRE(e && iters > 0, "invalid argument");
    oq_rhs *h = e->h;
    OQ_HIP(hipSetDevice(h->device));
    int rc;
    if ((rc = set_live(e))) return rc;
    OQ_HIP(hipEventRecord(h->ev0, h->stream));
    for (int i = 0; i < iters; ++i)
        if ((rc = ens_rhs(e, e->y.as<double>(), e->k.as<double>()))) return rc;
    OQ_HIP(hipEventRecord(h->ev1, h->stream));
    OQ_HIP(hipEventSynchronize(h->ev1));
    float ms = 0.f;
    OQ_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    if (ms_eval) *ms_eval = (double)ms / iters;
    return OQ_OK;
}

int oq_ensemble_destroy(oq_ensemble *e)
{
    if (e) {
        if (runtime_usable()) (void)hipSetDevice(e->h ? e->h->device : 0);
        ens_destroy(e);
    }
    return OQ_OK;
}

}  // extern "C"
