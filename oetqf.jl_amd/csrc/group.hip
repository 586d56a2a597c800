// group.hip -- grouped batches of geometries of uneven size, each with up to
// 16 frictional realizations.
//
// The reference runs one `assemble` + `solve` per (geometry, realization)
// (equation.jl:141-154).  Here G geometries with M <= 16 members each are
// one object on the device:
//
//   - every RHS is ONE grouped launch over a tile table that spans all
//     geometries: a 4-wavefront workgroup takes 64*RB consecutive receiver
//     rows of one geometry and all M members, so each operator K_g (FP32) is
//     streamed from HBM once per RHS.  Sixteen members fill the N dimension
//     of v_mfma_f32_16x16x4f32; the reduction index is permuted so that each
//     lane streams 16 consecutive columns of one row (64 contiguous bytes),
//     and the source block S_g is stored member-major so that the matching
//     B operands are contiguous too.  FP32 MFMA partials are folded into an
//     FP32 (hi, lo) two-sum pair every 256 columns and the row sums stored in
//     FP64; an element-wise epilogue then applies the rate-and-state update
//     (fault rows) or the stress-rate store (mantle rows) per (row, member);
//   - an LDS-staged variant stages 32 KB slices of K (row-contiguous global
//     loads) and the shared S slice through LDS;
//   - the grid is persistent (at most the resident workgroups) and walks the
//     tile table largest geometry first; the table is static, so every
//     workgroup ends after its last tile;
//   - tile shape (RB = rows per wavefront / 16, and whether the K stream is
//     software-pipelined, which doubles the load registers and may halve the
//     occupancy) is chosen per size bucket of N by timing every shape on the
//     bucket's geometries (oq_group_autotune);
//   - a lockstep Tsit5 advances all G*M members, each with its own dt, error
//     norm, PI controller state and acceptance.
//
// Layout: per geometry g the member fields are entry-major, x[soff_g + i*M
// + m]; geometries follow one another.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "adams.h"
#include "rhs_impl.h"

typedef float f4x __attribute__((ext_vector_type(4)));

namespace oq {

static constexpr int kGM = 16;       // member slots per geometry (MFMA N)
static constexpr int kGStep = 64;    // columns per wavefront pass (4 lane groups x 16)
static constexpr int kGWaves = 4;    // wavefronts per workgroup
static constexpr int kGChunks = 512;  // row chunks of the error norm (one workgroup row each)
static constexpr int kGBuckets = 8;  // size buckets: N < 2048, < 4096, ... (log2)

struct GrpGeo {
    const float *K;
    int64_t N, ld, nf, ne, n;
    int64_t soff;  // state entries before this geometry
    int64_t foff;  // fault-parameter entries before this geometry
    int64_t Soff;  // S floats before this geometry (S[Soff + m*ld + c])
    int64_t Yoff;  // row-sum entries before this geometry (Y[Yoff + row*M + m])
    StateLayout lay;
    const double *gamma, *nexp;
    int nlaws;
    double deps0[6];
};

struct GrpDev {
    const GrpGeo *geo;
    const int64_t *soffs;  // G+1 state offsets
    const int64_t *poffs;  // G+1 prologue offsets ((nf+ne)*M per geometry)
    const int64_t *yoffs;  // G+1 row-sum offsets (N*M per geometry)
    int G, M;
    const double *a, *b, *L, *sigma;    // per geometry nf x M, entry-major
    const double *eta, *vpl, *f0, *v0;  // G*M
    float *S;
    double *Y;  // K S row sums, FP64, entry-major per geometry
    int64_t nprolog, ntot, nrows;
    // 0 once every member of a lockstep run is done: the step's remaining
    // kernels return at once (1 outside lockstep runs)
    const int *live;
};

// segment g with off[g] <= x < off[g+1] (off ascending, off[0] = 0)
__device__ __forceinline__ int seg_find(const int64_t *__restrict__ off, int G, int64_t x)
{
    int lo = 0, hi = G - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The same for every thread of a block with x = this thread's index: one
// binary search per block (for the block's first index), then a forward walk
// (a block spans at most a few segments); called by all threads of the block
// before any of them returns.
__device__ __forceinline__ int seg_find_block(const int64_t *__restrict__ off, int G, int64_t x)
{
    __shared__ int g0;
    if (threadIdx.x == 0) g0 = seg_find(off, G, (int64_t)blockIdx.x * blockDim.x);
    __syncthreads();
    int g = g0;
    while (g + 1 < G && off[g + 1] <= x) ++g;
    return g;
}

// ---------------- prologue: S_g = [v - vpl ; eps_dot - deps0], member-major ----------------
__global__ void __launch_bounds__(256) k_grp_prologue(GrpDev d, const double *__restrict__ u,
                                                      double *__restrict__ du)
{
    if (!*d.live) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = seg_find_block(d.poffs, d.G, t);
    if (t >= d.nprolog) return;
    const GrpGeo &G = d.geo[g];
    const int64_t M = d.M, loc = t - d.poffs[g];
    const int64_t i = loc / M, m = loc % M, gm = (int64_t)g * M + m;
    const double *ug = u + G.soff;
    float *S = d.S + G.Soff + m * G.ld;
    if (i < G.nf) {
        S[i] = (float)(ug[(G.lay.v + i) * M + m] - d.vpl[gm]);
        return;
    }
    const int64_t c = i - G.nf, ne = G.ne;
    const double *sg = ug + G.lay.sig * M;
    const double sxx0 = sg[c * M + m], sxy = sg[(c + ne) * M + m], sxz = sg[(c + 2 * ne) * M + m];
    const double syy0 = sg[(c + 3 * ne) * M + m], syz = sg[(c + 4 * ne) * M + m];
    const double szz0 = sg[(c + 5 * ne) * M + m];
    const double skk = (sxx0 + syy0 + szz0) / 3.0;
    const double sxx = sxx0 - skk, syy = syy0 - skk, szz = szz0 - skk;
    const double tn = sqrt(sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz));
    double f = 0.0;
    for (int l = 0; l < G.nlaws; ++l) f += G.gamma[l * ne + c] * pow(tn, G.nexp[l * ne + c]);
    const double e[6] = {f * sxx, f * sxy, f * sxz, f * syy, f * syz, f * szz};
    double *dug = du + G.soff;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        dug[(G.lay.eps + c + k * ne) * M + m] = e[k];
        S[G.nf + c + k * ne] = (float)(e[k] - G.deps0[k]);
    }
}

// Global-address-space view of a device pointer: loads through it compile
// to global (not flat) loads.
typedef const __attribute__((address_space(1))) f4x gf4x;
__device__ __forceinline__ f4x gload4(const float *p)
{
    return *(gf4x *)p;
}

// Streaming (non-temporal) load of K: each operator entry is read once per
// RHS, while S is re-read by every row block of its geometry and should
// stay in the caches.
__device__ __forceinline__ f4x kload4(const float *p)
{
    return __builtin_nontemporal_load((gf4x *)p);
}

// s_waitcnt immediate that waits for at most n outstanding vector memory
// operations (vmcnt bits [3:0] and [15:14]) and leaves the LDS/export counts
// at their maxima
__host__ __device__ constexpr int vmcnt_imm(int n)
{
    return (n & 0xf) | ((n >> 4) << 14) | (0x7 << 4) | (0xf << 8);
}

__device__ __forceinline__ void grp_two_sum(float &hi, float &lo, float x)
{
    const float s = hi + x;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (x - bp);
    hi = s;
    lo += err;
}

// Row sums of one wavefront: element v of row block rb is row
// row0 + 16 rb + 4 lq + v, member lr (the 16x16x4 MFMA output layout).  The
// rate-and-state / stress-rate epilogue runs in k_grp_finish: its FP64
// transcendentals would otherwise set the register budget (and halve the
// occupancy) of the streaming loop.
template <int RB>
__device__ __forceinline__ void grp_store_rows(const GrpDev &d, const GrpGeo &G, int64_t row0,
                                               int lr, int lq, const float (&hi)[RB][4],
                                               const float (&lo)[RB][4])
{
    const int64_t M = d.M, N = G.N;
    if (lr >= M) return;
    double *Y = d.Y + G.Yoff + lr;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t row = row0 + rb * 16 + 4 * lq + v;
            if (row < N) Y[row * M] = (double)hi[rb][v] + (double)lo[rb][v];
        }
    }
}

// ---------------- epilogue: du from the row sums, per (row, member) ----------------
__global__ void __launch_bounds__(256) k_grp_finish(GrpDev d, const double *__restrict__ u,
                                                    double *__restrict__ du)
{
    if (!*d.live) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = seg_find_block(d.yoffs, d.G, t);
    if (t >= d.nrows) return;
    const GrpGeo &G = d.geo[g];
    const int64_t M = d.M, loc = t - G.Yoff;
    const int64_t row = loc / M, m = loc % M, gm = (int64_t)g * M + m;
    const double y = d.Y[t];
    const double *ug = u + G.soff;
    double *dug = du + G.soff;
    if (row < G.nf) {
        // update_fault! (equation.jl:237-245), member m's parameters
        const int64_t ix = G.foff + row * M + m;
        const double vv = ug[(G.lay.v + row) * M + m], th = ug[(G.lay.th + row) * M + m];
        const double a = d.a[ix], b = d.b[ix], L = d.L[ix];
        const double thp = th > 0.0 ? th : 0.0;
        const double psi1 = exp((d.f0[gm] + b * log(d.v0[gm] * thp / L)) / a) / (2.0 * d.v0[gm]);
        const double psi2 = d.sigma[ix] * psi1 / hypot(1.0, vv * psi1);
        const double dth = 1.0 - vv * th / L;
        dug[(G.lay.th + row) * M + m] = dth;
        dug[(G.lay.v + row) * M + m] = (y - b / th * vv * psi2 * dth) / (a * psi2 + d.eta[gm]);
        dug[(G.lay.del + row) * M + m] = vv;
    } else {
        dug[(G.lay.sig + row - G.nf) * M + m] = y;
    }
}

// ---------------- grouped GEMM + fused epilogue ----------------
// tiles[t] = (geometry, first row); a workgroup tile is kGWaves * RB * 16 rows.
template <int RB, bool PIPE>
__global__ void __launch_bounds__(256) k_grp_gemm(GrpDev d, const int2 *__restrict__ tiles,
                                                  int ntiles, const double *__restrict__ u,
                                                  double *__restrict__ du)
{
    if (!*d.live) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    for (int tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
        const int2 tl = tiles[tix];
        const GrpGeo &G = d.geo[tl.x];
        const int64_t N = G.N, ld = G.ld;
        const int64_t row0 = (int64_t)tl.y + (int64_t)wave * RB * 16;
        if (row0 >= N) continue;
        const float *kp[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            const int64_t r = row0 + rb * 16 + lr < N ? row0 + rb * 16 + lr : N - 1;
            kp[rb] = G.K + r * ld + 4 * lq;
        }
        const float *sp = d.S + G.Soff + (int64_t)lr * ld + 4 * lq;
        // NC independent accumulator chains per row block (RB * NC = 4), so
        // that consecutive MFMAs do not wait on each other's result
        constexpr int NC = 4 / RB;
        f4x acc[RB][NC];
        float hi[RB][4], lo[RB][4];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[rb][c] = (f4x){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int v = 0; v < 4; ++v) hi[rb][v] = lo[rb][v] = 0.f;
        }
        const int64_t nstep = ld / kGStep;
        // start each wavefront's sweep at a different pass (row pitches are
        // multiples of 2 KB, so aligned sweeps would load the same channels)
        int64_t p = (((int64_t)tix * kGWaves + wave) * 13) % nstep;
        // PIPE: software pipeline, the loads of the next pass are issued
        // before the MFMAs of the current one (twice the load registers)
        f4x sv[4], kv[RB][4];
        if (PIPE) {
#pragma unroll
            for (int t = 0; t < 4; ++t) sv[t] = gload4(sp + p * kGStep + 16 * t);
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                for (int t = 0; t < 4; ++t) kv[rb][t] = gload4(kp[rb] + p * kGStep + 16 * t);
        }
        for (int64_t it = 0; it < nstep; ++it) {
            const int64_t pn = (p + 1 == nstep) ? 0 : p + 1;
            const bool more = it + 1 < nstep;
            if (!PIPE) {
#pragma unroll
                for (int t = 0; t < 4; ++t) sv[t] = gload4(sp + p * kGStep + 16 * t);
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int t = 0; t < 4; ++t) kv[rb][t] = gload4(kp[rb] + p * kGStep + 16 * t);
            }
            // the prefetch is unconditional (pn wraps to the first pass after the
            // last one): a conditional load would make the compiler's wait
            // counts drain every load in flight at the branch join
            f4x sn[4], kn[RB][4];
            if (PIPE) {
#pragma unroll
                for (int t = 0; t < 4; ++t) sn[t] = gload4(sp + pn * kGStep + 16 * t);
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int t = 0; t < 4; ++t) kn[rb][t] = gload4(kp[rb] + pn * kGStep + 16 * t);
            }
            __builtin_amdgcn_sched_barrier(0);
            // lane (row lr, group lq) holds columns c0 + 16t + 4lq + e; MFMA
            // (t, e) reduces over lq, the 16 MFMAs over all 64 columns
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb)
                        acc[rb][t % NC] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                            kv[rb][t][e], sv[t][e], acc[rb][t % NC], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
            if ((it & 3) == 3 || !more) {
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) grp_two_sum(hi[rb][v], lo[rb][v], acc[rb][c][v]);
                        acc[rb][c] = (f4x){0.f, 0.f, 0.f, 0.f};
                    }
                }
            }
            if (PIPE) {
#pragma unroll
                for (int t = 0; t < 4; ++t) sv[t] = sn[t];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int t = 0; t < 4; ++t) kv[rb][t] = kn[rb][t];
            }
            p = pn;
        }
        grp_store_rows<RB>(d, G, row0, lr, lq, hi, lo);
    }
}

// LDS-staged variant: the workgroup stages a 64*RB-row x BK-column slice of
// K (32 KB, BK = 128/RB) and the matching 16 x BK slice of S through LDS.
// The global loads are row-contiguous (one wavefront instruction covers
// 2-4 rows x 256-512 bytes, whole cache lines) and the next slice is in
// flight in registers while the current one is reduced; the four
// wavefronts share the S slice.  Rows are padded by 4 floats so that the
// MFMA-layout reads (16 rows x 16 bytes per quarter wavefront) hit distinct
// banks.
template <int RB>
__global__ void __launch_bounds__(256) k_grp_gemm_lds(GrpDev d, const int2 *__restrict__ tiles,
                                                      int ntiles, const double *__restrict__ u,
                                                      double *__restrict__ du)
{
    if (!*d.live) return;
    constexpr int BR = 64 * RB, BK = 128 / RB, PAD = BK + 4, C4 = BK / 4;
    constexpr int KV = BR * C4 / 256;  // float4 of K per thread per slice (8)
    constexpr int SV = 16 * C4 / 256;  // float4 of S per thread per slice (2 or 1)
    constexpr int NC = 4 / RB;
    __shared__ __attribute__((aligned(16))) float As[BR * PAD];
    __shared__ __attribute__((aligned(16))) float Bs[16 * PAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const int trow = tid / C4, tc4 = tid % C4;  // this thread's slot in a slice
    for (int tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
        const int2 tl = tiles[tix];
        const GrpGeo &G = d.geo[tl.x];
        const int64_t N = G.N, ld = G.ld, r0 = tl.y;
        const int64_t nch = ld / BK;
        const float *kp[KV];
#pragma unroll
        for (int j = 0; j < KV; ++j) {
            const int64_t r = r0 + trow + j * (256 / C4);
            kp[j] = G.K + (r < N ? r : N - 1) * ld + 4 * tc4;
        }
        const float *sp[SV];
#pragma unroll
        for (int j = 0; j < SV; ++j)
            sp[j] = d.S + G.Soff + (int64_t)(trow + j * (256 / C4)) * ld + 4 * tc4;
        f4x acc[RB][NC];
        float hi[RB][4], lo[RB][4];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[rb][c] = (f4x){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int v = 0; v < 4; ++v) hi[rb][v] = lo[rb][v] = 0.f;
        }
        int64_t p = ((int64_t)tix * 13) % nch;
        f4x kr[KV], sr[SV];
#pragma unroll
        for (int j = 0; j < KV; ++j) kr[j] = kload4(kp[j] + p * BK);
#pragma unroll
        for (int j = 0; j < SV; ++j) sr[j] = gload4(sp[j] + p * BK);
        __syncthreads();  // the previous tile's reads of As/Bs are done
#pragma unroll
        for (int j = 0; j < KV; ++j)
            *reinterpret_cast<f4x *>(&As[(trow + j * (256 / C4)) * PAD + 4 * tc4]) = kr[j];
#pragma unroll
        for (int j = 0; j < SV; ++j)
            *reinterpret_cast<f4x *>(&Bs[(trow + j * (256 / C4)) * PAD + 4 * tc4]) = sr[j];
        __syncthreads();
        for (int64_t it = 0; it < nch; ++it) {
            const int64_t pn = (p + 1 == nch) ? 0 : p + 1;
            const bool more = it + 1 < nch;
            // unconditional prefetch and store (pn wraps; after the last
            // slice they are redundant and the next tile re-stages after a
            // barrier): conditional loads make the wait counts drain
#pragma unroll
            for (int j = 0; j < KV; ++j) kr[j] = kload4(kp[j] + pn * BK);
#pragma unroll
            for (int j = 0; j < SV; ++j) sr[j] = gload4(sp[j] + pn * BK);
#pragma unroll
            for (int sub = 0; sub < BK / 64; ++sub) {
                f4x a[RB][4], b[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    b[t] = *reinterpret_cast<const f4x *>(&Bs[lr * PAD + sub * 64 + 16 * t + 4 * lq]);
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb)
                        a[rb][t] = *reinterpret_cast<const f4x *>(
                            &As[(wave * 16 * RB + rb * 16 + lr) * PAD + sub * 64 + 16 * t + 4 * lq]);
                }
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
#pragma unroll
                        for (int rb = 0; rb < RB; ++rb)
                            acc[rb][t % NC] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                a[rb][t][e], b[t][e], acc[rb][t % NC], 0, 0, 0);
            }
            if (((it + 1) * BK) % 256 == 0 || !more) {
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) grp_two_sum(hi[rb][v], lo[rb][v], acc[rb][c][v]);
                        acc[rb][c] = (f4x){0.f, 0.f, 0.f, 0.f};
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < KV; ++j)
                *reinterpret_cast<f4x *>(&As[(trow + j * (256 / C4)) * PAD + 4 * tc4]) = kr[j];
#pragma unroll
            for (int j = 0; j < SV; ++j)
                *reinterpret_cast<f4x *>(&Bs[(trow + j * (256 / C4)) * PAD + 4 * tc4]) = sr[j];
            __syncthreads();
            p = pn;
        }
        grp_store_rows<RB>(d, G, r0 + (int64_t)wave * 16 * RB, lr, lq, hi, lo);
    }
}

// Depth-2 variant: two register sets alternate, so each workgroup has the
// next two 32 KB slices in flight while it reduces the one in LDS.
template <int RB>
__global__ void __launch_bounds__(256) k_grp_gemm_lds2(GrpDev d, const int2 *__restrict__ tiles,
                                                       int ntiles, const double *__restrict__ u,
                                                       double *__restrict__ du)
{
    if (!*d.live) return;
    constexpr int BR = 64 * RB, BK = 128 / RB, PAD = BK + 4, C4 = BK / 4;
    constexpr int KV = BR * C4 / 256, SV = 16 * C4 / 256, NC = 4 / RB, RS = 256 / C4;
    __shared__ __attribute__((aligned(16))) float As[BR * PAD];
    __shared__ __attribute__((aligned(16))) float Bs[16 * PAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const int trow = tid / C4, tc4 = tid % C4;
    for (int tix = blockIdx.x; tix < ntiles; tix += gridDim.x) {
        const int2 tl = tiles[tix];
        const GrpGeo &G = d.geo[tl.x];
        const int64_t N = G.N, ld = G.ld, r0 = tl.y;
        const int64_t nch = ld / BK;
        const int64_t pstart = ((int64_t)tix * 13) % nch;
        const float *kp[KV];
#pragma unroll
        for (int j = 0; j < KV; ++j) {
            const int64_t r = r0 + trow + j * RS;
            kp[j] = G.K + (r < N ? r : N - 1) * ld + 4 * tc4;
        }
        const float *sp[SV];
#pragma unroll
        for (int j = 0; j < SV; ++j) sp[j] = d.S + G.Soff + (int64_t)(trow + j * RS) * ld + 4 * tc4;
        f4x acc[RB][NC];
        float hi[RB][4], lo[RB][4];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[rb][c] = (f4x){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int v = 0; v < 4; ++v) hi[rb][v] = lo[rb][v] = 0.f;
        }
        auto col = [&](int64_t k) { const int64_t q = pstart + k; return (q >= nch ? q - nch : q) * BK; };
        f4x ka[KV], sa[SV], kb[KV], sb[SV];
        auto load = [&](f4x (&kr)[KV], f4x (&sr)[SV], int64_t k) {
            const int64_t c0 = col(k);
#pragma unroll
            for (int j = 0; j < KV; ++j) kr[j] = kload4(kp[j] + c0);
#pragma unroll
            for (int j = 0; j < SV; ++j) sr[j] = gload4(sp[j] + c0);
        };
        auto store = [&](const f4x (&kr)[KV], const f4x (&sr)[SV]) {
#pragma unroll
            for (int j = 0; j < KV; ++j) *reinterpret_cast<f4x *>(&As[(trow + j * RS) * PAD + 4 * tc4]) = kr[j];
#pragma unroll
            for (int j = 0; j < SV; ++j) *reinterpret_cast<f4x *>(&Bs[(trow + j * RS) * PAD + 4 * tc4]) = sr[j];
        };
        auto reduce = [&](int64_t k) {
#pragma unroll
            for (int sub = 0; sub < BK / 64; ++sub) {
                f4x a[RB][4], b[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    b[t] = *reinterpret_cast<const f4x *>(&Bs[lr * PAD + sub * 64 + 16 * t + 4 * lq]);
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb)
                        a[rb][t] = *reinterpret_cast<const f4x *>(
                            &As[(wave * 16 * RB + rb * 16 + lr) * PAD + sub * 64 + 16 * t + 4 * lq]);
                }
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
#pragma unroll
                        for (int rb = 0; rb < RB; ++rb)
                            acc[rb][t % NC] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                a[rb][t][e], b[t][e], acc[rb][t % NC], 0, 0, 0);
            }
            if (((k + 1) * BK) % 256 == 0 || k + 1 == nch) {
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) grp_two_sum(hi[rb][v], lo[rb][v], acc[rb][c][v]);
                        acc[rb][c] = (f4x){0.f, 0.f, 0.f, 0.f};
                    }
                }
            }
        };
        // Loads and stores are unconditional (slice indices clamp to the last
        // one; the redundant copies go to a buffer nobody reads and the next
        // tile re-stages after a barrier): conditional loads would make the
        // compiler's wait counts drain the younger slice too.  nch = ld / BK
        // is even (ld is a multiple of 512 columns).
        auto kc = [&](int64_t k) { return k < nch ? k : nch - 1; };
        load(ka, sa, 0);
        load(kb, sb, kc(1));
        __syncthreads();  // the previous tile's reads of As/Bs are done
        store(ka, sa);
        __syncthreads();
        load(ka, sa, kc(2));
        for (int64_t k = 0; k < nch; k += 2) {
            // LDS: slice k; kb: slice k+1; ka: slice k+2
            reduce(k);
            __syncthreads();
            store(kb, sb);
            __syncthreads();
            load(kb, sb, kc(k + 3));
            // LDS: slice k+1; ka: slice k+2; kb: slice k+3
            reduce(k + 1);
            __syncthreads();
            store(ka, sa);
            __syncthreads();
            load(ka, sa, kc(k + 4));
        }
        grp_store_rows<RB>(d, G, r0 + (int64_t)wave * 16 * RB, lr, lq, hi, lo);
    }
}

// ---------------- column-split work units on a balanced static schedule ----------------
// The row-block tiles above stream 0.6 MB (N = 1156) to 10.6 MB (N = 20692)
// each, and with ~1.1 tiles per resident workgroup the largest ones set the
// tail.  Here a tile's columns are split into chunks of ~2 MB ("units");
// the units are assigned to the resident workgroups by longest-first
// greedy balancing on the host (each workgroup gets a contiguous range of
// the unit list), every unit writes its FP64 partial row sums to a slot of
// P, and k_grp_sum_partials adds a row block's chunks in chunk order
// (deterministic; no atomics).
struct GrpUnit {
    int g, row0, k0, k1, slot, pad;  // slices [k0, k1) of BK columns
};
struct GrpRowBlock {
    int g, row0, slot0, nch;
};

template <int RB>
__global__ void __launch_bounds__(256, 3) k_grp_gemm_units(GrpDev d, const GrpUnit *__restrict__ units,
                                                        const int2 *__restrict__ wg_range,
                                                        double *__restrict__ P)
{
    if (!*d.live) return;
    constexpr int BR = 64 * RB, BK = 128 / RB, PAD = BK + 4, C4 = BK / 4;
    constexpr int KV = BR * C4 / 256, SV = 16 * C4 / 256, NC = 4 / RB, RS = 256 / C4;
    __shared__ __attribute__((aligned(16))) float As[BR * PAD];
    __shared__ __attribute__((aligned(16))) float Bs[16 * PAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const int trow = tid / C4, tc4 = tid % C4;
    const int2 range = wg_range[blockIdx.x];
    for (int ui = range.x; ui < range.y; ++ui) {
        const GrpUnit U = units[ui];
        const GrpGeo &G = d.geo[U.g];
        const int64_t N = G.N, ld = G.ld, r0 = U.row0;
        const int64_t nk = U.k1 - U.k0;  // even: chunks are multiples of 512 columns
        const int64_t pstart = ((int64_t)ui * 13) % nk;
        const float *kp[KV];
#pragma unroll
        for (int j = 0; j < KV; ++j) {
            const int64_t r = r0 + trow + j * RS;
            kp[j] = G.K + (r < N ? r : N - 1) * ld + 4 * tc4;
        }
        const float *sp[SV];
#pragma unroll
        for (int j = 0; j < SV; ++j) sp[j] = d.S + G.Soff + (int64_t)(trow + j * RS) * ld + 4 * tc4;
        f4x acc[RB][NC];
        float hi[RB][4], lo[RB][4];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[rb][c] = (f4x){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int v = 0; v < 4; ++v) hi[rb][v] = lo[rb][v] = 0.f;
        }
        auto col = [&](int64_t k) {
            const int64_t q = pstart + k;
            return (U.k0 + (q >= nk ? q - nk : q)) * BK;
        };
        f4x ka[KV], sa[SV], kb[KV], sb[SV];
        auto load = [&](f4x (&kr)[KV], f4x (&sr)[SV], int64_t k) {
            const int64_t c0 = col(k);
#pragma unroll
            for (int j = 0; j < KV; ++j) kr[j] = kload4(kp[j] + c0);
#pragma unroll
            for (int j = 0; j < SV; ++j) sr[j] = gload4(sp[j] + c0);
        };
        auto store = [&](const f4x (&kr)[KV], const f4x (&sr)[SV]) {
#pragma unroll
            for (int j = 0; j < KV; ++j) *reinterpret_cast<f4x *>(&As[(trow + j * RS) * PAD + 4 * tc4]) = kr[j];
#pragma unroll
            for (int j = 0; j < SV; ++j) *reinterpret_cast<f4x *>(&Bs[(trow + j * RS) * PAD + 4 * tc4]) = sr[j];
        };
        auto reduce = [&](int64_t k) {
#pragma unroll
            for (int sub = 0; sub < BK / 64; ++sub) {
                f4x a[RB][4], b[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    b[t] = *reinterpret_cast<const f4x *>(&Bs[lr * PAD + sub * 64 + 16 * t + 4 * lq]);
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb)
                        a[rb][t] = *reinterpret_cast<const f4x *>(
                            &As[(wave * 16 * RB + rb * 16 + lr) * PAD + sub * 64 + 16 * t + 4 * lq]);
                }
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
#pragma unroll
                        for (int rb = 0; rb < RB; ++rb)
                            acc[rb][t % NC] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                a[rb][t][e], b[t][e], acc[rb][t % NC], 0, 0, 0);
            }
            if (((k + 1) * BK) % 256 == 0 || k + 1 == nk) {
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) grp_two_sum(hi[rb][v], lo[rb][v], acc[rb][c][v]);
                        acc[rb][c] = (f4x){0.f, 0.f, 0.f, 0.f};
                    }
                }
            }
        };
        // as k_grp_gemm_lds2: two slices in flight in registers, loads and
        // stores unconditional (indices clamp to the last slice)
        auto kc = [&](int64_t k) { return k < nk ? k : nk - 1; };
        load(ka, sa, 0);
        load(kb, sb, kc(1));
        __syncthreads();  // the previous unit's reads of As/Bs are done
        store(ka, sa);
        __syncthreads();
        load(ka, sa, kc(2));
        for (int64_t k = 0; k < nk; k += 2) {
            reduce(k);
            __syncthreads();
            store(kb, sb);
            __syncthreads();
            load(kb, sb, kc(k + 3));
            reduce(k + 1);
            __syncthreads();
            store(ka, sa);
            __syncthreads();
            load(ka, sa, kc(k + 4));
        }
        // FP64 partial row sums of this unit: P[(slot * BR + local row) * M + member]
        if (lr < d.M) {
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int lrow = wave * 16 * RB + rb * 16 + 4 * lq + v;
                    if (r0 + lrow < N)
                        P[((int64_t)U.slot * BR + lrow) * d.M + lr] = (double)hi[rb][v] + (double)lo[rb][v];
                }
        }
    }
}

// Y = the sum of a row block's chunk partials, in chunk order; thread (local
// row, member) of block i of the row-block list.
template <int BR>
__global__ void __launch_bounds__(256) k_grp_sum_partials(GrpDev d, const GrpRowBlock *__restrict__ rbl,
                                                          const double *__restrict__ P)
{
    if (!*d.live) return;
    const GrpRowBlock B = rbl[blockIdx.y];
    const int M = d.M;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= BR * M) return;
    const int lrow = t / M, m = t % M;
    const GrpGeo &G = d.geo[B.g];
    if (B.row0 + lrow >= G.N) return;
    double y = 0.0;
    for (int c = 0; c < B.nch; ++c) y += P[((int64_t)(B.slot0 + c) * BR + lrow) * M + m];
    d.Y[G.Yoff + (int64_t)(B.row0 + lrow) * M + m] = y;
}

// ---------------- direct-to-LDS variant on column-split units ----------------
// The same units and partial sums as k_grp_gemm_units, on BR-row blocks,
// with every slice (BR rows x BK columns of K and 16 x BK of S) moved from
// HBM into LDS by global_load_lds_dwordx4: no register staging, so the
// depth of the NS-stage ring (NS - 1 slices in flight per workgroup) costs
// LDS, not registers.  One wavefront instruction moves 1 KB of whole row
// segments (BK/4 16-byte chunks per row); a lane's chunk j of row r lands
// in slot j but holds column chunk j ^ (r mod BK/4), so that the
// MFMA-layout reads (16 rows at one column chunk per quarter wavefront) fall
// in 16 distinct bank groups.  Wavefront w reduces columns [w BK/4, (w + 1)
// BK/4) of every slice for all BR rows; the four column partials are added
// through LDS at the end of a unit in a fixed order.
typedef __attribute__((address_space(3))) void lds_void;

template <int BR, int BK>
struct DmaShape {
    static constexpr int CH = BK / 4;                         // 16-byte chunks per row of a slice
    static constexpr int KSL = BR * BK, SSL = kGM * BK;       // floats of K and S per slice
    static constexpr int STAGE = KSL + SSL;
    static constexpr int KI = KSL / 4 / 256, SI = SSL / 4 / 256;  // DMA instructions per lane
    static constexpr int RB = BR / 16, T = CH / 16;           // row blocks; 4-chunk groups per wavefront
    static constexpr int FOLD = 1024 / BK;                    // slices per 256 columns of a wavefront
    static_assert(KI >= 1 && SI >= 1 && T >= 1 && RB >= 1, "slice shape");
};

template <bool NT>
__device__ __forceinline__ void dma16(const float *g, float *l)
{
    __builtin_amdgcn_global_load_lds(g, (lds_void *)l, 16, 0, NT ? 2 : 0);  // 2: the nt cache policy
}

// Workgroup barrier without the release fence of __syncthreads(): the
// fence would wait for every copy in flight (LDS DMA counts as a pending
// LDS write), draining the ring.  The copies a wavefront needs are waited
// for explicitly before it; the empty asm statements keep the compiler from
// moving memory accesses across it.
__device__ __forceinline__ void ring_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <int NS, int BR, int BK, int OCC>
__global__ void __launch_bounds__(256, OCC) k_grp_gemm_dma(GrpDev d, const GrpUnit *__restrict__ units,
                                                          const int2 *__restrict__ wg_range,
                                                          double *__restrict__ P)
{
    if (!*d.live) return;
    using Sh = DmaShape<BR, BK>;
    constexpr int CH = Sh::CH, KI = Sh::KI, SI = Sh::SI, RB = Sh::RB;
    extern __shared__ __attribute__((aligned(16))) float dsm[];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lr = lane & 15, lq = lane >> 4;
    const int2 range = wg_range[blockIdx.x];
    for (int ui = range.x; ui < range.y; ++ui) {
        const GrpUnit U = units[ui];
        const GrpGeo &G = d.geo[U.g];
        const int64_t N = G.N, ld = G.ld, r0 = U.row0;
        const int nk = U.k1 - U.k0;
        const float *kp[KI], *sp[SI];
#pragma unroll
        for (int i = 0; i < KI; ++i) {
            const int q = (wave * KI + i) * 64 + lane, r = q / CH, c = (q % CH) ^ (r % CH);
            const int64_t row = r0 + r < N ? r0 + r : N - 1;
            kp[i] = G.K + row * ld + (int64_t)U.k0 * BK + 4 * c;
        }
#pragma unroll
        for (int i = 0; i < SI; ++i) {
            const int q = (wave * SI + i) * 64 + lane, m = q / CH, c = (q % CH) ^ (m % CH);
            sp[i] = d.S + G.Soff + (int64_t)m * ld + (int64_t)U.k0 * BK + 4 * c;
        }
        // slice k (clamped to the last one: the copies past the end land in
        // stages nobody reads) into stage k mod NS; the sweep starts at a
        // different slice per unit (row pitches are multiples of 2 KB, so
        // aligned sweeps would load the same channels).  K is streamed with
        // the non-temporal policy, S (re-read by every row block) is not.
        const int rot = (int)(((int64_t)ui * 13) % nk);
        auto issue = [&](int k) {
            int kr = (k < nk ? k : nk - 1) + rot;
            kr = kr >= nk ? kr - nk : kr;
            const int64_t col = (int64_t)kr * BK;
            float *st = dsm + (k % NS) * Sh::STAGE;
#pragma unroll
            for (int i = 0; i < KI; ++i) dma16<true>(kp[i] + col, st + (wave * KI + i) * 256);
#pragma unroll
            for (int i = 0; i < SI; ++i) dma16<false>(sp[i] + col, st + Sh::KSL + (wave * SI + i) * 256);
        };
        f4x acc[RB];
        float hi[RB][4], lo[RB][4];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            acc[rb] = (f4x){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int v = 0; v < 4; ++v) hi[rb][v] = lo[rb][v] = 0.f;
        }
        __syncthreads();  // the previous unit's reads of the LDS are done
#pragma unroll
        for (int k = 0; k < NS - 1; ++k) issue(k);
        for (int k = 0; k < nk; ++k) {
            // slice k has landed once at most the NS - 2 younger slices
            // (KI + SI instructions each) are outstanding; then every
            // wavefront's part of it has
            __builtin_amdgcn_s_waitcnt(vmcnt_imm((NS - 2) * (KI + SI)));
            ring_barrier();
            issue(k + NS - 1);  // into the stage of slice k - 1, read by all
            const float *A = dsm + (k % NS) * Sh::STAGE, *B = A + Sh::KSL;
#pragma unroll
            for (int t = 0; t < Sh::T; ++t) {
                const int c = (CH / 4) * wave + 4 * t + lq;
                const f4x b = *reinterpret_cast<const f4x *>(&B[(lr * CH + (c ^ (lr % CH))) * 4]);
                f4x a[RB];
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    const int r = 16 * rb + lr;
                    a[rb] = *reinterpret_cast<const f4x *>(&A[(r * CH + (c ^ (r % CH))) * 4]);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb)
                        acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rb][e], b[e], acc[rb], 0, 0, 0);
            }
            if ((k + 1) % Sh::FOLD == 0 || k + 1 == nk) {  // every 256 columns of this wavefront
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) grp_two_sum(hi[rb][v], lo[rb][v], acc[rb][v]);
                    acc[rb] = (f4x){0.f, 0.f, 0.f, 0.f};
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
        __syncthreads();  // every copy has landed and every read is done
        double *red = reinterpret_cast<double *>(dsm);  // [wave][row][member]
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                red[(wave * BR + 16 * rb + 4 * lq + v) * kGM + lr] = (double)hi[rb][v] + (double)lo[rb][v];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < BR * kGM / 256; ++e) {
            const int x = tid + 256 * e, row = x / kGM, m = x % kGM;
            const double y = ((red[x] + red[BR * kGM + x]) + red[2 * BR * kGM + x]) + red[3 * BR * kGM + x];
            if (m < d.M && r0 + row < N) P[((int64_t)U.slot * BR + row) * d.M + m] = y;
        }
    }
}

// ---------------- lockstep Tsit5 over all members ----------------
__constant__ double kGA[7][6] = {
    {0, 0, 0, 0, 0, 0},
    {0.161, 0, 0, 0, 0, 0},
    {-0.008480655492356989, 0.335480655492357, 0, 0, 0, 0},
    {2.897153057105493, -6.359448489975075, 4.3622954328695815, 0, 0, 0},
    {5.325864828439257, -11.748883564062828, 7.4955393428898365, -0.09249506636175525, 0, 0},
    {5.86145544294642, -12.92096931784711, 8.159367898576159, -0.071584973281401,
     -0.028269050394068383, 0},
    {0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742, -3.290069515436081,
     2.324710524099774}};
__constant__ double kGBt[7] = {-0.00178001105222577714, -0.0008164344596567469,
                               0.007880878010261995, -0.1447110071732629, 0.5823571654525552,
                               -0.45808210592918697, 0.015151515151515152};

struct GrpCtrl {  // arrays of G*M
    double *t, *dt, *dtcur, *qold;
    long long *naccept, *nreject, *iter;
    int *done, *accept, *retcode;
    const int64_t *soffs;  // G+1
    const int64_t *nst;    // G: state entries per member
    double t1, reltol, abstol, dtmax, dtmin;
    long long maxiters;
    int G, M;
    int *live;  // the flag GrpDev::live reads (k_grp_live)
};

// After k_grp_begin: live = 1 while any member is still running.  One block.
__global__ void __launch_bounds__(1024) k_grp_live(GrpCtrl c)
{
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    int mine = 0;
    for (int m = threadIdx.x; m < c.G * c.M; m += blockDim.x) mine |= !c.done[m];
    if (mine) any = 1;  // benign race: every writer writes 1
    __syncthreads();
    if (threadIdx.x == 0) *c.live = any;
}

__global__ void k_grp_begin(GrpCtrl c)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= c.G * c.M) return;
    c.accept[q] = 0;
    c.dtcur[q] = 0.0;
    if (c.done[q]) return;
    if (!(c.t[q] < c.t1)) { c.done[q] = 1; c.retcode[q] = OQ_RET_SUCCESS; return; }
    if (c.iter[q] >= c.maxiters) { c.done[q] = 1; c.retcode[q] = OQ_RET_MAXITERS; return; }
    double dt = c.dt[q];
    if (dt > c.dtmax) dt = c.dtmax;
    if (c.t[q] + dt > c.t1) dt = c.t1 - c.t[q];
    if (!(dt >= c.dtmin) || !isfinite(dt)) {
        c.done[q] = 1;
        c.retcode[q] = isfinite(dt) ? OQ_RET_DTMIN : OQ_RET_NONFINITE;
        return;
    }
    c.dtcur[q] = dt;
    c.iter[q] += 1;
}

// member slot (g M + m) of state entry i; all threads of the block call it
__device__ __forceinline__ int64_t grp_member(const GrpCtrl &c, int64_t i)
{
    const int g = seg_find_block(c.soffs, c.G, i);
    return (int64_t)g * c.M + (i - c.soffs[g]) % c.M;
}

__global__ void __launch_bounds__(256) k_grp_stage(GrpCtrl c, int s, int64_t ntot,
                                                   const double *__restrict__ u,
                                                   const double *__restrict__ k,
                                                   double *__restrict__ y)
{
    if (!*c.live) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t q = grp_member(c, i);
    if (i >= ntot) return;
    const double dt = c.dtcur[q];
    double acc = 0.0;
    for (int j = 0; j < s; ++j) acc += kGA[s][j] * k[j * ntot + i];
    y[i] = u[i] + dt * acc;
}

__global__ void __launch_bounds__(256) k_grp_err(GrpCtrl c, int64_t ntot,
                                                 const double *__restrict__ u,
                                                 const double *__restrict__ y,
                                                 const double *__restrict__ k,
                                                 double *__restrict__ part)
{
    if (!*c.live) return;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    const int64_t GM = (int64_t)c.G * c.M;
    if (q >= GM) return;
    const int64_t g = q / c.M, m = q % c.M, M = c.M, n = c.nst[g], base = c.soffs[g];
    const int64_t r0 = n * chunk / kGChunks, r1 = n * (chunk + 1) / kGChunks;
    const double dt = c.dtcur[q];
    double acc = 0.0;
    for (int64_t r = r0; r < r1; ++r) {
        const int64_t i = base + r * M + m;
        double e = 0.0;
#pragma unroll
        for (int j = 0; j < 7; ++j) e += kGBt[j] * k[j * ntot + i];
        const double w = dt * e / (c.abstol + fmax(fabs(u[i]), fabs(y[i])) * c.reltol);
        acc += w * w;
    }
    part[chunk * GM + q] = acc;
}

__global__ void k_grp_controller(GrpCtrl c, const double *__restrict__ part)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int GM = c.G * c.M;
    if (q >= GM || c.done[q]) return;
    double acc = 0.0;
    for (int ch = 0; ch < kGChunks; ++ch) acc += part[ch * GM + q];
    const double beta1 = 7.0 / 50.0, beta2 = 2.0 / 25.0, qmin = 0.2, qmax = 10.0, gamma = 0.9,
                 qoldinit = 1e-4;
    const double eest = sqrt(acc / (double)c.nst[q / c.M]);
    double q11, qq;
    if (eest == 0.0) {
        q11 = 0.0;
        qq = 1.0 / qmax;
    } else {
        q11 = pow(eest, beta1);
        qq = q11 / pow(c.qold[q], beta2);
        qq = fmax(1.0 / qmax, fmin(1.0 / qmin, qq / gamma));
    }
    const double dtc = c.dtcur[q];
    if (eest <= 1.0) {
        c.t[q] = (c.t1 - (c.t[q] + dtc) <= 0.0) ? c.t1 : c.t[q] + dtc;
        c.qold[q] = fmax(eest, qoldinit);
        c.dt[q] = dtc / qq;
        c.naccept[q] += 1;
        c.accept[q] = 1;
    } else {
        c.nreject[q] += 1;
        if (!isfinite(eest)) {
            c.done[q] = 1;
            c.retcode[q] = OQ_RET_NONFINITE;
            return;
        }
        c.dt[q] = dtc / fmin(1.0 / qmin, q11 / gamma);
    }
}

__global__ void __launch_bounds__(256) k_grp_commit(GrpCtrl c, int64_t ntot,
                                                    double *__restrict__ u,
                                                    const double *__restrict__ y,
                                                    double *__restrict__ k)
{
    if (!*c.live) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t q = grp_member(c, i);
    if (i >= ntot || !c.accept[q]) return;
    u[i] = y[i];
    k[i] = k[6 * ntot + i];
}

// tile shapes: RB in {1, 2, 4} x {software-pipelined, plain} register
// kernels, and the LDS-staged kernels with RB in {1, 2}, one or two slices
// in flight (a third slice in flight measured the same as two: the kernel
// is not limited by its prefetch depth), and (shape 10) the two-slice LDS
// kernel with RB = 2 on column-split units balanced over the workgroups;
// (shapes 11-14) the direct-to-LDS kernel on units of BR rows with BK-column
// slices in an NS-stage ring (NS, BR, BK, workgroups per CU in kDma; a
// 4-stage ring of the 64 x 128 slices measured the same as 3 stages, 32 x
// 256 slices and a 6-stage ring of 64 x 64 slices (one workgroup per CU)
// 2-7 % slower than two workgroups of 3 stages, and the same kernel without
// the non-temporal policy on K 8 % slower; the register-streaming kernels,
// whose lanes load 64-byte row pieces, were slower with it)
static constexpr int kGShapes = 15;
static constexpr int kGUnitShape = 10;  // first of the unit shapes 10..14
static constexpr int kGUnitShapes = 5;
static constexpr int kGUnitRB = 2, kGUnitBR = 64 * kGUnitRB, kGUnitBK = 128 / kGUnitRB;
static constexpr int64_t kGUnitBytes = (int64_t)2 << 20;  // target bytes of K per unit
struct DmaCfg {
    int ns, br, bk, occ;
};
static constexpr DmaCfg kDma[4] = {{3, 64, 128, 1}, {3, 32, 128, 2}, {3, 64, 64, 2}, {2, 64, 64, 4}};
static inline bool is_unit_shape(int s) { return s >= kGUnitShape; }
static inline const DmaCfg &dma_cfg(int s) { return kDma[s - kGUnitShape - 1]; }
static inline int unit_br(int s) { return s == kGUnitShape ? kGUnitBR : dma_cfg(s).br; }
static inline int unit_bk(int s) { return s == kGUnitShape ? kGUnitBK : dma_cfg(s).bk; }
static inline size_t dma_lds(int s)
{
    const DmaCfg &c = dma_cfg(s);
    return (size_t)c.ns * (size_t)(c.br + kGM) * (size_t)c.bk * 4;
}
static inline int shape_rb(int s)
{
    if (is_unit_shape(s)) return s == kGUnitShape ? kGUnitRB : 1;
    return s < 6 ? (s / 2 == 0 ? 1 : (s / 2 == 1 ? 2 : 4)) : (s % 2 == 0 ? 1 : 2);
}

// Units of a unit shape over a set of geometries, with their schedule.
struct GrpUnitTable {
    oq::DevBuf units, wg, rbl;
    int nunits = 0, nwg = 0, nrbl = 0, nslots = 0;
};

// ---------------- lockstep VCABM5 over all members ----------------
// As the ensemble's lockstep VCABM5, over the members of every geometry:
// phi = k[0, 6 ntot), f(p) = k[6 ntot, 7 ntot), p = y; yn, fn in ab.
struct GrpAb {
    int *k, *kcur;            // G*M
    double *hist, *g, *beta;  // G*M x (kAbK + 1)
    double *yn, *fn;          // ntot each
};

__global__ void k_gab_coeffs(GrpCtrl c, GrpAb a)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= c.G * c.M || c.done[q] || !(c.dtcur[q] > 0.0)) return;
    const int k = a.k[q];
    a.kcur[q] = k;
    ab_coeffs(k, c.dtcur[q], a.hist + q * (kAbK + 1), a.g + q * (kAbK + 1), a.beta + q * (kAbK + 1));
}

__global__ void __launch_bounds__(256) k_gab_predict(GrpCtrl c, GrpAb a, int64_t ntot,
                                                     const double *__restrict__ u,
                                                     const double *__restrict__ phi, double *__restrict__ p)
{
    if (!*c.live) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t q = grp_member(c, i);
    if (i >= ntot) return;
    const double h = c.dtcur[q];
    if (!(h > 0.0)) {
        p[i] = u[i];
        return;
    }
    const int k = a.kcur[q];
    const double *g = a.g + q * (kAbK + 1), *be = a.beta + q * (kAbK + 1);
    double acc = 0.0;
    for (int j = 0; j < k; ++j) acc += g[j] * be[j] * phi[j * ntot + i];
    p[i] = u[i] + h * acc;
}

__global__ void __launch_bounds__(256) k_gab_correct(GrpCtrl c, GrpAb a, int64_t ntot,
                                                     const double *__restrict__ u,
                                                     const double *__restrict__ p,
                                                     const double *__restrict__ fp,
                                                     const double *__restrict__ phi,
                                                     double *__restrict__ part)
{
    if (!*c.live) return;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    const int64_t GM = (int64_t)c.G * c.M;
    if (q >= GM) return;
    const int64_t g = q / c.M, m = q % c.M, M = c.M, n = c.nst[g], base = c.soffs[g];
    const int64_t r0 = n * chunk / kGChunks, r1 = n * (chunk + 1) / kGChunks;
    const double h = c.dtcur[q];
    const int k = a.kcur[q];
    const double *gg = a.g + q * (kAbK + 1), *be = a.beta + q * (kAbK + 1);
    const double gk = h > 0.0 ? gg[k] : 0.0, dg = h > 0.0 ? gg[k] - gg[k - 1] : 0.0;
    double acc = 0.0;
    for (int64_t r = r0; r < r1; ++r) {
        const int64_t i = base + r * M + m;
        double ph = fp[i];
        if (h > 0.0)
            for (int j = 0; j < k; ++j) ph -= be[j] * phi[j * ntot + i];
        const double yn = p[i] + h * gk * ph;
        a.yn[i] = yn;
        const double w = h * dg * ph / (c.abstol + fmax(fabs(u[i]), fabs(yn)) * c.reltol);
        acc += w * w;
    }
    part[chunk * GM + q] = acc;
}

__global__ void k_gab_controller(GrpCtrl c, GrpAb a, const double *__restrict__ part)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int GM = c.G * c.M;
    if (q >= GM || c.done[q] || !(c.dtcur[q] > 0.0)) return;
    double acc = 0.0;
    for (int ch = 0; ch < kGChunks; ++ch) acc += part[ch * GM + q];
    const double eest = sqrt(acc / (double)c.nst[q / c.M]);
    const int k = a.kcur[q];
    const double f = (eest == 0.0) ? 0.5 : fmax(0.5, fmin(5.0, pow(eest, 1.0 / (k + 1)) / 0.9));
    const double dtc = c.dtcur[q];
    if (eest <= 1.0) {
        c.t[q] = (c.t1 - (c.t[q] + dtc) <= 0.0) ? c.t1 : c.t[q] + dtc;
        double *hist = a.hist + q * (kAbK + 1);
        for (int j = kAbK; j > 0; --j) hist[j] = hist[j - 1];
        hist[0] = dtc;
        if (a.k[q] < kAbK) a.k[q] += 1;
        c.dt[q] = dtc / f;
        c.naccept[q] += 1;
        c.accept[q] = 1;
    } else {
        c.nreject[q] += 1;
        if (!isfinite(eest)) {
            c.done[q] = 1;
            c.retcode[q] = OQ_RET_NONFINITE;
            return;
        }
        c.dt[q] = dtc / f;
    }
}

__global__ void __launch_bounds__(256) k_gab_commit(GrpCtrl c, GrpAb a, int64_t ntot, double *__restrict__ u,
                                                    double *__restrict__ phi)
{
    if (!*c.live) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t q = grp_member(c, i);
    if (i >= ntot || !c.accept[q]) return;
    const int k = a.kcur[q];
    const double *be = a.beta + q * (kAbK + 1);
    double tmp = a.fn[i];
    for (int j = 0; j < k; ++j) {
        const double old = phi[j * ntot + i];
        phi[j * ntot + i] = tmp;
        tmp -= be[j] * old;
    }
    phi[k * ntot + i] = tmp;
    u[i] = a.yn[i];
}

}  // namespace oq

struct oq_group {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int64_t G = 0, M = 0, ntot = 0, nprolog = 0;
    std::vector<oq_rhs *> rhs;  // per geometry: K (FP32), viscosity arrays
    std::vector<oq::GrpGeo> geo;
    std::vector<int64_t> soffs, nst;
    int bucket_shape[oq::kGBuckets];
    std::vector<int> geo_bucket;
    std::vector<std::vector<int2>> tiles_of_shape;  // host tables per shape
    int64_t nrows = 0;
    oq::DevBuf dgeo, dsoffs, dpoffs, dyoffs, dnst, a, b, L, sigma, eta, vpl, f0, v0, S, Y;
    oq::DevBuf dtiles[oq::kGShapes];
    int ntiles[oq::kGShapes] = {0};
    int grid[oq::kGShapes] = {0};
    oq::GrpUnitTable ut[oq::kGUnitShapes];  // unit shapes over the geometries that use them
    oq::DevBuf P;         // partial row sums of the units (capacity pcap doubles)
    size_t pcap = 0;
    oq::DevBuf u, y, k, part;
    oq::DevBuf ct, cdt, cdtcur, cqold, cna, cnr, citer, cdone, cacc, cret;
    oq::DevBuf live;  // GrpDev::live
    oq::GrpDev dev{};
    oq::GrpCtrl ctrl{};
    bool stepping = false;
    bool started = false;  // the controller arrays hold a run (get_ctrl)
    int alg = 0;           // 0 Tsit5, 1 VCABM5 (oq_group_start_vcabm5)
    oq::DevBuf abk, abkcur, abhist, abg, abbeta, ab;  // VCABM5 state (ab: yn, fn)
    oq::GrpAb gab{};
    // the captured lockstep step: [0] one step, [1] kGraphSteps steps back
    // to back (fewer launches; steps past the end of the run are no-ops)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
};

namespace oq {

static constexpr int kGraphSteps = 8;  // steps per launch of the multi-step graph

static void drop_graphs(oq_group *e)
{
    for (int q = 0; q < 2; ++q) {
        if (e->exec[q]) { (void)hipGraphExecDestroy(e->exec[q]); e->exec[q] = nullptr; }
        if (e->graph[q]) { (void)hipGraphDestroy(e->graph[q]); e->graph[q] = nullptr; }
    }
}

static inline unsigned gbk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

static int bucket_of(int64_t N)
{
    int b = 0;
    while (b + 1 < kGBuckets && N >= ((int64_t)2048 << b)) ++b;
    return b;
}

template <class F>
static int occupancy_grid(F kern, int *grid, size_t dyn_lds = 0)
{
    int nb = 0, dev = 0;
    OQ_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    OQ_HIP(hipGetDeviceProperties(&prop, dev));
    OQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 256, dyn_lds));
    *grid = (nb > 0 ? nb : 1) * prop.multiProcessorCount;
    return OQ_OK;
}

// host tile table of shape s over the geometries in `sel`, largest first
static std::vector<int2> build_tiles(const oq_group *e, int s, const std::vector<int> &sel)
{
    const int rows = kGWaves * shape_rb(s) * 16;
    std::vector<int> order(sel);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return e->geo[x].ld > e->geo[y].ld; });
    std::vector<int2> t;
    for (int g : order)
        for (int64_t r = 0; r < e->geo[g].N; r += rows) t.push_back(make_int2(g, (int)r));
    return t;
}

// Units of unit shape s over the geometries in `sel`: every row block's
// slices split into chunks of ~kGUnitBytes (multiples of 512 columns), then
// longest-first greedy assignment to `nwg` workgroups.  Partials of row
// block i go to slots [slot0, slot0 + nch).
static int build_units(oq_group *e, int s, const std::vector<int> &sel, GrpUnitTable &T)
{
    const int BR = unit_br(s), BK = unit_bk(s);
    std::vector<GrpUnit> units;
    std::vector<GrpRowBlock> rbl;
    int slot = 0;
    for (int g : sel) {
        const GrpGeo &G = e->geo[g];
        const int64_t nslices = G.ld / BK;                       // multiple of 4
        const int64_t per = std::max<int64_t>(8, kGUnitBytes / ((int64_t)BR * BK * 4) / 8 * 8);
        const int64_t nch = (nslices + per - 1) / per;
        const int64_t cs = (nslices / nch + 7) / 8 * 8;          // slices per chunk
        for (int64_t r0 = 0; r0 < G.N; r0 += BR) {
            int c = 0;
            for (int64_t k0 = 0; k0 < nslices; k0 += cs, ++c) {
                const int64_t k1 = std::min(nslices, k0 + cs);
                units.push_back(GrpUnit{g, (int)r0, (int)k0, (int)k1, slot + c, 0});
            }
            rbl.push_back(GrpRowBlock{g, (int)r0, slot, c});
            slot += c;
        }
    }
    // longest first onto the least loaded workgroup (cost = rows x columns)
    const int W = std::max(1, std::min(e->grid[s], (int)units.size()));
    auto cost = [&](const GrpUnit &u) {
        const int64_t rows = std::min<int64_t>(BR, e->geo[u.g].N - u.row0);
        return rows * (int64_t)(u.k1 - u.k0);
    };
    std::vector<int> order(units.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return cost(units[a]) > cost(units[b]); });
    std::vector<std::pair<int64_t, int>> heap;  // (load, workgroup), min-heap
    for (int w = 0; w < W; ++w) heap.push_back({0, w});
    std::vector<std::vector<int>> bins((size_t)W);
    auto cmp = [](const std::pair<int64_t, int> &a, const std::pair<int64_t, int> &b) { return a > b; };
    for (int i : order) {
        std::pop_heap(heap.begin(), heap.end(), cmp);
        heap.back().first += cost(units[i]);
        bins[(size_t)heap.back().second].push_back(i);
        std::push_heap(heap.begin(), heap.end(), cmp);
    }
    std::vector<GrpUnit> sched;
    std::vector<int2> wg((size_t)W);
    for (int w = 0; w < W; ++w) {
        wg[(size_t)w].x = (int)sched.size();
        for (int i : bins[(size_t)w]) sched.push_back(units[i]);
        wg[(size_t)w].y = (int)sched.size();
    }
    T.units.release(); T.wg.release(); T.rbl.release();
    T.nunits = (int)sched.size();
    T.nwg = W;
    T.nrbl = (int)rbl.size();
    T.nslots = slot;
    if (sched.empty()) return OQ_OK;
    int rc = T.units.upload(sched.data(), sched.size() * sizeof(GrpUnit), e->stream);
    if (!rc) rc = T.wg.upload(wg.data(), wg.size() * sizeof(int2), e->stream);
    if (!rc) rc = T.rbl.upload(rbl.data(), rbl.size() * sizeof(GrpRowBlock), e->stream);
    const size_t need = (size_t)slot * BR * (size_t)e->M;
    if (!rc && need > e->pcap) {
        rc = e->P.alloc(need * 8);
        e->pcap = rc ? 0 : need;
    }
    return rc;
}

static int launch_units(oq_group *e, int s, const GrpUnitTable &T)
{
    if (T.nunits == 0) return OQ_OK;
    const dim3 grid((unsigned)T.nwg), blk(256);
    if (s == kGUnitShape) {
        hipLaunchKernelGGL((k_grp_gemm_units<kGUnitRB>), grid, blk, 0, e->stream, e->dev,
                           T.units.as<GrpUnit>(), T.wg.as<int2>(), e->P.as<double>());
        hipLaunchKernelGGL((k_grp_sum_partials<kGUnitBR>), dim3(gbk(kGUnitBR * e->M, 256), (unsigned)T.nrbl),
                           blk, 0, e->stream, e->dev, T.rbl.as<GrpRowBlock>(), e->P.as<double>());
    } else {
        const GrpUnit *un = T.units.as<GrpUnit>();
        const int2 *wr = T.wg.as<int2>();
        double *P = e->P.as<double>();
        switch (s - kGUnitShape - 1) {
        case 0: hipLaunchKernelGGL((k_grp_gemm_dma<3, 64, 128, 1>), grid, blk, dma_lds(s), e->stream, e->dev, un, wr, P); break;
        case 1: hipLaunchKernelGGL((k_grp_gemm_dma<3, 32, 128, 2>), grid, blk, dma_lds(s), e->stream, e->dev, un, wr, P); break;
        case 2: hipLaunchKernelGGL((k_grp_gemm_dma<3, 64, 64, 2>), grid, blk, dma_lds(s), e->stream, e->dev, un, wr, P); break;
        default: hipLaunchKernelGGL((k_grp_gemm_dma<2, 64, 64, 4>), grid, blk, dma_lds(s), e->stream, e->dev, un, wr, P); break;
        }
        if (unit_br(s) == 32)
            hipLaunchKernelGGL((k_grp_sum_partials<32>), dim3(gbk(32 * e->M, 256), (unsigned)T.nrbl), blk, 0,
                               e->stream, e->dev, T.rbl.as<GrpRowBlock>(), P);
        else
            hipLaunchKernelGGL((k_grp_sum_partials<64>), dim3(gbk(64 * e->M, 256), (unsigned)T.nrbl), blk, 0,
                               e->stream, e->dev, T.rbl.as<GrpRowBlock>(), P);
    }
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

static int upload_tiles(oq_group *e)
{
    for (int s = 0; s < kGShapes; ++s) {
        std::vector<int> sel;
        for (int g = 0; g < (int)e->G; ++g)
            if (e->bucket_shape[e->geo_bucket[g]] == s) sel.push_back(g);
        if (is_unit_shape(s)) {
            int rc = build_units(e, s, sel, e->ut[s - kGUnitShape]);
            if (rc) return rc;
            continue;
        }
        const std::vector<int2> t = build_tiles(e, s, sel);
        e->ntiles[s] = (int)t.size();
        e->dtiles[s].release();
        if (!t.empty()) {
            int rc = e->dtiles[s].upload(t.data(), t.size() * sizeof(int2), e->stream);
            if (rc) return rc;
        }
    }
    OQ_HIP(hipStreamSynchronize(e->stream));
    return OQ_OK;
}

static int launch_gemm(oq_group *e, int s, const int2 *tiles, int ntiles, const double *u,
                       double *du)
{
    if (ntiles == 0) return OQ_OK;
    const unsigned grid = (unsigned)std::min(ntiles, e->grid[s]);
    switch (s) {
    case 0: hipLaunchKernelGGL((k_grp_gemm<1, true>), dim3(grid), dim3(256), 0, e->stream, e->dev, tiles, ntiles, u, du); break;
    case 1: hipLaunchKernelGGL((k_grp_gemm<1, false>), dim3(grid), dim3(256), 0, e->stream, e->dev, tiles, ntiles, u, du); break;
    case 2: hipLaunchKernelGGL((k_grp_gemm<2, true>), dim3(grid), dim3(256), 0, e->stream, e->dev, tiles, ntiles, u, du); break;
    case 3: hipLaunchKernelGGL((k_grp_gemm<2, false>), dim3(grid), dim3(256), 0, e->stream, e->dev, tiles, ntiles, u, du); break;
    case 4: hipLaunchKernelGGL((k_grp_gemm<4, true>), dim3(grid), dim3(256), 0, e->stream, e->dev, tiles, ntiles, u, du); break;
    case 5: hipLaunchKernelGGL((k_grp_gemm<4, false>), dim3(grid), dim3(256), 0, e->stream, e->dev, tiles, ntiles, u, du); break;
    case 6: hipLaunchKernelGGL((k_grp_gemm_lds<1>), dim3(grid), dim3(256), 0, e->stream, e->dev, tiles, ntiles, u, du); break;
    case 7: hipLaunchKernelGGL((k_grp_gemm_lds<2>), dim3(grid), dim3(256), 0, e->stream, e->dev, tiles, ntiles, u, du); break;
    case 8: hipLaunchKernelGGL((k_grp_gemm_lds2<1>), dim3(grid), dim3(256), 0, e->stream, e->dev, tiles, ntiles, u, du); break;
    default: hipLaunchKernelGGL((k_grp_gemm_lds2<2>), dim3(grid), dim3(256), 0, e->stream, e->dev, tiles, ntiles, u, du); break;
    }
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

static int grp_rhs(oq_group *e, const double *u, double *du)
{
    hipLaunchKernelGGL(k_grp_prologue, dim3(gbk(e->nprolog, 256)), dim3(256), 0, e->stream, e->dev,
                       u, du);
    OQ_HIP(hipGetLastError());
    for (int s = 0; s < kGShapes; ++s) {
        int rc = is_unit_shape(s) ? launch_units(e, s, e->ut[s - kGUnitShape])
                                  : launch_gemm(e, s, e->dtiles[s].as<int2>(), e->ntiles[s], u, du);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_grp_finish, dim3(gbk(e->nrows, 256)), dim3(256), 0, e->stream, e->dev, u,
                       du);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

static int grp_step_ab(oq_group *e);

static int grp_step(oq_group *e)
{
    if (e->alg == 1) return grp_step_ab(e);
    const int64_t nt = e->ntot;
    const int GM = (int)(e->G * e->M);
    double *u = e->u.as<double>(), *y = e->y.as<double>(), *k = e->k.as<double>();
    int rc;
    hipLaunchKernelGGL(k_grp_begin, dim3(gbk(GM, 256)), dim3(256), 0, e->stream, e->ctrl);
    hipLaunchKernelGGL(k_grp_live, dim3(1), dim3(1024), 0, e->stream, e->ctrl);
    for (int s = 1; s < 7; ++s) {
        hipLaunchKernelGGL(k_grp_stage, dim3(gbk(nt, 256)), dim3(256), 0, e->stream, e->ctrl, s, nt,
                           u, k, y);
        if ((rc = grp_rhs(e, y, k + s * nt))) return rc;
    }
    hipLaunchKernelGGL(k_grp_err, dim3(gbk(GM, 256), kGChunks), dim3(256), 0, e->stream, e->ctrl, nt,
                       u, y, k, e->part.as<double>());
    hipLaunchKernelGGL(k_grp_controller, dim3(gbk(GM, 256)), dim3(256), 0, e->stream, e->ctrl,
                       e->part.as<double>());
    hipLaunchKernelGGL(k_grp_commit, dim3(gbk(nt, 256)), dim3(256), 0, e->stream, e->ctrl, nt, u, y,
                       k);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

// One attempted lockstep VCABM5 step: two grouped RHS.
static int grp_step_ab(oq_group *e)
{
    const int64_t nt = e->ntot;
    const int GM = (int)(e->G * e->M);
    double *u = e->u.as<double>(), *p = e->y.as<double>(), *phi = e->k.as<double>(), *fp = phi + 6 * nt;
    const dim3 gq(gbk(GM, 256)), gn(gbk(nt, 256)), blk(256);
    int rc;
    hipLaunchKernelGGL(k_grp_begin, gq, blk, 0, e->stream, e->ctrl);
    hipLaunchKernelGGL(k_gab_coeffs, gq, blk, 0, e->stream, e->ctrl, e->gab);
    hipLaunchKernelGGL(k_grp_live, dim3(1), dim3(1024), 0, e->stream, e->ctrl);
    hipLaunchKernelGGL(k_gab_predict, gn, blk, 0, e->stream, e->ctrl, e->gab, nt, u, phi, p);
    if ((rc = grp_rhs(e, p, fp))) return rc;
    hipLaunchKernelGGL(k_gab_correct, dim3(gbk(GM, 256), kGChunks), blk, 0, e->stream, e->ctrl, e->gab, nt, u, p,
                       fp, phi, e->part.as<double>());
    hipLaunchKernelGGL(k_gab_controller, gq, blk, 0, e->stream, e->ctrl, e->gab, e->part.as<double>());
    if ((rc = grp_rhs(e, e->gab.yn, e->gab.fn))) return rc;
    hipLaunchKernelGGL(k_gab_commit, gn, blk, 0, e->stream, e->ctrl, e->gab, nt, u, phi);
    OQ_HIP(hipGetLastError());
    return OQ_OK;
}

// Outside lockstep runs the RHS kernels must run: live = 1 (a copy on the
// stream, ordered before what follows).
static int set_live(oq_group *e)
{
    static const int one = 1;
    OQ_HIP(hipMemcpyAsync(e->live.p, &one, sizeof(int), hipMemcpyHostToDevice, e->stream));
    return OQ_OK;
}

// Drops the captured step: its kernel arguments point at tile tables and
// buffers that have just been replaced (oq_group_start captures anew).
static void drop_step(oq_group *e)
{
    if (e->exec[0] || e->graph[0]) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    e->stepping = false;
}

static void grp_destroy(oq_group *e)
{
    if (!e) return;
    const bool ok = teardown_sync(e->stream);
    if (ok) {
        for (int q = 0; q < 2; ++q) {
            if (e->exec[q]) (void)hipGraphExecDestroy(e->exec[q]);
            if (e->graph[q]) (void)hipGraphDestroy(e->graph[q]);
        }
    }
    for (oq_rhs *h : e->rhs)
        if (h) oq_rhs_destroy(h);
    if (ok) {
        if (e->ev0) (void)hipEventDestroy(e->ev0);
        if (e->ev1) (void)hipEventDestroy(e->ev1);
        if (e->stream) (void)hipStreamDestroy(e->stream);
    }
    delete e;
}

// host member-major states (geometry g: M members of n_g entries) <-> device
// entry-major segments
static void to_entry_major(const oq_group *e, const double *src, std::vector<double> &dst)
{
    dst.resize((size_t)e->ntot);
    const int64_t M = e->M;
    int64_t hoff = 0;
    for (int64_t g = 0; g < e->G; ++g) {
        const int64_t n = e->nst[g], base = e->soffs[g];
        for (int64_t m = 0; m < M; ++m)
            for (int64_t i = 0; i < n; ++i) dst[(size_t)(base + i * M + m)] = src[hoff + m * n + i];
        hoff += n * M;
    }
}

static void from_entry_major(const oq_group *e, const std::vector<double> &src, double *dst)
{
    const int64_t M = e->M;
    int64_t hoff = 0;
    for (int64_t g = 0; g < e->G; ++g) {
        const int64_t n = e->nst[g], base = e->soffs[g];
        for (int64_t m = 0; m < M; ++m)
            for (int64_t i = 0; i < n; ++i) dst[hoff + m * n + i] = src[(size_t)(base + i * M + m)];
        hoff += n * M;
    }
}

}  // namespace oq

using namespace oq;

extern "C" {

int oq_group_create(int device, int64_t ngeo, const oq_fault_mesh *mf,
                    const oq_hex8_mesh *const *ma, const oq_visc_props *const *visc,
                    double lambda, double mu, int ftype, int64_t nrept, double buffer_ratio,
                    const oq_quadrature *quad, int64_t nmembers, const oq_rs_props *members,
                    int kdtype, oq_group **out)
{
    OQ_REQUIRE(out != nullptr && mf != nullptr && members != nullptr, "NULL argument");
    *out = nullptr;
    OQ_REQUIRE(ngeo >= 1 && ngeo <= (1 << 20), "ngeo must be in [1, 2^20]");
    OQ_REQUIRE(nmembers >= 1 && nmembers <= kGM, "nmembers must be in [1, 16]");
    OQ_REQUIRE(kdtype == OQ_K_FP32, "the grouped operator is FP32");
    for (int64_t q = 0; q < ngeo * nmembers; ++q) {
        const oq_rs_props &p = members[q];
        OQ_REQUIRE(p.a && p.b && p.L && p.sigma, "member property array is NULL");
        OQ_REQUIRE(p.f0 > 0.0 && p.v0 > 0.0 && p.eta > 0.0 && p.vpl > 0.0,
                   "RateStateQuasiDynamicProperty requires f₀ > 0, v₀ > 0, η > 0, vpl > 0");
    }
    for (int64_t g = 0; g < ngeo; ++g)
        if (ma && ma[g]) OQ_REQUIRE(visc && visc[g], "a geometry with a mantle needs its viscosity");
    oq_group *e = new oq_group();
    e->device = device;
    e->G = ngeo;
    e->M = nmembers;
    for (int b = 0; b < kGBuckets; ++b) e->bucket_shape[b] = 1;  // RB = 1, plain
    int rc = OQ_OK;
    for (int64_t g = 0; g < ngeo && !rc; ++g) {
        oq_rhs *h = nullptr;
        const oq_hex8_mesh *mg = ma ? ma[g] : nullptr;
        rc = oq_rhs_create_from_mesh(device, &mf[g], mg, lambda, mu, ftype, nrept, buffer_ratio,
                                     mg ? quad : nullptr, &members[g * nmembers],
                                     mg ? visc[g] : nullptr, OQ_K_FP32, &h);
        e->rhs.push_back(h);
    }
    if (rc) {
        grp_destroy(e);
        return rc;
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&e->stream) != hipSuccess ||
        hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess) {
        set_error("stream/event creation failed");
        grp_destroy(e);
        return OQ_ERR_HIP;
    }
    const int64_t M = nmembers;
    int64_t soff = 0, foff = 0, Soff = 0, poff = 0, yoff = 0;
    std::vector<int64_t> poffs, yoffs;
    for (int64_t g = 0; g < ngeo; ++g) {
        const oq_rhs *h = e->rhs[g];
        GrpGeo q{};
        q.K = static_cast<const float *>(h->K.p);
        q.N = h->N; q.ld = h->ld; q.nf = h->nf; q.ne = h->ne; q.n = h->lay.n;
        q.soff = soff; q.foff = foff; q.Soff = Soff; q.Yoff = yoff;
        q.lay = h->lay;
        q.gamma = h->dev.gamma; q.nexp = h->dev.nexp; q.nlaws = h->dev.nlaws;
        for (int k = 0; k < 6; ++k) q.deps0[k] = h->dev.deps0[k];
        e->geo.push_back(q);
        e->soffs.push_back(soff);
        e->nst.push_back(h->lay.n);
        e->geo_bucket.push_back(bucket_of(h->N));
        poffs.push_back(poff);
        yoffs.push_back(yoff);
        yoff += h->N * M;
        soff += h->lay.n * M;
        foff += h->nf * M;
        Soff += (int64_t)kGM * h->ld;
        poff += (h->nf + h->ne) * M;
    }
    e->soffs.push_back(soff);
    poffs.push_back(poff);
    yoffs.push_back(yoff);
    e->ntot = soff;
    e->nprolog = poff;
    e->nrows = yoff;
    // per-member friction, entry-major per geometry
    std::vector<double> va((size_t)foff), vb((size_t)foff), vL((size_t)foff), vs((size_t)foff);
    std::vector<double> eta(ngeo * M), vpl(ngeo * M), f0(ngeo * M), v0(ngeo * M);
    for (int64_t g = 0; g < ngeo; ++g) {
        const int64_t nf = e->geo[g].nf, fo = e->geo[g].foff;
        for (int64_t m = 0; m < M; ++m) {
            const oq_rs_props &p = members[g * M + m];
            for (int64_t i = 0; i < nf; ++i) {
                const size_t ix = (size_t)(fo + i * M + m);
                va[ix] = p.a[i]; vb[ix] = p.b[i]; vL[ix] = p.L[i]; vs[ix] = p.sigma[i];
            }
            eta[g * M + m] = p.eta; vpl[g * M + m] = p.vpl; f0[g * M + m] = p.f0; v0[g * M + m] = p.v0;
        }
    }
    const size_t f8 = (size_t)foff * 8, m8 = (size_t)(ngeo * M) * 8;
    if (!rc) rc = e->dgeo.upload(e->geo.data(), e->geo.size() * sizeof(GrpGeo), e->stream);
    if (!rc) rc = e->dsoffs.upload(e->soffs.data(), e->soffs.size() * 8, e->stream);
    if (!rc) rc = e->dpoffs.upload(poffs.data(), poffs.size() * 8, e->stream);
    if (!rc) rc = e->dyoffs.upload(yoffs.data(), yoffs.size() * 8, e->stream);
    if (!rc) rc = e->Y.alloc((size_t)yoff * 8);
    if (!rc) rc = e->dnst.upload(e->nst.data(), e->nst.size() * 8, e->stream);
    if (!rc) rc = e->a.upload(va.data(), f8, e->stream);
    if (!rc) rc = e->b.upload(vb.data(), f8, e->stream);
    if (!rc) rc = e->L.upload(vL.data(), f8, e->stream);
    if (!rc) rc = e->sigma.upload(vs.data(), f8, e->stream);
    if (!rc) rc = e->eta.upload(eta.data(), m8, e->stream);
    if (!rc) rc = e->vpl.upload(vpl.data(), m8, e->stream);
    if (!rc) rc = e->f0.upload(f0.data(), m8, e->stream);
    if (!rc) rc = e->v0.upload(v0.data(), m8, e->stream);
    if (!rc) rc = e->S.alloc((size_t)Soff * 4);
    if (!rc && hipMemsetAsync(e->S.p, 0, (size_t)Soff * 4, e->stream) != hipSuccess) rc = OQ_ERR_HIP;
    const int64_t GM = ngeo * M;
    if (!rc) rc = e->u.alloc((size_t)soff * 8);
    if (!rc) rc = e->y.alloc((size_t)soff * 8);
    if (!rc) rc = e->k.alloc((size_t)(7 * soff) * 8);
    if (!rc) rc = e->part.alloc((size_t)(kGChunks * GM) * 8);
    {
        const int one = 1;
        if (!rc) rc = e->live.upload(&one, sizeof(int), e->stream);
    }
    DevBuf *cb8[] = {&e->ct, &e->cdt, &e->cdtcur, &e->cqold, &e->cna, &e->cnr, &e->citer};
    for (DevBuf *bp : cb8)
        if (!rc) rc = bp->alloc((size_t)GM * 8);
    DevBuf *cb4[] = {&e->cdone, &e->cacc, &e->cret};
    for (DevBuf *bp : cb4)
        if (!rc) rc = bp->alloc((size_t)GM * 4);
    if (!rc) {
        int gr[kGShapes];
        rc = occupancy_grid(k_grp_gemm<1, true>, &gr[0]);
        if (!rc) rc = occupancy_grid(k_grp_gemm<1, false>, &gr[1]);
        if (!rc) rc = occupancy_grid(k_grp_gemm<2, true>, &gr[2]);
        if (!rc) rc = occupancy_grid(k_grp_gemm<2, false>, &gr[3]);
        if (!rc) rc = occupancy_grid(k_grp_gemm<4, true>, &gr[4]);
        if (!rc) rc = occupancy_grid(k_grp_gemm<4, false>, &gr[5]);
        if (!rc) rc = occupancy_grid(k_grp_gemm_lds<1>, &gr[6]);
        if (!rc) rc = occupancy_grid(k_grp_gemm_lds<2>, &gr[7]);
        if (!rc) rc = occupancy_grid(k_grp_gemm_lds2<1>, &gr[8]);
        if (!rc) rc = occupancy_grid(k_grp_gemm_lds2<2>, &gr[9]);
        if (!rc) rc = occupancy_grid(k_grp_gemm_units<kGUnitRB>, &gr[10]);
        // the direct-to-LDS shapes need 60-144 KB of LDS per workgroup: a
        // device that cannot give it leaves the shape out (grid 0)
        const void *dk[4] = {(const void *)k_grp_gemm_dma<3, 64, 128, 1>, (const void *)k_grp_gemm_dma<3, 32, 128, 2>,
                             (const void *)k_grp_gemm_dma<3, 64, 64, 2>, (const void *)k_grp_gemm_dma<2, 64, 64, 4>};
        for (int i = 0; i < 4 && !rc; ++i) {
            const int sh = kGUnitShape + 1 + i;
            gr[sh] = 0;
            if (hipFuncSetAttribute(dk[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)dma_lds(sh)) !=
                hipSuccess) {
                (void)hipGetLastError();
                continue;
            }
            switch (i) {
            case 0: rc = occupancy_grid(k_grp_gemm_dma<3, 64, 128, 1>, &gr[sh], dma_lds(sh)); break;
            case 1: rc = occupancy_grid(k_grp_gemm_dma<3, 32, 128, 2>, &gr[sh], dma_lds(sh)); break;
            case 2: rc = occupancy_grid(k_grp_gemm_dma<3, 64, 64, 2>, &gr[sh], dma_lds(sh)); break;
            default: rc = occupancy_grid(k_grp_gemm_dma<2, 64, 64, 4>, &gr[sh], dma_lds(sh)); break;
            }
        }
        for (int s = 0; s < kGShapes; ++s) e->grid[s] = gr[s];
    }
    if (!rc) rc = upload_tiles(e);
    if (!rc && hipStreamSynchronize(e->stream) != hipSuccess) rc = OQ_ERR_HIP;
    if (rc) {
        grp_destroy(e);
        return rc;
    }
    GrpDev &d = e->dev;
    d.geo = e->dgeo.as<GrpGeo>(); d.soffs = e->dsoffs.as<int64_t>(); d.poffs = e->dpoffs.as<int64_t>();
    d.G = (int)ngeo; d.M = (int)M;
    d.a = e->a.as<double>(); d.b = e->b.as<double>(); d.L = e->L.as<double>(); d.sigma = e->sigma.as<double>();
    d.eta = e->eta.as<double>(); d.vpl = e->vpl.as<double>(); d.f0 = e->f0.as<double>(); d.v0 = e->v0.as<double>();
    d.S = e->S.as<float>();
    d.Y = e->Y.as<double>();
    d.live = e->live.as<int>();
    d.yoffs = e->dyoffs.as<int64_t>();
    d.nrows = e->nrows;
    d.nprolog = e->nprolog;
    d.ntot = e->ntot;
    GrpCtrl &c = e->ctrl;
    c.t = e->ct.as<double>(); c.dt = e->cdt.as<double>(); c.dtcur = e->cdtcur.as<double>();
    c.qold = e->cqold.as<double>(); c.naccept = e->cna.as<long long>(); c.nreject = e->cnr.as<long long>();
    c.iter = e->citer.as<long long>(); c.done = e->cdone.as<int>(); c.accept = e->cacc.as<int>();
    c.retcode = e->cret.as<int>();
    c.soffs = e->dsoffs.as<int64_t>(); c.nst = e->dnst.as<int64_t>();
    c.G = (int)ngeo; c.M = (int)M;
    c.live = e->live.as<int>();
    *out = e;
    return OQ_OK;
}

int oq_group_info(oq_group *e, int64_t *ntot, int64_t *nstate, int64_t *kbytes)
{
    OQ_REQUIRE(e != nullptr, "NULL argument");
    if (ntot) *ntot = e->ntot;
    int64_t kb = 0;
    for (int64_t g = 0; g < e->G; ++g) {
        if (nstate) nstate[g] = e->nst[g];
        kb += e->geo[g].N * e->geo[g].N * 4;
    }
    if (kbytes) *kbytes = kb;
    return OQ_OK;
}

int oq_group_set_shapes(oq_group *e, const int *shape_of_bucket)
{
    OQ_REQUIRE(e != nullptr && shape_of_bucket != nullptr, "NULL argument");
    for (int b = 0; b < kGBuckets; ++b)
        OQ_REQUIRE(shape_of_bucket[b] >= -1 && shape_of_bucket[b] < kGShapes &&
                       (shape_of_bucket[b] < 0 || e->grid[shape_of_bucket[b]] > 0),
                   "tile shape out of range or not available on this device");
    OQ_HIP(hipSetDevice(e->device));
    for (int b = 0; b < kGBuckets; ++b)
        if (shape_of_bucket[b] >= 0) e->bucket_shape[b] = shape_of_bucket[b];
    int rc = upload_tiles(e);
    if (rc) return rc;
    drop_step(e);  // a captured step holds the old tile tables
    return OQ_OK;
}

int oq_group_autotune(oq_group *e, int iters, int *shape_of_bucket, double *ms_of_bucket,
                      double *ms_table)
{
    OQ_REQUIRE(e != nullptr && iters > 0, "invalid argument");
    OQ_HIP(hipSetDevice(e->device));
    if (int rc0 = set_live(e)) return rc0;
    // S = f(state 0) is irrelevant to the timing; the output goes to scratch
    double *u = e->y.as<double>(), *du = e->k.as<double>();
    std::vector<double> tab((size_t)(kGBuckets * kGShapes), 0.0);
    bool used[kGBuckets] = {false};
    for (int b = 0; b < kGBuckets; ++b) {
        std::vector<int> sel;
        for (int g = 0; g < (int)e->G; ++g)
            if (e->geo_bucket[g] == b) sel.push_back(g);
        if (sel.empty()) continue;
        used[b] = true;
        double best = 1e300;
        int bs = e->bucket_shape[b];
        for (int s = 0; s < kGShapes; ++s) {
            if (e->grid[s] == 0) continue;  // not available on this device (timing left 0)
            const std::vector<int2> t = is_unit_shape(s) ? std::vector<int2>() : build_tiles(e, s, sel);
            DevBuf dt;
            GrpUnitTable ut;
            int rc = is_unit_shape(s) ? build_units(e, s, sel, ut)
                                      : dt.upload(t.data(), t.size() * sizeof(int2), e->stream);
            if (rc) return rc;
            auto run = [&]() {
                return is_unit_shape(s) ? launch_units(e, s, ut)
                                        : launch_gemm(e, s, dt.as<int2>(), (int)t.size(), u, du);
            };
            if ((rc = run())) return rc;
            OQ_HIP(hipEventRecord(e->ev0, e->stream));
            for (int i = 0; i < iters; ++i)
                if ((rc = run())) return rc;
            OQ_HIP(hipEventRecord(e->ev1, e->stream));
            OQ_HIP(hipEventSynchronize(e->ev1));
            float ms = 0.f;
            OQ_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
            tab[(size_t)(b * kGShapes + s)] = ms / iters;
            if (ms / iters < best) {
                best = ms / iters;
                bs = s;
            }
        }
        e->bucket_shape[b] = bs;
    }
    // Per-bucket timings miss what one launch per shape in use costs (each
    // launch drains before the next starts, so every bucket pays its own
    // tail).  The per-bucket choice is therefore compared with each single
    // shape for all buckets (one launch, largest tiles first) on the whole
    // set of GEMM launches, and the faster configuration is kept.
    auto time_all = [&](double *ms_out) -> int {
        int rc = upload_tiles(e);
        for (int rep = 0; rep <= iters && !rc; ++rep) {  // rep 0 warms up
            if (rep == 1) OQ_HIP(hipEventRecord(e->ev0, e->stream));
            for (int s = 0; s < kGShapes && !rc; ++s)
                rc = is_unit_shape(s) ? launch_units(e, s, e->ut[s - kGUnitShape])
                                      : launch_gemm(e, s, e->dtiles[s].as<int2>(), e->ntiles[s], u, du);
        }
        if (rc) return rc;
        OQ_HIP(hipEventRecord(e->ev1, e->stream));
        OQ_HIP(hipEventSynchronize(e->ev1));
        float ms = 0.f;
        OQ_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        *ms_out = ms / iters;
        return OQ_OK;
    };
    int mix[kGBuckets];
    memcpy(mix, e->bucket_shape, sizeof(mix));
    double best = 0.0;
    int rc = time_all(&best), pick = -1;
    for (int s = 0; s < kGShapes && !rc; ++s) {
        if (e->grid[s] == 0) continue;
        for (int b = 0; b < kGBuckets; ++b) e->bucket_shape[b] = s;
        double ms = 0.0;
        rc = time_all(&ms);
        if (!rc && ms < best) {
            best = ms;
            pick = s;
        }
    }
    if (rc) return rc;
    for (int b = 0; b < kGBuckets; ++b) e->bucket_shape[b] = pick >= 0 ? pick : mix[b];
    for (int b = 0; b < kGBuckets; ++b) {
        if (shape_of_bucket) shape_of_bucket[b] = used[b] ? e->bucket_shape[b] : -1;
        if (ms_of_bucket) ms_of_bucket[b] = used[b] ? tab[(size_t)(b * kGShapes + e->bucket_shape[b])] : 0.0;
    }
    if (ms_table) memcpy(ms_table, tab.data(), tab.size() * sizeof(double));
    rc = upload_tiles(e);
    if (rc) return rc;
    drop_step(e);  // a captured step holds the old tile tables
    return OQ_OK;
}

int oq_group_eval(oq_group *e, const double *u, double *du)
{
    OQ_REQUIRE(e && u && du, "NULL argument");
    OQ_HIP(hipSetDevice(e->device));
    std::vector<double> ui;
    to_entry_major(e, u, ui);
    OQ_HIP(hipMemcpyAsync(e->y.p, ui.data(), ui.size() * 8, hipMemcpyHostToDevice, e->stream));
    int rc = set_live(e);
    if (!rc) rc = grp_rhs(e, e->y.as<double>(), e->k.as<double>());
    if (rc) return rc;
    std::vector<double> o((size_t)e->ntot);
    OQ_HIP(hipMemcpyAsync(o.data(), e->k.p, o.size() * 8, hipMemcpyDeviceToHost, e->stream));
    OQ_HIP(hipStreamSynchronize(e->stream));
    from_entry_major(e, o, du);
    return OQ_OK;
}

int oq_group_time(oq_group *e, int iters, double *ms_eval)
{
    OQ_REQUIRE(e && iters > 0, "invalid argument");
    OQ_HIP(hipSetDevice(e->device));
    int rc;
    if ((rc = set_live(e))) return rc;
    OQ_HIP(hipEventRecord(e->ev0, e->stream));
    for (int i = 0; i < iters; ++i)
        if ((rc = grp_rhs(e, e->y.as<double>(), e->k.as<double>()))) return rc;
    OQ_HIP(hipEventRecord(e->ev1, e->stream));
    OQ_HIP(hipEventSynchronize(e->ev1));
    float ms = 0.f;
    OQ_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    if (ms_eval) *ms_eval = (double)ms / iters;
    return OQ_OK;
}

// Uploads the members' states (host layout of oq_group_start) and
// controller arrays, seeds k1 = f(u) and captures the step graph.
static int grp_begin(oq_group *e, const double *u0, const double *tv, const double *dtv, const double *qv,
                     const long long *na, const long long *nr, const long long *it, const int *done,
                     const int *ret, double t1, const oq_solver_opts *o, int alg)
{
    OQ_HIP(hipSetDevice(e->device));
    const int64_t GM = e->G * e->M;
    if (alg == 1) {  // VCABM5: order 1, no step history; phi_0 = f(u) is the seed below
        int rc = OQ_OK;
        if (!e->ab.p) {
            const size_t h6 = (size_t)(GM * (kAbK + 1)) * 8;
            rc = e->abk.alloc((size_t)GM * 4);
            if (!rc) rc = e->abkcur.alloc((size_t)GM * 4);
            if (!rc) rc = e->abhist.alloc(h6);
            if (!rc) rc = e->abg.alloc(h6);
            if (!rc) rc = e->abbeta.alloc(h6);
            if (!rc) rc = e->ab.alloc((size_t)(2 * e->ntot) * 8);
            if (rc) return rc;
        }
        const std::vector<int> one(GM, 1);
        const std::vector<double> zh((size_t)(GM * (kAbK + 1)), 0.0);
        OQ_HIP(hipMemcpyAsync(e->abk.p, one.data(), GM * 4, hipMemcpyHostToDevice, e->stream));
        OQ_HIP(hipMemcpyAsync(e->abkcur.p, one.data(), GM * 4, hipMemcpyHostToDevice, e->stream));
        OQ_HIP(hipMemcpyAsync(e->abhist.p, zh.data(), zh.size() * 8, hipMemcpyHostToDevice, e->stream));
        OQ_HIP(hipMemcpyAsync(e->abg.p, zh.data(), zh.size() * 8, hipMemcpyHostToDevice, e->stream));
        OQ_HIP(hipMemcpyAsync(e->abbeta.p, zh.data(), zh.size() * 8, hipMemcpyHostToDevice, e->stream));
        e->gab = GrpAb{e->abk.as<int>(), e->abkcur.as<int>(), e->abhist.as<double>(), e->abg.as<double>(),
                       e->abbeta.as<double>(), e->ab.as<double>(), e->ab.as<double>() + e->ntot};
        OQ_HIP(hipStreamSynchronize(e->stream));  // the host vectors are read before they go
    }
    e->alg = alg;
    std::vector<double> ui;
    to_entry_major(e, u0, ui);
    const std::vector<int> zi(GM, 0);
    OQ_HIP(hipMemcpyAsync(e->u.p, ui.data(), ui.size() * 8, hipMemcpyHostToDevice, e->stream));
    OQ_HIP(hipMemcpyAsync(e->ct.p, tv, GM * 8, hipMemcpyHostToDevice, e->stream));
    OQ_HIP(hipMemcpyAsync(e->cdt.p, dtv, GM * 8, hipMemcpyHostToDevice, e->stream));
    OQ_HIP(hipMemcpyAsync(e->cqold.p, qv, GM * 8, hipMemcpyHostToDevice, e->stream));
    OQ_HIP(hipMemcpyAsync(e->cna.p, na, GM * 8, hipMemcpyHostToDevice, e->stream));
    OQ_HIP(hipMemcpyAsync(e->cnr.p, nr, GM * 8, hipMemcpyHostToDevice, e->stream));
    OQ_HIP(hipMemcpyAsync(e->citer.p, it, GM * 8, hipMemcpyHostToDevice, e->stream));
    OQ_HIP(hipMemcpyAsync(e->cdone.p, done, GM * 4, hipMemcpyHostToDevice, e->stream));
    OQ_HIP(hipMemcpyAsync(e->cacc.p, zi.data(), GM * 4, hipMemcpyHostToDevice, e->stream));
    OQ_HIP(hipMemcpyAsync(e->cret.p, ret, GM * 4, hipMemcpyHostToDevice, e->stream));
    GrpCtrl &c = e->ctrl;
    c.t1 = t1;
    c.reltol = o->reltol;
    c.abstol = o->abstol;
    c.dtmax = o->dtmax > 0.0 ? o->dtmax : INFINITY;
    c.dtmin = o->dtmin;
    c.maxiters = o->maxiters > 0 ? o->maxiters : (long long)1e18;
    // FSAL seed: k1 = f(u) (on a resume with the same tile shapes, the same
    // bits as the k7 of the last accepted step)
    int rc = set_live(e);
    if (!rc) rc = grp_rhs(e, e->u.as<double>(), e->k.as<double>());
    if (rc) return rc;
    drop_graphs(e);
    OQ_HIP(hipStreamSynchronize(e->stream));
    for (int q = 0; q < 2 && !rc; ++q) {
        OQ_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < (q ? kGraphSteps : 1) && !rc; ++i) rc = grp_step(e);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(e->stream, &g);
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

int oq_group_start(oq_group *e, const double *u0, double t0, double t1, const oq_solver_opts *o)
{
    OQ_REQUIRE(e && u0 && o, "NULL argument");
    OQ_REQUIRE(t1 > t0, "tspan must satisfy t1 > t0");
    OQ_REQUIRE(o->dt0 > 0.0, "the grouped stepper needs an initial step dt0 > 0");
    const int64_t GM = e->G * e->M;
    const std::vector<double> tv(GM, t0), dtv(GM, o->dt0), qv(GM, 1e-4);
    const std::vector<long long> zl(GM, 0);
    const std::vector<int> zi(GM, 0);
    return grp_begin(e, u0, tv.data(), dtv.data(), qv.data(), zl.data(), zl.data(), zl.data(), zi.data(),
                     zi.data(), t1, o, 0);
}

int oq_group_start_vcabm5(oq_group *e, const double *u0, double t0, double t1, const oq_solver_opts *o)
{
    OQ_REQUIRE(e && u0 && o, "NULL argument");
    OQ_REQUIRE(t1 > t0, "tspan must satisfy t1 > t0");
    OQ_REQUIRE(o->dt0 > 0.0, "the grouped stepper needs an initial step dt0 > 0");
    const int64_t GM = e->G * e->M;
    const std::vector<double> tv(GM, t0), dtv(GM, o->dt0), qv(GM, 1e-4);
    const std::vector<long long> zl(GM, 0);
    const std::vector<int> zi(GM, 0);
    return grp_begin(e, u0, tv.data(), dtv.data(), qv.data(), zl.data(), zl.data(), zl.data(), zi.data(),
                     zi.data(), t1, o, 1);
}

int oq_group_get_ctrl(oq_group *e, oq_lockstep_ctrl *c)
{
    OQ_REQUIRE(e && c, "NULL argument");
    OQ_REQUIRE(e->started, "the grouped stepper has not been started");
    OQ_HIP(hipSetDevice(e->device));
    const int64_t GM = e->G * e->M;
    const std::pair<void *, const DevBuf *> d8[] = {{c->t, &e->ct}, {c->dt, &e->cdt}, {c->qold, &e->cqold},
                                                     {c->naccept, &e->cna}, {c->nreject, &e->cnr},
                                                     {c->iter, &e->citer}};
    for (const auto &x : d8)
        if (x.first) OQ_HIP(hipMemcpyAsync(x.first, x.second->p, GM * 8, hipMemcpyDeviceToHost, e->stream));
    if (c->done) OQ_HIP(hipMemcpyAsync(c->done, e->cdone.p, GM * 4, hipMemcpyDeviceToHost, e->stream));
    if (c->retcode) OQ_HIP(hipMemcpyAsync(c->retcode, e->cret.p, GM * 4, hipMemcpyDeviceToHost, e->stream));
    OQ_HIP(hipStreamSynchronize(e->stream));
    return OQ_OK;
}

int oq_group_resume(oq_group *e, const double *u, const oq_lockstep_ctrl *c, double t1, const oq_solver_opts *o)
{
    OQ_REQUIRE(e && u && c && o, "NULL argument");
    OQ_REQUIRE(c->t && c->dt && c->qold && c->naccept && c->nreject && c->iter && c->done && c->retcode,
               "every controller array is needed to resume");
    for (int64_t q = 0; q < e->G * e->M; ++q)
        OQ_REQUIRE(c->done[q] || (c->t[q] < t1 && c->dt[q] > 0.0), "a running member needs t < t1 and dt > 0");
    static_assert(sizeof(long long) == sizeof(int64_t), "int64_t counters");
    return grp_begin(e, u, c->t, c->dt, c->qold, reinterpret_cast<const long long *>(c->naccept),
                     reinterpret_cast<const long long *>(c->nreject), reinterpret_cast<const long long *>(c->iter),
                     c->done, c->retcode, t1, o, 0);
}

int oq_group_run(oq_group *e, int64_t max_steps, int64_t *nrunning)
{
    OQ_REQUIRE(e && max_steps > 0, "invalid argument");
    OQ_REQUIRE(e->stepping, "call oq_group_start first");
    OQ_HIP(hipSetDevice(e->device));
    const int64_t GM = e->G * e->M;
    std::vector<int> done(GM);
    int64_t left = GM;
    // steps queued back to back; past the end they return at once (GrpDev::live)
    constexpr int64_t kRunSync = 256;
    for (int64_t s = 0; s < max_steps; s += kRunSync) {
        const int64_t m = (max_steps - s) < kRunSync ? (max_steps - s) : kRunSync;
        for (int64_t i = 0; i + kGraphSteps <= m; i += kGraphSteps) OQ_HIP(hipGraphLaunch(e->exec[1], e->stream));
        for (int64_t i = m / kGraphSteps * kGraphSteps; i < m; ++i) OQ_HIP(hipGraphLaunch(e->exec[0], e->stream));
        OQ_HIP(hipMemcpyAsync(done.data(), e->cdone.p, GM * 4, hipMemcpyDeviceToHost, e->stream));
        OQ_HIP(hipStreamSynchronize(e->stream));
        left = 0;
        for (int v : done) left += v ? 0 : 1;
        if (left == 0) break;
    }
    if (nrunning) *nrunning = left;
    return OQ_OK;
}

int oq_group_state(oq_group *e, double *t, double *u, int64_t *naccept, int64_t *nreject,
                   int *retcode)
{
    OQ_REQUIRE(e != nullptr, "NULL argument");
    OQ_HIP(hipSetDevice(e->device));
    const int64_t GM = e->G * e->M;
    if (t) OQ_HIP(hipMemcpyAsync(t, e->ct.p, GM * 8, hipMemcpyDeviceToHost, e->stream));
    if (naccept) OQ_HIP(hipMemcpyAsync(naccept, e->cna.p, GM * 8, hipMemcpyDeviceToHost, e->stream));
    if (nreject) OQ_HIP(hipMemcpyAsync(nreject, e->cnr.p, GM * 8, hipMemcpyDeviceToHost, e->stream));
    if (retcode) OQ_HIP(hipMemcpyAsync(retcode, e->cret.p, GM * 4, hipMemcpyDeviceToHost, e->stream));
    std::vector<double> o;
    if (u) {
        o.resize((size_t)e->ntot);
        OQ_HIP(hipMemcpyAsync(o.data(), e->u.p, o.size() * 8, hipMemcpyDeviceToHost, e->stream));
    }
    OQ_HIP(hipStreamSynchronize(e->stream));
    if (u) from_entry_major(e, o, u);
    return OQ_OK;
}

int oq_group_destroy(oq_group *e)
{
    if (e) {
        if (runtime_usable()) (void)hipSetDevice(e->device);
        grp_destroy(e);
    }
    return OQ_OK;
}

}  // extern "C"
