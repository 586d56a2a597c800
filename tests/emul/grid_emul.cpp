// CPU emulation of the grid model's K_mm tiles: k_mm_grid_table, the
// convolution blocks of k_grid_fused (mm_grid_conv_block) and
// k_expand_mm_grid, with the shapes of mm_grid_shape
// (oetqf.jl_amd/csrc/rhs.hip), their index arithmetic restated over
// bounds-checked buffers (vector::at, LDS staging filled with NaN), against a
// dense operator built from a synthetic translation-invariant pair function;
// and the wavefront mapping of the dense K_fm / K_mf streams
// (grid_matvec_wave, mm_grid_chunks): every row and column chunk covered
// exactly once.  The restatement follows the kernels by hand, so a change to
// their indexing has to be made here as well.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>
typedef long long i64;

static double pairf(int k, int p, i64 dx, i64 dy, i64 jz, i64 iz)
{
    return std::sin(1.3 * k + 0.7 * p + 0.11 * dx + 0.37 * dy + 1.9 * jz + 2.3 * iz) + 0.1 * k * p;
}

struct Shape { int r, tp, nchunk, threads, nxp; i64 lds; };
static Shape shape(i64 nx, i64 ny, i64 nz)
{
    Shape g;
    g.r = ny * ((nx + 7) / 8) >= 256 ? 8 : 4;
    g.nxp = (int)((nx + g.r - 1) / g.r * g.r);
    const i64 Wy = 2 * ny - 1, Wp = 2 * g.nxp + 1, nxg = g.nxp / g.r;
    const i64 U = ny * nxg, per = Wy * Wp + 2 * ny * g.nxp;
    if (per * 8 > 160 * 1024) { printf("too big\n"); exit(1); }
    i64 tp = 256 / U;
    if (tp < 1) tp = 1;
    if (tp > 6 * nz) tp = 6 * nz;
    while (tp > 1 && tp * per * 8 > 64 * 1024) --tp;
    g.tp = (int)tp;
    g.nchunk = (int)((6 * nz + tp - 1) / tp);
    const i64 t = (tp * U + 63) / 64 * 64;
    g.threads = 256; (void)t;
    g.lds = tp * per;  // doubles
    return g;
}

template <int R>
static void conv_block(const std::vector<double> &T, const std::vector<double> &s, int nx, int ny, int nz, int tp,
                       std::vector<double> &part, int PO, int c, int threads, i64 lds)
{
    std::vector<double> gsh((size_t)lds, NAN);
    const int npl = 6 * nz, nxy = nx * ny;
    const int pi0 = c * tp, np = (pi0 + tp < npl ? pi0 + tp : npl) - pi0;
    const int nxp = (nx + R - 1) / R * R, Wp = 2 * nxp + 1, Wy = 2 * ny - 1, tile = Wy * Wp;
    const i64 oTs = 0, oSs = (i64)tp * tile, oOs = oSs + (i64)tp * ny * nxp;
    for (int tid = 0; tid < threads; ++tid) {
        for (int x = tid; x < np * tile; x += threads) gsh.at(oTs + x) = T.at((i64)(PO * npl + pi0) * tile + x);
        for (int x = tid; x < np * ny * nxp; x += threads) {
            const int row = x / nxp, ix = x - row * nxp;
            gsh.at(oSs + x) = ix < nx ? s.at((size_t)pi0 * nxy + row * nx + ix) : 0.0;
        }
    }
    const int nxg = nxp / R, U = ny * nxg;
    for (int tid = 0; tid < threads; ++tid)
        for (int t = tid; t < np * U; t += threads) {
            const int q = t / U, g = t - q * U, jy = g / nxg, jx0 = (g - jy * nxg) * R;
            double acc[R];
            for (int r = 0; r < R; ++r) acc[r] = 0.0;
            for (int iy = 0; iy < ny; ++iy) {
                const i64 rowb = oTs + (i64)(q * Wy + (jy - iy + ny - 1)) * Wp;
                const i64 Tr = rowb + nxp;
                const i64 Sr = oSs + (i64)(q * ny + iy) * nxp;
                double w[R];
                for (int r = 0; r < R; ++r) w[r] = gsh.at(Tr + jx0 + r);
                for (int ix = 0; ix < nxp; ix += R)
                    for (int kk = 0; kk < R; ++kk) {
                        const double sv = gsh.at(Sr + ix + kk);
                        if (std::isnan(sv)) { printf("nan s\n"); exit(1); }
                        for (int r = 0; r < R; ++r) acc[r] += w[(r - kk + R) % R] * sv;
                        const i64 a = Tr + jx0 - (ix + kk) - 1;
                        if (a < rowb || a >= rowb + Wp) { printf("row overrun\n"); exit(1); }
                        w[(R - 1 - kk) % R] = gsh.at(a);
                    }
            }
            for (int r = 0; r < R; ++r) {
                if (std::isnan(acc[r])) { printf("nan acc\n"); exit(1); }
                gsh.at(oOs + (i64)(q * ny + jy) * nxp + jx0 + r) = acc[r];
            }
        }
    for (int tid = 0; tid < threads; ++tid)
        for (int o = tid; o < nxy; o += threads) {
            const int jy = o / nx, jx = o - jy * nx;
            double a = 0.0;
            for (int q = 0; q < np; ++q) a += gsh.at(oOs + (i64)(q * ny + jy) * nxp + jx);
            part.at(((i64)c * npl + PO) * nxy + o) = a;
        }
}

// grid_matvec_wave mapping (8 rows per wavefront, 4 wavefronts per block):
// every (fault row, column) of K_fm and every mantle row covered exactly
// once; chunk sizes as mm_grid_chunks.
static int matvec_cover(i64 nf, i64 ne, i64 ldm, i64 ldf)
{
    const int R = 8, W = 4;
    i64 S = (ldm + ldf / 2) / ldf;
    if (S < 1) S = 1;
    const i64 cw = ((ldm + S - 1) / S + 127) / 128 * 128;
    S = (ldm + cw - 1) / cw;
    const i64 m6 = 6 * ne, nfw = (nf + R - 1) / R * S;
    const i64 waves = nfw + (m6 + R - 1) / R, blocks = (waves + W - 1) / W;
    std::vector<int> fcov((size_t)(nf * (ldm / 128)), 0), mcov((size_t)m6, 0);
    for (i64 gw = 0; gw < blocks * W; ++gw) {
        if (gw < nfw) {
            const i64 c = gw % S, row0 = (gw / S) * R, c0 = c * cw, ncol = (c0 + cw < ldm ? c0 + cw : ldm) - c0;
            if (ncol <= 0 || ncol % 128) { printf("bad chunk\n"); return 1; }
            for (int r = 0; r < R; ++r)
                if (row0 + r < nf)
                    for (i64 p = 0; p < ncol / 128; ++p) fcov.at((row0 + r) * (ldm / 128) + c0 / 128 + p)++;
        } else {
            const i64 row0 = (gw - nfw) * R;
            if (row0 >= m6) continue;
            for (int r = 0; r < R; ++r)
                if (row0 + r < m6) mcov.at(row0 + r)++;
        }
    }
    for (int v : fcov) if (v != 1) { printf("fault coverage %d\n", v); return 1; }
    for (int v : mcov) if (v != 1) { printf("mantle coverage %d\n", v); return 1; }
    printf("matvec cover nf %lld ne %lld ldm %lld ldf %lld: S %lld cw %lld waves %lld ok\n", nf, ne, ldm, ldf, S, cw,
           waves);
    return 0;
}

static int run(i64 nx, i64 ny, i64 nz)
{
    const i64 ne = nx * ny * nz, nxy = nx * ny, npl = 6 * nz, Wx = 2 * nx - 1, Wy = 2 * ny - 1;
    const i64 nrec = Wx * Wy * nz, m6 = 6 * ne;
    std::vector<double> raw((size_t)(6 * nrec * npl));
    for (int k = 0; k < 6; ++k)
        for (int p = 0; p < 6; ++p)
            for (i64 jz = 0; jz < nz; ++jz)
                for (i64 yy = 0; yy < Wy; ++yy)
                    for (i64 xx = 0; xx < Wx; ++xx)
                        for (i64 iz = 0; iz < nz; ++iz)
                            raw.at((k * nrec + xx + Wx * (yy + Wy * jz)) * npl + p * nz + iz) =
                                pairf(k, p, xx - (nx - 1), yy - (ny - 1), jz, iz);
    const Shape g = shape(nx, ny, nz);
    const i64 nxp = g.nxp, Wp = 2 * nxp + 1;
    const i64 nt = npl * npl * Wy * Wp;
    std::vector<double> T((size_t)nt);
    for (i64 t = 0; t < nt; ++t) {  // k_mm_grid_table
        const i64 xp = t % Wp, r1 = t / Wp, yy = r1 % Wy, r2 = r1 / Wy;
        const i64 PI = r2 % npl, PO = r2 / npl, k = PO / nz, jz = PO % nz;
        const i64 m = xp - nxp;
        T.at(t) = (m > -nx && m < nx) ? raw.at((k * nrec + (m + nx - 1) + Wx * (yy + Wy * jz)) * npl + PI) : 0.0;
    }
    // dense reference and k_expand_mm_grid, row by row
    std::vector<double> s((size_t)m6);
    srand(7);
    for (auto &v : s) v = (double)rand() / RAND_MAX - 0.5;
    std::vector<double> ref((size_t)m6, 0.0);
    double maxexp = 0.0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : maxexp)
    for (i64 row = 0; row < m6; ++row) {
        const i64 k = row / ne, cr = row % ne, jx = cr % nx, jy = (cr / nx) % ny, jz = cr / nxy;
        for (i64 col = 0; col < m6; ++col) {
            const i64 p = col / ne, cc = col % ne, ix = cc % nx, iy = (cc / nx) % ny, iz = cc / nxy;
            const double kd = pairf((int)k, (int)p, jx - ix, jy - iy, jz, iz);
            ref[row] += kd * s[col];
            const i64 PO = row / nxy, o = row % nxy, PI = col / nxy, o2 = col % nxy;
            const i64 ey = o / nx, ex = o % nx, fy = o2 / nx, fx = o2 % nx;
            const double ke = T.at(((PO * npl + PI) * Wy + (ey - fy + ny - 1)) * Wp + (ex - fx + nxp));
            maxexp = fmax(maxexp, fabs(ke - kd));
        }
    }
    std::vector<double> part((size_t)(g.nchunk * m6), NAN);
    for (i64 PO = 0; PO < npl; ++PO)
        for (i64 c = 0; c < g.nchunk; ++c) {
            if (g.r == 8) conv_block<8>(T, s, (int)nx, (int)ny, (int)nz, g.tp, part, (int)PO, (int)c, g.threads, g.lds);
            else conv_block<4>(T, s, (int)nx, (int)ny, (int)nz, g.tp, part, (int)PO, (int)c, g.threads, g.lds);
        }
    double err = 0.0, mx = 0.0;
    for (i64 r = 0; r < m6; ++r) {
        double a = 0.0;
        for (int c = 0; c < g.nchunk; ++c) a += part.at(c * m6 + r);
        if (std::isnan(a)) { printf("nan out\n"); return 1; }
        err = fmax(err, fabs(a - ref[r]));
        mx = fmax(mx, fabs(ref[r]));
    }
    printf("nx %lld ny %lld nz %lld  R %d tp %d nchunk %d threads %d lds %lld B  expand maxdiff %g  conv relerr %g\n",
           nx, ny, nz, g.r, g.tp, g.nchunk, g.threads, g.lds * 8, maxexp, err / mx);
    return (maxexp == 0.0 && err / mx < 1e-13) ? 0 : 1;
}

int main()
{
    const i64 dims[][3] = {{5, 3, 2}, {9, 4, 1}, {3, 7, 2}, {1, 1, 2}, {25, 2, 1}, {4, 4, 3},
                           {16, 8, 2}, {9, 100, 1}, {16, 16, 8}, {33, 9, 2}, {32, 32, 8}, {41, 60, 1}};
    int bad = 0;
    for (auto &d : dims) bad |= run(d[0], d[1], d[2]);
    // the configurations' shapes (nf, ne, ldm = round_up(6ne, 512), ldf = round_up(nf, 512)) and odd ones
    const i64 mv[][4] = {{4096, 2048, 12288, 4096}, {16384, 8192, 49152, 16384}, {384, 45, 512, 512},
                         {1000, 3000, 18432, 1024}, {5000, 10, 512, 5120}, {96, 24, 512, 512}};
    for (auto &m : mv) bad |= matvec_cover(m[0], m[1], m[2], m[3]);
    printf(bad ? "FAIL\n" : "ok\n");
    return bad;
}
