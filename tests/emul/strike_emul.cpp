// CPU emulation of the grid model's along-strike tables
// (oetqf.jl_amd/csrc/rhs.hip: the virtual geometry of mm_strike_build, the
// raw layouts of k_gf_fm / k_gf_mf, k_grid_ftable / k_grid_gtable), the F
// and G block loops of k_grid_fused, k_expand_strike and the role table of
// mm_roles_build, their index arithmetic restated over bounds-checked
// buffers, against dense products of synthetic translation-invariant pair
// functions.  The shapes follow mm_strike_shape.  The restatement follows the
// kernels by hand, so a change to their indexing has to be made here as well.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <utility>
#include <vector>
typedef long long i64;

static double fm_pair(double dx, i64 b, i64 jy, i64 jz, int p) { return std::sin(1e-3 * dx + 0.3 * b + 0.7 * jy + 1.1 * jz + 0.5 * p); }
static double mf_pair(double dx, i64 jy, i64 jz, int k, i64 b) { return std::cos(1.3e-3 * dx + 0.2 * b + 0.5 * jy + 0.9 * jz + 0.4 * k); }

struct Chk {
    std::vector<double> a;
    explicit Chk(size_t n) : a(n, NAN) {}
    double &at(i64 i) { if (i < 0 || (size_t)i >= a.size()) { printf("OOB %lld of %zu\n", i, a.size()); exit(1); } return a[(size_t)i]; }
};

static double wave_sum64(std::vector<double> &v) { double s = 0; for (double x : v) s += x; return s; }

static int run(i64 nx, i64 nxi, i64 nxc, i64 nyc, i64 nzc, i64 m, i64 lds_doubles)
{
    const double dxf = 625.0, x0 = -40000.0 + 312.5, X0 = -40000.0 + 0.5 * m * dxf;
    const i64 D = nx + m * (nxc - 1), nyz = nyc * nzc, Q = 6 * nyz, nf = nx * nxi, ne = nxc * nyz, m6 = 6 * ne;
    const i64 nxy = nxc * nyc;
    // shapes (mm_strike_shape)
    const int kRW = 8;
    const i64 cap = lds_doubles;
    const i64 nr = (nx + m - 1) / m, fseg = (nr + kRW - 1) / kRW, gseg = (nxc + kRW - 1) / kRW;
    const auto pitch8 = [](i64 w) { return w + ((8 - w % 32) + 32) % 32; };
    const i64 Fp = m * (fseg * kRW + nxc - 1), Gp = pitch8(nx + m * gseg * kRW), swp = pitch8(m * nr);
    if (m * fseg > 256 || gseg > 4 || m > 64 * (4 / gseg)) { printf("no fit (shape): dense blocks\n"); return 0; }
    i64 lpp = 256 / (m * fseg);
    if (lpp * nx > cap) lpp = cap / nx;
    const i64 fqb = cap / (Fp + nxc), gbb = cap / (Gp + swp);
    if (fqb < 1 || gbb < 1 || lpp < 1 || 4 * kRW > cap) { printf("no fit\n"); return 1; }
    const i64 nfc0 = Q < 16 ? Q : 16, fqs = (Q + nfc0 - 1) / nfc0, nfc = (Q + fqs - 1) / fqs;
    const i64 gbs = nxi, ngc = 1;
    // raw F: k_gf_fm over virtual patches (receiver i = d + D b, rs = Q, cs = 1) and cells (0, jy, jz) (j = jy + ny jz)
    std::vector<double> rawF((size_t)(nxi * Q * D));
    for (i64 b = 0; b < nxi; ++b)
        for (i64 d = 0; d < D; ++d)
            for (int p = 0; p < 6; ++p)
                for (i64 jz = 0; jz < nzc; ++jz)
                    for (i64 jy = 0; jy < nyc; ++jy) {
                        const double xr = x0 + (double)(d - m * (nxc - 1)) * dxf, xs = X0;
                        const i64 i = d + D * b, j = jy + nyc * jz;
                        rawF.at((size_t)(i * Q + (p * nyz + j))) = fm_pair(xr - xs, b, jy, jz, p);
                    }
    // raw G: k_gf_mf over virtual cells (receiver i = e + D (jy + ny jz), ne' = D nyz) and patches (0, b) (j = b)
    const i64 nev = D * nyz;
    std::vector<double> rawG((size_t)(nxi * Q * D));
    for (i64 jz = 0; jz < nzc; ++jz)
        for (i64 jy = 0; jy < nyc; ++jy)
            for (i64 e = 0; e < D; ++e)
                for (int k = 0; k < 6; ++k)
                    for (i64 b = 0; b < nxi; ++b) {
                        const double xr = X0 + (double)(e - (nx - 1)) * dxf, xs = x0;
                        const i64 i = e + D * (jy + nyc * jz);
                        rawG.at((size_t)((k * nev + i) * nxi + b)) = mf_pair(xr - xs, jy, jz, k, b);
                    }
    // tables
    std::vector<double> F((size_t)(nxi * Q * D)), G((size_t)(nxi * Q * D));
    for (i64 t = 0; t < nxi * Q * D; ++t) {
        { const i64 dd = t % D, r = t / D, q = r % Q, b = r / Q; F.at(t) = rawF.at((dd + D * b) * Q + q); }
        { const i64 e = t % D, r = t / D, b = r % nxi, qq = r / nxi;
          const i64 k = qq / nyz, rem = qq % nyz, jz = rem / nyc, jy = rem % nyc;
          G.at(t) = rawG.at((k * nev + e + D * (jy + nyc * jz)) * nxi + b); }
    }
    // dense reference products and the expansion
    std::vector<double> sf((size_t)nf), sm((size_t)m6);
    srand(3);
    for (auto &v : sf) v = (double)rand() / RAND_MAX - 0.5;
    for (auto &v : sm) v = (double)rand() / RAND_MAX - 0.5;
    std::vector<double> tau((size_t)nf, 0.0), sig((size_t)m6, 0.0);
    double maxexp = 0;
    for (i64 f = 0; f < nf; ++f) {
        const i64 a = f % nx, b = f / nx;
        for (i64 col = 0; col < m6; ++col) {
            const int p = (int)(col / ne);
            const i64 c = col % ne, jx = c % nxc, jy = (c / nxc) % nyc, jz = c / nxy;
            const double kd = fm_pair((x0 + a * dxf) - (X0 + jx * m * dxf), b, jy, jz, p);
            tau[f] += kd * sm[col];
            const i64 q = col / nxc;  // = (p nz + jz) ny + jy
            if (q != (p * nzc + jz) * nyc + jy) { printf("q order\n"); return 1; }
            const double ke = F.at((b * Q + q) * D + a - m * jx + m * (nxc - 1));
            maxexp = std::max(maxexp, std::fabs(ke - kd));
        }
    }
    for (i64 row = 0; row < m6; ++row) {
        const int k = (int)(row / ne);
        const i64 c = row % ne, jx = c % nxc, jy = (c / nxc) % nyc, jz = c / nxy, qq = row / nxc;
        for (i64 f = 0; f < nf; ++f) {
            const i64 a = f % nx, b = f / nx;
            const double kd = mf_pair((X0 + jx * m * dxf) - (x0 + a * dxf), jy, jz, k, b);
            sig[row] += kd * sf[f];
            const double ke = G.at((qq * nxi + b) * D + m * jx - a + nx - 1);
            maxexp = std::max(maxexp, std::fabs(ke - kd));
        }
    }
    // F blocks (grid_f_block) with 256 threads
    Chk fpart((size_t)((1 + nfc) * nf));
    for (i64 f = 0; f < nf; ++f) fpart.at(f) = 0.0;
    for (i64 b = 0; b < nxi; ++b)
        for (i64 c = 0; c < nfc; ++c) {
            Chk sh((size_t)cap);
            const i64 q0 = c * fqs, q1 = std::min(q0 + fqs, Q);
            const i64 tpl = m * fseg;
            std::vector<std::vector<double>> acc(256, std::vector<double>(kRW, 0.0));
            for (i64 qb = q0; qb < q1; qb += fqb) {
                const i64 nq = std::min(fqb, q1 - qb);
                for (int t = 0; t < 256; ++t) {
                    for (i64 x = t; x < nq * Fp; x += 256) {
                        const i64 q = x / Fp, dd = x - q * Fp;
                        sh.at(x) = dd < D ? F.at((b * Q + qb + q) * D + dd) : 0.0;
                    }
                    for (i64 x = t; x < nq * nxc; x += 256) sh.at(fqb * Fp + x) = sm.at(qb * nxc + x);
                }
                for (int t = 0; t < 256; ++t) {
                    const i64 li = t / tpl, rr = t - li * tpl, sg = rr / m, a0 = rr - sg * m;
                    if (li >= lpp) continue;
                    for (i64 q = li; q < nq; q += lpp) {
                        const i64 rowb = q * Fp;
                        const i64 Fq = rowb + a0 + m * (sg * kRW) + m * (nxc - 1);
                        const i64 Sq = fqb * Fp + q * nxc;
                        auto LF = [&](i64 i) { if (i < rowb || i >= rowb + Fp) { printf("F row overrun\n"); exit(1); } return sh.at(i); };
                        std::vector<double> w(kRW);
                        for (int r = 0; r < kRW; ++r) w[r] = LF(Fq + m * r);
                        for (i64 j0 = 0; j0 < nxc; j0 += kRW)
                            for (int kk = 0; kk < kRW; ++kk) {
                                const i64 jx = j0 + kk;
                                if (jx < nxc) {
                                    const double sv = sh.at(Sq + jx);
                                    for (int r = 0; r < kRW; ++r) acc[t][r] += w[(r - kk + kRW) % kRW] * sv;
                                    if (jx + 1 < nxc) w[(kRW - 1 - kk) % kRW] = LF(Fq - m * (jx + 1));
                                }
                            }
                    }
                }
            }
            Chk R((size_t)cap);
            for (int t = 0; t < 256; ++t) {
                const i64 li = t / tpl, rr = t - li * tpl, sg = rr / m, a0 = rr - sg * m;
                if (li >= lpp) continue;
                for (int r = 0; r < kRW; ++r) {
                    const i64 a = a0 + m * (sg * kRW + r);
                    if (a < nx) R.at(li * nx + a) = acc[t][r];
                }
            }
            for (i64 a = 0; a < nx; ++a) {
                double y = 0;
                for (i64 l = 0; l < lpp; ++l) y += R.at(l * nx + a);
                fpart.at((1 + c) * nf + a + nx * b) = y;
            }
        }
    // G blocks (grid_g_block)
    Chk gpart((size_t)(ngc * m6));
    const i64 sw = swp;
    for (i64 qq = 0; qq < Q; ++qq)
        for (i64 c = 0; c < ngc; ++c) {
            Chk sh((size_t)cap);
            const i64 b0 = c * gbs, b1 = std::min(b0 + gbs, nxi);
            const i64 wps = 4 / gseg, tps = 64 * wps, bp = tps / m;
            std::vector<std::vector<double>> acc(256, std::vector<double>(kRW, 0.0));
            for (i64 bb = b0; bb < b1; bb += gbb) {
                const i64 nb = std::min(gbb, b1 - bb);
                for (int t = 0; t < 256; ++t) {
                    for (i64 x = t; x < nb * Gp; x += 256) {
                        const i64 bq = x / Gp, e = x - bq * Gp - m;
                        sh.at(x) = (e >= 0 && e < D) ? G.at((qq * nxi + bb + bq) * D + e) : 0.0;
                    }
                    for (i64 x = t; x < nb * sw; x += 256) {
                        const i64 bq = x / sw, a = x - bq * sw;
                        sh.at(gbb * Gp + x) = a < nx ? sf.at((bb + bq) * nx + a) : 0.0;
                    }
                }
                for (int t = 0; t < 256; ++t) {
                    const i64 sg = t / tps, x0 = t - sg * tps, bi = x0 / m, a0 = x0 - bi * m;
                    const bool active = sg < gseg && bi < bp;
                    for (i64 b = bi; active && b < nb; b += bp) {
                        const i64 rowb = b * Gp;
                        const i64 Gq = rowb + m + (nx - 1 - a0) + m * (sg * kRW);
                        const i64 Sq = gbb * Gp + b * sw + a0;
                        auto LG = [&](i64 i) { if (i < rowb || i >= rowb + Gp) { printf("G row overrun\n"); exit(1); } return sh.at(i); };
                        std::vector<double> wv(kRW);
                        for (int r = 0; r < kRW; ++r) wv[r] = LG(Gq + m * r);
                        for (i64 k0 = 0; k0 < nr; k0 += kRW)
                            for (int kk = 0; kk < kRW; ++kk) {
                                const i64 k = k0 + kk;
                                if (k < nr) {
                                    const i64 si = Sq + m * k;
                                    if (si >= gbb * Gp + b * sw + m * nr) { printf("S row overrun\n"); exit(1); }
                                    const double sv = sh.at(si);
                                    for (int r = 0; r < kRW; ++r) acc[t][r] += wv[(r - kk + kRW) % kRW] * sv;
                                    if (k + 1 < nr) wv[(kRW - 1 - kk) % kRW] = LG(Gq - m * (k + 1));
                                }
                            }
                    }
                }
            }
            std::vector<double> Rs(4 * kRW, 0.0);
            for (int w = 0; w < 4; ++w)
                for (int r = 0; r < kRW; ++r) {
                    std::vector<double> v;
                    for (int l = 0; l < 64; ++l) v.push_back(acc[w * 64 + l][r]);
                    Rs[w * kRW + r] = wave_sum64(v);
                }
            for (i64 o = 0; o < gseg * kRW; ++o) {
                const i64 s2 = o / kRW, r = o - s2 * kRW, jx = s2 * kRW + r;
                if (jx < nxc) {
                    double y = 0;
                    for (i64 v = 0; v < wps; ++v) y += Rs[(s2 * wps + v) * kRW + r];
                    gpart.at(c * m6 + qq * nxc + jx) = y;
                }
            }
        }
    // finish sums against the dense products
    double ef = 0, mf = 0, eg = 0, mg = 0;
    for (i64 f = 0; f < nf; ++f) {
        double y = 0;
        for (i64 c = 0; c < 1 + nfc; ++c) y += fpart.at(c * nf + f);
        ef = std::max(ef, std::fabs(y - tau[f]));
        mf = std::max(mf, std::fabs(tau[f]));
    }
    for (i64 r = 0; r < m6; ++r) {
        double y = 0;
        for (i64 c = 0; c < ngc; ++c) y += gpart.at(c * m6 + r);
        eg = std::max(eg, std::fabs(y - sig[r]));
        mg = std::max(mg, std::fabs(sig[r]));
    }
    printf("nx %lld nxi %lld cells %lldx%lldx%lld m %lld D %lld: F %lld/%lld x %lld (seg %lld), G %lld/%lld x %lld (seg %lld); "
           "table maxdiff %.1e, tau relerr %.1e, sigma relerr %.1e\n", nx, nxi, nxc, nyc, nzc, m, D, fqs, fqb, nfc, fseg,
           gbs, gbb, ngc, gseg, maxexp, ef / mf, eg / mg);
    return (maxexp < 1e-12 && ef / mf < 1e-13 && eg / mg < 1e-13) ? 0 : 1;
}

// role table: every block of every role exactly once
static int roles(i64 c0, i64 c1, i64 c2, i64 c3)
{
    const i64 cnt[4] = {c0, c1, c2, c3};
    std::vector<std::pair<double, unsigned>> e;
    for (int r = 0; r < 4; ++r)
        for (i64 b = 0; b < cnt[r]; ++b) e.push_back({((double)b + 0.5) / (double)cnt[r], ((unsigned)r << 28) | (unsigned)b});
    std::stable_sort(e.begin(), e.end(), [](const std::pair<double, unsigned> &x, const std::pair<double, unsigned> &y) { return x.first < y.first; });
    std::vector<std::vector<int>> seen(4);
    for (int r = 0; r < 4; ++r) seen[r].assign((size_t)cnt[r], 0);
    for (auto &x : e) seen[x.second >> 28].at(x.second & 0x0fffffffu)++;
    for (auto &v : seen) for (int s : v) if (s != 1) { printf("role coverage\n"); return 1; }
    return 0;
}

int main()
{
    int bad = 0;
    // composite (cells 16x16x8 would take long densely; a thin slab keeps the dense check cheap)
    bad |= run(128, 4, 16, 2, 2, 8, 5120);
    bad |= run(32, 12, 4, 4, 2, 8, 6143);     // the small GPU test geometry (m = 8)
    bad |= run(300, 3, 5, 3, 1, 7, 6143);     // nx > 256: two outputs per thread
    bad |= run(40, 5, 1, 3, 2, 1, 3000);      // one cell along x
    bad |= run(64, 9, 9, 1, 1, 3, 1200);      // small LDS: several batches
    bad |= run(256, 3, 32, 2, 1, 8, 5120);    // the 65536 geometry along x: two F segments, two G segments
    bad |= run(90, 3, 40, 2, 1, 3, 6143);     // 40 cells along x: five G segments do not fit, blocks stay dense
    bad |= run(17, 4, 2, 2, 2, 5, 6143);      // nx not a multiple of m
    bad |= roles(576, 0, 256, 1536) | roles(576, 768, 0, 0) | roles(1, 0, 7, 3);
    printf(bad ? "FAIL\n" : "ok\n");
    return bad;
}
