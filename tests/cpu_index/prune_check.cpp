// Host check of the pruned 2D / 3D FFT's data movement (fft_prune.hip): the
// library's transpose parameters (prune_* in nufft_launch.h) and offset maps,
// driven element by element as the GPU transposes apply them, with naive 1D
// DFTs in place of hipFFT, against direct 2D / 3D sums.  Every read and write is
// bounds-checked against its buffer; every output element of a transpose must
// be written exactly once (scatters: every cell, zeros included; gathers:
// every mode).  Built with AddressSanitizer and UndefinedBehaviorSanitizer.
//
// usage: prune_check ms0 ms1 ms2 nf0 nf1 nf2 ntr iflag   (ms2 = nf2 = 1: the 2D
// passes)
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../cufinufft_amd/csrc/nufft_launch.h"

using cd = std::complex<double>;
using cfa::PruneGather;
using cfa::PruneScatter;

static int g_fail = 0;
#define CHECK(c, ...)                                   \
    do {                                                \
        if (!(c)) {                                     \
            if (g_fail < 20) std::printf(__VA_ARGS__);  \
            ++g_fail;                                   \
        }                                               \
    } while (0)

static const double kTwoPi = 6.283185307179586476925286766559;

// In-place unnormalised DFTs of `lines` contiguous lines of length n, sign iflag.
static void dft_lines(std::vector<cd> &a, long long lines, int n, int iflag)
{
    CHECK((long long)a.size() >= lines * n, "DFT over %lld lines of %d exceeds %zu\n", lines, n,
          a.size());
    std::vector<cd> tmp(n);
    for (long long l = 0; l < lines; ++l) {
        cd *x = a.data() + l * n;
        for (int k = 0; k < n; ++k) {
            cd s = 0;
            for (int j = 0; j < n; ++j) s += x[j] * std::polar(1.0, iflag * kTwoPi * j * k / n);
            tmp[k] = s;
        }
        for (int k = 0; k < n; ++k) x[k] = tmp[k];
    }
}

static void gather(const std::vector<cd> &in, std::vector<cd> &out, const PruneGather &d, int ntr,
                   const double *kr, const double *kp, const double *kc, const char *name)
{
    std::vector<int> hits(out.size(), 0);
    for (int t = 0; t < ntr; ++t)
        for (int p = 0; p < d.P; ++p)
            for (int r = 0; r < d.R; ++r)
                for (int q = 0; q < d.Nc; ++q) {
                    const long long s = cfa::prune_gather_src(d, t, p, r, q);
                    const long long o = cfa::prune_gather_dst(d, t, p, r, q);
                    if (s < 0 || s >= (long long)in.size() || o < 0 ||
                        o >= (long long)out.size()) {
                        CHECK(false, "%s: t %d p %d r %d q %d reads %lld of %zu, writes %lld of %zu\n",
                              name, t, p, r, q, s, in.size(), o, out.size());
                        continue;
                    }
                    cd v = in[s];
                    if (kr)
                        v /= kr[std::abs(r - d.hr)] *
                             ((kp ? kp[std::abs(p - d.hp)] : 1.0) * kc[std::abs(q - d.kneg)]);
                    out[o] = v;
                    ++hits[o];
                }
    const long long want = (long long)ntr * d.P * d.R * d.Nc;
    long long once = 0;
    for (int h : hits) {
        CHECK(h <= 1, "%s: an element written %d times\n", name, h);
        once += h == 1;
    }
    CHECK(once == want, "%s: %lld distinct elements written, want %lld\n", name, once, want);
}

static void scatter(const std::vector<cd> &in, std::vector<cd> &out, const PruneScatter &d,
                    int ntr, const double *kr, const double *kp, const double *kc, int wsel,
                    int wsign, const char *name, bool whole)
{
    std::vector<int> hits(out.size(), 0);
    for (int t = 0; t < ntr; ++t)
        for (int p = 0; p < d.P; ++p)
            for (int r = 0; r < d.R; ++r)
                for (int c = 0; c < d.C; ++c) {
                    const long long o = cfa::prune_scatter_dst(d, t, p, r, c);
                    if (o < 0 || o >= (long long)out.size()) {
                        CHECK(false, "%s: t %d p %d r %d c %d writes %lld of %zu\n", name, t, p,
                              r, c, o, out.size());
                        continue;
                    }
                    cd v = 0;
                    const int q = cfa::prune_mode(c, d.Nc, d.kneg, d.C);
                    if (q >= 0) {
                        CHECK(cfa::prune_cell(q, d.kneg, d.C) == c, "%s: cell %d -> mode %d -> cell %d\n",
                              name, c, q, cfa::prune_cell(q, d.kneg, d.C));
                        const long long s = cfa::prune_scatter_src(d, t, p, r, q);
                        if (s < 0 || s >= (long long)in.size()) {
                            CHECK(false, "%s: reads %lld of %zu\n", name, s, in.size());
                            continue;
                        }
                        v = in[s];
                        if (kr) {
                            v /= kr[std::abs(r - d.hr)] *
                                 ((kp ? kp[std::abs(p - d.hp)] : 1.0) * kc[std::abs(q - d.kneg)]);
                            if (wsel >= 0) {
                                const int f = wsel == 0 ? r - d.hr : (wsel == 1 ? p - d.hp : q - d.kneg);
                                v *= cd(0, (double)(wsign * f));
                            }
                        }
                    }
                    out[o] = v;
                    ++hits[o];
                }
    const long long want = (long long)ntr * d.P * d.R * d.C;
    long long once = 0;
    for (int h : hits) {
        CHECK(h <= 1, "%s: an element written %d times\n", name, h);
        once += h == 1;
    }
    CHECK(once == want, "%s: %lld distinct elements written, want %lld\n", name, once, want);
    if (whole) CHECK(want == (long long)out.size(), "%s: %lld of %zu cells written\n", name, want,
                     out.size());
}

static double rel_err(const std::vector<cd> &a, const std::vector<cd> &b)
{
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += std::norm(a[i] - b[i]);
        den += std::norm(b[i]);
    }
    return std::sqrt(num / (den > 0 ? den : 1));
}

int main(int argc, char **argv)
{
    if (argc < 9) {
        std::printf("usage: %s ms0 ms1 ms2 nf0 nf1 nf2 ntr iflag\n", argv[0]);
        return 2;
    }
    const int ms[3] = {std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3])};
    const int nf[3] = {std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6])};
    const int ntr = std::atoi(argv[7]);
    const int iflag = std::atoi(argv[8]) >= 0 ? 1 : -1;
    for (int a = 0; a < 3; ++a)
        if (ms[a] < 1 || nf[a] < ms[a]) {
            std::printf("need 1 <= ms <= nf on every axis\n");
            return 2;
        }
    const long long nftot = (long long)nf[0] * nf[1] * nf[2];
    const long long ntot = (long long)ms[0] * ms[1] * ms[2];
    std::mt19937_64 rng(1234);
    std::normal_distribution<double> nd;
    std::uniform_real_distribution<double> ud(0.5, 2.0);
    std::vector<double> ker[3];
    for (int a = 0; a < 3; ++a) {
        ker[a].resize(ms[a] / 2 + 1);
        for (auto &v : ker[a]) v = ud(rng);
    }
    auto freq = [&](int a, int k) { return k - ms[a] / 2; };

    // ---- type 1: fw -> fk -------------------------------------------------
    std::vector<cd> fw(nftot * ntr), fw0;
    for (auto &v : fw) v = cd(nd(rng), nd(rng));
    fw0 = fw;
    std::vector<cd> b1((long long)ms[0] * nf[2] * nf[1] * ntr);
    std::vector<cd> fk(ntot * ntr);
    const bool two = ms[2] == 1 && nf[2] == 1;
    if (two) ker[2][0] = 1.0;  // no z factor in 2D
    dft_lines(fw, (long long)nf[1] * nf[2] * ntr, nf[0], iflag);
    if (two) {
        gather(fw, b1, cfa::prune2_fw_to_b1(ms, nf), ntr, nullptr, nullptr, nullptr, "fw->B1 2D");
        dft_lines(b1, (long long)ms[0] * ntr, nf[1], iflag);
        gather(b1, fk, cfa::prune2_b1_to_fk(ms, nf), ntr, ker[0].data(), nullptr, ker[1].data(),
               "B1->fk 2D");
    } else {
        gather(fw, b1, cfa::prune_fw_to_b1(ms, nf), ntr, nullptr, nullptr, nullptr, "fw->B1");
        dft_lines(b1, (long long)ms[0] * nf[2] * ntr, nf[1], iflag);
        gather(b1, fw, cfa::prune_b1_to_b2(ms, nf), ntr, nullptr, nullptr, nullptr, "B1->B2");
        dft_lines(fw, (long long)ms[0] * ms[1] * ntr, nf[2], iflag);  // B2 at the start of fw
        gather(fw, fk, cfa::prune_b2_to_fk(ms, nf), ntr, ker[0].data(), ker[1].data(),
               ker[2].data(), "B2->fk");
    }
    std::vector<cd> ref(ntot * ntr);
    for (int t = 0; t < ntr; ++t)
        for (int kz = 0; kz < ms[2]; ++kz)
            for (int ky = 0; ky < ms[1]; ++ky)
                for (int kx = 0; kx < ms[0]; ++kx) {
                    const int fx = freq(0, kx), fy = freq(1, ky), fz = freq(2, kz);
                    cd s = 0;
                    for (int z = 0; z < nf[2]; ++z)
                        for (int y = 0; y < nf[1]; ++y)
                            for (int x = 0; x < nf[0]; ++x) {
                                const double ph = (double)x * fx / nf[0] + (double)y * fy / nf[1] +
                                                  (double)z * fz / nf[2];
                                s += fw0[t * nftot + ((long long)z * nf[1] + y) * nf[0] + x] *
                                     std::polar(1.0, iflag * kTwoPi * ph);
                            }
                    s /= ker[0][std::abs(fx)] * (ker[1][std::abs(fy)] * ker[2][std::abs(fz)]);
                    ref[t * ntot + ((long long)kz * ms[1] + ky) * ms[0] + kx] = s;
                }
    const double e1 = rel_err(fk, ref);
    std::printf("type 1: rel err vs direct sum %.3e\n", e1);
    CHECK(e1 < 1e-12, "type 1 modes differ from the direct sum: %.3e\n", e1);

    // ---- type 2: fk -> fw, plain and with each derivative weight ----------
    std::vector<cd> fkin(ntot * ntr);
    for (auto &v : fkin) v = cd(nd(rng), nd(rng));
    for (int wsel = -1; wsel < (two ? 2 : 3); ++wsel) {
        const int wsign = iflag;
        std::vector<cd> g(nftot * ntr, cd(99, 99)), h(b1.size(), cd(99, 99));
        if (two) {  // the library maps axis y to the scatter's c axis (wsel 2)
            scatter(fkin, h, cfa::prune2_fk_to_b1(ms, nf), ntr, ker[0].data(), nullptr,
                    ker[1].data(), wsel == 1 ? 2 : wsel, wsign, "fk->B1 2D", true);
            dft_lines(h, (long long)ms[0] * ntr, nf[1], iflag);
            scatter(h, g, cfa::prune2_b1_to_fw(ms, nf), ntr, nullptr, nullptr, nullptr, -1, 1,
                    "B1->fw 2D", true);
        } else {
            scatter(fkin, g, cfa::prune_fk_to_b2(ms, nf), ntr, ker[0].data(), ker[1].data(),
                    ker[2].data(), wsel, wsign, "fk->B2", false);
            dft_lines(g, (long long)ms[0] * ms[1] * ntr, nf[2], iflag);
            scatter(g, h, cfa::prune_b2_to_b1(ms, nf), ntr, nullptr, nullptr, nullptr, -1, 1,
                    "B2->B1", true);
            dft_lines(h, (long long)ms[0] * nf[2] * ntr, nf[1], iflag);
            scatter(h, g, cfa::prune_b1_to_fw(ms, nf), ntr, nullptr, nullptr, nullptr, -1, 1,
                    "B1->fw", true);
        }
        dft_lines(g, (long long)nf[1] * nf[2] * ntr, nf[0], iflag);
        std::vector<cd> r2(nftot * ntr);
        for (int t = 0; t < ntr; ++t)
            for (int z = 0; z < nf[2]; ++z)
                for (int y = 0; y < nf[1]; ++y)
                    for (int x = 0; x < nf[0]; ++x) {
                        cd s = 0;
                        for (int kz = 0; kz < ms[2]; ++kz)
                            for (int ky = 0; ky < ms[1]; ++ky)
                                for (int kx = 0; kx < ms[0]; ++kx) {
                                    const int fx = freq(0, kx), fy = freq(1, ky), fz = freq(2, kz);
                                    cd v = fkin[t * ntot + ((long long)kz * ms[1] + ky) * ms[0] + kx];
                                    v /= ker[0][std::abs(fx)] *
                                         (ker[1][std::abs(fy)] * ker[2][std::abs(fz)]);
                                    if (wsel >= 0)
                                        v *= cd(0, (double)(wsign * (wsel == 0 ? fx : (wsel == 1 ? fy : fz))));
                                    const double ph = (double)x * fx / nf[0] +
                                                      (double)y * fy / nf[1] +
                                                      (double)z * fz / nf[2];
                                    s += v * std::polar(1.0, iflag * kTwoPi * ph);
                                }
                        r2[t * nftot + ((long long)z * nf[1] + y) * nf[0] + x] = s;
                    }
        const double e2 = rel_err(g, r2);
        std::printf("type 2 (weight axis %d): rel err vs direct sum %.3e\n", wsel, e2);
        CHECK(e2 < 1e-12, "type 2 grid (weight axis %d) differs from the direct sum: %.3e\n", wsel,
              e2);
    }
    std::printf("OK: %d failed checks\n", g_fail);
    return g_fail ? 1 : 0;
}
