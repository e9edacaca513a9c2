// Host-side index check of the setpts and spread index arithmetic at full
// workload sizes (test infrastructure; built for the CPU with AddressSanitizer
// and UndefinedBehaviorSanitizer by tests/cpu_index/Makefile).
//
// The library's index helpers are compiled here for the host from the same
// source as the kernels (nufft_device.cuh: rescale_exact, bin1, Tile,
// for_tile_cells, subprob_job, tile_window, rr_slot; nufft_launch.h:
// bin_sort_shape, spread_cfg, fx32_bound, sm_tile_cells).  The program draws
// the points of a workload, then walks every index the device code forms:
//   * the counting sort: for every pass, every (digit, tile) count index and
//     every destination the scatter computes (tile, wave, round, lane order
//     and equal-digit ranks as the kernel forms them) -- each in [0, M) and
//     hit exactly once, the result sorted and stable;
//   * the bin counts, starts and subproblem map against the buffer sizes the
//     plan allocates (nbins + ceil(M / maxsub) subproblems);
//   * every SM spread job: its point range, each point's tap window, start
//     cell, bank residue and round-robin slot, the last tap's tile cell, the
//     LDS bytes touched against the launch's allocation, and the fine-grid
//     offset of every flushed tile cell.
// Signed overflow anywhere in that arithmetic stops the run (UBSan), as does
// any access outside an allocation (ASan).  Usage:
//   index_check f|d dim nf1 nf2 nf3 ns M bx by bz maxsub tol [uniform|mixed]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../cufinufft_amd/csrc/nufft_device.cuh"

using namespace cfa;

static int g_fail = 0;
#define CHECK(cond, ...)                                                        \
    do {                                                                        \
        if (!(cond)) {                                                          \
            if (g_fail < 20) {                                                  \
                std::fprintf(stderr, "FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond); \
                std::fprintf(stderr, __VA_ARGS__);                              \
                std::fprintf(stderr, "\n");                                     \
            }                                                                   \
            ++g_fail;                                                           \
        }                                                                       \
    } while (0)

template <typename T>
static void draw_points(int dim, long long M, bool mixed, std::vector<std::vector<T>> &P)
{
    std::mt19937_64 rng(3);
    std::uniform_real_distribution<double> U(-M_PI, M_PI);
    std::normal_distribution<double> G(0.0, 0.2);
    double cen[4][3];
    for (auto &c : cen)
        for (double &v : c) v = U(rng);
    P.assign(dim, std::vector<T>((size_t)M));
    const long long Mu = mixed ? M / 2 : M;
    for (long long i = 0; i < M; ++i) {
        const int k = (int)(rng() % 4);
        for (int d = 0; d < dim; ++d) {
            double v = i < Mu ? U(rng) : cen[k][d] + G(rng);
            v = std::fmod(v + M_PI, 2 * M_PI);
            if (v < 0) v += 2 * M_PI;
            P[d][(size_t)i] = (T)(v - M_PI);
        }
    }
}

template <typename T, int DIM, int NS>
static int run(const std::vector<int> &nf, long long M, const int *bs, int maxsub, double tol,
               bool mixed)
{
    std::vector<std::vector<T>> P;
    draw_points<T>(DIM, M, mixed, P);
    BinGeom g{};
    for (int d = 0; d < 3; ++d) {
        g.nf[d] = d < DIM ? nf[d] : 1;
        g.bs[d] = d < DIM ? bs[d] : 1;
        g.nb[d] = d < DIM ? (int)std::ceil((T)g.nf[d] / (T)g.bs[d]) : 1;
    }
    const long long nbins = (long long)g.nb[0] * g.nb[1] * g.nb[2];
    const long long nftot = (long long)g.nf[0] * g.nf[1] * g.nf[2];
    std::printf("M %lld dim %d nf %d %d %d ns %d bins %d %d %d (%lld) maxsub %d\n", M, DIM,
                g.nf[0], g.nf[1], g.nf[2], NS, g.bs[0], g.bs[1], g.bs[2], nbins, maxsub);

    // ---- keys (the sort key of the bin sort: bin of each axis, combined)
    std::vector<int> key((size_t)M);
    for (long long i = 0; i < M; ++i) {
        long long k = 0, mul = 1;
        for (int d = 0; d < DIM; ++d) {
            const int b = bin1<T>(rescale_exact(P[d][(size_t)i], g.nf[d]), g.bs[d], g.nb[d]);
            CHECK(b >= 0 && b < g.nb[d], "point %lld axis %d bin %d", i, d, b);
            k += mul * b;
            mul *= g.nb[d];
        }
        CHECK(k >= 0 && k < nbins, "key %lld", k);
        key[(size_t)i] = (int)k;
    }

    // ---- counting sort passes, destinations formed as the scatter kernel does
    const BinSortShape sh = bin_sort_shape((int)M, nbins);
    std::printf("sort: %d key bits, %d passes of %d bits, %d tiles, %lld counts\n", sh.bits,
                sh.passes, sh.digit_bits, sh.tiles, sh.counts);
    std::vector<int> cur_key = key, nxt_key((size_t)M);
    std::vector<int> cur_idx((size_t)M), nxt_idx((size_t)M);
    for (long long i = 0; i < M; ++i) cur_idx[(size_t)i] = (int)i;
    std::vector<int> counts((size_t)sh.counts);
    std::vector<unsigned char> hit((size_t)M);
    for (int p = 0; p < sh.passes; ++p) {
        const int shift = p * sh.digit_bits;
        const int dbits = std::min(sh.digit_bits, sh.bits - shift);
        const int nd = 1 << dbits;
        std::fill(counts.begin(), counts.end(), 0);
        for (long long j = 0; j < M; ++j) {
            const long long t = j / kSortTile;
            const int d = (cur_key[(size_t)j] >> shift) & (nd - 1);
            const long long ci = (long long)d * sh.tiles + t;
            CHECK(ci >= 0 && ci < sh.counts, "count index %lld", ci);
            counts[(size_t)ci] += 1;
        }
        long long run = 0;  // exclusive scan, digit-major
        for (long long q = 0; q < ((long long)sh.tiles << dbits); ++q) {
            const int c = counts[(size_t)q];
            counts[(size_t)q] = (int)run;
            run += c;
        }
        CHECK(run == M, "scan total %lld", run);
        std::fill(hit.begin(), hit.end(), 0);
        std::vector<int> wb((size_t)kSortWaves * nd);
        for (long long t = 0; t < sh.tiles; ++t) {
            std::fill(wb.begin(), wb.end(), 0);
            for (int w = 0; w < kSortWaves; ++w)
                for (int r = 0; r < kSortRounds; ++r)
                    for (int l = 0; l < 64; ++l) {
                        const long long j = t * kSortTile + w * (kSortRounds * 64) + r * 64 + l;
                        if (j < M) wb[(size_t)w * nd + ((cur_key[(size_t)j] >> shift) & (nd - 1))]++;
                    }
            for (int d = 0; d < nd; ++d) {
                int base = counts[(size_t)((long long)d * sh.tiles + t)];
                for (int w = 0; w < kSortWaves; ++w) {
                    const int c = wb[(size_t)w * nd + d];
                    wb[(size_t)w * nd + d] = base;
                    base += c;
                }
            }
            for (int w = 0; w < kSortWaves; ++w)
                for (int r = 0; r < kSortRounds; ++r)
                    for (int l = 0; l < 64; ++l) {  // lane order = rank order among equal digits
                        const long long j = t * kSortTile + w * (kSortRounds * 64) + r * 64 + l;
                        if (j >= M) continue;
                        const int d = (cur_key[(size_t)j] >> shift) & (nd - 1);
                        const int dst = wb[(size_t)w * nd + d]++;
                        CHECK(dst >= 0 && dst < M, "pass %d dst %d", p, dst);
                        if (dst < 0 || dst >= M) continue;
                        CHECK(!hit[(size_t)dst], "pass %d dst %d twice", p, dst);
                        hit[(size_t)dst] = 1;
                        nxt_key[(size_t)dst] = cur_key[(size_t)j];
                        nxt_idx[(size_t)dst] = cur_idx[(size_t)j];
                    }
        }
        cur_key.swap(nxt_key);
        cur_idx.swap(nxt_idx);
    }
    for (long long s = 1; s < M; ++s) {
        const int a = cur_key[(size_t)s - 1], b = cur_key[(size_t)s];
        CHECK(a < b || (a == b && cur_idx[(size_t)s - 1] < cur_idx[(size_t)s]),
              "sorted order broken at %lld", s);
    }

    // ---- bin counts, starts, subproblem map and the plan's allocation
    std::vector<int> binsize((size_t)nbins, 0), binstart((size_t)nbins), substart((size_t)nbins + 1);
    for (long long s = 0; s < M; ++s) binsize[(size_t)cur_key[(size_t)s]]++;
    long long acc = 0, subs = 0;
    for (long long b = 0; b < nbins; ++b) {
        binstart[(size_t)b] = (int)acc;
        substart[(size_t)b] = (int)subs;
        acc += binsize[(size_t)b];
        subs += (binsize[(size_t)b] + maxsub - 1) / maxsub;
    }
    substart[(size_t)nbins] = (int)subs;
    const long long sub_bound = nbins + (M + maxsub - 1) / maxsub;  // s2b allocation
    CHECK(subs <= sub_bound, "subproblems %lld > allocated %lld", subs, sub_bound);
    std::vector<int> s2b((size_t)std::max(subs, 1ll));
    for (long long b = 0; b < nbins; ++b)
        for (int s = substart[(size_t)b]; s < substart[(size_t)b + 1]; ++s) s2b[(size_t)s] = (int)b;
    const int maxbin = *std::max_element(binsize.begin(), binsize.end());
    std::printf("bins: fullest %d points, %lld subproblems (bound %lld)\n", maxbin, subs,
                sub_bound);

    // ---- SM spread jobs
    const int elem = (int)sizeof(T);
    const int fx32 = elem == 4 ? fx32_bound(tol) : 0;
    const SpreadCfg cfg = spread_cfg(DIM, NS, g.bs, elem, fx32, maxsub, false);
    const int BT = cfg.bt;
    const int NK = cfg.acc_bytes == 4 ? 64 : 32;
    std::printf("spread: %d threads, %d-byte cells, %zu LDS bytes (limit %zu)\n", BT,
                cfg.acc_bytes, cfg.lds, kMaxLdsBytes);
    CHECK(cfg.lds <= kMaxLdsBytes, "LDS %zu", cfg.lds);
    const size_t cells = sm_tile_cells(DIM, NS, g.bs);
    const size_t stage_end = cells * 2 * cfg.acc_bytes + (size_t)BT * (4 + (DIM + 2) * elem) +
                             64 * 4 + (size_t)(BT / 64) * elem;
    SubprobArgs sa{s2b.data(), substart.data(), binstart.data(), binsize.data(), nullptr, maxsub};
    std::vector<int> cnt(64), slot_hit((size_t)BT);
    long long flushed = 0;
    for (long long job = 0; job < subs; ++job) {
        const SubprobJob J = subprob_job(job, (int)subs, sa);
        CHECK(J.t == 0 && J.np >= 1 && J.np <= maxsub && J.p0 >= 0 && (long long)J.p0 + J.np <= M,
              "job %lld: t %d p0 %d np %d", job, J.t, J.p0, J.np);
        const Tile<DIM, NS> tl(g, J.b);
        CHECK((size_t)tl.n == cells, "tile cells %d", tl.n);
        const int nch = (J.np + BT - 1) / BT;
        const bool drain = cfg.acc_bytes == 4 && nch > 1;
        CHECK(!drain || stage_end + cells * 8 <= cfg.lds, "drain planes outside LDS (job %lld)", job);
        CHECK(stage_end <= cfg.lds, "staging outside LDS");
        for (int ch = 0; ch < nch; ++ch) {
            const int nc = std::min(BT, J.np - ch * BT);
            std::fill(cnt.begin(), cnt.end(), 0);
            std::vector<int> keys(nc), ranks(nc);
            for (int q = 0; q < nc; ++q) {
                const long long si = J.p0 + (long long)ch * BT + q;
                const int i = cur_idx[(size_t)si];
                int l[3] = {0, 0, 0};
                for (int d = 0; d < DIM; ++d) {
                    const double v = tap_pos(rescale_exact(P[d][(size_t)i], g.nf[d]), g.nf[d]);
                    double x1;
                    l[d] = tile_window<NS>(v, tl.o[d], tl.tw[d], x1);
                    CHECK(l[d] >= 0 && l[d] + NS <= tl.tw[d], "window %d", l[d]);
                    CHECK(x1 > -0.5 * NS - 1.0 && x1 < 1.0, "offset %g", x1);
                }
                const long long s = l[0] + (long long)tl.tw[0] * (l[1] + (long long)tl.tw[1] * l[2]);
                // last tap: (dz, dy, a) = (NS-1, NS-1, NS-1) on the axes in use
                long long last = s + (NS - 1);
                if (DIM > 1) last += (long long)(NS - 1) * tl.tw[0];
                if (DIM > 2) last += (long long)(NS - 1) * tl.tw[0] * tl.tw[1];
                CHECK(s >= 0 && last < tl.n, "taps %lld..%lld of %d", s, last, tl.n);
                keys[q] = (int)(s & (NK - 1));
                ranks[q] = cnt[(size_t)keys[q]]++;
            }
            std::fill(slot_hit.begin(), slot_hit.end(), 0);
            for (int q = 0; q < nc; ++q) {
                const int slot = rr_slot(cnt.data(), NK, keys[q], ranks[q]);
                CHECK(slot >= 0 && slot < nc, "slot %d of %d", slot, nc);
                if (slot >= 0 && slot < nc) {
                    CHECK(!slot_hit[(size_t)slot], "slot %d twice", slot);
                    slot_hit[(size_t)slot] = 1;
                }
            }
        }
        for (int tid = 0; tid < BT; ++tid)
            for_tile_cells(tl, g, tid, BT, [&](int q, long long off) {
                CHECK(q >= 0 && q < tl.n && off >= 0 && off < nftot, "flush cell %d off %lld", q,
                      off);
                ++flushed;
            });
    }
    std::printf("spread: %lld jobs, %lld tile cells flushed\n", subs, flushed);

    // ---- tile slots and their gather (plans where every axis has >= 3 bins,
    // a halo no wider than a bin and a tile no wider than the grid): along
    // each axis, the (bin, tile coordinate) pairs the gather lists for a grid
    // index must map to it, and must include every pair that maps to it and
    // that a point of the bin can reach (within P of the bin's cells)
    {
        const int Ph = (NS + 1) / 2;
        bool slots = true;
        for (int d = 0; d < DIM; ++d)
            slots = slots && g.nb[d] >= 3 && Ph <= g.bs[d] && g.bs[d] + 2 * Ph <= g.nf[d];
        long long pairs = 0;
        for (int d = 0; slots && d < DIM; ++d) {
            const int nb = g.nb[d], bs = g.bs[d], nfd = g.nf[d], tw = bs + 2 * Ph;
            for (int gc = 0; gc < nfd; ++gc) {
                int cb[5], cq[5];
                const int n = tile_cover_axis(gc, nb, bs, nfd, Ph, cb, cq);
                CHECK(n >= 1 && n <= 5, "axis %d cell %d: %d covering tiles", d, gc, n);
                for (int k = 0; k < n; ++k) {
                    const int c = cb[k] * bs - Ph + cq[k];
                    CHECK(cb[k] >= 0 && cb[k] < nb && cq[k] >= 0 && cq[k] < tw &&
                              ((c % nfd) + nfd) % nfd == gc,
                          "axis %d cell %d: tile (%d, %d) maps elsewhere", d, gc, cb[k], cq[k]);
                    for (int k2 = 0; k2 < k; ++k2)
                        CHECK(cb[k2] != cb[k] || cq[k2] != cq[k], "axis %d cell %d: pair twice", d,
                              gc);
                }
                for (int u = 0; u < nb; ++u) {
                    const int end = std::min((u + 1) * bs, nfd);
                    for (int q = 0; q < tw; ++q) {
                        const int c = u * bs - Ph + q;
                        if (((c % nfd) + nfd) % nfd != gc || c >= end + Ph) continue;
                        bool listed = false;
                        for (int k = 0; k < n; ++k) listed = listed || (cb[k] == u && cq[k] == q);
                        CHECK(listed, "axis %d cell %d: reachable tile (%d, %d) not gathered", d,
                              gc, u, q);
                        ++pairs;
                    }
                }
            }
        }
        std::printf("tile slots: %s, %lld reachable (bin, cell) pairs checked\n",
                    slots ? "used" : "not used", pairs);
    }
    std::printf("%s: %d failed checks\n", g_fail ? "FAILED" : "OK", g_fail);
    return g_fail ? 1 : 0;
}

template <typename T, int DIM>
static int dispatch_ns(int ns, const std::vector<int> &nf, long long M, const int *bs, int maxsub,
                       double tol, bool mixed)
{
    switch (ns) {
    case 4: return run<T, DIM, 4>(nf, M, bs, maxsub, tol, mixed);
    case 6: return run<T, DIM, 6>(nf, M, bs, maxsub, tol, mixed);
    case 7: return run<T, DIM, 7>(nf, M, bs, maxsub, tol, mixed);
    case 9: return run<T, DIM, 9>(nf, M, bs, maxsub, tol, mixed);
    case 10: return run<T, DIM, 10>(nf, M, bs, maxsub, tol, mixed);
    case 11: return run<T, DIM, 11>(nf, M, bs, maxsub, tol, mixed);
    case 15: return run<T, DIM, 15>(nf, M, bs, maxsub, tol, mixed);
    }
    std::fprintf(stderr, "ns %d not instantiated\n", ns);
    return 2;
}

int main(int argc, char **argv)
{
    if (argc < 13) {
        std::fprintf(stderr,
                     "usage: %s f|d dim nf1 nf2 nf3 ns M bx by bz maxsub tol [uniform|mixed]\n",
                     argv[0]);
        return 2;
    }
    const bool dbl = argv[1][0] == 'd';
    const int dim = std::atoi(argv[2]);
    const std::vector<int> nf = {std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5])};
    const int ns = std::atoi(argv[6]);
    const long long M = std::atoll(argv[7]);
    const int bs[3] = {std::atoi(argv[8]), std::atoi(argv[9]), std::atoi(argv[10])};
    const int maxsub = std::atoi(argv[11]);
    const double tol = std::atof(argv[12]);
    const bool mixed = argc > 13 && std::string(argv[13]) == "mixed";
    if (dim == 2) return dbl ? dispatch_ns<double, 2>(ns, nf, M, bs, maxsub, tol, mixed)
                             : dispatch_ns<float, 2>(ns, nf, M, bs, maxsub, tol, mixed);
    if (dim == 3) return dbl ? dispatch_ns<double, 3>(ns, nf, M, bs, maxsub, tol, mixed)
                             : dispatch_ns<float, 3>(ns, nf, M, bs, maxsub, tol, mixed);
    return dbl ? dispatch_ns<double, 1>(ns, nf, M, bs, maxsub, tol, mixed)
               : dispatch_ns<float, 1>(ns, nf, M, bs, maxsub, tol, mixed);
}
