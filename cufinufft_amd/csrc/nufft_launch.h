// Host-side launch interface between the plan logic (plan.cpp) and the HIP
// kernels (nufft_device.cuh).  Every function returns 0 or a library error
// code and launches asynchronously on the given stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace cfa {

struct BinGeom;
template <typename T> struct KerParams;
template <typename T> struct cpx;

constexpr int kErrHip = 10;
constexpr int kJobqWords = 9;  // SM job queue: one counter per XCD range + finished workgroups
constexpr int kErrFft = 11;
constexpr int kErrLds = 12;
constexpr int kErrNoAlloc = 13;  // no-allocation entry: input beyond what is allocated
// Dynamic LDS of one gfx950 workgroup: its 160 KiB less room for the
// kernels' static job slot.
constexpr size_t kMaxLdsBytes = 160 * 1024 - 64;

// Device arrays written by setpts.
struct SortBufs {
    int *binsize = nullptr;     // [nbins]
    int *binstart = nullptr;    // [nbins]
    int *substart = nullptr;    // [nbins + 1]
    int *s2b = nullptr;         // [nbins + ceil(M / maxsub)] upper bound
    int *skey = nullptr;        // [M] sort keys, pass ping-pong buffer 0
    int *skey_out = nullptr;    // [M] sort keys, pass ping-pong buffer 1
    void *rec[2] = {nullptr, nullptr};  // [M] point records between sort passes
    int *sort_counts = nullptr; // (digit, tile) counts of one sort pass
    void *sort_part = nullptr;  // int2 partial sums of their scan
    int *idx = nullptr;         // [M] caller index of each sorted point
    double *pos = nullptr;      // [dim][pos_stride] sorted rescaled positions
    long long pos_stride = 0;   // capacity in points of each pos plane
    int *counters = nullptr;    // [2]: points, subproblems
    // SM spread (type 1): the padded tile of every bin that one subproblem
    // covers, stored whole instead of added to the fine grid
    // ([transform][bin][tile cell] complex; null when not allocated)
    void *slot = nullptr;
    unsigned long long *jobq = nullptr;  // [kJobqWords] SM job queue words, zero between launches
    void *scan_part = nullptr;  // int2[ceil(nbins / chunk)]
    // method 4: point copies counted per block
    int *kcount = nullptr;       // [nkeys]
    int *kstart = nullptr;       // [nkeys]
    void *kscan_part = nullptr;  // int2[ceil(nkeys / chunk)]
};

struct Launch {
    hipStream_t stream = nullptr;
    int ncu = 256;
};

// Bin sort of M points: LSD counting sort of the bin keys (fine: bin and
// fine-grid cell inside the bin) carrying the points, per-bin counts, scans,
// sorted index and positions, subproblem map.
template <typename T>
int run_bin_sort(int dim, const BinGeom &g, long long nbins, int M, const T *x, const T *y,
                 const T *z, int maxsub, bool fine, bool need_subprob, SortBufs &sb,
                 const Launch &L);

// Sort tiles: one workgroup of kSortWaves waves, kSortRounds rounds of 64
// points per wave; digits of at most kMaxDigitBits bits.
constexpr int kSortWaves = 4;
constexpr int kSortRounds = 16;
constexpr int kSortTile = 4096;
constexpr int kMaxDigitBits = 10;

// Pass structure of that sort for M points and nkeys distinct keys: key
// bits, digit bits per pass, passes, tiles, (digit, tile) counts per pass.
// The fewest passes of at most kMaxDigitBits bits, with one pass's count
// array within 2^27 entries.
struct BinSortShape {
    int bits = 1, digit_bits = 1, passes = 1, tiles = 0;
    long long counts = 0;
};
inline BinSortShape bin_sort_shape(int M, long long nkeys)
{
    BinSortShape s;
    int bits = 1;
    while ((1ll << bits) < nkeys) ++bits;
    s.bits = bits;
    s.tiles = (int)(((long long)M + kSortTile - 1) / kSortTile);
    int dmax = kMaxDigitBits;
    while (dmax > 4 && ((long long)s.tiles << dmax) > (1ll << 27)) --dmax;
    s.passes = (bits + dmax - 1) / dmax;
    s.digit_bits = (bits + s.passes - 1) / s.passes;
    s.counts = (long long)s.tiles << s.digit_bits;
    return s;
}
// Bytes of one point record moved between passes: dim coordinates of elem
// bytes and an int index, aligned to elem.
inline size_t bin_sort_rec_bytes(int dim, int elem)
{
    const size_t raw = (size_t)dim * elem + 4;
    return (raw + elem - 1) / elem * elem;
}
// Natural order (gpu_sort = 0): identity index and positions.
template <typename T>
int run_natural_order(int dim, const BinGeom &g, int M, const T *x, const T *y, const T *z,
                      SortBufs &sb, const Launch &L);
// Method 4 (owner-computes blocks, g.bs = block size): count and scan the
// point copies per block (total copies -> counters[0]), then scatter them and
// build the per-block subproblems.
template <typename T>
int run_blk_count(int dim, const BinGeom &g, int ns, int M, const T *x, const T *y, const T *z,
                  SortBufs &sb, const Launch &L);
template <typename T>
int run_blk_scatter(int dim, const BinGeom &g, int ns, int M, int ncopies, const T *x,
                    const T *y, const T *z, int maxsub, SortBufs &sb, const Launch &L);

// Spread / interp of ntr transforms over the sorted points of sb.
// method 1 = GM(-sort), 2 = SM, 4 = owner-computes blocks (spread only).
template <typename T>
int run_spread(int method, int dim, int ns, int M, const SortBufs &sb, const cpx<T> *c,
               cpx<T> *fw, const BinGeom &g, const KerParams<T> &kp, int maxsub,
               long long sub_bound, int ntr, long long nftot, const Launch &L);
template <typename T>
int run_interp(int method, int dim, int ns, int M, const SortBufs &sb, cpx<T> *c,
               const cpx<T> *fw, const BinGeom &g, const KerParams<T> &kp, int maxsub,
               long long sub_bound, int ntr, long long nftot, const Launch &L);

// Cells of the SM tile of a bin: the bin plus a halo of ceil(ns/2) per side.
inline size_t sm_tile_cells(int dim, int ns, const int *bs)
{
    const int P = (ns + 1) / 2;
    size_t cells = (size_t)(bs[0] + 2 * P);
    if (dim > 1) cells *= (size_t)(bs[1] + 2 * P);
    if (dim > 2) cells *= (size_t)(bs[2] + 2 * P);
    return cells;
}
// Bytes of LDS the SM interp needs for a tile (a complex T per cell).
inline size_t sm_tile_bytes(int dim, int ns, const int *bs, size_t elem)
{
    return 2 * elem * sm_tile_cells(dim, ns, bs);
}

// LDS bytes of one SM spread job with bt threads: re / im accumulator planes
// (acc_bytes per cell each), the staged chunk of bt points (start cell, dim
// first-tap offsets and the complex strength, elem = sizeof(T)), 64 bank
// counters, one reduction value per wave, and last the FP32 drain planes,
// which only jobs of more than bt points use (32-bit cells with subproblems
// larger than one chunk).
inline size_t spread_job_lds(size_t cells, int acc_bytes, int dim, int elem, int bt, bool drain)
{
    return cells * 2 * acc_bytes + (size_t)bt * (4 + (size_t)(dim + 2) * elem) + 64 * 4 +
           (size_t)(bt / 64) * elem + (drain ? cells * 8 : 0);
}
// Largest number of points one 32-bit fixed-point spread scale may cover at
// tolerance tol: 4 n 2^-30 <= tol / 5 (536 at tol 1e-5).
inline int fx32_bound(double tol)
{
    double n = tol / 5.0 / (4.0 / 1073741824.0);
    n = n < 0 ? 0 : (n > 2e9 ? 2e9 : n);
    return (int)n;
}

// Launch shape of the SM spread: threads per workgroup, accumulator bytes
// per cell (4: 32-bit fixed point drained between chunks, 8: 64-bit) and
// dynamic LDS bytes, for a plan (group = false) or a plan group.
struct SpreadCfg {
    int bt = 256;
    int acc_bytes = 8;
    size_t lds = 0;
};
inline SpreadCfg spread_cfg(int dim, int ns, const int *bs, size_t elem, int fx32max,
                            int maxsub, bool group)
{
    const size_t cells = sm_tile_cells(dim, ns, bs);
    SpreadCfg cfg;
    // 32-bit cells when a whole workgroup-wide chunk stays within the plan's
    // 32-bit bound (BT = 512, or 256 in plan groups, whose kernels run 256
    // threads); 64-bit cells otherwise
    const int chunk = group ? 256 : 512;
    if (elem == 4 && fx32max >= chunk) {
        cfg.acc_bytes = 4;
        cfg.bt = chunk;
    } else {
        cfg.acc_bytes = 8;
        // 1024 threads when the tile leaves one workgroup per CU at any
        // width (over half the LDS) and subproblems fill it: the wider
        // workgroup adds waves without taking concurrent jobs away
        const size_t l256 = spread_job_lds(cells, 8, dim, (int)elem, 256, false);
        const size_t l1024 = spread_job_lds(cells, 8, dim, (int)elem, 1024, false);
        cfg.bt = !group && maxsub >= 1024 && l256 > kMaxLdsBytes / 2 && l1024 <= kMaxLdsBytes
                     ? 1024
                     : 256;
    }
    cfg.lds = spread_job_lds(cells, cfg.acc_bytes, dim, (int)elem, cfg.bt,
                             cfg.acc_bytes == 4 && maxsub > cfg.bt);
    return cfg;
}
// LDS bytes of a method-4 (owner-computes block) spread job over blocks ob.
size_t blk_lds_bytes(int dim, const int *ob, size_t elem, int maxsub, int fxwide);

template <typename T>
int run_deconvolve(int dim, const int *ms, const int *nf, const cpx<T> *fw, cpx<T> *fk,
                   const T *const *ker, int ntr, const Launch &L);
// wax >= 0 also multiplies the modes by i * wsign * k_wax (derivative modes).
template <typename T>
int run_amplify(int dim, const int *ms, const int *nf, cpx<T> *fw, const cpx<T> *fk,
                const T *const *ker, int ntr, const Launch &L, int wax = -1, int wsign = 1);
template <typename T>
int run_zero(cpx<T> *p, long long n, const Launch &L);

// Pruned 3D FFT transposes (fft_prune.hip).  Gather: out[t][c'][p][r] =
// in[t][p][r][src(c')] over the Nc mode positions c' of an axis of C cells
// (c' < kneg: the kneg negative frequencies at the top of the axis), r the
// output's unit-stride axis, c the input's.  Scatter: the inverse, writing
// every c < C (zeros between the modes).  hr / hp: the mode-index offsets
// (ms/2) of r and p where the Fourier-series tables are applied.
struct PruneGather {
    long long in_t = 0, ip = 0, ir = 0;    // input strides (c unit)
    long long out_t = 0, oc = 0, op = 0;   // output strides (r unit)
    int R = 0, P = 0, Nc = 0, C = 0, kneg = 0, hr = 0, hp = 0;
};
struct PruneScatter {
    long long in_t = 0, ic = 0, ip = 0;    // input strides (r unit)
    long long out_t = 0, op = 0, orr = 0;  // output strides (c unit)
    int R = 0, P = 0, Nc = 0, C = 0, kneg = 0, hr = 0, hp = 0;
};
// Cell of mode index q (frequency q - kneg) on an axis of C cells, and the
// mode index of cell c (-1 where the cell holds no mode); Nc modes, the kneg
// negative frequencies at the top of the axis.
__host__ __device__ inline int prune_cell(int q, int kneg, int C)
{
    return q < kneg ? C - kneg + q : q - kneg;
}
__host__ __device__ inline int prune_mode(int c, int Nc, int kneg, int C)
{
    return c < Nc - kneg ? c + kneg : (c >= C - kneg ? c - C + kneg : -1);
}

// The transposes of the pruned 3D FFT for modes ms and fine grid nf (x
// fastest everywhere).  Buffers: fw [z][y][x] (nf0 nf1 nf2 per transform),
// B1 [x'][z][y] (ms0 nf2 nf1), B2 [x'][y'][z] (ms0 ms1 nf2, held in fw),
// fk [kz][ky][kx] (ms0 ms1 ms2); x', y', kx, ky, kz are mode indices.
// Type 1: fw -> B1 (x pass done), B1 -> B2 (y pass done), B2 -> fk (z pass
// done; with the Fourier-series tables: the deconvolve).
inline PruneGather prune_fw_to_b1(const int *ms, const int *nf)
{
    PruneGather d;
    d.in_t = (long long)nf[0] * nf[1] * nf[2];
    d.ip = (long long)nf[1] * nf[0];
    d.ir = nf[0];
    d.out_t = (long long)ms[0] * nf[2] * nf[1];
    d.oc = (long long)nf[2] * nf[1];
    d.op = nf[1];
    d.R = nf[1];
    d.P = nf[2];
    d.Nc = ms[0];
    d.C = nf[0];
    d.kneg = ms[0] / 2;
    return d;
}
inline PruneGather prune_b1_to_b2(const int *ms, const int *nf)
{
    PruneGather d;
    d.in_t = (long long)ms[0] * nf[2] * nf[1];
    d.ip = (long long)nf[2] * nf[1];
    d.ir = nf[1];
    d.out_t = (long long)ms[0] * ms[1] * nf[2];
    d.oc = nf[2];
    d.op = (long long)ms[1] * nf[2];
    d.R = nf[2];
    d.P = ms[0];
    d.Nc = ms[1];
    d.C = nf[1];
    d.kneg = ms[1] / 2;
    return d;
}
inline PruneGather prune_b2_to_fk(const int *ms, const int *nf)
{
    PruneGather d;
    d.in_t = (long long)ms[0] * ms[1] * nf[2];
    d.ip = nf[2];
    d.ir = (long long)ms[1] * nf[2];
    d.out_t = (long long)ms[0] * ms[1] * ms[2];
    d.oc = (long long)ms[1] * ms[0];
    d.op = ms[0];
    d.R = ms[0];
    d.P = ms[1];
    d.Nc = ms[2];
    d.C = nf[2];
    d.kneg = ms[2] / 2;
    d.hr = ms[0] / 2;
    d.hp = ms[1] / 2;
    return d;
}
// Type 2: fk -> B2 (with the tables: the amplify; z then transformed),
// B2 -> B1 (y then transformed), B1 -> fw (x then transformed).
inline PruneScatter prune_fk_to_b2(const int *ms, const int *nf)
{
    PruneScatter d;
    d.in_t = (long long)ms[0] * ms[1] * ms[2];
    d.ic = (long long)ms[1] * ms[0];
    d.ip = ms[0];
    d.out_t = (long long)ms[0] * ms[1] * nf[2];
    d.op = nf[2];
    d.orr = (long long)ms[1] * nf[2];
    d.R = ms[0];
    d.P = ms[1];
    d.Nc = ms[2];
    d.C = nf[2];
    d.kneg = ms[2] / 2;
    d.hr = ms[0] / 2;
    d.hp = ms[1] / 2;
    return d;
}
inline PruneScatter prune_b2_to_b1(const int *ms, const int *nf)
{
    PruneScatter d;
    d.in_t = (long long)ms[0] * ms[1] * nf[2];
    d.ic = nf[2];
    d.ip = (long long)ms[1] * nf[2];
    d.out_t = (long long)ms[0] * nf[2] * nf[1];
    d.op = (long long)nf[2] * nf[1];
    d.orr = nf[1];
    d.R = nf[2];
    d.P = ms[0];
    d.Nc = ms[1];
    d.C = nf[1];
    d.kneg = ms[1] / 2;
    return d;
}
inline PruneScatter prune_b1_to_fw(const int *ms, const int *nf)
{
    PruneScatter d;
    d.in_t = (long long)ms[0] * nf[2] * nf[1];
    d.ic = (long long)nf[2] * nf[1];
    d.ip = nf[1];
    d.out_t = (long long)nf[0] * nf[1] * nf[2];
    d.op = (long long)nf[1] * nf[0];
    d.orr = nf[0];
    d.R = nf[1];
    d.P = nf[2];
    d.Nc = ms[0];
    d.C = nf[0];
    d.kneg = ms[0] / 2;
    return d;
}
// 2D (one plane, P = 1): fw [y][x] (nf0 nf1), B1 [x'][y] (ms0 nf1, the
// mode columns), fk [ky][kx].  Type 1: fw -> B1 after the x pass, B1 -> fk
// after the y pass (the deconvolve, tables kr = x and kc = y); type 2:
// fk -> B1 (the amplify), y pass, B1 -> fw, x pass.
inline PruneGather prune2_fw_to_b1(const int *ms, const int *nf)
{
    PruneGather d;
    d.in_t = (long long)nf[0] * nf[1];
    d.ir = nf[0];
    d.out_t = (long long)ms[0] * nf[1];
    d.oc = nf[1];
    d.R = nf[1];
    d.P = 1;
    d.Nc = ms[0];
    d.C = nf[0];
    d.kneg = ms[0] / 2;
    return d;
}
inline PruneGather prune2_b1_to_fk(const int *ms, const int *nf)
{
    PruneGather d;
    d.in_t = (long long)ms[0] * nf[1];
    d.ir = nf[1];
    d.out_t = (long long)ms[0] * ms[1];
    d.oc = ms[0];
    d.R = ms[0];
    d.P = 1;
    d.Nc = ms[1];
    d.C = nf[1];
    d.kneg = ms[1] / 2;
    d.hr = ms[0] / 2;
    return d;
}
inline PruneScatter prune2_fk_to_b1(const int *ms, const int *nf)
{
    PruneScatter d;
    d.in_t = (long long)ms[0] * ms[1];
    d.ic = ms[0];
    d.out_t = (long long)ms[0] * nf[1];
    d.orr = nf[1];
    d.R = ms[0];
    d.P = 1;
    d.Nc = ms[1];
    d.C = nf[1];
    d.kneg = ms[1] / 2;
    d.hr = ms[0] / 2;
    return d;
}
inline PruneScatter prune2_b1_to_fw(const int *ms, const int *nf)
{
    PruneScatter d;
    d.in_t = (long long)ms[0] * nf[1];
    d.ic = nf[1];
    d.out_t = (long long)nf[0] * nf[1];
    d.orr = nf[0];
    d.R = nf[1];
    d.P = 1;
    d.Nc = ms[0];
    d.C = nf[0];
    d.kneg = ms[0] / 2;
    return d;
}
// Element offsets the transposes read and write: gather (t, p, r, mode c'),
// scatter (t, p, r, mode c') read / (t, p, r, cell c) write.
__host__ __device__ inline long long prune_gather_src(const PruneGather &d, long long t, int p,
                                                      int r, int q)
{
    return t * d.in_t + p * d.ip + r * d.ir + prune_cell(q, d.kneg, d.C);
}
__host__ __device__ inline long long prune_gather_dst(const PruneGather &d, long long t, int p,
                                                      int r, int q)
{
    return t * d.out_t + q * d.oc + p * d.op + r;
}
__host__ __device__ inline long long prune_scatter_src(const PruneScatter &d, long long t, int p,
                                                       int r, int q)
{
    return t * d.in_t + q * d.ic + p * d.ip + r;
}
__host__ __device__ inline long long prune_scatter_dst(const PruneScatter &d, long long t, int p,
                                                       int r, int c)
{
    return t * d.out_t + p * d.op + r * d.orr + c;
}

template <typename T>
int run_prune_gather(const cpx<T> *in, cpx<T> *out, const PruneGather &d, int ntr,
                     const T *kr, const T *kp, const T *kc, const Launch &L);
template <typename T>
int run_prune_scatter(const cpx<T> *in, cpx<T> *out, const PruneScatter &d, int ntr,
                      const T *kr, const T *kp, const T *kc, int wsel, int wsign,
                      const Launch &L);

// Plan groups: one persistent spread (type 1) or interp (type 2) launch over
// the subproblems of every member.  items is a device array of n
// GroupItem<T> (filled on the host with group_item_fill and copied), off a
// device array of n+1 job offsets (written by the launch itself).
size_t group_item_bytes(bool dbl);
template <typename T, int DIM>
int group_dim(int type, int ns, const void *items, int n, const long long *off,
              const KerParams<T> &kp, const int *bs, int maxsub, long long jobs_bound,
              const Launch &L);
template <typename T, int DIM>
void group_item_fill(void *dst, const BinGeom &g, const SortBufs &sb, int maxsub, const void *in,
                     void *out, long long nftot, int M, int ntr);

// Per-(T, DIM) spread/interp launchers (one translation unit each).
template <typename T, int DIM>
int spread_dim(int method, int ns, int M, const SortBufs &sb, const cpx<T> *c, cpx<T> *fw,
               const BinGeom &g, const KerParams<T> &kp, int maxsub, long long sub_bound,
               int ntr, long long nftot, const Launch &L);
template <typename T, int DIM>
int interp_dim(int method, int ns, int M, const SortBufs &sb, cpx<T> *c, const cpx<T> *fw,
               const BinGeom &g, const KerParams<T> &kp, int maxsub, long long sub_bound,
               int ntr, long long nftot, const Launch &L);

}  // namespace cfa
