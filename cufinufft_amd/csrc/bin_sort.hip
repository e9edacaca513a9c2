// setpts bin sort: a hand-written LSD counting sort of the points' bin keys.
//
// The reference sorts with an atomic histogram (one atomicAdd per point on
// its bin, whose old value is the point's rank, CalcBinSize_noghost_*,
// spreadinterp2d.cu:102-126, spreadinterp3d.cu:15-42), an exclusive scan
// (spread2d_wrapper.cu:268-300) and a scatter of the point indices
// (CalcInvertofGlobalSortIdx_*, spreadinterp3d.cu:44-70).  On clustered
// points its histogram atomics pile onto a few hot bins and serialise in L2
// (11 ms for the 1e8 points of the 3D type-1 workload, against ~0.3 ms of
// memory traffic), and its execute gathers coordinates through the index.
//
// Here the key (the bin, or bin and fine cell for gpu_sort = 2) is sorted
// in one to three counting passes of at most 10 digit bits each, least
// significant digit first.  Every pass is a per-tile digit histogram in LDS,
// one exclusive scan of the (digit, tile) counts over the whole array, and a
// stable scatter in which each wave ranks its own items among equal digits
// with ballots (no same-address atomics at any point density).  The payload
// is the point itself -- its coordinates and input index -- so the last pass
// writes the sorted index and the rescaled positions the execute kernels
// stream, with no gather through the index afterwards.  The result is the
// stable sort of the keys: bins, bin starts and subproblems are the
// reference's exactly, and inside a bin the points keep their input order.
#include <algorithm>

#include "nufft_device.cuh"
#include "nufft_launch.h"

namespace cfa {

namespace {
inline int ok_or_err() { return hipGetLastError() == hipSuccess ? 0 : kErrHip; }

inline int grid_pts(long long n, int ncu)
{
    long long b = (n + 255) / 256;
    long long cap = (long long)ncu * 32;
    return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace

static_assert(kSortTile == kSortWaves * kSortRounds * 64, "tile = 4 waves x 16 rounds x 64");

// One point as the passes move it: coordinates and input index.
template <typename T, int DIM> struct SortRec {
    T v[DIM];
    int i;
};
static_assert(sizeof(SortRec<float, 3>) == 16 && sizeof(SortRec<double, 3>) == 32 &&
                  sizeof(SortRec<float, 2>) == 12 && sizeof(SortRec<double, 2>) == 24 &&
                  sizeof(SortRec<float, 1>) == 8 && sizeof(SortRec<double, 1>) == 16,
              "record sizes as bin_sort_rec_bytes computes them");

// Sort key of point i: the bin (point_bin), or for FINE the bin, then the
// fine-grid cell inside it (bin * cells per bin + cell).
template <typename T, int DIM, bool FINE>
__device__ __forceinline__ int sort_key(const T *v, const BinGeom &g)
{
    if constexpr (FINE) {
        int b0, c0, b1 = 0, c1 = 0, b2 = 0, c2 = 0;
        bin_and_cell<T>(rescale_exact(v[0], g.nf[0]), g.bs[0], g.nb[0], b0, c0);
        if (DIM > 1) bin_and_cell<T>(rescale_exact(v[1], g.nf[1]), g.bs[1], g.nb[1], b1, c1);
        if (DIM > 2) bin_and_cell<T>(rescale_exact(v[2], g.nf[2]), g.bs[2], g.nb[2], b2, c2);
        const int bin = b0 + g.nb[0] * (b1 + g.nb[1] * b2);
        const int vol = g.bs[0] * g.bs[1] * g.bs[2];
        return bin * vol + c0 + g.bs[0] * (c1 + g.bs[1] * c2);
    } else {
        int k = bin1<T>(rescale_exact(v[0], g.nf[0]), g.bs[0], g.nb[0]);
        if (DIM > 1) k += g.nb[0] * bin1<T>(rescale_exact(v[1], g.nf[1]), g.bs[1], g.nb[1]);
        if (DIM > 2)
            k += g.nb[0] * g.nb[1] * bin1<T>(rescale_exact(v[2], g.nf[2]), g.bs[2], g.nb[2]);
        return k;
    }
}

// Record of item j of the pass input: the caller's coordinates and the
// index in the first pass, the previous pass's output later.
template <typename T, int DIM, bool FIRST>
__device__ __forceinline__ void load_rec(long long j, const T *__restrict__ x,
                                         const T *__restrict__ y, const T *__restrict__ z,
                                         const SortRec<T, DIM> *__restrict__ rin,
                                         SortRec<T, DIM> &r)
{
    if constexpr (FIRST) {
        r.v[0] = x[j];
        if constexpr (DIM > 1) r.v[1] = y[j];
        if constexpr (DIM > 2) r.v[2] = z[j];
        r.i = (int)j;
    } else {
        r = rin[j];
    }
}

// Pass histogram: counts[d * tiles + t] = items of tile t whose digit is d.
template <typename T, int DIM, bool FINE, bool FIRST>
__global__ void __launch_bounds__(256)
k_sort_hist(int M, const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
            BinGeom g, const int *__restrict__ kin, int shift, int dbits,
            int *__restrict__ counts, int tiles)
{
    __shared__ int h[1 << kMaxDigitBits];
    const int nd = 1 << dbits;
    for (int d = threadIdx.x; d < nd; d += 256) h[d] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kSortTile;
    // all of a thread's loads in flight before the first wait, then the keys
    // and their counts
    constexpr int kPer = kSortTile / 256;
    if constexpr (FIRST) {
        T v[kPer][3];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const long long j = base + k * 256 + threadIdx.x;
            v[k][0] = v[k][1] = v[k][2] = T(0);
            if (j < M) {
                v[k][0] = x[j];
                if (DIM > 1) v[k][1] = y[j];
                if (DIM > 2) v[k][2] = z[j];
            }
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (base + k * 256 + threadIdx.x < M)
                atomicAdd(&h[(sort_key<T, DIM, FINE>(v[k], g) >> shift) & (nd - 1)], 1);
    } else {
        int key[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const long long j = base + k * 256 + threadIdx.x;
            key[k] = j < M ? kin[j] : 0;
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (base + k * 256 + threadIdx.x < M) atomicAdd(&h[(key[k] >> shift) & (nd - 1)], 1);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < nd; d += 256) counts[(long long)d * tiles + blockIdx.x] = h[d];
}

// Pass scatter.  counts holds the exclusive scan of the histogram, so
// counts[d * tiles + t] is where tile t's first item of digit d goes.  Wave w
// of the tile owns items [t * tile + w * 1024, + 1024) in 16 rounds of 64:
// first it counts its digits; the tile's digit totals are scanned into the
// tile-local start of every digit, and each wave's base within a digit
// follows in wave order; then each round ranks its lanes among equal digits
// with ballots (the lowest lane of a group advances the base), so equal keys
// keep item order.  The ranked items are first placed in LDS in their sorted
// order inside the tile (key and input position), and the tile is written
// out from there: consecutive threads take consecutive sorted items, whose
// destinations are consecutive within each digit's run, so the records of a
// run leave as whole cache lines instead of one line per lane.  LAST writes
// the sorted key, index and rescaled positions instead of the records.
template <typename T, int DIM, bool FINE, bool FIRST, bool LAST>
__global__ void __launch_bounds__(256)
k_sort_scatter(int M, const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
               BinGeom g, const int *__restrict__ kin, const SortRec<T, DIM> *__restrict__ rin,
               int shift, int dbits, const int *__restrict__ counts, int tiles,
               int *__restrict__ kout, SortRec<T, DIM> *__restrict__ rout,
               int *__restrict__ idx, double *__restrict__ pos, long long stride)
{
    constexpr int kMaxDigits = 1 << kMaxDigitBits;
    constexpr int kPer = kMaxDigits / 256;  // digits per thread in the tile scan
    __shared__ int wb[kSortWaves][kMaxDigits];
    __shared__ int dl[kMaxDigits];  // tile-local start of each digit's items
    __shared__ int gd[kMaxDigits];  // global start of them
    __shared__ int sk[kSortTile];   // the tile's keys in sorted order
    __shared__ int sj[kSortTile];   // and their input positions
    const int nd = 1 << dbits;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int d = threadIdx.x; d < nd; d += 256)
#pragma unroll
        for (int q = 0; q < kSortWaves; ++q) wb[q][d] = 0;
    __syncthreads();
    const long long tbase = (long long)blockIdx.x * kSortTile;
    const long long wbase = tbase + w * (kSortRounds * 64);
    int keys[kSortRounds];  // this lane's keys, kept for the ranking rounds
    // all of the lane's loads in flight before the first wait, then the keys
    // and their counts
    if constexpr (FIRST) {
        T v[kSortRounds][3];
#pragma unroll
        for (int r = 0; r < kSortRounds; ++r) {
            const long long j = wbase + r * 64 + lane;
            v[r][0] = v[r][1] = v[r][2] = T(0);
            if (j < M) {
                v[r][0] = x[j];
                if (DIM > 1) v[r][1] = y[j];
                if (DIM > 2) v[r][2] = z[j];
            }
        }
#pragma unroll
        for (int r = 0; r < kSortRounds; ++r) keys[r] = sort_key<T, DIM, FINE>(v[r], g);
    } else {
#pragma unroll
        for (int r = 0; r < kSortRounds; ++r) {
            const long long j = wbase + r * 64 + lane;
            keys[r] = j < M ? kin[j] : 0;
        }
    }
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        if (wbase + r * 64 + lane < M)
            atomicAdd(&wb[w][(keys[r] >> shift) & (nd - 1)], 1);
        else
            keys[r] = 0;  // items past M are never ranked or written
    }
    __syncthreads();
    // digit totals of the tile (thread t owns digits kPer t .. kPer t + kPer-1),
    // their exclusive scan, and each wave's tile-local base per digit
    int tot[kPer];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int d = threadIdx.x * kPer + k;
        tot[k] = 0;
        if (d < nd)
#pragma unroll
            for (int q = 0; q < kSortWaves; ++q) tot[k] += wb[q][d];
        sum += tot[k];
    }
    int2 ex;
    block_exscan2(make_int2(sum, 0), ex);
    int run = ex.x;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int d = threadIdx.x * kPer + k;
        if (d < nd) {
            dl[d] = run;
            gd[d] = counts[(long long)d * tiles + blockIdx.x];
            int b = run;
#pragma unroll
            for (int q = 0; q < kSortWaves; ++q) {
                const int cq = wb[q][d];
                wb[q][d] = b;
                b += cq;
            }
            run += tot[k];
        }
    }
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const long long j = wbase + r * 64 + lane;
        const bool act = j < M;
        const int d = act ? (keys[r] >> shift) & (nd - 1) : 0;
        unsigned long long peers = __ballot(act);
        for (int b = 0; b < dbits; ++b) {
            const bool bit = (d >> b) & 1;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        int lp = 0;
        if (act) lp = wb[w][d] + rank;
        if (act && rank == 0) wb[w][d] += __popcll(peers);
        if (act) {
            sk[lp] = keys[r];
            sj[lp] = (int)j;
        }
    }
    __syncthreads();
    // write-out in batches of kBatch items per thread: all of a batch's
    // record loads are issued before the first store, so their latencies
    // overlap instead of adding up item by item
    const int nt = (int)(M - tbase < kSortTile ? M - tbase : kSortTile);
    constexpr int kBatch = 8;
    for (int b0 = 0; b0 < kSortTile; b0 += 256 * kBatch) {
        SortRec<T, DIM> rec[kBatch];
        int key[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const int lp = b0 + 256 * k + threadIdx.x;
            key[k] = 0;
            if (lp < nt) {
                key[k] = sk[lp];
                load_rec<T, DIM, FIRST>(sj[lp], x, y, z, rin, rec[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            const int lp = b0 + 256 * k + threadIdx.x;
            if (lp < nt) {
                const int d = (key[k] >> shift) & (nd - 1);
                const int dst = gd[d] + (lp - dl[d]);
                kout[dst] = key[k];
                if constexpr (LAST) {
                    idx[dst] = rec[k].i;
#pragma unroll
                    for (int a = 0; a < DIM; ++a)
                        pos[a * stride + dst] =
                            tap_pos(rescale_exact(rec[k].v[a], g.nf[a]), g.nf[a]);
                } else {
                    rout[dst] = rec[k];
                }
            }
        }
    }
}

// binsize[b] = end - start of b's run in the sorted keys (binsize zeroed
// first; the first and last element of each run add -start and end + 1).
// Four keys per thread (one 16-byte load, skey 16-byte aligned) and their
// two neighbours; the division by the cells per bin only for (bin, cell)
// keys (DIV).  A bin of -1 marks "no neighbour".
template <bool DIV>
__global__ void __launch_bounds__(256)
k_run_sizes(int M, int vol, const int *__restrict__ skey, int *__restrict__ binsize)
{
    const long long n4 = ((long long)M + 3) / 4;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n4;
         q += (long long)gridDim.x * blockDim.x) {
        const int s0 = (int)(q * 4);
        int k[6];
        if (s0 + 3 < M) {
            const int4 v = reinterpret_cast<const int4 *>(skey)[q];
            k[1] = v.x;
            k[2] = v.y;
            k[3] = v.z;
            k[4] = v.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) k[1 + j] = s0 + j < M ? skey[s0 + j] : 0;
        }
        k[0] = s0 > 0 ? skey[s0 - 1] : 0;
        k[5] = s0 + 4 < M ? skey[s0 + 4] : 0;
        int b[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) b[j] = DIV ? k[j] / vol : k[j];
        if (s0 == 0) b[0] = -1;
#pragma unroll
        for (int j = 1; j <= 4; ++j) {
            const int s = s0 + j - 1;
            if (s >= M) break;
            if (s + 1 >= M) b[j + 1] = -1;
            if (b[j - 1] != b[j]) atomicAdd(&binsize[b[j]], -s);
            if (b[j + 1] != b[j]) atomicAdd(&binsize[b[j]], s + 1);
        }
    }
}

namespace {

// Exclusive scan of n ints in place (the int2 scan kernels with the second
// component unused).
void scan_inplace(int *a, long long n, int2 *part, const Launch &L)
{
    const long long nparts = (n + kScanChunk - 1) / kScanChunk;
    k_scan_reduce<<<(unsigned)nparts, kScanThreads, 0, L.stream>>>(a, n, 1, part);
    k_scan_parts<<<1, kScanThreads, 0, L.stream>>>(part, (int)nparts, nullptr, nullptr);
    k_scan_apply<<<(unsigned)nparts, kScanThreads, 0, L.stream>>>(a, n, 1, part, a, nullptr);
}

template <typename T, int DIM, bool FINE>
int sort_passes(int M, const T *x, const T *y, const T *z, const BinGeom &g,
                const BinSortShape &sh, SortBufs &sb, const Launch &L)
{
    using Rec = SortRec<T, DIM>;
    Rec *rec[2] = {reinterpret_cast<Rec *>(sb.rec[0]), reinterpret_cast<Rec *>(sb.rec[1])};
    int *keys[2] = {sb.skey, sb.skey_out};
    const long long st = sb.pos_stride;
    int2 *part = static_cast<int2 *>(sb.sort_part);
    const int tiles = sh.tiles;
    for (int p = 0; p < sh.passes; ++p) {
        const int shift = p * sh.digit_bits;
        const int dbits = std::min(sh.digit_bits, sh.bits - shift);
        const int in = (p + 1) & 1, out = p & 1;  // pass p writes buffer p % 2
        const bool first = p == 0, last = p + 1 == sh.passes;
        if (first)
            k_sort_hist<T, DIM, FINE, true><<<tiles, 256, 0, L.stream>>>(
                M, x, y, z, g, nullptr, shift, dbits, sb.sort_counts, tiles);
        else
            k_sort_hist<T, DIM, FINE, false><<<tiles, 256, 0, L.stream>>>(
                M, x, y, z, g, keys[in], shift, dbits, sb.sort_counts, tiles);
        scan_inplace(sb.sort_counts, (long long)tiles << dbits, part, L);
#define CFA_SCATTER(F, LST)                                                                   \
    k_sort_scatter<T, DIM, FINE, F, LST><<<tiles, 256, 0, L.stream>>>(                        \
        M, x, y, z, g, keys[in], rec[in], shift, dbits, sb.sort_counts, tiles, keys[out],     \
        rec[out], sb.idx, sb.pos, st)
        if (first && last)
            CFA_SCATTER(true, true);
        else if (first)
            CFA_SCATTER(true, false);
        else if (last)
            CFA_SCATTER(false, true);
        else
            CFA_SCATTER(false, false);
#undef CFA_SCATTER
    }
    return ok_or_err();
}

}  // namespace

template <typename T>
int run_bin_sort(int dim, const BinGeom &g, long long nbins, int M, const T *x, const T *y,
                 const T *z, int maxsub, bool fine, bool need_subprob, SortBufs &sb,
                 const Launch &L)
{
    const int vol = fine ? g.bs[0] * g.bs[1] * g.bs[2] : 1;
    k_fill_int<<<grid_pts(nbins, L.ncu), 256, 0, L.stream>>>(nbins, 0, sb.binsize);
    int e = 0;
    if (M > 0) {
        const BinSortShape sh = bin_sort_shape(M, nbins * vol);
#define CFA_PASSES(D)                                                                  \
    e = fine ? sort_passes<T, D, true>(M, x, y, z, g, sh, sb, L)                       \
             : sort_passes<T, D, false>(M, x, y, z, g, sh, sb, L)
        if (dim == 1)
            CFA_PASSES(1);
        else if (dim == 2)
            CFA_PASSES(2);
        else
            CFA_PASSES(3);
#undef CFA_PASSES
        if (e) return e;
        // the last pass wrote buffer (passes - 1) % 2
        const int *final_keys = ((sh.passes - 1) & 1) ? sb.skey_out : sb.skey;
        if (reinterpret_cast<uintptr_t>(final_keys) % 16 != 0) return kErrHip;
        const int gq = grid_pts(((long long)M + 3) / 4, L.ncu);
        if (vol > 1)
            k_run_sizes<true><<<gq, 256, 0, L.stream>>>(M, vol, final_keys, sb.binsize);
        else
            k_run_sizes<false><<<gq, 256, 0, L.stream>>>(M, vol, final_keys, sb.binsize);
    }
    const long long nparts = (nbins + kScanChunk - 1) / kScanChunk;
    int2 *part = static_cast<int2 *>(sb.scan_part);
    k_scan_reduce<<<(unsigned)nparts, kScanThreads, 0, L.stream>>>(sb.binsize, nbins, maxsub,
                                                                   part);
    k_scan_parts<<<1, kScanThreads, 0, L.stream>>>(part, (int)nparts, sb.counters,
                                                   sb.substart + nbins);
    k_scan_apply<<<(unsigned)nparts, kScanThreads, 0, L.stream>>>(sb.binsize, nbins, maxsub, part,
                                                                  sb.binstart, sb.substart);
    if (need_subprob)
        k_map_subprob<<<grid_pts(nbins, L.ncu), 256, 0, L.stream>>>(nbins, sb.substart, sb.s2b);
    return ok_or_err();
}

template int run_bin_sort<float>(int, const BinGeom &, long long, int, const float *,
                                 const float *, const float *, int, bool, bool, SortBufs &,
                                 const Launch &);
template int run_bin_sort<double>(int, const BinGeom &, long long, int, const double *,
                                  const double *, const double *, int, bool, bool, SortBufs &,
                                  const Launch &);

}  // namespace cfa
