// setpts bin sort by a stable radix sort of the points' bin keys.
//
// The reference sorts with an atomic histogram (one atomicAdd per point on
// its bin, whose old value is the point's in-bin rank, CalcBinSize_noghost_*,
// spreadinterp2d.cu:102-126), an exclusive scan and a scatter of the point
// indices (CalcInvertofGlobalSortIdx_*, spreadinterp2d.cu:128-149).  On
// clustered points the histogram atomics pile onto a few hot bins and
// serialise in L2 (measured: 11 ms for the 1e8 points of the 3D type-1
// workload, against ~0.3 ms of memory traffic), and the scatter writes four
// values per point to random addresses.  Here the keys (bin, or bin and fine
// cell for gpu_sort = 2) are sorted with the point index as payload by the
// hipCUB device radix sort, which has no same-address contention whatever
// the distribution; per-bin counts come from the run boundaries of the sorted
// keys (two atomics per occupied bin), the bin starts and subproblem map from
// the same scans as before, and the positions are gathered in sorted order so
// that their writes are coalesced; the key pass packs each point's
// coordinates into one record, so that the gather reads one cache line per
// point rather than one per axis.  Bins, bin starts and subproblems are the
// reference's exactly (the same key of every point, counted); inside a bin
// the points are in input order (stable sort) instead of atomic arrival order.
#include <hipcub/hipcub.hpp>

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

int key_bits(long long nkeys)
{
    int b = 1;
    while ((1ll << b) < nkeys) ++b;
    return b;
}
}  // namespace

// Coordinates of one point packed into one aligned record (x, y[, z, pad]),
// so that the sorted-order gather reads one cache line per point instead of
// one per axis.
template <typename T, int DIM> struct alignas(sizeof(T) * (DIM == 3 ? 4 : 2)) PackRec {
    T v[DIM == 3 ? 4 : 2];
};

// Bytes of one packed record (0 in 1D, where the gather reads x alone).
size_t pack_rec_bytes(int dim, int elem)
{
    return dim == 1 ? 0 : (size_t)elem * (dim == 3 ? 4 : 2);
}

// key[i] = bin of point i (FINE: bin * vol + cell in bin), val[i] = i, and
// (2D/3D) the point's coordinates packed in input order.
template <typename T, int DIM, bool FINE>
__global__ void __launch_bounds__(256)
k_sort_keys(int M, const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
            BinGeom g, int *__restrict__ key, int *__restrict__ val,
            PackRec<T, DIM> *__restrict__ pack)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        const T xi = x[i];
        const T yi = DIM > 1 ? y[i] : T(0);
        const T zi = DIM > 2 ? z[i] : T(0);
        int k;
        if constexpr (FINE) {
            int b0, c0, b1 = 0, c1 = 0, b2 = 0, c2 = 0;
            bin_and_cell<T>(rescale_exact(xi, g.nf[0]), g.bs[0], g.nb[0], b0, c0);
            if (DIM > 1) bin_and_cell<T>(rescale_exact(yi, g.nf[1]), g.bs[1], g.nb[1], b1, c1);
            if (DIM > 2) bin_and_cell<T>(rescale_exact(zi, g.nf[2]), g.bs[2], g.nb[2], b2, c2);
            const int bin = b0 + g.nb[0] * (b1 + g.nb[1] * b2);
            const int vol = g.bs[0] * g.bs[1] * g.bs[2];
            k = bin * vol + c0 + g.bs[0] * (c1 + g.bs[1] * c2);
        } else {
            k = bin1<T>(rescale_exact(xi, g.nf[0]), g.bs[0], g.nb[0]);
            if (DIM > 1) k += g.nb[0] * bin1<T>(rescale_exact(yi, g.nf[1]), g.bs[1], g.nb[1]);
            if (DIM > 2)
                k += g.nb[0] * g.nb[1] * bin1<T>(rescale_exact(zi, g.nf[2]), g.bs[2], g.nb[2]);
        }
        key[i] = k;
        val[i] = i;
        if constexpr (DIM > 1) {
            PackRec<T, DIM> r;
            r.v[0] = xi;
            r.v[1] = yi;
            if constexpr (DIM > 2) {
                r.v[2] = zi;
                r.v[3] = T(0);
            }
            pack[i] = r;
        }
    }
}

// binsize[b] = end - start of b's run in the sorted keys (binsize zeroed
// first; the first and last element of each run add -start and end + 1).
__global__ void __launch_bounds__(256)
k_run_sizes(int M, int vol, const int *__restrict__ skey, int *__restrict__ binsize)
{
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < M; s += gridDim.x * blockDim.x) {
        const int b = skey[s] / vol;
        if (s == 0 || skey[s - 1] / vol != b) atomicAdd(&binsize[b], -s);
        if (s == M - 1 || skey[s + 1] / vol != b) atomicAdd(&binsize[b], s + 1);
    }
}

// pos[d*stride + s] = rescaled coordinate of sorted point s (coalesced
// writes; one random record read per point in 2D/3D, the x array in 1D).
template <typename T, int DIM>
__global__ void __launch_bounds__(256)
k_gather_pos(int M, const T *__restrict__ x, const PackRec<T, DIM> *__restrict__ pack,
             BinGeom g, const int *__restrict__ idx, double *__restrict__ pos, long long stride)
{
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < M; s += gridDim.x * blockDim.x) {
        const int i = idx[s];
        if constexpr (DIM == 1) {
            pos[s] = tap_pos(rescale_exact(x[i], g.nf[0]), g.nf[0]);
        } else {
            const PackRec<T, DIM> r = pack[i];
            pos[s] = tap_pos(rescale_exact(r.v[0], g.nf[0]), g.nf[0]);
            pos[stride + s] = tap_pos(rescale_exact(r.v[1], g.nf[1]), g.nf[1]);
            if constexpr (DIM > 2)
                pos[2 * stride + s] = tap_pos(rescale_exact(r.v[2], g.nf[2]), g.nf[2]);
        }
    }
}

size_t radix_tmp_bytes(int M, long long nkeys)
{
    size_t bytes = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const int *)nullptr, (int *)nullptr,
                                       (const int *)nullptr, (int *)nullptr, M, 0,
                                       key_bits(nkeys));
    return bytes;
}

template <typename T>
int run_radix_sort(int dim, const BinGeom &g, long long nbins, int M, const T *x, const T *y,
                   const T *z, int maxsub, bool fine, bool need_subprob, SortBufs &sb,
                   const Launch &L)
{
    const int vol = fine ? g.bs[0] * g.bs[1] * g.bs[2] : 1;
    k_fill_int<<<grid_pts(nbins, L.ncu), 256, 0, L.stream>>>(nbins, 0, sb.binsize);
    if (M > 0) {
        const int gp = grid_pts(M, L.ncu);
#define CFA_KEYS(D)                                                                      \
    do {                                                                                 \
        auto *pk = reinterpret_cast<PackRec<T, D> *>(sb.pack);                           \
        if (fine)                                                                        \
            k_sort_keys<T, D, true>                                                      \
                <<<gp, 256, 0, L.stream>>>(M, x, y, z, g, sb.skey, sb.sval, pk);         \
        else                                                                             \
            k_sort_keys<T, D, false>                                                     \
                <<<gp, 256, 0, L.stream>>>(M, x, y, z, g, sb.skey, sb.sval, pk);         \
    } while (0)
        if (dim == 1)
            CFA_KEYS(1);
        else if (dim == 2)
            CFA_KEYS(2);
        else
            CFA_KEYS(3);
#undef CFA_KEYS
        size_t bytes = sb.sort_tmp_bytes;
        if (hipcub::DeviceRadixSort::SortPairs(sb.sort_tmp, bytes, sb.skey, sb.skey_out, sb.sval,
                                               sb.idx, M, 0, key_bits(nbins * vol),
                                               L.stream) != hipSuccess)
            return kErrHip;
        k_run_sizes<<<gp, 256, 0, L.stream>>>(M, vol, sb.skey_out, sb.binsize);
    }
    const long long nparts = (nbins + kScanChunk - 1) / kScanChunk;
    int2 *part = static_cast<int2 *>(sb.scan_part);
    k_scan_reduce<<<(unsigned)nparts, kScanThreads, 0, L.stream>>>(sb.binsize, nbins, maxsub,
                                                                   part);
    k_scan_parts<<<1, kScanThreads, 0, L.stream>>>(part, (int)nparts, sb.counters,
                                                   sb.substart + nbins);
    k_scan_apply<<<(unsigned)nparts, kScanThreads, 0, L.stream>>>(sb.binsize, nbins, maxsub, part,
                                                                  sb.binstart, sb.substart);
    if (M > 0) {
        const int gp = grid_pts(M, L.ncu);
        const long long st = sb.pos_stride;
        if (dim == 1)
            k_gather_pos<T, 1><<<gp, 256, 0, L.stream>>>(
                M, x, reinterpret_cast<const PackRec<T, 1> *>(sb.pack), g, sb.idx, sb.pos, st);
        else if (dim == 2)
            k_gather_pos<T, 2><<<gp, 256, 0, L.stream>>>(
                M, x, reinterpret_cast<const PackRec<T, 2> *>(sb.pack), g, sb.idx, sb.pos, st);
        else
            k_gather_pos<T, 3><<<gp, 256, 0, L.stream>>>(
                M, x, reinterpret_cast<const PackRec<T, 3> *>(sb.pack), g, sb.idx, sb.pos, st);
    }
    if (need_subprob)
        k_map_subprob<<<grid_pts(nbins, L.ncu), 256, 0, L.stream>>>(nbins, sb.substart, sb.s2b);
    return ok_or_err();
}

template int run_radix_sort<float>(int, const BinGeom &, long long, int, const float *,
                                   const float *, const float *, int, bool, bool, SortBufs &,
                                   const Launch &);
template int run_radix_sort<double>(int, const BinGeom &, long long, int, const double *,
                                    const double *, const double *, int, bool, bool, SortBufs &,
                                    const Launch &);

}  // namespace cfa
