// Runtime-ns dispatch of the spread/interp kernels for one (T, DIM); included
// by the spreadinterp_<prec><dim>.hip translation units.
#include "nufft_device.cuh"
#include "nufft_launch.h"

namespace cfa {

namespace {

inline int grid_for_points(int M, int ncu)
{
    long long b = (M + 255) / 256;
    long long cap = (long long)ncu * 32;
    return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}

// Persistent grid for a subproblem kernel: as many workgroups as fit on the
// device at the tile's LDS footprint, never more than there are jobs.
template <typename K>
int sm_grid(K kern, size_t lds, long long jobs, int ncu, int *grid)
{
    static_assert(sizeof(K) > 0, "");
    if (lds > 64 * 1024) {
        if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess)
            return kErrLds;
    }
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, 256, lds) != hipSuccess ||
        occ < 1)
        return kErrLds;
    long long g = (long long)occ * ncu;
    if (g > jobs) g = jobs;
    *grid = (int)(g < 1 ? 1 : g);
    return 0;
}

template <typename T, int DIM, int NS>
int spread_ns(int method, int M, const T *x, const T *y, const T *z, const cpx<T> *c, cpx<T> *fw,
              const BinGeom &g, const KerParams<T> &kp, const SortBufs &sb, int maxsub,
              long long sub_bound, int ntr, long long nftot, const Launch &L)
{
    if (method == 1) {
        k_spread_gm<T, DIM, NS><<<grid_for_points(M, L.ncu), 256, 0, L.stream>>>(
            M, x, y, z, sb.idx, c, fw, g, kp, ntr, nftot);
    } else {
        const size_t lds = sm_tile_bytes(DIM, NS, g.bs, sizeof(T));
        if (lds > kMaxLdsBytes) return kErrLds;
        int grid = 0;
        int e = sm_grid(k_spread_sm<T, DIM, NS>, lds, sub_bound * ntr, L.ncu, &grid);
        if (e) return e;
        SubprobArgs sa{sb.s2b, sb.substart, sb.binstart, sb.binsize, sb.counters, maxsub};
        k_spread_sm<T, DIM, NS><<<grid, 256, lds, L.stream>>>(M, x, y, z, sb.idx, c, fw, g, kp,
                                                              sa, ntr, nftot);
    }
    return hipGetLastError() == hipSuccess ? 0 : kErrHip;
}

template <typename T, int DIM, int NS>
int interp_ns(int method, int M, const T *x, const T *y, const T *z, cpx<T> *c,
              const cpx<T> *fw, const BinGeom &g, const KerParams<T> &kp, const SortBufs &sb,
              int maxsub, long long sub_bound, int ntr, long long nftot, const Launch &L)
{
    if (method == 1) {
        k_interp_gm<T, DIM, NS><<<grid_for_points(M, L.ncu), 256, 0, L.stream>>>(
            M, x, y, z, sb.idx, c, fw, g, kp, ntr, nftot);
    } else {
        const size_t lds = sm_tile_bytes(DIM, NS, g.bs, sizeof(T));
        if (lds > kMaxLdsBytes) return kErrLds;
        int grid = 0;
        int e = sm_grid(k_interp_sm<T, DIM, NS>, lds, sub_bound * ntr, L.ncu, &grid);
        if (e) return e;
        SubprobArgs sa{sb.s2b, sb.substart, sb.binstart, sb.binsize, sb.counters, maxsub};
        k_interp_sm<T, DIM, NS><<<grid, 256, lds, L.stream>>>(M, x, y, z, sb.idx, c, fw, g, kp,
                                                              sa, ntr, nftot);
    }
    return hipGetLastError() == hipSuccess ? 0 : kErrHip;
}

}  // namespace

#define CFA_NS_CASES(F)                                                                      \
    F(2) F(3) F(4) F(5) F(6) F(7) F(8) F(9) F(10) F(11) F(12) F(13) F(14) F(15) F(16)

template <typename T, int DIM>
int spread_dim(int method, int ns, int M, const T *x, const T *y, const T *z, const cpx<T> *c,
               cpx<T> *fw, const BinGeom &g, const KerParams<T> &kp, const SortBufs &sb,
               int maxsub, long long sub_bound, int ntr, long long nftot, const Launch &L)
{
    switch (ns) {
#define CFA_CASE(N)                                                                          \
    case N:                                                                                  \
        return spread_ns<T, DIM, N>(method, M, x, y, z, c, fw, g, kp, sb, maxsub, sub_bound, \
                                    ntr, nftot, L);
        CFA_NS_CASES(CFA_CASE)
#undef CFA_CASE
    }
    return 1;
}

template <typename T, int DIM>
int interp_dim(int method, int ns, int M, const T *x, const T *y, const T *z, cpx<T> *c,
               const cpx<T> *fw, const BinGeom &g, const KerParams<T> &kp, const SortBufs &sb,
               int maxsub, long long sub_bound, int ntr, long long nftot, const Launch &L)
{
    switch (ns) {
#define CFA_CASE(N)                                                                          \
    case N:                                                                                  \
        return interp_ns<T, DIM, N>(method, M, x, y, z, c, fw, g, kp, sb, maxsub, sub_bound, \
                                    ntr, nftot, L);
        CFA_NS_CASES(CFA_CASE)
#undef CFA_CASE
    }
    return 1;
}

}  // namespace cfa

#define CFA_INSTANTIATE_SPREADINTERP(T, DIM)                                                  \
    template int cfa::spread_dim<T, DIM>(int, int, int, const T *, const T *, const T *,     \
                                         const cfa::cpx<T> *, cfa::cpx<T> *,                 \
                                         const cfa::BinGeom &, const cfa::KerParams<T> &,    \
                                         const cfa::SortBufs &, int, long long, int,         \
                                         long long, const cfa::Launch &);                    \
    template int cfa::interp_dim<T, DIM>(int, int, int, const T *, const T *, const T *,     \
                                         cfa::cpx<T> *, const cfa::cpx<T> *,                 \
                                         const cfa::BinGeom &, const cfa::KerParams<T> &,    \
                                         const cfa::SortBufs &, int, long long, int,         \
                                         long long, const cfa::Launch &);
