// Runtime-ns dispatch of the spread/interp kernels for one (T, DIM); included
// by the spread_<prec><dim>.hip and interp_<prec><dim>.hip translation units.
#include "nufft_device.cuh"
#include "nufft_launch.h"

namespace cfa {

namespace {

// Point-driven kernels: enough 256-point chunks to fill the device a few
// times over, rounded to whole XCD groups.
inline int grid_for_points(int M, int ncu)
{
    long long chunks = (M + kPtChunk - 1) / kPtChunk;
    long long cap = (long long)ncu * 16;
    long long g = chunks < cap ? chunks : cap;
    if (g >= kNumXcd) g -= g % kNumXcd;
    return (int)(g < 1 ? 1 : g);
}

// Persistent grid for a subproblem kernel: as many workgroups as fit on the
// device at the tile's LDS footprint, never more than there are jobs, rounded
// to whole XCD groups.  The LDS attribute and occupancy are cached per kernel
// (the template is instantiated once per kernel, so kernels of one signature
// keep separate caches) and LDS size, so launches after the first make no
// runtime queries (which keeps them legal inside a stream capture).
template <auto Kern>
int sm_grid(size_t lds, long long jobs, int ncu, int *grid, int threads = 256)
{
    const auto kern = Kern;
    static thread_local size_t attr_lds = 64 * 1024, occ_lds = 0;
    static thread_local int occ_cached = 0;
    if (lds > attr_lds) {
        if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess)
            return kErrLds;
        attr_lds = lds;
    }
    int occ = occ_cached;
    if (lds != occ_lds || occ < 1) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, threads, lds) != hipSuccess ||
            occ < 1)
            return kErrLds;
        occ_cached = occ;
        occ_lds = lds;
    }
    long long g = (long long)occ * ncu;
    if (g > jobs) g = jobs;
    if (g >= kNumXcd) g -= g % kNumXcd;
    *grid = (int)(g < 1 ? 1 : g);
    return 0;
}

inline SortedPts sorted_pts(const SortBufs &sb) { return SortedPts{sb.idx, sb.pos, sb.pos_stride}; }

template <typename T, int DIM, int NS>
int spread_ns(int method, int M, const SortBufs &sb, const cpx<T> *c, cpx<T> *fw,
              const BinGeom &g, const KerParams<T> &kp, int maxsub, long long sub_bound, int ntr,
              long long nftot, const Launch &L)
{
    if (method == 1) {
        k_spread_gm<T, DIM, NS><<<grid_for_points(M, L.ncu), 256, 0, L.stream>>>(
            M, sorted_pts(sb), c, fw, g, kp, ntr, nftot);
    } else if (method == 4) {
        const size_t lds = blk_lds_bytes(DIM, g.bs, sizeof(T), maxsub, kp.fxwide);
        if (lds > kMaxLdsBytes) return kErrLds;
        int grid = 0;
        int e = sm_grid<k_spread_blk<T, DIM, NS>>(lds, sub_bound * ntr, L.ncu, &grid);
        if (e) return e;
        SubprobArgs sa{sb.s2b, sb.substart, sb.binstart, sb.binsize, sb.counters, maxsub};
        k_spread_blk<T, DIM, NS><<<grid, 256, lds, L.stream>>>(M, sorted_pts(sb), c, fw, g, kp,
                                                               sa, ntr, nftot);
    } else {
        const SpreadCfg cfg = spread_cfg(DIM, NS, g.bs, sizeof(T), kp.fx32max, maxsub, false);
        if (cfg.lds > kMaxLdsBytes) return kErrLds;
        int grid = 0;
        SubprobArgs sa{sb.s2b, sb.substart, sb.binstart, sb.binsize, sb.counters, maxsub};
#define CFA_SM_SPREAD(ACC, BT)                                                                \
    do {                                                                                      \
        int e = sm_grid<k_spread_sm<T, ACC, DIM, NS, BT>>(cfg.lds, sub_bound * ntr, L.ncu,    \
                                                           &grid, BT);                        \
        if (e) return e;                                                                      \
        k_spread_sm<T, ACC, DIM, NS, BT><<<grid, BT, cfg.lds, L.stream>>>(                    \
            M, sorted_pts(sb), c, fw, g, kp, sa, ntr, nftot, sb.jobq, slot);                  \
    } while (0)
        cpx<T> *slot = static_cast<cpx<T> *>(sb.slot);
        if constexpr (sizeof(T) == 4) {
            if (cfg.acc_bytes == 4)
                CFA_SM_SPREAD(int, 512);
            else if (cfg.bt == 1024)
                CFA_SM_SPREAD(long long, 1024);
            else
                CFA_SM_SPREAD(long long, 256);
        } else {
            if (cfg.bt == 1024)
                CFA_SM_SPREAD(long long, 1024);
            else
                CFA_SM_SPREAD(long long, 256);
        }
#undef CFA_SM_SPREAD
        if (slot) {
            // add the stored tiles of the single-subproblem bins: one
            // one-wave workgroup per grid row and transform (measured on the
            // 3D type-1 workload at 512^3: 64 threads 11.6 ms for spread and
            // gather, 128 11.7, 256 12.1, 512 13.2)
            const long long rows = nftot / g.nf[0] * ntr;
            k_tile_gather<T, DIM><<<(unsigned)rows, 64, 0, L.stream>>>(
                fw, slot, g, (NS + 1) / 2, sb.substart, nftot);
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : kErrHip;
}

template <typename T, int DIM, int NS>
int interp_ns(int method, int M, const SortBufs &sb, cpx<T> *c, const cpx<T> *fw,
              const BinGeom &g, const KerParams<T> &kp, int maxsub, long long sub_bound, int ntr,
              long long nftot, const Launch &L)
{
    if (method == 1) {
        k_interp_gm<T, DIM, NS><<<grid_for_points(M, L.ncu), 256, 0, L.stream>>>(
            M, sorted_pts(sb), c, fw, g, kp, ntr, nftot);
    } else {
        const size_t lds = sm_tile_bytes(DIM, NS, g.bs, sizeof(T));
        if (lds > kMaxLdsBytes) return kErrLds;
        int grid = 0;
        SubprobArgs sa{sb.s2b, sb.substart, sb.binstart, sb.binsize, sb.counters, maxsub};
        if (maxsub >= 1024) {
            int e = sm_grid<k_interp_sm<T, DIM, NS, 1024>>(lds, sub_bound * ntr, L.ncu, &grid,
                                                           1024);
            if (e) return e;
            k_interp_sm<T, DIM, NS, 1024><<<grid, 1024, lds, L.stream>>>(
                M, sorted_pts(sb), c, fw, g, kp, sa, ntr, nftot, sb.jobq);
        } else if (maxsub >= 512) {
            int e = sm_grid<k_interp_sm<T, DIM, NS, 512>>(lds, sub_bound * ntr, L.ncu, &grid, 512);
            if (e) return e;
            k_interp_sm<T, DIM, NS, 512><<<grid, 512, lds, L.stream>>>(
                M, sorted_pts(sb), c, fw, g, kp, sa, ntr, nftot, sb.jobq);
        } else {
            int e = sm_grid<k_interp_sm<T, DIM, NS, 256>>(lds, sub_bound * ntr, L.ncu, &grid);
            if (e) return e;
            k_interp_sm<T, DIM, NS, 256><<<grid, 256, lds, L.stream>>>(
                M, sorted_pts(sb), c, fw, g, kp, sa, ntr, nftot, sb.jobq);
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : kErrHip;
}

template <typename T, int DIM, int NS>
int group_ns(int type, const GroupItem<T> *items, int n, const long long *off,
             const KerParams<T> &kp, const int *bs, int maxsub, long long jobs_bound,
             const Launch &L)
{
    // the job queue words follow the n + 1 offsets in the same buffer
    unsigned long long *jq =
        reinterpret_cast<unsigned long long *>(const_cast<long long *>(off) + n + 1);
    if (type == 1) {
        const SpreadCfg cfg = spread_cfg(DIM, NS, bs, sizeof(T), kp.fx32max, maxsub, true);
        if (cfg.lds > kMaxLdsBytes) return kErrLds;
        int grid = 0;
        if (sizeof(T) == 4 && cfg.acc_bytes == 4) {
            int e = sm_grid<k_spread_sm_group<T, int, DIM, NS>>(cfg.lds, jobs_bound, L.ncu, &grid);
            if (e) return e;
            k_spread_sm_group<T, int, DIM, NS><<<grid, 256, cfg.lds, L.stream>>>(items, n, off, kp,
                                                                                jq);
        } else {
            int e = sm_grid<k_spread_sm_group<T, long long, DIM, NS>>(cfg.lds, jobs_bound, L.ncu,
                                                                      &grid);
            if (e) return e;
            k_spread_sm_group<T, long long, DIM, NS><<<grid, 256, cfg.lds, L.stream>>>(
                items, n, off, kp, jq);
        }
    } else {
        const size_t lds = sm_tile_bytes(DIM, NS, bs, sizeof(T));
        if (lds > kMaxLdsBytes) return kErrLds;
        int grid = 0;
        int e = sm_grid<k_interp_sm_group<T, DIM, NS>>(lds, jobs_bound, L.ncu, &grid);
        if (e) return e;
        k_interp_sm_group<T, DIM, NS><<<grid, 256, lds, L.stream>>>(items, n, off, kp, jq);
    }
    return hipGetLastError() == hipSuccess ? 0 : kErrHip;
}

}  // namespace

#ifndef CFA_NS_CASES
#define CFA_NS_CASES(F)                                                                      \
    F(2) F(3) F(4) F(5) F(6) F(7) F(8) F(9) F(10) F(11) F(12) F(13) F(14) F(15) F(16)
#endif

template <typename T, int DIM>
int spread_dim(int method, int ns, int M, const SortBufs &sb, const cpx<T> *c, cpx<T> *fw,
               const BinGeom &g, const KerParams<T> &kp, int maxsub, long long sub_bound,
               int ntr, long long nftot, const Launch &L)
{
    switch (ns) {
#define CFA_CASE(N)                                                                          \
    case N:                                                                                  \
        return spread_ns<T, DIM, N>(method, M, sb, c, fw, g, kp, maxsub, sub_bound, ntr, nftot, \
                                    L);
        CFA_NS_CASES(CFA_CASE)
#undef CFA_CASE
    }
    return 1;
}

template <typename T, int DIM>
int interp_dim(int method, int ns, int M, const SortBufs &sb, cpx<T> *c, const cpx<T> *fw,
               const BinGeom &g, const KerParams<T> &kp, int maxsub, long long sub_bound,
               int ntr, long long nftot, const Launch &L)
{
    switch (ns) {
#define CFA_CASE(N)                                                                          \
    case N:                                                                                  \
        return interp_ns<T, DIM, N>(method, M, sb, c, fw, g, kp, maxsub, sub_bound, ntr, nftot, \
                                    L);
        CFA_NS_CASES(CFA_CASE)
#undef CFA_CASE
    }
    return 1;
}

template <typename T, int DIM>
int group_dim(int type, int ns, const void *items, int n, const long long *off,
              const KerParams<T> &kp, const int *bs, int maxsub, long long jobs_bound,
              const Launch &L)
{
    const GroupItem<T> *it = static_cast<const GroupItem<T> *>(items);
    k_group_offsets<T><<<1, 64, 0, L.stream>>>(it, n, const_cast<long long *>(off));
    if (hipGetLastError() != hipSuccess) return kErrHip;
    switch (ns) {
#define CFA_CASE(N)                                                                          \
    case N:                                                                                  \
        return group_ns<T, DIM, N>(type, it, n, off, kp, bs, maxsub, jobs_bound, L);
        CFA_NS_CASES(CFA_CASE)
#undef CFA_CASE
    }
    return 1;
}

template <typename T, int DIM>
void group_item_fill(void *dst, const BinGeom &g, const SortBufs &sb, int maxsub, const void *in,
                     void *out, long long nftot, int M, int ntr)
{
    GroupItem<T> &it = *static_cast<GroupItem<T> *>(dst);
    it.g = g;
    it.sp = SortedPts{sb.idx, sb.pos, sb.pos_stride};
    it.sa = SubprobArgs{sb.s2b, sb.substart, sb.binstart, sb.binsize, sb.counters, maxsub};
    it.in = static_cast<const cpx<T> *>(in);
    it.out = static_cast<cpx<T> *>(out);
    it.nftot = nftot;
    it.M = M;
    it.ntr = ntr;
}

}  // namespace cfa

// Spread side (spread kernels and plan groups) and interp side instantiate in separate translation units so
// the two halves of one precision and dimension compile in parallel.
#define CFA_INSTANTIATE_SPREAD(T, DIM)                                                        \
    template int cfa::group_dim<T, DIM>(int, int, const void *, int, const long long *,      \
                                        const cfa::KerParams<T> &, const int *, int,         \
                                        long long, const cfa::Launch &);                     \
    template void cfa::group_item_fill<T, DIM>(void *, const cfa::BinGeom &,                 \
                                               const cfa::SortBufs &, int, const void *,     \
                                               void *, long long, int, int);                 \
    template int cfa::spread_dim<T, DIM>(int, int, int, const cfa::SortBufs &,               \
                                         const cfa::cpx<T> *, cfa::cpx<T> *,                 \
                                         const cfa::BinGeom &, const cfa::KerParams<T> &,    \
                                         int, long long, int, long long, const cfa::Launch &);

#define CFA_INSTANTIATE_INTERP(T, DIM)                                                        \
    template int cfa::interp_dim<T, DIM>(int, int, int, const cfa::SortBufs &, cfa::cpx<T> *, \
                                         const cfa::cpx<T> *, const cfa::BinGeom &,          \
                                         const cfa::KerParams<T> &, int, long long, int,     \
                                         long long, const cfa::Launch &);
