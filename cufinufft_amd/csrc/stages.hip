// Launchers of the bin sort, the dimension dispatch of spread/interp, and the
// deconvolve / amplify / zero-fill passes.
#include "nufft_device.cuh"
#include "nufft_launch.h"

namespace cfa {

namespace {
inline int ok_or_err() { return hipGetLastError() == hipSuccess ? 0 : kErrHip; }

inline int grid_1d(long long n, int ncu, int per_cu = 32)
{
    long long b = (n + 255) / 256;
    long long cap = (long long)ncu * per_cu;
    return (int)(b < 1 ? 1 : (b > cap ? cap : b));
}
}  // namespace

size_t group_item_bytes(bool dbl)
{
    return dbl ? sizeof(GroupItem<double>) : sizeof(GroupItem<float>);
}

size_t blk_lds_bytes(int dim, const int *ob, size_t elem, int maxsub, int fxwide)
{
    size_t cells = (size_t)ob[0];
    if (dim > 1) cells *= (size_t)ob[1];
    if (dim > 2) cells *= (size_t)ob[2];
    (void)maxsub;
    const int acc = elem == 4 ? (fxwide ? 8 : 4) : 8;
    return blk_lds(cells, acc);
}

// Method 4, first half of setpts: count the copies of every point per
// block and scan the counts; the total lands in counters[0] for the host to
// size the copy arrays.
template <typename T>
int run_blk_count(int dim, const BinGeom &g, int ns, int M, const T *x, const T *y, const T *z,
                  SortBufs &sb, const Launch &L)
{
    const long long nkeys = (long long)g.nb[0] * g.nb[1] * g.nb[2];
    k_fill_int<<<grid_1d(nkeys, L.ncu, 64), 256, 0, L.stream>>>(nkeys, 0, sb.kcount);
    if (M > 0) {
        const int gp = grid_1d(M, L.ncu);
        if (dim == 1)
            k_blk_count<T, 1><<<gp, 256, 0, L.stream>>>(M, x, y, z, g, ns, sb.kcount);
        else if (dim == 2)
            k_blk_count<T, 2><<<gp, 256, 0, L.stream>>>(M, x, y, z, g, ns, sb.kcount);
        else
            k_blk_count<T, 3><<<gp, 256, 0, L.stream>>>(M, x, y, z, g, ns, sb.kcount);
    }
    const long long kparts = (nkeys + kScanChunk - 1) / kScanChunk;
    int2 *kp = static_cast<int2 *>(sb.kscan_part);
    k_scan_reduce<<<(unsigned)kparts, kScanThreads, 0, L.stream>>>(sb.kcount, nkeys, 1, kp);
    k_scan_parts<<<1, kScanThreads, 0, L.stream>>>(kp, (int)kparts, sb.counters, sb.counters + 1);
    k_scan_apply<<<(unsigned)kparts, kScanThreads, 0, L.stream>>>(sb.kcount, nkeys, 1, kp,
                                                                  sb.kstart, nullptr);
    return ok_or_err();
}

// Method 4, second half: scatter the ncopies copies (index and shifted
// positions), derive per-block sizes / starts and the subproblem map.
template <typename T>
int run_blk_scatter(int dim, const BinGeom &g, int ns, int M, int ncopies, const T *x,
                    const T *y, const T *z, int maxsub, SortBufs &sb, const Launch &L)
{
    const long long nblocks = (long long)g.nb[0] * g.nb[1] * g.nb[2];
    if (M > 0) {
        const int gp = grid_1d(M, L.ncu);
        const long long st = sb.pos_stride;
        if (dim == 1)
            k_blk_scatter<T, 1><<<gp, 256, 0, L.stream>>>(M, x, y, z, g, ns, sb.kstart, sb.kcount,
                                                          sb.idx, sb.pos, st);
        else if (dim == 2)
            k_blk_scatter<T, 2><<<gp, 256, 0, L.stream>>>(M, x, y, z, g, ns, sb.kstart, sb.kcount,
                                                          sb.idx, sb.pos, st);
        else
            k_blk_scatter<T, 3><<<gp, 256, 0, L.stream>>>(M, x, y, z, g, ns, sb.kstart, sb.kcount,
                                                          sb.idx, sb.pos, st);
    }
    k_bins_from_keys<<<grid_1d(nblocks, L.ncu), 256, 0, L.stream>>>(nblocks, 1, ncopies,
                                                                    sb.kstart, sb.binsize);
    const long long nparts = (nblocks + kScanChunk - 1) / kScanChunk;
    int2 *part = static_cast<int2 *>(sb.scan_part);
    k_scan_reduce<<<(unsigned)nparts, kScanThreads, 0, L.stream>>>(sb.binsize, nblocks, maxsub,
                                                                   part);
    k_scan_parts<<<1, kScanThreads, 0, L.stream>>>(part, (int)nparts, sb.counters,
                                                   sb.substart + nblocks);
    k_scan_apply<<<(unsigned)nparts, kScanThreads, 0, L.stream>>>(sb.binsize, nblocks, maxsub,
                                                                  part, sb.binstart, sb.substart);
    k_map_subprob<<<grid_1d(nblocks, L.ncu), 256, 0, L.stream>>>(nblocks, sb.substart, sb.s2b);
    return ok_or_err();
}

template <typename T>
int run_natural_order(int dim, const BinGeom &g, int M, const T *x, const T *y, const T *z,
                      SortBufs &sb, const Launch &L)
{
    if (M == 0) return 0;
    const int gp = grid_1d(M, L.ncu);
    const long long st = sb.pos_stride;
    if (dim == 1)
        k_natural_order<T, 1><<<gp, 256, 0, L.stream>>>(M, x, y, z, g, sb.idx, sb.pos, st);
    else if (dim == 2)
        k_natural_order<T, 2><<<gp, 256, 0, L.stream>>>(M, x, y, z, g, sb.idx, sb.pos, st);
    else
        k_natural_order<T, 3><<<gp, 256, 0, L.stream>>>(M, x, y, z, g, sb.idx, sb.pos, st);
    return ok_or_err();
}

template <typename T>
int run_spread(int method, int dim, int ns, int M, const SortBufs &sb, const cpx<T> *c,
               cpx<T> *fw, const BinGeom &g, const KerParams<T> &kp, int maxsub,
               long long sub_bound, int ntr, long long nftot, const Launch &L)
{
    if (M == 0) return 0;
    if (dim == 1)
        return spread_dim<T, 1>(method, ns, M, sb, c, fw, g, kp, maxsub, sub_bound, ntr, nftot, L);
    if (dim == 2)
        return spread_dim<T, 2>(method, ns, M, sb, c, fw, g, kp, maxsub, sub_bound, ntr, nftot, L);
    return spread_dim<T, 3>(method, ns, M, sb, c, fw, g, kp, maxsub, sub_bound, ntr, nftot, L);
}

template <typename T>
int run_interp(int method, int dim, int ns, int M, const SortBufs &sb, cpx<T> *c,
               const cpx<T> *fw, const BinGeom &g, const KerParams<T> &kp, int maxsub,
               long long sub_bound, int ntr, long long nftot, const Launch &L)
{
    if (M == 0) return 0;
    if (dim == 1)
        return interp_dim<T, 1>(method, ns, M, sb, c, fw, g, kp, maxsub, sub_bound, ntr, nftot, L);
    if (dim == 2)
        return interp_dim<T, 2>(method, ns, M, sb, c, fw, g, kp, maxsub, sub_bound, ntr, nftot, L);
    return interp_dim<T, 3>(method, ns, M, sb, c, fw, g, kp, maxsub, sub_bound, ntr, nftot, L);
}

template <typename T>
int run_deconvolve(int dim, const int *ms, const int *nf, const cpx<T> *fw, cpx<T> *fk,
                   const T *const *ker, int ntr, const Launch &L)
{
    const long long rows = (long long)ms[1] * ms[2] * ntr;
    const long long nftot = (long long)nf[0] * nf[1] * nf[2];
    const int grid = (int)(rows < (long long)L.ncu * 64 ? rows : (long long)L.ncu * 64);
    if (dim == 1)
        k_deconvolve<T, 1><<<grid, 256, 0, L.stream>>>(ms[0], ms[1], ms[2], nf[0], nf[1], nf[2],
                                                       fw, fk, ker[0], ker[1], ker[2], rows,
                                                       nftot);
    else if (dim == 2)
        k_deconvolve<T, 2><<<grid, 256, 0, L.stream>>>(ms[0], ms[1], ms[2], nf[0], nf[1], nf[2],
                                                       fw, fk, ker[0], ker[1], ker[2], rows,
                                                       nftot);
    else
        k_deconvolve<T, 3><<<grid, 256, 0, L.stream>>>(ms[0], ms[1], ms[2], nf[0], nf[1], nf[2],
                                                       fw, fk, ker[0], ker[1], ker[2], rows,
                                                       nftot);
    return ok_or_err();
}

template <typename T>
int run_amplify(int dim, const int *ms, const int *nf, cpx<T> *fw, const cpx<T> *fk,
                const T *const *ker, int ntr, const Launch &L, int wax, int wsign)
{
    const long long rows = (long long)nf[1] * nf[2] * ntr;
    const long long ntot = (long long)ms[0] * ms[1] * ms[2];
    const int grid = (int)(rows < (long long)L.ncu * 64 ? rows : (long long)L.ncu * 64);
    if (dim == 1)
        k_amplify<T, 1><<<grid, 256, 0, L.stream>>>(ms[0], ms[1], ms[2], nf[0], nf[1], nf[2], fw,
                                                    fk, ker[0], ker[1], ker[2], rows, ntot, wax, wsign);
    else if (dim == 2)
        k_amplify<T, 2><<<grid, 256, 0, L.stream>>>(ms[0], ms[1], ms[2], nf[0], nf[1], nf[2], fw,
                                                    fk, ker[0], ker[1], ker[2], rows, ntot, wax, wsign);
    else
        k_amplify<T, 3><<<grid, 256, 0, L.stream>>>(ms[0], ms[1], ms[2], nf[0], nf[1], nf[2], fw,
                                                    fk, ker[0], ker[1], ker[2], rows, ntot, wax, wsign);
    return ok_or_err();
}

template <typename T>
int run_zero(cpx<T> *p, long long n, const Launch &L)
{
    if (n > 0) k_zero<T><<<grid_1d(n, L.ncu, 64), 256, 0, L.stream>>>(p, n);
    return ok_or_err();
}

#define CFA_INST(T)                                                                          \
    template int run_natural_order<T>(int, const BinGeom &, int, const T *, const T *,       \
                                      const T *, SortBufs &, const Launch &);                \
    template int run_blk_count<T>(int, const BinGeom &, int, int, const T *, const T *,      \
                                  const T *, SortBufs &, const Launch &);                    \
    template int run_blk_scatter<T>(int, const BinGeom &, int, int, int, const T *,          \
                                    const T *, const T *, int, SortBufs &, const Launch &);  \
    template int run_spread<T>(int, int, int, int, const SortBufs &, const cpx<T> *,         \
                               cpx<T> *, const BinGeom &, const KerParams<T> &, int,         \
                               long long, int, long long, const Launch &);                   \
    template int run_interp<T>(int, int, int, int, const SortBufs &, cpx<T> *,               \
                               const cpx<T> *, const BinGeom &, const KerParams<T> &, int,   \
                               long long, int, long long, const Launch &);                   \
    template int run_deconvolve<T>(int, const int *, const int *, const cpx<T> *, cpx<T> *,  \
                                   const T *const *, int, const Launch &);                   \
    template int run_amplify<T>(int, const int *, const int *, cpx<T> *, const cpx<T> *,     \
                                const T *const *, int, const Launch &, int, int);            \
    template int run_zero<T>(cpx<T> *, long long, const Launch &);

CFA_INST(float)
CFA_INST(double)

}  // namespace cfa
