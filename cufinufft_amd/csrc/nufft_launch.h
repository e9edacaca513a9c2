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
// Dynamic LDS of one gfx950 workgroup: its 160 KiB less room for the
// kernels' static job slot.
constexpr size_t kMaxLdsBytes = 160 * 1024 - 64;

// Device arrays written by setpts.
struct SortBufs {
    int *binsize = nullptr;     // [nbins]
    int *binstart = nullptr;    // [nbins]
    int *substart = nullptr;    // [nbins + 1]
    int *s2b = nullptr;         // [nbins + ceil(M / maxsub)] upper bound
    int *skey = nullptr;        // [M] sort keys (bin, or bin and cell)
    int *sval = nullptr;        // [M] point indices fed to the sort
    int *skey_out = nullptr;    // [M] sorted keys
    void *sort_tmp = nullptr;   // radix sort scratch
    size_t sort_tmp_bytes = 0;
    int *idx = nullptr;         // [M] caller index of each sorted point
    double *pos = nullptr;      // [dim][pos_stride] sorted rescaled positions
    double *pack = nullptr;     // [M] input-order coordinates, one record per point (2D/3D)
    long long pos_stride = 0;   // capacity in points of each pos plane
    int *counters = nullptr;    // [2]: points, subproblems
    unsigned long long *jobq = nullptr;  // [kJobqWords] SM job queue words, zero between launches
    void *scan_part = nullptr;  // int2[ceil(nbins / chunk)]
    // fine sort (key = bin * cells-per-bin + cell in bin), when enabled
    int *kcount = nullptr;       // [nkeys]
    int *kstart = nullptr;       // [nkeys]
    void *kscan_part = nullptr;  // int2[ceil(nkeys / chunk)]
};

struct Launch {
    hipStream_t stream = nullptr;
    int ncu = 256;
};

// Bin sort of M points: stable radix sort of the bin keys (fine: bin and
// fine-grid cell inside the bin), per-bin counts, scans, positions gathered
// into sorted order, subproblem map.
template <typename T>
int run_radix_sort(int dim, const BinGeom &g, long long nbins, int M, const T *x, const T *y,
                   const T *z, int maxsub, bool fine, bool need_subprob, SortBufs &sb,
                   const Launch &L);
// Scratch bytes of that sort for M points and nkeys distinct keys.
size_t radix_tmp_bytes(int M, long long nkeys);
// Bytes per point of sb.pack for that sort (elem = sizeof(T); 0 in 1D).
size_t pack_rec_bytes(int dim, int elem);
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

// Bytes of LDS the SM kernels need for a tile (re and im of every cell).
size_t sm_tile_bytes(int dim, int ns, const int *bs, size_t elem);
// LDS bytes of the SM spread: 32- or 64-bit fixed-point tile plus staged
// strengths for single precision (elem 4), FP64 tile for double precision.
size_t spread_lds_bytes(int dim, int ns, const int *bs, size_t elem, int maxsub, int fxwide);
// LDS bytes of a method-4 (owner-computes block) spread job over blocks ob.
size_t blk_lds_bytes(int dim, const int *ob, size_t elem, int maxsub, int fxwide);

template <typename T>
int run_deconvolve(int dim, const int *ms, const int *nf, const cpx<T> *fw, cpx<T> *fk,
                   const T *const *ker, int ntr, const Launch &L);
template <typename T>
int run_amplify(int dim, const int *ms, const int *nf, cpx<T> *fw, const cpx<T> *fk,
                const T *const *ker, int ntr, const Launch &L);
template <typename T>
int run_zero(cpx<T> *p, long long n, const Launch &L);

// Interp of the value and its DIM coordinate derivatives (differentiable
// form): dc holds DIM planes of dc_plane complex values.
template <typename T>
int run_interp_grad(int dim, int ns, int M, const SortBufs &sb, cpx<T> *c, cpx<T> *dc,
                    const cpx<T> *fw, const BinGeom &g, const KerParams<T> &kp, int ntr,
                    long long nftot, long long dc_plane, const Launch &L);
template <typename T, int DIM>
int interp_grad_dim(int ns, int M, const SortBufs &sb, cpx<T> *c, cpx<T> *dc, const cpx<T> *fw,
                    const BinGeom &g, const KerParams<T> &kp, int ntr, long long nftot,
                    long long dc_plane, const Launch &L);

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
