// Pruned 3D FFT: the data movement between its three 1D passes.
//
// A type-1 transform needs only the ms0 x ms1 x ms2 block of modes out of the
// nf0 x nf1 x nf2 fine-grid FFT (a quarter or less per axis pair at
// upsampling 2), and a type-2 transform starts from that block zero-padded.
// The full 3D FFT runs its y and z passes as strided transforms over the
// whole grid; here each 1D pass is a contiguous batched hipFFT (x fastest,
// then y, then z fastest) and only the lines that hold modes go on:
//
//   type 1: x-FFT of fw [z][y][x]          (all nf1 nf2 lines)
//           gather  fw -> B1 [x'][z][y]    (mode columns x' only, y fastest)
//           y-FFT of B1                    (ms0 nf2 lines)
//           gather  B1 -> B2 [x'][y'][z]   (mode rows y' only, z fastest)
//           z-FFT of B2                    (ms0 ms1 lines)
//           gather  B2 -> fk [kz][ky][kx]  (mode planes kz, divided by the
//                                           kernel's Fourier series: the deconvolve)
//   type 2: the same in reverse with scatters that write zeros at the
//           non-mode positions (the amplify is the first scatter).
//
// Both kernels are LDS-tiled transposes of one plane: a W x W tile is read
// along the input's unit-stride axis and written along the output's, so
// every global access is a coalesced row segment.
#include "nufft_device.cuh"
#include "nufft_launch.h"

namespace cfa {

namespace {
inline int ok_or_err() { return hipGetLastError() == hipSuccess ? 0 : kErrHip; }

// Tile edge: 64 complex FP32 (512-byte rows, 33 KiB of LDS) or 32 complex
// FP64 (the same row bytes, 17 KiB).
template <typename T> constexpr int prune_w() { return sizeof(T) == 4 ? 64 : 32; }
}  // namespace

// out[t][c'][p][r] = s * in[t][p][r][src(c')] for the kept c' < Nc, where
// src(c') = C - kneg + c' for c' < kneg (negative frequencies, at the top of
// the axis) and c' - kneg otherwise; s = 1 / (kr[|r-hr|] kp[|p-hp|] kc[|c'-kneg|])
// when the tables are given (the type-1 deconvolve, as k_deconvolve; kp may
// be null: a factor 1, the 2D case), else 1.
template <typename T, int W, bool SCALE>
__global__ void __launch_bounds__(256)
k_prune_gather(const cpx<T> *__restrict__ in, cpx<T> *__restrict__ out, PruneGather d,
               const T *__restrict__ kr, const T *__restrict__ kp, const T *__restrict__ kc)
{
    __shared__ cpx<T> tile[W][W + 1];
    constexpr int ROWS = 256 / W, PER = W / ROWS;
    const int nrt = (d.R + W - 1) / W;
    const int r0 = (int)(blockIdx.x % nrt) * W, c0 = (int)(blockIdx.x / nrt) * W;
    const int p = blockIdx.y;
    const long long t = blockIdx.z;
    const int tx = threadIdx.x % W, ty = threadIdx.x / W;
    // all PER loads of a thread in flight together, then into the tile
    {
        const int cq = c0 + tx;
        cpx<T> v[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int r = r0 + ty + k * ROWS;
            v[k].x = 0;
            v[k].y = 0;
            if (cq < d.Nc && r < d.R) v[k] = in[prune_gather_src(d, t, p, r, cq)];
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) tile[tx][ty + k * ROWS] = v[k];
    }
    __syncthreads();
    const int r = r0 + tx;
    if (r >= d.R) return;
    T sr = 1, sp = 1;
    if (SCALE) {
        sr = kr[abs(r - d.hr)];
        sp = kp ? kp[abs(p - d.hp)] : T(1);
    }
    cpx<T> v[PER];
    T kv[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int c = c0 + ty + k * ROWS;
        v[k] = tile[ty + k * ROWS][tx];
        kv[k] = 1;
        // the mode index is clamped so every table read is in bounds
        if (SCALE) kv[k] = sr * (sp * kc[abs(min(c, d.Nc - 1) - d.kneg)]);
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int c = c0 + ty + k * ROWS;
        if (c < d.Nc) {
            cpx<T> o = v[k];
            if (SCALE) {
                o.x = o.x / kv[k];
                o.y = o.y / kv[k];
            }
            out[prune_gather_dst(d, t, p, r, c)] = o;
        }
    }
}

// out[t][p][r][c] for every c < C: s * w * in[t][c'][p][r] where c holds a
// mode (c' as in k_prune_gather), 0 elsewhere.  With the tables given, s is
// 1 / the kernel's Fourier series product (the type-2 amplify, as k_amplify)
// and, for wsel 0 / 1 / 2, w = i * wsign * the frequency along r / p / c (the
// derivative modes of execute_grad); else s = w = 1.
template <typename T, int W, bool SCALE>
__global__ void __launch_bounds__(256)
k_prune_scatter(const cpx<T> *__restrict__ in, cpx<T> *__restrict__ out, PruneScatter d,
                const T *__restrict__ kr, const T *__restrict__ kp, const T *__restrict__ kc,
                int wsel, int wsign)
{
    __shared__ cpx<T> tile[W][W + 1];
    constexpr int ROWS = 256 / W, PER = W / ROWS;
    const int nrt = (d.R + W - 1) / W;
    const int r0 = (int)(blockIdx.x % nrt) * W, c0 = (int)(blockIdx.x / nrt) * W;
    const int p = blockIdx.y;
    const long long t = blockIdx.z;
    const int tx = threadIdx.x % W, ty = threadIdx.x / W;
    {
        const int r = r0 + tx;
        T sr = 1, sp = 1;
        if (SCALE && r < d.R) {
            sr = kr[abs(r - d.hr)];
            sp = kp ? kp[abs(p - d.hp)] : T(1);
        }
        cpx<T> v[PER];
        int q[PER];
        T kv[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int c = c0 + ty + k * ROWS;
            q[k] = c < d.C && r < d.R ? prune_mode(c, d.Nc, d.kneg, d.C) : -1;
            v[k].x = 0;
            v[k].y = 0;
            kv[k] = 1;
            if (q[k] >= 0) {
                v[k] = in[prune_scatter_src(d, t, p, r, q[k])];
                if (SCALE) kv[k] = sr * (sp * kc[abs(q[k] - d.kneg)]);
            }
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            cpx<T> o = v[k];
            if (SCALE && q[k] >= 0) {
                o.x = o.x / kv[k];
                o.y = o.y / kv[k];
                if (wsel >= 0) {
                    const int f =
                        wsel == 0 ? r - d.hr : (wsel == 1 ? p - d.hp : q[k] - d.kneg);
                    const T w = (T)(wsign * f);
                    const T re = -o.y * w;
                    o.y = o.x * w;
                    o.x = re;
                }
            }
            tile[ty + k * ROWS][tx] = o;
        }
    }
    __syncthreads();
    const int c = c0 + tx;
    if (c >= d.C) return;
    cpx<T> v[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) v[k] = tile[tx][ty + k * ROWS];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int r = r0 + ty + k * ROWS;
        if (r < d.R) out[prune_scatter_dst(d, t, p, r, c)] = v[k];
    }
}

template <typename T>
int run_prune_gather(const cpx<T> *in, cpx<T> *out, const PruneGather &d, int ntr,
                     const T *kr, const T *kp, const T *kc, const Launch &L)
{
    constexpr int W = prune_w<T>();
    if (d.R <= 0 || d.Nc <= 0 || d.P <= 0 || ntr <= 0) return 0;
    if (d.P > 65535 || ntr > 65535) return kErrHip;
    const long long nt = (long long)((d.R + W - 1) / W) * ((d.Nc + W - 1) / W);
    dim3 grid((unsigned)nt, (unsigned)d.P, (unsigned)ntr);
    if (kr && kc)
        k_prune_gather<T, W, true><<<grid, 256, 0, L.stream>>>(in, out, d, kr, kp, kc);
    else
        k_prune_gather<T, W, false><<<grid, 256, 0, L.stream>>>(in, out, d, kr, kp, kc);
    return ok_or_err();
}

template <typename T>
int run_prune_scatter(const cpx<T> *in, cpx<T> *out, const PruneScatter &d, int ntr,
                      const T *kr, const T *kp, const T *kc, int wsel, int wsign,
                      const Launch &L)
{
    constexpr int W = prune_w<T>();
    if (d.R <= 0 || d.C <= 0 || d.P <= 0 || ntr <= 0) return 0;
    if (d.P > 65535 || ntr > 65535) return kErrHip;
    const long long nt = (long long)((d.R + W - 1) / W) * ((d.C + W - 1) / W);
    dim3 grid((unsigned)nt, (unsigned)d.P, (unsigned)ntr);
    if (kr && kc)
        k_prune_scatter<T, W, true><<<grid, 256, 0, L.stream>>>(in, out, d, kr, kp, kc, wsel,
                                                                wsign);
    else
        k_prune_scatter<T, W, false><<<grid, 256, 0, L.stream>>>(in, out, d, kr, kp, kc, wsel,
                                                                 wsign);
    return ok_or_err();
}

#define CFA_PRUNE_INST(T)                                                                   \
    template int run_prune_gather<T>(const cpx<T> *, cpx<T> *, const PruneGather &, int,    \
                                     const T *, const T *, const T *, const Launch &);      \
    template int run_prune_scatter<T>(const cpx<T> *, cpx<T> *, const PruneScatter &, int,  \
                                      const T *, const T *, const T *, int, int,            \
                                      const Launch &);
CFA_PRUNE_INST(float)
CFA_PRUNE_INST(double)
#undef CFA_PRUNE_INST

}  // namespace cfa
