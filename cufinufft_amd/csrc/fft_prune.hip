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
// when the tables are given (the type-1 deconvolve, as k_deconvolve), else 1.
template <typename T, int W>
__global__ void __launch_bounds__(256)
k_prune_gather(const cpx<T> *__restrict__ in, cpx<T> *__restrict__ out, PruneGather d,
               const T *__restrict__ kr, const T *__restrict__ kp, const T *__restrict__ kc)
{
    __shared__ cpx<T> tile[W][W + 1];
    constexpr int ROWS = 256 / W;
    const int nrt = (d.R + W - 1) / W;
    const int r0 = (int)(blockIdx.x % nrt) * W, c0 = (int)(blockIdx.x / nrt) * W;
    const int p = blockIdx.y;
    const long long t = blockIdx.z;
    const int tx = threadIdx.x % W, ty = threadIdx.x / W;
    {
        const int cq = c0 + tx;
        if (cq < d.Nc) {
            for (int i = ty; i < W; i += ROWS) {
                const int r = r0 + i;
                if (r < d.R) tile[tx][i] = in[prune_gather_src(d, t, p, r, cq)];
            }
        }
    }
    __syncthreads();
    const int r = r0 + tx;
    if (r >= d.R) return;
    T sr = 1, sp = 1;
    if (kr) {
        sr = kr[abs(r - d.hr)];
        sp = kp[abs(p - d.hp)];
    }
    for (int j = ty; j < W; j += ROWS) {
        const int c = c0 + j;
        if (c >= d.Nc) break;
        cpx<T> v = tile[j][tx];
        if (kr) {
            const T kv = sr * (sp * kc[abs(c - d.kneg)]);
            v.x = v.x / kv;
            v.y = v.y / kv;
        }
        out[prune_gather_dst(d, t, p, r, c)] = v;
    }
}

// out[t][p][r][c] for every c < C: s * w * in[t][c'][p][r] where c holds a
// mode (c' as in k_prune_gather), 0 elsewhere.  With the tables given, s is
// 1 / the kernel's Fourier series product (the type-2 amplify, as k_amplify)
// and, for wsel 0 / 1 / 2, w = i * wsign * the frequency along r / p / c (the
// derivative modes of execute_grad); else s = w = 1.
template <typename T, int W>
__global__ void __launch_bounds__(256)
k_prune_scatter(const cpx<T> *__restrict__ in, cpx<T> *__restrict__ out, PruneScatter d,
                const T *__restrict__ kr, const T *__restrict__ kp, const T *__restrict__ kc,
                int wsel, int wsign)
{
    __shared__ cpx<T> tile[W][W + 1];
    constexpr int ROWS = 256 / W;
    const int nrt = (d.R + W - 1) / W;
    const int r0 = (int)(blockIdx.x % nrt) * W, c0 = (int)(blockIdx.x / nrt) * W;
    const int p = blockIdx.y;
    const long long t = blockIdx.z;
    const int tx = threadIdx.x % W, ty = threadIdx.x / W;
    {
        const int r = r0 + tx;
        T sr = 1, sp = 1;
        if (kr && r < d.R) {
            sr = kr[abs(r - d.hr)];
            sp = kp[abs(p - d.hp)];
        }
        for (int j = ty; j < W; j += ROWS) {
            const int c = c0 + j;
            cpx<T> v;
            v.x = 0;
            v.y = 0;
            const int cq = c < d.C ? prune_mode(c, d.Nc, d.kneg, d.C) : -1;
            if (cq >= 0 && r < d.R) {
                v = in[prune_scatter_src(d, t, p, r, cq)];
                if (kr) {
                    const T kv = sr * (sp * kc[abs(cq - d.kneg)]);
                    v.x = v.x / kv;
                    v.y = v.y / kv;
                    if (wsel >= 0) {
                        const int f = wsel == 0 ? r - d.hr : (wsel == 1 ? p - d.hp : cq - d.kneg);
                        const T w = (T)(wsign * f);
                        const T re = -v.y * w;
                        v.y = v.x * w;
                        v.x = re;
                    }
                }
            }
            tile[j][tx] = v;
        }
    }
    __syncthreads();
    const int c = c0 + tx;
    if (c >= d.C) return;
    for (int i = ty; i < W; i += ROWS) {
        const int r = r0 + i;
        if (r >= d.R) break;
        out[prune_scatter_dst(d, t, p, r, c)] = tile[tx][i];
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
    k_prune_gather<T, W><<<grid, 256, 0, L.stream>>>(in, out, d, kr, kp, kc);
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
    k_prune_scatter<T, W><<<grid, 256, 0, L.stream>>>(in, out, d, kr, kp, kc, wsel, wsign);
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
