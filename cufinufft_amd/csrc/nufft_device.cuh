// Device-side building blocks of the NUFFT path for gfx950: coordinate
// rescaling, bin indices, ES-kernel evaluation, and every kernel of the
// setpts / execute stages.  Templated on the plan precision T, the dimension
// DIM and (for the spread/interp kernels) the kernel width NS, so that the
// tap loops are fully unrolled for each width.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstddef>
#include <type_traits>

#include "nufft_launch.h"

namespace cfa {

// Index arithmetic shared by the kernels and the host-side index checks
// (tests/cpu_index): compiled for both sides.
#define CFA_HD __host__ __device__ __forceinline__

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

// 1/(2 pi) as the double literal the reference uses (contrib/utils.h:18).
constexpr double kInv2Pi = 0.159154943091895336;

// Geometry of the fine grid and its bins.  Unused axes have nf = bs = nb = 1.
struct BinGeom {
    int nf[3];
    int bs[3];
    int nb[3];
};

template <typename T>
struct KerParams {
    T beta;  // ES beta
    T c;     // ES c = 4/ns^2
    T hw;    // half width ns/2
    // Piecewise-polynomial kernel (gpu_kerevalmeth = 1): hdeg > 0 selects it.
    // hc[k*ns + i] is the z^k coefficient of tap i, z = 2(start - v) + ns - 1
    // in [-1, 1); a second table of the same size follows at hc + (hdeg+1)*ns
    // holding the kernel divided by e^beta (peak 1).
    const T *hc;
    int hdeg;
    // FP32 fixed-point spread tiles: fx32max is the largest number of points
    // one 32-bit scale may cover at the plan's tolerance; fxwide (method 4)
    // selects 64-bit block tiles when its subproblems exceed it
    int fxwide;
    int fx32max;
    // beta * log2(e) as a sum bl2e + bl2e_lo of two T values (the FP64
    // product split): FP32 kernels evaluate e^(beta s) as 2^(beta log2(e) s)
    // with the hardware exp2, forming the exponent as
    // fma(bl2e, s, bl2e_lo * s) so that it carries one rounding of its own
    // and none of the constant (whose rounding, shared by every tap, would
    // scale the whole grid by up to ~1e-6)
    T bl2e;
    T bl2e_lo;
};

// 2^(beta log2(e) s) in FP32 (see KerParams::bl2e).
__device__ __forceinline__ float exp_beta(float s, const KerParams<float> &k)
{
    return __builtin_amdgcn_exp2f(fmaf(k.bl2e, s, k.bl2e_lo * s));
}

// Points as the execute kernels read them, in bin-sorted order: idx[i] is
// the caller's index of sorted point i (for strengths / outputs) and
// pos[d*stride + i] its rescaled grid coordinate along axis d (the unrounded
// double RESCALE value clamped to [-1, nf+1]), written once by setpts so that
// execute streams positions instead of gathering and rescaling coordinates.
struct SortedPts {
    const int *idx;
    const double *pos;
    long long stride;
};

// Workgroups are dispatched round-robin over the XCDs of the device (8 on
// gfx950, each with its own L2).  Persistent kernels map blockIdx so that each
// XCD walks a contiguous range of jobs: neighbouring bins, whose halo tiles
// overlap, then share one L2.  A grid that is a multiple of kNumXcd keeps the
// mapping exact; any other grid is still correct (all jobs are covered).
constexpr int kNumXcd = 8;

struct JobRange {
    long long begin, end, step;
};

__device__ __forceinline__ JobRange xcd_jobs(long long njobs)
{
    const int G = gridDim.x;
    if (G % kNumXcd != 0) return {(long long)blockIdx.x, njobs, (long long)G};
    const int xcd = blockIdx.x % kNumXcd;
    const int per = G / kNumXcd;
    const long long b = njobs * xcd / kNumXcd, e = njobs * (xcd + 1) / kNumXcd;
    return {b + blockIdx.x / kNumXcd, e, (long long)per};
}

// Job scheduling of the SM kernels.  The jobs are split into kNumXcd
// contiguous ranges as above.  A workgroup first runs its static share of
// the first half of its XCD's range (every (grid/kNumXcd)-th job, no
// atomics), then takes jobs one by one from an atomic counter over the rest
// of the range, and when that runs dry, from the counters of the next
// ranges.  Job costs follow the points per subproblem, so on clustered points
// equal-count ranges carry very unequal work (1.3-1.6x the mean on the
// densest range for the 3D workloads): the dynamic halves balance that while
// neighbouring bins stay on one XCD.  The static half keeps the counters out
// of the way of balanced workloads (on uniform points a counter per job cost
// ~10%: every workgroup of an XCD hits one address at the same moments).
// Launches with fewer than 4 jobs per workgroup, or a grid that is not a
// multiple of kNumXcd, keep the static split alone (the probes of a dry
// queue cost more than any balance gained there).
//
// jq holds kNumXcd counters of dynamic jobs taken (0-based) and a count of
// finished workgroups; they are zero between launches: the last workgroup
// to finish clears them, so a launch needs no set-up kernel (and a captured
// graph replays correctly).  Launches sharing one jq must be serialised (one
// plan's stream, or one group's).
static_assert(kJobqWords == kNumXcd + 1, "job queue: a counter per XCD range + a done count");

CFA_HD long long jobq_static(long long njobs, int x, int per)
{
    const long long n = njobs * (x + 1) / kNumXcd - njobs * x / kNumXcd;
    return n / 2 / per * per;
}

// Next dynamic job from queue q onwards (thread 0 only); q and left (queues
// not yet found empty, counting q) persist across calls, and -1 means all are
// empty.  A counter only grows, so a queue found empty stays empty.
__device__ __forceinline__ long long jobq_take(unsigned long long *jq, long long njobs, int per,
                                               int &q, int &left)
{
    while (left > 0) {
        const long long t = (long long)atomicAdd(&jq[q], 1ull);
        const long long j = njobs * q / kNumXcd + jobq_static(njobs, q, per) + t;
        if (j < njobs * (q + 1) / kNumXcd) return j;
        q = q + 1 == kNumXcd ? 0 : q + 1;
        --left;
    }
    return -1;
}

// Runs job(j) for every job handed to this workgroup: the static sequence,
// then the dynamic jobs.  The counter of the next dynamic job is bumped
// before the current one runs and checked after it, so the atomic's latency
// overlaps the job; job(j) must end with a barrier (every SM job function
// does).  job is called from one place only, so the job body is inlined
// once.
//
// pos(j) (optional) gives the sorted-point position of job j's first point
// over all transforms (monotone in j; npts = the position past the last
// job).  With it the workgroups first estimate the work of each XCD range
// (points plus a per-job overhead of 8 points, from 9 lookups spread over
// the lanes of wave 0) and keep the static split when no range exceeds the
// mean by more than 10%: uniform points pay no counters at all.
template <class F, class Pos = std::nullptr_t>
__device__ __forceinline__ void jobq_run(unsigned long long *jq, long long njobs, F &&job,
                                         Pos &&pos = nullptr, long long npts = 0)
{
    __shared__ long long s_job;
    const int G = gridDim.x;
    bool dyn = G % kNumXcd == 0 && njobs >= 4ll * G;
    if constexpr (!std::is_same_v<std::decay_t<Pos>, std::nullptr_t>) {
        if (dyn) {
            if (threadIdx.x < 64) {
                const int k = threadIdx.x;
                long long w = 0;
                if (k < kNumXcd) {
                    const long long b = njobs * k / kNumXcd, e = njobs * (k + 1) / kNumXcd;
                    if (e > b) w = (e < njobs ? pos(e) : npts) - pos(b) + 8 * (e - b);
                }
                long long mx = w, sum = w;
#pragma unroll
                for (int d = 4; d >= 1; d >>= 1) {
                    mx = max(mx, (long long)__shfl_xor(mx, d, 64));
                    sum += (long long)__shfl_xor(sum, d, 64);
                }
                if (threadIdx.x == 0) s_job = 80 * mx > 11 * sum ? 1 : 0;
            }
            __syncthreads();
            dyn = s_job != 0;
            __syncthreads();
        }
    }
    const int x = blockIdx.x % kNumXcd, per = G / kNumXcd;
    long long j, se, step;  // static sequence j, j + step, ... < se
    if (dyn) {
        const long long b = njobs * x / kNumXcd;
        j = b + blockIdx.x / kNumXcd;
        se = b + jobq_static(njobs, x, per);
        step = per;
    } else {
        const JobRange jr = xcd_jobs(njobs);
        j = jr.begin;
        se = jr.end;
        step = jr.step;
    }
    int q = x, left = dyn ? kNumXcd : 0;  // dynamic state, used by thread 0
    bool stat = j < se;
    if (!stat) {
        if (threadIdx.x == 0) s_job = jobq_take(jq, njobs, per, q, left);
        __syncthreads();
        j = s_job;
    }
    while (j >= 0) {
        unsigned long long raw = 0;
        if (!stat && threadIdx.x == 0) raw = atomicAdd(&jq[q], 1ull);
        job(j);
        if (stat) {
            j += step;
            if (j < se) continue;
            stat = false;
            if (threadIdx.x == 0) s_job = jobq_take(jq, njobs, per, q, left);
        } else if (threadIdx.x == 0) {
            long long nx = njobs * q / kNumXcd + jobq_static(njobs, q, per) + (long long)raw;
            if (nx >= njobs * (q + 1) / kNumXcd) {
                q = q + 1 == kNumXcd ? 0 : q + 1;
                --left;
                nx = jobq_take(jq, njobs, per, q, left);
            }
            s_job = nx;
        }
        __syncthreads();
        j = s_job;
    }
    // every queue is dry for this workgroup; the last one to get here clears
    // the words for the next launch (all others are done with them)
    if (dyn && threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&jq[kNumXcd], 1ull) == (unsigned long long)G - 1)
            for (int k = 0; k < kJobqWords; ++k) atomicExch(&jq[k], 0ull);
    }
}

// Subproblem bookkeeping written by setpts, read by the SM kernels.
struct SubprobArgs {
    const int *s2b;       // subproblem -> bin
    const int *substart;  // first subproblem of each bin (nbins+1)
    const int *binstart;  // first sorted point of each bin
    const int *binsize;   // points per bin
    const int *counters;  // [0] points, [1] subproblems
    int maxsub;
};

// ---------------------------------------------------------------------------
// RESCALE (pirange = 1), contrib/spreadinterp.h:36-38: x*M_1_2PI + off in
// double (one fused multiply-add, as nvcc contracts it by default), times N in
// double.  Returned unrounded: the caller rounds to T for the bin index (the
// reference's value, bit for bit) and keeps the double for tap offsets.
CFA_HD double rescale_exact(float x, int n)
{
    const float pi = 3.14159265358979323846f;  // (float)M_PI
    const double off = x < -pi ? 1.5 : (x >= pi ? -0.5 : 0.5);
    return __builtin_fma((double)x, kInv2Pi, off) * (double)n;
}
CFA_HD double rescale_exact(double x, int n)
{
    const double pi = 3.14159265358979323846;
    const double off = x < -pi ? 1.5 : (x >= pi ? -0.5 : 0.5);
    return __builtin_fma(x, kInv2Pi, off) * (double)n;
}

// Bin of one axis from the T-rounded rescaled coordinate:
// b = floor(xr/bs) in T; b == nb -> nb-1; b < 0 -> 0 (spreadinterp2d.cu:
// 113-118).  Written so that NaN or far out-of-range input (which the
// reference indexes out of bounds with) lands in a valid bin instead.
template <typename T>
CFA_HD int bin1(double v, int bs, int nb)
{
    const T xt = (T)v;
    const T q = floor(xt / (T)bs);
    return q >= (T)0 ? (q < (T)nb ? (int)q : nb - 1) : 0;
}

template <typename T, int DIM>
__device__ __forceinline__ int point_bin(const T *__restrict__ x, const T *__restrict__ y,
                                         const T *__restrict__ z, int i, const BinGeom &g)
{
    int b = bin1<T>(rescale_exact(x[i], g.nf[0]), g.bs[0], g.nb[0]);
    if (DIM > 1) b += g.nb[0] * bin1<T>(rescale_exact(y[i], g.nf[1]), g.bs[1], g.nb[1]);
    if (DIM > 2)
        b += g.nb[0] * g.nb[1] * bin1<T>(rescale_exact(z[i], g.nf[2]), g.bs[2], g.nb[2]);
    return b;
}

// ES kernel phi(x) = exp(beta sqrt(1 - c x^2)) for |x| < ns/2, else 0
// (src/cuspreadinterp.h:6-16).  The argument of sqrt is floored at 0 so that
// rounding at |x| -> ns/2 cannot produce NaN.
//
// FP32 evaluates the exponential with the hardware exp2 (v_exp_f32, within
// 0.8 ulp measured; the library's expf is ~12 VALU instructions more per
// tap) on an exponent formed without rounding the constant (exp_beta).  The
// square root stays the correctly rounded sqrtf: its error is multiplied by
// beta (~14-20) in the exponent, and the hardware v_sqrt_f32 (within 0.9 ulp,
// not correctly rounded) moved the least-squares scale of a 3D FP32 spread
// by 4e-7.
__device__ __forceinline__ float es_eval(float x, const KerParams<float> &k)
{
    const float a = fmaxf(1.0f - k.c * x * x, 0.0f);
    return fabsf(x) < k.hw ? exp_beta(sqrtf(a), k) : 0.0f;
}
__device__ __forceinline__ double es_eval(double x, const KerParams<double> &k)
{
    const double a = fmax(1.0 - k.c * x * x, 0.0);
    return fabs(x) < k.hw ? exp(k.beta * sqrt(a)) : 0.0;
}

// Position used for tap offsets: the unrounded rescale, clamped to
// [-1, n+1] so that invalid input cannot index outside one period of wrap.
CFA_HD double tap_pos(double v, int n)
{
    return fmin(fmax(v, -1.0), (double)n + 1.0);
}

// Horner evaluation of the ns per-tap polynomials at z (all taps of one
// point share z); coefficient addresses are lane-uniform.
// The table is read through the constant address space: it is read-only for
// the whole launch and every lane reads the same address, so the loads are
// scalar (s_load into SGPRs) instead of a chain of vector loads per degree.
template <typename T, int NS>
__device__ __forceinline__ void horner_taps(T z, const T *__restrict__ hc, int deg, T *ker)
{
    using CT = const __attribute__((address_space(4))) T;
    CT *h = (CT *)hc;
#pragma unroll
    for (int i = 0; i < NS; ++i) ker[i] = h[deg * NS + i];
    for (int k = deg - 1; k >= 0; --k) {
#pragma unroll
        for (int i = 0; i < NS; ++i) ker[i] = ker[i] * z + h[k * NS + i];
    }
}

// First tap index and the ns kernel values of one axis:
// start = ceil(v - ns/2) (exact in double), ker[i] = phi(start + i - v),
// by exp(sqrt()) or by the polynomial table.
template <typename T, int NS>
__device__ __forceinline__ int taps_global(double v, const KerParams<T> &kp, T *ker)
{
    const int start = (int)ceil(v - 0.5 * NS);
    if (kp.hdeg > 0) {
        horner_taps<T, NS>((T)(2.0 * ((double)start - v) + (NS - 1)), kp.hc, kp.hdeg, ker);
        return start;
    }
    const T x1 = (T)((double)start - v);
#pragma unroll
    for (int i = 0; i < NS; ++i) ker[i] = es_eval(x1 + (T)i, kp);
    return start;
}

// The same for a point inside a tile whose cell 0 sits at global index o and
// which has width tw: the local first tap is clamped into [0, tw-ns] and the
// kernel values are evaluated at the true distances of the clamped taps, so
// a tap that falls outside the tile (possible only within rounding of the
// kernel edge, where phi is 0) simply gets weight phi(.) of its real offset.
template <typename T, int NS>
__device__ __forceinline__ int taps_local(double v, int o, int tw, const KerParams<T> &kp,
                                          T *ker)
{
    const int l0 = (int)ceil(v - 0.5 * NS) - o;
    const int l = l0 < 0 ? 0 : (l0 > tw - NS ? tw - NS : l0);
    if (kp.hdeg > 0 && l == l0) {
        horner_taps<T, NS>((T)(2.0 * ((double)(o + l) - v) + (NS - 1)), kp.hc, kp.hdeg, ker);
        return l;
    }
    const T x1 = (T)((double)(o + l) - v);
#pragma unroll
    for (int i = 0; i < NS; ++i) ker[i] = es_eval(x1 + (T)i, kp);
    return l;
}

CFA_HD int wrap1(int i, int n)
{
    return i < 0 ? i + n : (i >= n ? i - n : i);
}
CFA_HD int wrapmod(int i, int n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

// Fire-and-forget device-scope floating-point atomic add to global memory
// (global_atomic_add_f32 / _f64; the fine grid is coarse-grained hipMalloc
// memory, where the hardware float atomics are exact).
__device__ __forceinline__ void gadd(float *p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void gadd(double *p, double v) { unsafeAtomicAdd(p, v); }

// LDS atomic add (ds_add_f32 / ds_add_f64).
__device__ __forceinline__ void ladd(float *p, float v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void ladd(double *p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// ===========================================================================
// setpts: bin sort (keys here; the radix sort itself is in sort_radix.hip)
// ===========================================================================

// Fine sort key: the bin (point_bin) then the fine-grid cell of the point
// inside its bin, key = bin * vol + local cell (vol = cells per bin).  Sorting
// by this key orders each bin's points by cell, so neighbouring lanes of a
// wave touch neighbouring LDS cells (or the same cell, which the LDS
// broadcasts); the per-bin counts and starts are unchanged.
template <typename T>
CFA_HD void bin_and_cell(double v, int bs, int nb, int &b, int &lc)
{
    const T xt = (T)v;
    const T q = floor(xt / (T)bs);
    b = q >= (T)0 ? (q < (T)nb ? (int)q : nb - 1) : 0;
    const T f = floor(xt);
    int c = f >= (T)0 ? (f < (T)2147483647 ? (int)f : 2147483647) : 0;
    c -= b * bs;
    lc = c < 0 ? 0 : (c >= bs ? bs - 1 : c);
}

// Per-bin size and start from the key starts: the bin's keys are contiguous.
static __global__ void __launch_bounds__(256)
k_bins_from_keys(long long nbins, int vol, int M, const int *__restrict__ keystart,
                 int *__restrict__ binsize)
{
    for (long long b = blockIdx.x * (long long)blockDim.x + threadIdx.x; b < nbins;
         b += (long long)gridDim.x * blockDim.x) {
        const int s0 = keystart[b * vol];
        const int s1 = b + 1 < nbins ? keystart[(b + 1) * vol] : M;
        binsize[b] = s1 - s0;
    }
}

// Natural order (gpu_sort = 0, TrivialGlobalSortIdx): idx[i] = i and positions.
template <typename T, int DIM>
__global__ void __launch_bounds__(256)
k_natural_order(int M, const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
                BinGeom g, int *__restrict__ idx, double *__restrict__ pos, long long stride)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        idx[i] = i;
        pos[i] = tap_pos(rescale_exact(x[i], g.nf[0]), g.nf[0]);
        if (DIM > 1) pos[stride + i] = tap_pos(rescale_exact(y[i], g.nf[1]), g.nf[1]);
        if (DIM > 2) pos[2 * stride + i] = tap_pos(rescale_exact(z[i], g.nf[2]), g.nf[2]);
    }
}

// Exclusive scans of (binsize, numsubprob = ceil(binsize/maxsub)) in three
// passes: per-chunk sums, one-block scan of the chunk sums, per-chunk scan.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanChunk = kScanThreads * kScanItems;

__device__ __forceinline__ int2 add2(int2 a, int2 b) { return make_int2(a.x + b.x, a.y + b.y); }

// Block-wide exclusive scan of one int2 per thread; returns the block total.
__device__ __forceinline__ int2 block_exscan2(int2 v, int2 &excl)
{
    __shared__ int2 wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int2 inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int2 o;
        o.x = __shfl_up(inc.x, d, 64);
        o.y = __shfl_up(inc.y, d, 64);
        if (lane >= d) inc = add2(inc, o);
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int2 off = make_int2(0, 0), tot = make_int2(0, 0);
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        if (w < wid) off = add2(off, wsum[w]);
        tot = add2(tot, wsum[w]);
    }
    excl = make_int2(off.x + inc.x - v.x, off.y + inc.y - v.y);
    __syncthreads();
    return tot;
}

__device__ __forceinline__ int2 load_pair(const int *__restrict__ cnt, long long i, long long n,
                                          int maxsub)
{
    const int c = i < n ? cnt[i] : 0;
    return make_int2(c, (c + maxsub - 1) / maxsub);
}

static __global__ void __launch_bounds__(kScanThreads)
k_scan_reduce(const int *__restrict__ cnt, long long n, int maxsub, int2 *__restrict__ part)
{
    const long long base = (long long)blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
    int2 s = make_int2(0, 0);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s = add2(s, load_pair(cnt, base + k, n, maxsub));
    int2 ex;
    const int2 tot = block_exscan2(s, ex);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// One block: exclusive scan of the chunk sums in place; writes the totals to
// counters[0..1] and substart[n].
static __global__ void __launch_bounds__(kScanThreads)
k_scan_parts(int2 *__restrict__ part, int nparts, int *__restrict__ counters,
             int *__restrict__ substart_end)
{
    int2 carry = make_int2(0, 0);
    for (int base = 0; base < nparts; base += kScanChunk) {
        int2 v[kScanItems];
        int2 s = make_int2(0, 0);
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int i = base + threadIdx.x * kScanItems + k;
            v[k] = i < nparts ? part[i] : make_int2(0, 0);
            s = add2(s, v[k]);
        }
        int2 ex;
        const int2 tot = block_exscan2(s, ex);
        int2 run = add2(carry, ex);
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int i = base + threadIdx.x * kScanItems + k;
            if (i < nparts) part[i] = run;
            run = add2(run, v[k]);
        }
        carry = add2(carry, tot);
        __syncthreads();
    }
    if (threadIdx.x == 0 && counters) {
        counters[0] = carry.x;
        counters[1] = carry.y;
        *substart_end = carry.y;
    }
}

static __global__ void __launch_bounds__(kScanThreads)
k_scan_apply(const int *__restrict__ cnt, long long n, int maxsub, const int2 *__restrict__ part,
             int *__restrict__ binstart, int *__restrict__ substart)
{
    const long long base = (long long)blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
    int2 v[kScanItems];
    int2 s = make_int2(0, 0);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = load_pair(cnt, base + k, n, maxsub);
        s = add2(s, v[k]);
    }
    int2 ex;
    block_exscan2(s, ex);
    int2 run = add2(part[blockIdx.x], ex);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) {
            binstart[base + k] = run.x;
            if (substart) substart[base + k] = run.y;
        }
        run = add2(run, v[k]);
    }
}

// subprob_to_bin[substart[b] + j] = b for j < numsubprob[b]
// (MapBintoSubProb, precision_independent.cu:75-85,117-127).
static __global__ void __launch_bounds__(256)
k_map_subprob(long long nbins, const int *__restrict__ substart, int *__restrict__ s2b)
{
    for (long long b = blockIdx.x * (long long)blockDim.x + threadIdx.x; b < nbins;
         b += (long long)gridDim.x * blockDim.x) {
        const int s0 = substart[b], s1 = substart[b + 1];
        for (int s = s0; s < s1; ++s) s2b[s] = (int)b;
    }
}

static __global__ void __launch_bounds__(256) k_fill_int(long long n, int v, int *__restrict__ p)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// ===========================================================================
// execute: spreading (type 1) and interpolation (type 2)
// ===========================================================================

// Sorted-point chunks of 256 for the point-driven kernels: each XCD takes a
// contiguous range of the bin-sorted order, so nearby points share its L2.
constexpr int kPtChunk = 256;

// GM / GM-sort spread: one thread per point, in idxnupts order, global float
// atomics for each of the ns^DIM taps; all ntr transforms of the batch reuse
// the kernel values (Spread_{2,3}d_NUptsdriven, spreadinterp2d.cu:16-58).
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(256)
k_spread_gm(int M, SortedPts sp, const cpx<T> *__restrict__ c, cpx<T> *__restrict__ fw,
            BinGeom g, KerParams<T> kp, int ntr, long long nftot)
{
    const int n1 = g.nf[0], n2 = g.nf[1], n3 = g.nf[2];
    const JobRange jr = xcd_jobs((M + kPtChunk - 1) / kPtChunk);
    for (long long ch = jr.begin; ch < jr.end; ch += jr.step) {
        const int i = (int)(ch * kPtChunk) + (int)threadIdx.x;
        if (i >= M) continue;
        const int j = sp.idx[i];
        T k1[NS], k2[NS], k3[NS];
        const int s1 = taps_global<T, NS>(sp.pos[i], kp, k1);
        int s2 = 0, s3 = 0;
        if (DIM > 1) s2 = taps_global<T, NS>(sp.pos[sp.stride + i], kp, k2);
        if (DIM > 2) s3 = taps_global<T, NS>(sp.pos[2 * sp.stride + i], kp, k3);
        int ix[NS];
#pragma unroll
        for (int a = 0; a < NS; ++a) ix[a] = wrap1(s1 + a, n1);
        for (int t = 0; t < ntr; ++t) {
            const cpx<T> cv = c[(long long)t * M + j];
            cpx<T> *f = fw + (long long)t * nftot;
#pragma unroll
            for (int dz = 0; dz < (DIM > 2 ? NS : 1); ++dz) {
                const int iz = DIM > 2 ? wrap1(s3 + dz, n3) : 0;
                const T kz = DIM > 2 ? k3[dz] : (T)1;
#pragma unroll
                for (int dy = 0; dy < (DIM > 1 ? NS : 1); ++dy) {
                    const int iy = DIM > 1 ? wrap1(s2 + dy, n2) : 0;
                    const T kyz = (DIM > 1 ? k2[dy] : (T)1) * kz;
                    const T cr = cv.x * kyz, ci = cv.y * kyz;
                    cpx<T> *row = f + ((long long)iz * n2 + iy) * n1;
#pragma unroll
                    for (int a = 0; a < NS; ++a) {
                        gadd(&row[ix[a]].x, cr * k1[a]);
                        gadd(&row[ix[a]].y, ci * k1[a]);
                    }
                }
            }
        }
    }
}

// GM / GM-sort interp: one thread per point gathers ns^DIM taps
// (Interp_{2,3}d_NUptsdriven, spreadinterp2d.cu:516-554).
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(256)
k_interp_gm(int M, SortedPts sp, cpx<T> *__restrict__ c, const cpx<T> *__restrict__ fw,
            BinGeom g, KerParams<T> kp, int ntr, long long nftot)
{
    const int n1 = g.nf[0], n2 = g.nf[1], n3 = g.nf[2];
    const JobRange jr = xcd_jobs((M + kPtChunk - 1) / kPtChunk);
    for (long long ch = jr.begin; ch < jr.end; ch += jr.step) {
        const int i = (int)(ch * kPtChunk) + (int)threadIdx.x;
        if (i >= M) continue;
        const int j = sp.idx[i];
        T k1[NS], k2[NS], k3[NS];
        const int s1 = taps_global<T, NS>(sp.pos[i], kp, k1);
        int s2 = 0, s3 = 0;
        if (DIM > 1) s2 = taps_global<T, NS>(sp.pos[sp.stride + i], kp, k2);
        if (DIM > 2) s3 = taps_global<T, NS>(sp.pos[2 * sp.stride + i], kp, k3);
        int ix[NS];
#pragma unroll
        for (int a = 0; a < NS; ++a) ix[a] = wrap1(s1 + a, n1);
        for (int t = 0; t < ntr; ++t) {
            const cpx<T> *f = fw + (long long)t * nftot;
            T ar = 0, ai = 0;
#pragma unroll
            for (int dz = 0; dz < (DIM > 2 ? NS : 1); ++dz) {
                const int iz = DIM > 2 ? wrap1(s3 + dz, n3) : 0;
                const T kz = DIM > 2 ? k3[dz] : (T)1;
#pragma unroll
                for (int dy = 0; dy < (DIM > 1 ? NS : 1); ++dy) {
                    const int iy = DIM > 1 ? wrap1(s2 + dy, n2) : 0;
                    const T kyz = (DIM > 1 ? k2[dy] : (T)1) * kz;
                    const cpx<T> *row = f + ((long long)iz * n2 + iy) * n1;
                    T rr = 0, ri = 0;
#pragma unroll
                    for (int a = 0; a < NS; ++a) {
                        const cpx<T> v = row[ix[a]];
                        rr += v.x * k1[a];
                        ri += v.y * k1[a];
                    }
                    ar += rr * kyz;
                    ai += ri * kyz;
                }
            }
            cpx<T> o;
            o.x = ar;
            o.y = ai;
            c[(long long)t * M + j] = o;
        }
    }
}

// Tile of one subproblem: the bin's cells plus a halo of P = ceil(ns/2) on
// each side; cell 0 of the tile is global index o[d] = bin*bs - P.
template <int DIM, int NS>
struct Tile {
    static constexpr int P = (NS + 1) / 2;
    int tw[3];
    int o[3];
    int n;
    CFA_HD Tile(const BinGeom &g, int b)
    {
        const int bx = b % g.nb[0];
        const int by = (b / g.nb[0]) % g.nb[1];
        const int bz = b / (g.nb[0] * g.nb[1]);
        tw[0] = g.bs[0] + 2 * P;
        tw[1] = DIM > 1 ? g.bs[1] + 2 * P : 1;
        tw[2] = DIM > 2 ? g.bs[2] + 2 * P : 1;
        o[0] = bx * g.bs[0] - P;
        o[1] = DIM > 1 ? by * g.bs[1] - P : 0;
        o[2] = DIM > 2 ? bz * g.bs[2] - P : 0;
        n = tw[0] * tw[1] * tw[2];
    }
    // fine-grid offset of tile cell q (periodic wrap of every axis)
    CFA_HD long long grid_offset(int q, const BinGeom &g) const
    {
        const int qx = q % tw[0];
        const int qy = (q / tw[0]) % tw[1];
        const int qz = q / (tw[0] * tw[1]);
        const int gx = wrapmod(o[0] + qx, g.nf[0]);
        const int gy = DIM > 1 ? wrapmod(o[1] + qy, g.nf[1]) : 0;
        const int gz = DIM > 2 ? wrapmod(o[2] + qz, g.nf[2]) : 0;
        return ((long long)gz * g.nf[1] + gy) * g.nf[0] + gx;
    }
    // Every tile coordinate o + q of every axis lies in [-nf, 2 nf), so one
    // compare each way wraps it (false only for tiles wider than the grid).
    CFA_HD bool walkable(const BinGeom &g) const
    {
        bool ok = o[0] >= -g.nf[0] && o[0] + tw[0] <= 2 * g.nf[0];
        if (DIM > 1) ok = ok && o[1] >= -g.nf[1] && o[1] + tw[1] <= 2 * g.nf[1];
        if (DIM > 2) ok = ok && o[2] >= -g.nf[2] && o[2] + tw[2] <= 2 * g.nf[2];
        return ok;
    }
};

// Calls f(q, fine-grid offset of tile cell q) for q = q0, q0 + step, ... <
// tl.n.  On walkable tiles the cell's tile coordinates advance by the step's
// decomposition with carries and wrap with compares, so the loop runs no
// integer division (six per cell in grid_offset, which dominated the index
// arithmetic of the tile loads and flushes).
template <int DIM, int NS, class F>
CFA_HD void for_tile_cells(const Tile<DIM, NS> &tl, const BinGeom &g, int q0,
                                               int step, F &&f)
{
    if (!tl.walkable(g)) {
        for (int q = q0; q < tl.n; q += step) f(q, tl.grid_offset(q, g));
        return;
    }
    const int t0 = tl.tw[0], t1 = tl.tw[1];
    int x = q0 % t0, y = (q0 / t0) % t1, z = q0 / (t0 * t1);
    const int sx = step % t0, sy = (step / t0) % t1, sz = step / (t0 * t1);
    for (int q = q0; q < tl.n; q += step) {
        int gx = tl.o[0] + x;
        gx += gx < 0 ? g.nf[0] : 0;
        gx -= gx >= g.nf[0] ? g.nf[0] : 0;
        int gy = 0, gz = 0;
        if (DIM > 1) {
            gy = tl.o[1] + y;
            gy += gy < 0 ? g.nf[1] : 0;
            gy -= gy >= g.nf[1] ? g.nf[1] : 0;
        }
        if (DIM > 2) {
            gz = tl.o[2] + z;
            gz += gz < 0 ? g.nf[2] : 0;
            gz -= gz >= g.nf[2] ? g.nf[2] : 0;
        }
        f(q, ((long long)gz * g.nf[1] + gy) * g.nf[0] + gx);
        // x, y < their widths before the step, so one carry each
        x += sx;
        y += sy;
        z += sz;
        if (x >= t0) {
            x -= t0;
            ++y;
        }
        if (y >= t1) {
            y -= t1;
            ++z;
        }
    }
}

struct SubprobJob {
    int t, b, p0, np;
};

CFA_HD SubprobJob subprob_job(long long job, int total, const SubprobArgs &sa)
{
    SubprobJob J;
    J.t = (int)(job / total);
    const int s = (int)(job - (long long)J.t * total);
    J.b = sa.s2b[s];
    const int kk = s - sa.substart[J.b];
    J.p0 = sa.binstart[J.b] + kk * sa.maxsub;
    J.np = min(sa.maxsub, sa.binsize[J.b] - kk * sa.maxsub);
    return J;
}

// Position of job's first point in the sorted points of all transforms
// (transform-major), for the work estimate of the job scheduler.
__device__ __forceinline__ long long subprob_pos(long long job, int total, int M,
                                                 const SubprobArgs &sa)
{
    const int t = (int)(job / total);
    const int s = (int)(job - (long long)t * total);
    const int b = sa.s2b[s];
    return (long long)t * M + sa.binstart[b] + (long long)(s - sa.substart[b]) * sa.maxsub;
}

// Sum over the NW waves of the workgroup (returned to all); scratch holds NW
// values.
template <typename T, int NW>
__device__ __forceinline__ T block_sum(T v, T *scratch)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) scratch[wid] = v;
    __syncthreads();
    T s = scratch[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) s += scratch[w];
    __syncthreads();
    return s;
}
template <typename T>
__device__ __forceinline__ T block_sum256(T v, T *scratch)
{
    return block_sum<T, 4>(v, scratch);
}

// Fixed-point accumulator widths: headroom bits of the tile scale, the
// round-to-nearest conversion and the LDS atomic (ds_add_u32 / ds_add_u64;
// two's-complement sums are exact modulo 2^bits, and the scale keeps every
// final cell value inside the signed range).
template <typename Acc> struct FxAcc;
template <> struct FxAcc<int> {
    static constexpr int kBits = 30;
    // Round to nearest even (rndne + cvt).  The one-instruction
    // floor(v + 1/2) (v_cvt_rpi_i32_f32) rounds exact halves up: a bias of
    // ~2^-25 of each tap's value, always of one sign, which n taps on one cell
    // sum coherently against the sqrt(n) growth of the cell value -- measured
    // 4e-6 relative on 2D points packed ~15000 to a cell, twice the tol/5 bar
    __device__ static int cvt(float v) { return __float2int_rn(v); }
    __device__ static void add(int *p, int v) { atomicAdd(p, v); }
};
template <> struct FxAcc<long long> {
    static constexpr int kBits = 62;
    __device__ static long long cvt(float v) { return __float2ll_rn(v); }
    __device__ static long long cvt(double v) { return __double2ll_rn(v); }
    __device__ static void add(long long *p, long long v)
    {
        atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v);
    }
};

// Kernel values phi(x1 + i) / e^beta (peak 1) of the ns taps of one axis from
// the offset x1 = start - v of the first tap (the polynomial table when it
// was fitted and x1 lies in its interval [-ns/2, -ns/2 + 1), i.e. the tap
// window was not clamped to the tile).
template <typename T, int NS>
__device__ __forceinline__ void taps_from_offset(T x1, const KerParams<T> &kp, T *ker)
{
    if (kp.hdeg > 0 && x1 >= -kp.hw && x1 < -kp.hw + T(1)) {
        horner_taps<T, NS>(T(2) * x1 + T(NS - 1), kp.hc + (kp.hdeg + 1) * NS, kp.hdeg, ker);
        return;
    }
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        const T t = x1 + (T)i;
        const T a = fmax(T(1) - kp.c * t * t, T(0));
        if constexpr (sizeof(T) == 4)
            ker[i] = fabsf(t) < kp.hw ? exp_beta(sqrtf(a) - 1.0f, kp) : 0.0f;
        else
            ker[i] = fabs(t) < kp.hw ? exp(kp.beta * (sqrt(a) - T(1))) : T(0);
    }
}

// Tap window of a point at grid coordinate v along one axis of a tile whose
// cell 0 is global index o and which is tw cells wide: the first tap
// ceil(v - ns/2), made tile-local and clamped into [0, tw - ns] (a tap
// outside the tile is possible only within rounding of the kernel edge,
// where the kernel is 0); x1 receives the offset (o + l) - v of the window's
// first tap, at which the kernel values are evaluated.
template <int NS>
CFA_HD int tile_window(double v, int o, int tw, double &x1)
{
    const int l0 = (int)ceil(v - 0.5 * NS) - o;
    const int l = l0 < 0 ? 0 : (l0 > tw - NS ? tw - NS : l0);
    x1 = (double)(o + l) - v;
    return l;
}

// Slot of the rank-th point of residue key in the round-robin order over nk
// residues with counts cnt: the points of earlier rounds, then those of
// earlier residues in round rank.
CFA_HD int rr_slot(const int *cnt, int nk, int key, int rank)
{
    int slot = 0;
    for (int k = 0; k < nk; ++k) {
        const int ck = cnt[k];
        slot += (ck < rank ? ck : rank) + (k < key && ck > rank ? 1 : 0);
    }
    return slot;
}

// rr_slot for the lanes of one wave (every lane of the wave calls it; key and
// rank are those of the lane's point, anything for lanes without one).  The
// counts are read once per lane: lane k < nk holds cnt[k].  With R the
// largest count, rounds r < min(R, 64) are tabulated in lanes: lane r keeps
// S(r) = sum_k min(cnt_k, r) (built as the running sum of the rounds' sizes)
// and the ballot of {k : cnt_k > r}; a point of rank r < 64 then reads its
// round's entries by lane shuffle and counts the residues before its own.
__device__ __forceinline__ int rr_slot_wave(const int *cnt, int nk, int key, int rank)
{
    const int lane = threadIdx.x & 63;
    const int ck = lane < nk ? cnt[lane] : 0;
    int mx = ck;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) mx = max(mx, __shfl_xor(mx, d, 64));
    const int R = mx < 64 ? mx : 64;
    int myS = 0, runS = 0;
    unsigned long long mym = 0;
    for (int r = 0; r < R; ++r) {
        const unsigned long long m = __ballot(ck > r);
        if (lane == r) {
            myS = runS;
            mym = m;
        }
        runS += __popcll(m);
    }
    const int rr = rank < 64 ? rank : 0;
    const int S = __shfl(myS, rr, 64);
    const unsigned lo = __shfl((unsigned)mym, rr, 64);
    const unsigned hi = __shfl((unsigned)(mym >> 32), rr, 64);
    if (rank >= 64) return rr_slot(cnt, nk, key, rank);
    const unsigned long long m = ((unsigned long long)hi << 32) | lo;
    const unsigned long long below = key > 0 ? m & (~0ull >> (64 - key)) : 0ull;
    return S + __popcll(below);
}

// SM spread of one (subproblem, transform) job (Spread_{2,3}d_Subprob,
// spreadinterp2d.cu:152-231, spreadinterp3d.cu:285-389): the workgroup
// accumulates every tap of the subproblem's points in an LDS tile of the
// bin plus a halo of ceil(ns/2), then flushes the non-zero cells to the fine
// grid with global float atomics, row by row (the halo overlaps neighbouring
// bins).  Two measured properties of the gfx950 LDS shape it.
//
// Integer atomics.  ds_add_u32 runs at the LDS store rate, ds_add_f32 far
// below it, so the tile holds fixed-point re / im planes.  With the kernel
// normalised to phi / e^beta <= 1 per axis, no cell can exceed
// A = sum_p (|re c_p| + |im c_p|) over the points accumulated, so the
// contributions are scaled by the power of two S = 2^(kBits - e), A < 2^e,
// and rounded to nearest: exact sums, an error of at most A 2^-(kBits+1) per
// tap.  64-bit cells (Acc = long long; FP64 plans, and FP32 plans at tight
// tolerances) take one scale for the whole job.  32-bit cells keep that
// error within the plan's budget only for a bounded A, so the job runs in
// chunks of BT points, at most the plan's 32-bit bound (536 at tol 1e-5):
// each chunk takes its own scale, and between chunks the tile is drained
// into FP32 planes (plain LDS reads and writes, each thread on its own
// cells).  The flush multiplies by e^(DIM beta) / S in FP64 (in FP32 the
// rounding of DIM beta ~ 50 alone would scale the tile by up to 2e-6).
//
// Bank-balanced order.  Tap (dz, dy, a) of a point whose first tap is tile
// cell s lands on s + (dz tw1 + dy) tw0 + a, an offset common to all points,
// so two lanes of one wave-wide atomic collide on an LDS bank exactly when
// their start cells agree modulo the bank count, and then they collide at
// every tap (random start cells: 8-14 lane-ops per clock instead of the
// full rate).  Each chunk is therefore staged in LDS in round-robin order of
// the residue s mod NK (NK = 64 for 4-byte cells, 32 for 8-byte cells):
// round r holds the r-th point of every residue that has one, in residue
// order, so the lanes of a wave carry distinct residues except where rounds
// run short.  Staging also evaluates the tap windows once (start cell and
// the first-tap offset of each axis), so the tap pass reads LDS only.
template <typename T, typename Acc, int DIM, int NS, int BT>
__device__ __forceinline__ void spread_sm_job(const SubprobJob &J, int M, const SortedPts &sp,
                                              const cpx<T> *__restrict__ c,
                                              cpx<T> *__restrict__ fw, const BinGeom &g,
                                              const KerParams<T> &kp, long long nftot,
                                              unsigned char *smem,
                                              cpx<T> *__restrict__ slot = nullptr)
{
    constexpr bool DRAIN = sizeof(Acc) == 4;
    constexpr int NK = sizeof(Acc) == 4 ? 64 : 32;
    const Tile<DIM, NS> tl(g, J.b);
    const int n = tl.n;
    Acc *sre = reinterpret_cast<Acc *>(smem);
    Acc *sim = sre + n;
    int *rs = reinterpret_cast<int *>(sim + n);
    T *rx = reinterpret_cast<T *>(rs + BT);  // DIM planes of first-tap offsets
    T *rcr = rx + DIM * BT;
    T *rci = rcr + BT;
    int *cnt = reinterpret_cast<int *>(rci + BT);
    T *red = reinterpret_cast<T *>(cnt + 64);
    // drain planes: present in the allocation only when jobs can exceed one
    // chunk (J.np > BT), and touched only then
    float *dre = reinterpret_cast<float *>(red + BT / 64);
    float *dimg = dre + n;
    const int tid = threadIdx.x;
    const int nch = (J.np + BT - 1) / BT;
    for (int p = tid; p < n; p += BT) {
        sre[p] = 0;
        sim[p] = 0;
        if (DRAIN && nch > 1) {
            dre[p] = 0.0f;
            dimg[p] = 0.0f;
        }
    }
    const cpx<T> *cj = c + (long long)J.t * M;
    int sh = 0;
    if (!DRAIN) {  // one scale for the whole job: its strength mass first
        T part = 0;
        for (int p = tid; p < J.np; p += BT) {
            const cpx<T> v = cj[sp.idx[J.p0 + p]];
            part += fabs(v.x) + fabs(v.y);
        }
        const T A = block_sum<T, BT / 64>(part, red);
        int ex = 0;
        frexp(A, &ex);
        sh = A > T(0) ? FxAcc<Acc>::kBits - ex : 0;
    }
    const double eb = exp((double)DIM * (double)kp.beta);
    for (int ch = 0; ch < nch; ++ch) {
        const int nc = min(BT, J.np - ch * BT);
        if (tid < NK) cnt[tid] = 0;
        __syncthreads();  // orders the zeroing / the previous drain and tap pass
        int key = 0, rank = 0, s = 0;
        T x1[3] = {T(0), T(0), T(0)};
        cpx<T> cv;
        cv.x = T(0);
        cv.y = T(0);
        if (tid < nc) {
            const int i = J.p0 + ch * BT + tid;
            cv = cj[sp.idx[i]];
            int l[3] = {0, 0, 0};
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                double off;
                l[d] = tile_window<NS>(sp.pos[d * sp.stride + i], tl.o[d], tl.tw[d], off);
                x1[d] = (T)off;
            }
            s = l[0] + tl.tw[0] * (l[1] + tl.tw[1] * l[2]);
            key = s & (NK - 1);
            rank = atomicAdd(&cnt[key], 1);
        }
        if (DRAIN) {  // per-chunk scale; the reduction also completes the counts
            const T A = block_sum<T, BT / 64>(tid < nc ? fabs(cv.x) + fabs(cv.y) : T(0), red);
            int ex = 0;
            frexp(A, &ex);
            sh = A > T(0) ? FxAcc<Acc>::kBits - ex : 0;
        } else {
            __syncthreads();
        }
        // Slot from a per-wave table of the rounds: lane r holds S(r), the
        // points of all rounds before r, and the mask of the residues with a
        // point in round r (one ballot per round, rounds < 64); a rank of 64
        // or more (a chunk crowded onto few residues) walks the counts.
        const int slot = (tid & ~63) < nc ? rr_slot_wave(cnt, NK, key, tid < nc ? rank : 0) : 0;
        if (tid < nc) {
            rs[slot] = s;
#pragma unroll
            for (int d = 0; d < DIM; ++d) rx[d * BT + slot] = x1[d];
            rcr[slot] = cv.x;
            rci[slot] = cv.y;
        }
        __syncthreads();
        if (tid < nc) {
            const int s0 = rs[tid];
            T k1[NS], k2[NS], k3[NS];
            taps_from_offset<T, NS>(rx[tid], kp, k1);
            if (DIM > 1) taps_from_offset<T, NS>(rx[BT + tid], kp, k2);
            if (DIM > 2) taps_from_offset<T, NS>(rx[2 * BT + tid], kp, k3);
            const T sr = ldexp(rcr[tid], sh), si = ldexp(rci[tid], sh);
#pragma unroll
            for (int dz = 0; dz < (DIM > 2 ? NS : 1); ++dz) {
                const T kz = DIM > 2 ? k3[dz] : T(1);
#pragma unroll
                for (int dy = 0; dy < (DIM > 1 ? NS : 1); ++dy) {
                    const T kyz = (DIM > 1 ? k2[dy] : T(1)) * kz;
                    const T cr = sr * kyz, ci = si * kyz;
                    // one base per row: the taps are constant LDS offsets
                    Acc *pr = sre + (s0 + (dz * tl.tw[1] + dy) * tl.tw[0]);
                    Acc *pi = pr + n;
#pragma unroll
                    for (int a = 0; a < NS; ++a) {
                        FxAcc<Acc>::add(pr + a, FxAcc<Acc>::cvt(cr * k1[a]));
                        FxAcc<Acc>::add(pi + a, FxAcc<Acc>::cvt(ci * k1[a]));
                    }
                }
            }
        }
        if (DRAIN && ch + 1 < nch) {
            __syncthreads();
            // 32-bit cells convert exactly to FP32 up to 2^24 and to within
            // one rounding beyond, like the FP32 factor: FP32 drains and
            // flushes match the FP32 grid they feed
            const float bf = (float)ldexp(eb, -sh);
            for (int p = tid; p < n; p += BT) {
                const Acc ir = sre[p], ii = sim[p];
                dre[p] += (float)ir * bf;
                dimg[p] += (float)ii * bf;
                sre[p] = 0;
                sim[p] = 0;
            }
        }
    }
    __syncthreads();
    const double back = ldexp(eb, -sh);
    cpx<T> *f = fw + (long long)J.t * nftot;
    if (slot) {
        // the bin's only subproblem: store the whole tile, in tile order
        // (coalesced plain stores; the gather pass adds it to the grid)
        cpx<T> *sl = slot + ((long long)J.t * g.nb[0] * g.nb[1] * g.nb[2] + J.b) * n;
        const float bf = (float)back;
        const bool drained = DRAIN && nch > 1;
        for (int p = tid; p < n; p += BT) {
            cpx<T> v;
            if constexpr (DRAIN) {
                v.x = (float)sre[p] * bf;
                v.y = (float)sim[p] * bf;
                if (drained) {
                    v.x += dre[p];
                    v.y += dimg[p];
                }
            } else {
                v.x = (T)((double)sre[p] * back);
                v.y = (T)((double)sim[p] * back);
            }
            sl[p] = v;
        }
    } else if constexpr (DRAIN) {
        const float bf = (float)back;
        const bool drained = nch > 1;
        for_tile_cells(tl, g, tid, BT, [&](int p, long long off) {
            float vr = (float)sre[p] * bf, vi = (float)sim[p] * bf;
            if (drained) {
                vr += dre[p];
                vi += dimg[p];
            }
            if (vr != 0.0f || vi != 0.0f) {
                cpx<T> *q = f + off;
                gadd(&q->x, (T)vr);
                gadd(&q->y, (T)vi);
            }
        });
    } else {
        for_tile_cells(tl, g, tid, BT, [&](int p, long long off) {
            const double vr = (double)sre[p] * back, vi = (double)sim[p] * back;
            if (vr != 0.0 || vi != 0.0) {
                cpx<T> *q = f + off;
                gadd(&q->x, (T)vr);
                gadd(&q->y, (T)vi);
            }
        });
    }
    __syncthreads();
}

// SM spread kernel: persistent workgroups of BT threads take jobs from the
// per-XCD queues.  Acc = int: 32-bit cells in chunks of BT points (BT within
// the plan's 32-bit bound); Acc = long long: 64-bit cells.
template <typename T, typename Acc, int DIM, int NS, int BT>
__global__ void __launch_bounds__(BT)
k_spread_sm(int M, SortedPts sp, const cpx<T> *__restrict__ c, cpx<T> *__restrict__ fw,
            BinGeom g, KerParams<T> kp, SubprobArgs sa, int ntr, long long nftot,
            unsigned long long *jq, cpx<T> *__restrict__ slot)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int total = sa.counters[1];
    jobq_run(
        jq, (long long)total * ntr,
        [&](long long job) {
            const SubprobJob J = subprob_job(job, total, sa);
            // a bin of one subproblem stores its tile to its slot (when the
            // plan has them); bins of several keep adding into the grid
            cpx<T> *sl = slot && sa.substart[J.b + 1] - sa.substart[J.b] == 1 ? slot : nullptr;
            spread_sm_job<T, Acc, DIM, NS, BT>(J, M, sp, c, fw, g, kp, nftot, smem, sl);
        },
        [&](long long job) { return subprob_pos(job, total, M, sa); }, (long long)M * ntr);
}

// Tiles covering grid index gc along one axis of nb bins of bs cells (grid
// nf, halo P, tile bs + 2P wide, tile of bin u starting at u bs - P): bins
// up to two away from the cell's own (two, because the halo of the bin
// before a partial last bin narrower than P reaches past it, across the
// period), a wrapped bin's tile one period away.  Writes up to five (bin,
// tile coordinate) pairs, returns how many.  Complete (every tile cell a
// point can reach that maps to gc is listed, none twice) when nb >= 3,
// P <= bs and the tile is no wider than the grid; the far end of a partial
// last bin's tile, which no point reaches, may be left out.
CFA_HD int tile_cover_axis(int gc, int nb, int bs, int nf, int P, int *cb, int *cq)
{
    const int tw = bs + 2 * P;
    const int b = gc / bs < nb - 1 ? gc / bs : nb - 1;
    int n = 0;
    for (int k = -2; k <= 2; ++k) {
        int u = b + k, shift = 0;
        if (u < 0) {
            u += nb;
            shift = -nf;
        } else if (u >= nb) {
            u -= nb;
            shift = nf;
        }
        const int q = gc - (u * bs - P + shift);
        if (q >= 0 && q < tw) {
            cb[n] = u;
            cq[n] = q;
            ++n;
        }
    }
    return n;
}

// Tile gather after an SM spread with slots: every fine-grid cell adds the
// stored tiles of the single-subproblem bins that cover it (tile_cover_axis
// along each axis), in a fixed order.  Cells of bins with several
// subproblems already hold those bins' atomic flushes.  One one-wave
// workgroup per grid row (y, z) and transform: the row's covers along y and
// z are found once, and the lanes walk x, so the wave reads runs of one tile
// row per covering bin and the grid row in order.
template <typename T, int DIM>
__global__ void __launch_bounds__(64)
k_tile_gather(cpx<T> *__restrict__ fw, const cpx<T> *__restrict__ slot, BinGeom g, int P,
              const int *__restrict__ substart, long long nftot)
{
    const long long nbins = (long long)g.nb[0] * g.nb[1] * g.nb[2];
    const int tw0 = g.bs[0] + 2 * P;
    const int tw1 = DIM > 1 ? g.bs[1] + 2 * P : 1;
    const int tw2 = DIM > 2 ? g.bs[2] + 2 * P : 1;
    const long long tcells = (long long)tw0 * tw1 * tw2;
    const int rows = (DIM > 1 ? g.nf[1] : 1) * (DIM > 2 ? g.nf[2] : 1);
    const int t = blockIdx.x / rows;
    const int row = blockIdx.x - t * rows;
    const int gy = DIM > 1 ? row % g.nf[1] : 0;
    const int gz = DIM > 2 ? row / g.nf[1] : 0;
    int yb[5], yq[5], zb[5], zq[5];
    int ny = 1, nz = 1;
    yb[0] = yq[0] = zb[0] = zq[0] = 0;
    if (DIM > 1) ny = tile_cover_axis(gy, g.nb[1], g.bs[1], g.nf[1], P, yb, yq);
    if (DIM > 2) nz = tile_cover_axis(gz, g.nb[2], g.bs[2], g.nf[2], P, zb, zq);
    // which candidate bins of the row hold one subproblem: a byte per
    // (y/z cover, x bin) in LDS when they fit, read back without a memory
    // round trip, so each cell's slot loads (all x covers of a y/z cover,
    // under a lane mask) go out together instead of one chain per tile
    __shared__ unsigned char sflag[4096];
    const int nbx = g.nb[0];
    const bool lflag = ny * nz * nbx <= 4096;
    if (lflag) {
        for (int i = threadIdx.x; i < ny * nz * nbx; i += blockDim.x) {
            const int combo = i / nbx, bx = i - combo * nbx;
            const long long bin =
                bx + (long long)nbx * (yb[combo % ny] + (long long)g.nb[1] * zb[combo / ny]);
            sflag[i] = substart[bin + 1] - substart[bin] == 1;
        }
        __syncthreads();
    }
    cpx<T> *f = fw + (long long)t * nftot + (long long)row * g.nf[0];
    const cpx<T> *sl = slot + (long long)t * nbins * tcells;
    for (int gx = threadIdx.x; gx < g.nf[0]; gx += blockDim.x) {
        int xb[5], xq[5];
        const int nx = tile_cover_axis(gx, nbx, g.bs[0], g.nf[0], P, xb, xq);
        cpx<T> acc = f[gx];
        for (int iz = 0; iz < nz; ++iz)
            for (int iy = 0; iy < ny; ++iy) {
                const long long brow = (long long)nbx * (yb[iy] + (long long)g.nb[1] * zb[iz]);
                const long long qrow = (long long)tw0 * (yq[iy] + (long long)tw1 * zq[iz]);
                const int frow = (iz * ny + iy) * nbx;
                cpx<T> v[5];
#pragma unroll
                for (int ix = 0; ix < 5; ++ix) {
                    v[ix].x = T(0);
                    v[ix].y = T(0);
                    if (ix < nx) {
                        const long long bin = brow + xb[ix];
                        const bool single = lflag ? sflag[frow + xb[ix]] != 0
                                                  : substart[bin + 1] - substart[bin] == 1;
                        if (single) v[ix] = sl[bin * tcells + qrow + xq[ix]];
                    }
                }
#pragma unroll
                for (int ix = 0; ix < 5; ++ix) {
                    acc.x += v[ix].x;
                    acc.y += v[ix].y;
                }
            }
        f[gx] = acc;
    }
}

// One member of a plan group (many independent plans of one type, dimension,
// precision and kernel width executed together): everything a subproblem job
// of that plan needs, in device memory.
template <typename T>
struct GroupItem {
    BinGeom g;
    SortedPts sp;
    SubprobArgs sa;
    const cpx<T> *in;  // c (type 1) or fine grid (type 2)
    cpx<T> *out;       // fine grid (type 1) or c (type 2)
    long long nftot;
    int M;
    int ntr;
};

// Job offsets of a plan group: off[p] = sum over q < p of subproblems(q) *
// ntr(q), computed on the device from each plan's setpts counters (so the
// whole group launch sequence needs no host round trip and can be captured).
template <typename T>
__global__ void k_group_offsets(const GroupItem<T> *__restrict__ items, int n,
                                long long *__restrict__ off)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        long long acc = 0;
        for (int p = 0; p < n; ++p) {
            off[p] = acc;
            acc += (long long)items[p].sa.counters[1] * items[p].ntr;
        }
        off[n] = acc;
    }
}

__device__ __forceinline__ int group_member(const long long *off, int n, long long job)
{
    int lo = 0, hi = n - 1;  // last p with off[p] <= job
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= job) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// SM spread of a whole plan group in one persistent launch over the global
// job queue of all members (256-thread workgroups).
template <typename T, typename Acc, int DIM, int NS>
__global__ void __launch_bounds__(256)
k_spread_sm_group(const GroupItem<T> *__restrict__ items, int n,
                  const long long *__restrict__ off, KerParams<T> kp, unsigned long long *jq)
{
    extern __shared__ __align__(16) unsigned char smem[];
    jobq_run(jq, off[n], [&](long long job) {
        const int p = group_member(off, n, job);
        const GroupItem<T> &it = items[p];
        const SubprobJob J = subprob_job(job - off[p], it.sa.counters[1], it.sa);
        spread_sm_job<T, Acc, DIM, NS, 256>(J, it.M, it.sp, it.in, it.out, it.g, kp, it.nftot,
                                            smem);
    });
}

// ===========================================================================
// Method 4: owner-computes block spread.  The role is that of the reference's
// block-gather method (spread3d_wrapper.cu:393-889: output bins, points
// replicated into ghost regions, no halo atomics); the design here is this
// library's own.  The fine grid is cut into blocks of ob_x x ob_y x ob_z
// cells (nf % ob == 0, as the reference requires).  setpts places one copy
// of each point in every block that its ns-wide support touches, sorted by
// (block, in-block cell) so that a block's copies are in cell order; each
// copy stores its position shifted by a multiple of nf into the block's own
// periodic frame.  execute accumulates, per block, only the taps inside the
// block in a halo-free LDS tile and writes the block with plain stores when
// one subproblem covers it (no other job touches those cells), else adds
// the bounding box its taps touched with global atomics.  Inside a block the
// copies stay in arrival order: cell order would put the 64 lanes of a wave
// on the same cells in dense clusters, and same-address LDS atomics
// serialise (measured: 15x the address conflicts of arrival order).
// ===========================================================================

CFA_HD int floor_div(int a, int b)
{
    return a >= 0 ? a / b : -((-a + b - 1) / b);
}

// Along one axis: first tap and the unwrapped block frames kb0..kb1 touched.
CFA_HD void blk_frames(double v, int ns, int ob, int &start, int &kb0,
                                           int &kb1)
{
    start = (int)ceil(v - 0.5 * ns);
    kb0 = floor_div(start, ob);
    kb1 = floor_div(start + ns - 1, ob);
}

// Largest number of copies of one point: blocks touched per axis.
__host__ __device__ inline int blk_max_copies(int dim, int ns, const int *ob)
{
    int c = 1;
    for (int d = 0; d < dim; ++d) c *= (ns - 2 + ob[d]) / ob[d] + 1;
    return c;
}

// Visits every copy of point i: f(block, axis positions shifted into the
// block's frame).
template <int DIM, typename F>
__device__ __forceinline__ void blk_copies(const double *v, int ns, const BinGeom &g, F f)
{
    int st[3] = {0, 0, 0}, k0[3] = {0, 0, 0}, k1[3] = {0, 0, 0};
    for (int d = 0; d < DIM; ++d) blk_frames(v[d], ns, g.bs[d], st[d], k0[d], k1[d]);
    for (int kz = k0[2]; kz <= k1[2]; ++kz)
        for (int ky = k0[1]; ky <= k1[1]; ++ky)
            for (int kx = k0[0]; kx <= k1[0]; ++kx) {
                const int kb[3] = {kx, ky, kz};
                int B[3] = {0, 0, 0};
                double sv[3] = {0, 0, 0};
                for (int d = 0; d < DIM; ++d) {
                    B[d] = wrapmod(kb[d], g.nb[d]);
                    sv[d] = v[d] + (double)(B[d] - kb[d]) * g.bs[d];  // exact: a multiple of nf
                }
                f(B[0] + g.nb[0] * (B[1] + g.nb[1] * B[2]), sv);
            }
}

template <typename T, int DIM>
__global__ void __launch_bounds__(256)
k_blk_count(int M, const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
            BinGeom g, int ns, int *__restrict__ blkcount)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        double v[3];
        v[0] = tap_pos(rescale_exact(x[i], g.nf[0]), g.nf[0]);
        if (DIM > 1) v[1] = tap_pos(rescale_exact(y[i], g.nf[1]), g.nf[1]);
        if (DIM > 2) v[2] = tap_pos(rescale_exact(z[i], g.nf[2]), g.nf[2]);
        blk_copies<DIM>(v, ns, g, [&](int blk, const double *) { atomicAdd(&blkcount[blk], 1); });
    }
}

// Scatter of the copies: slot = blkstart[blk] + (remaining count - 1); the
// counts are consumed (each reaches 0).
template <typename T, int DIM>
__global__ void __launch_bounds__(256)
k_blk_scatter(int M, const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
              BinGeom g, int ns, const int *__restrict__ blkstart, int *__restrict__ blkcount,
              int *__restrict__ idx, double *__restrict__ pos, long long stride)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x) {
        double v[3];
        v[0] = tap_pos(rescale_exact(x[i], g.nf[0]), g.nf[0]);
        if (DIM > 1) v[1] = tap_pos(rescale_exact(y[i], g.nf[1]), g.nf[1]);
        if (DIM > 2) v[2] = tap_pos(rescale_exact(z[i], g.nf[2]), g.nf[2]);
        blk_copies<DIM>(v, ns, g, [&](int blk, const double *sv) {
            const int s = blkstart[blk] + atomicSub(&blkcount[blk], 1) - 1;
            idx[s] = i;
            for (int d = 0; d < DIM; ++d) pos[d * stride + s] = sv[d];
        });
    }
}

// Kernel values of one axis at the taps start..start+NS-1 of a point at v,
// optionally divided by e^beta (peak 1, for the fixed-point tiles).
template <typename T, int NS, bool NORM>
__device__ __forceinline__ int taps_blk(double v, const KerParams<T> &kp, T *ker)
{
    const int start = (int)ceil(v - 0.5 * NS);
    if (kp.hdeg > 0) {
        horner_taps<T, NS>((T)(2.0 * ((double)start - v) + (NS - 1)),
                           NORM ? kp.hc + (kp.hdeg + 1) * NS : kp.hc, kp.hdeg, ker);
        return start;
    }
    const T x1 = (T)((double)start - v);
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        const T t = x1 + (T)i;
        const T a = fmax(T(1) - kp.c * t * t, T(0));
        ker[i] = fabs(t) < kp.hw ? exp(kp.beta * (NORM ? sqrt(a) - T(1) : sqrt(a))) : T(0);
    }
    return start;
}

// LDS bytes of a method-4 job: re / im accumulators per block cell (Acc of
// accsize bytes), reduction scratch and the bounding box.  Strengths are not
// staged (the fixed-point scale pass gathers them once more, from L2), so
// large subproblems cost no LDS.
__host__ __device__ inline size_t blk_lds(size_t cells, int accsize)
{
    return cells * 2 * accsize + 48;
}

template <typename Acc> struct BlkAcc;
template <> struct BlkAcc<double> {
    __device__ static double cvt(double v) { return v; }
    __device__ static void add(double *p, double v) { ladd(p, v); }
};
template <> struct BlkAcc<int> {
    __device__ static int cvt(float v) { return __float2int_rn(v); }
    __device__ static void add(int *p, int v) { atomicAdd(p, v); }
};
template <> struct BlkAcc<long long> {
    __device__ static long long cvt(float v) { return __float2ll_rn(v); }
    __device__ static void add(long long *p, long long v)
    {
        atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v);
    }
};

// One method-4 job: the copies J.p0 .. J.p0+J.np-1 of block J.b.  Acc is
// double for FP64 plans, int / long long (fixed point, as in the SM spread)
// for FP32 plans.
template <typename T, typename Acc, int DIM, int NS>
__device__ __forceinline__ void spread_blk_job(const SubprobJob &J, int nsub, int M,
                                               const SortedPts &sp,
                                               const cpx<T> *__restrict__ c,
                                               cpx<T> *__restrict__ fw, const BinGeom &g,
                                               const KerParams<T> &kp, long long nftot,
                                               unsigned char *smem, int maxsub)
{
    constexpr bool FX = sizeof(T) == 4;
    const int wx = g.bs[0], wy = g.bs[1], wz = g.bs[2];
    const int ox = (J.b % g.nb[0]) * wx;
    const int oy = ((J.b / g.nb[0]) % g.nb[1]) * wy;
    const int oz = (J.b / (g.nb[0] * g.nb[1])) * wz;
    const int n = wx * wy * wz;
    Acc *sre = reinterpret_cast<Acc *>(smem);
    Acc *sim = sre + n;
    float *red = reinterpret_cast<float *>(sim + n);
    int *box = reinterpret_cast<int *>(red + 4);
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        sre[q] = 0;
        sim[q] = 0;
    }
    if (threadIdx.x < 3) {
        box[threadIdx.x] = 1 << 30;      // min corner
        box[3 + threadIdx.x] = -1;       // max corner
    }
    int sh = 0;
    if constexpr (FX) {
        float part = 0.0f;
        for (int q = threadIdx.x; q < J.np; q += blockDim.x) {
            const cpx<float> cv = c[(long long)J.t * M + sp.idx[J.p0 + q]];
            part += fabsf(cv.x) + fabsf(cv.y);
        }
        const float A = block_sum256(part, red);  // also orders the tile zeroing
        int ex = 0;
        frexpf(A, &ex);
        sh = A > 0.0f ? (sizeof(Acc) == 4 ? 30 : 62) - ex : 0;
    } else {
        __syncthreads();
    }
    const int w[3] = {wx, wy, wz}, o[3] = {ox, oy, oz};
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};  // touched cells
    for (int q = threadIdx.x; q < J.np; q += blockDim.x) {
        const int i = J.p0 + q;
        T k[3][NS];
        int l[3] = {0, 0, 0}, a0[3] = {0, 0, 0}, a1[3] = {1, 1, 1};
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            l[d] = taps_blk<T, NS, FX>(sp.pos[d * sp.stride + i], kp, k[d]) - o[d];
            a0[d] = l[d] < 0 ? -l[d] : 0;
            a1[d] = w[d] - l[d] < NS ? w[d] - l[d] : NS;
            lo[d] = min(lo[d], l[d] + a0[d]);
            hi[d] = max(hi[d], l[d] + a1[d] - 1);
        }
        const cpx<T> cv = c[(long long)J.t * M + sp.idx[i]];
        T sr = cv.x, si = cv.y;
        if constexpr (FX) {
            sr = ldexpf(cv.x, sh);
            si = ldexpf(cv.y, sh);
        }
#pragma unroll
        for (int dz = 0; dz < (DIM > 2 ? NS : 1); ++dz) {
            if (dz < a0[2] || dz >= a1[2]) continue;
            const T kz = DIM > 2 ? k[2][dz] : T(1);
#pragma unroll
            for (int dy = 0; dy < (DIM > 1 ? NS : 1); ++dy) {
                if (dy < a0[1] || dy >= a1[1]) continue;
                const T kyz = (DIM > 1 ? k[1][dy] : T(1)) * kz;
                const T cr = sr * kyz, ci = si * kyz;
                const int row = ((l[2] + dz) * wy + (l[1] + dy)) * wx + l[0];
#pragma unroll
                for (int a = 0; a < NS; ++a) {
                    if (a >= a0[0] && a < a1[0]) {
                        BlkAcc<Acc>::add(&sre[row + a], BlkAcc<Acc>::cvt(cr * k[0][a]));
                        BlkAcc<Acc>::add(&sim[row + a], BlkAcc<Acc>::cvt(ci * k[0][a]));
                    }
                }
            }
        }
    }
    if (nsub > 1) {
        // bounding box of the touched cells: wave reductions, then one LDS
        // atomic per wave (same-address LDS atomics serialise across lanes)
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                lo[d] = min(lo[d], __shfl_xor(lo[d], off, 64));
                hi[d] = max(hi[d], __shfl_xor(hi[d], off, 64));
            }
            if ((threadIdx.x & 63) == 0) {
                atomicMin(&box[d], lo[d]);
                atomicMax(&box[3 + d], hi[d]);
            }
        }
    }
    __syncthreads();
    // e^(DIM*beta) / 2^sh in FP64 for the fixed-point tiles (see the SM spread)
    const double back = FX ? ldexp(exp((double)DIM * (double)kp.beta), -sh) : 1.0;
    cpx<T> *f = fw + (long long)J.t * nftot;
    if (nsub == 1) {
        // the only job of this block: plain coalesced stores of every cell
        for (int q = threadIdx.x; q < n; q += blockDim.x) {
            const int qx = q % wx, qy = (q / wx) % wy, qz = q / (wx * wy);
            cpx<T> v;
            v.x = (T)((double)sre[q] * back);
            v.y = (T)((double)sim[q] * back);
            f[((long long)(oz + qz) * g.nf[1] + (oy + qy)) * g.nf[0] + ox + qx] = v;
        }
    } else {
        const int x0 = box[0], y0 = DIM > 1 ? box[1] : 0, z0 = DIM > 2 ? box[2] : 0;
        const int bw = box[3] - x0 + 1;
        const int bh = DIM > 1 ? box[4] - y0 + 1 : 1;
        const int bd = DIM > 2 ? box[5] - z0 + 1 : 1;
        const int nb = bw > 0 && bh > 0 && bd > 0 ? bw * bh * bd : 0;
        for (int q = threadIdx.x; q < nb; q += blockDim.x) {
            const int qx = x0 + q % bw, qy = y0 + (q / bw) % bh, qz = z0 + q / (bw * bh);
            const int t = (qz * wy + qy) * wx + qx;
            const Acc ir = sre[t], ii = sim[t];
            if (ir != 0 || ii != 0) {
                cpx<T> *p = f + ((long long)(oz + qz) * g.nf[1] + (oy + qy)) * g.nf[0] + ox + qx;
                gadd(&p->x, (T)((double)ir * back));
                gadd(&p->y, (T)((double)ii * back));
            }
        }
    }
    __syncthreads();
}

template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(256)
k_spread_blk(int M, SortedPts sp, const cpx<T> *__restrict__ c, cpx<T> *__restrict__ fw,
             BinGeom g, KerParams<T> kp, SubprobArgs sa, int ntr, long long nftot)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int total = sa.counters[1];
    const JobRange jr = xcd_jobs((long long)total * ntr);
    for (long long job = jr.begin; job < jr.end; job += jr.step) {
        const SubprobJob J = subprob_job(job, total, sa);
        const int nsub = sa.substart[J.b + 1] - sa.substart[J.b];
        if constexpr (sizeof(T) == 4) {
            if (kp.fxwide)
                spread_blk_job<T, long long, DIM, NS>(J, nsub, M, sp, c, fw, g, kp, nftot, smem,
                                                      sa.maxsub);
            else
                spread_blk_job<T, int, DIM, NS>(J, nsub, M, sp, c, fw, g, kp, nftot, smem,
                                                sa.maxsub);
        } else {
            spread_blk_job<T, double, DIM, NS>(J, nsub, M, sp, c, fw, g, kp, nftot, smem,
                                               sa.maxsub);
        }
    }
}

// SM interp (Interp_{2,3}d_Subprob, spreadinterp2d.cu:596-669): load the
// padded tile of the subproblem's bin into LDS once (complex cells, so one
// tap is one LDS read), then each thread gathers the taps of its points.
template <typename T, int DIM, int NS>
__device__ __forceinline__ void interp_sm_job(const SubprobJob &J, int M, const SortedPts &sp,
                                              cpx<T> *__restrict__ c,
                                              const cpx<T> *__restrict__ fw, const BinGeom &g,
                                              const KerParams<T> &kp, long long nftot,
                                              unsigned char *smem)
{
    const Tile<DIM, NS> tl(g, J.b);
    cpx<T> *tile = reinterpret_cast<cpx<T> *>(smem);
    const cpx<T> *f = fw + (long long)J.t * nftot;
    for_tile_cells(tl, g, threadIdx.x, blockDim.x,
                   [&](int q, long long off) { tile[q] = f[off]; });
    __syncthreads();
    for (int q = threadIdx.x; q < J.np; q += blockDim.x) {
        const int i = J.p0 + q;
        const int j = sp.idx[i];
        T k1[NS], k2[NS], k3[NS];
        const int l1 = taps_local<T, NS>(sp.pos[i], tl.o[0], tl.tw[0], kp, k1);
        int l2 = 0, l3 = 0;
        if (DIM > 1) l2 = taps_local<T, NS>(sp.pos[sp.stride + i], tl.o[1], tl.tw[1], kp, k2);
        if (DIM > 2) l3 = taps_local<T, NS>(sp.pos[2 * sp.stride + i], tl.o[2], tl.tw[2], kp, k3);
        T ar = 0, ai = 0;
#pragma unroll
        for (int dz = 0; dz < (DIM > 2 ? NS : 1); ++dz) {
            const T kz = DIM > 2 ? k3[dz] : (T)1;
#pragma unroll
            for (int dy = 0; dy < (DIM > 1 ? NS : 1); ++dy) {
                const T kyz = (DIM > 1 ? k2[dy] : (T)1) * kz;
                const int row = ((l3 + dz) * tl.tw[1] + (l2 + dy)) * tl.tw[0] + l1;
                T rr = 0, ri = 0;
#pragma unroll
                for (int a = 0; a < NS; ++a) {
                    const cpx<T> v = tile[row + a];
                    rr += v.x * k1[a];
                    ri += v.y * k1[a];
                }
                ar += rr * kyz;
                ai += ri * kyz;
            }
        }
        cpx<T> o;
        o.x = ar;
        o.y = ai;
        c[(long long)J.t * M + j] = o;
    }
    __syncthreads();
}

// BT threads per workgroup: 512 or 1024 (8 or 16 waves over one LDS tile)
// multiply the waves per CU at the same LDS footprint when subproblems hold
// that many points.  The FP64 interp is bound by LDS latency at 1-2 tiles per
// CU (measured: sharing taps between points that read the same window, which
// cuts LDS reads ~40% but needs ~200 VGPRs, ran slower), so waves that hide
// the latency pay more than fewer reads.
template <typename T, int DIM, int NS, int BT>
__global__ void __launch_bounds__(BT)
k_interp_sm(int M, SortedPts sp, cpx<T> *__restrict__ c, const cpx<T> *__restrict__ fw,
            BinGeom g, KerParams<T> kp, SubprobArgs sa, int ntr, long long nftot,
            unsigned long long *jq)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int total = sa.counters[1];
    jobq_run(
        jq, (long long)total * ntr,
        [&](long long job) {
            interp_sm_job<T, DIM, NS>(subprob_job(job, total, sa), M, sp, c, fw, g, kp, nftot,
                                      smem);
        },
        [&](long long job) { return subprob_pos(job, total, M, sa); }, (long long)M * ntr);
}

// SM interp of a whole plan group in one persistent launch.
template <typename T, int DIM, int NS>
__global__ void __launch_bounds__(256)
k_interp_sm_group(const GroupItem<T> *__restrict__ items, int n,
                  const long long *__restrict__ off, KerParams<T> kp, unsigned long long *jq)
{
    extern __shared__ __align__(16) unsigned char smem[];
    jobq_run(jq, off[n], [&](long long job) {
        const int p = group_member(off, n, job);
        const GroupItem<T> &it = items[p];
        const SubprobJob J = subprob_job(job - off[p], it.sa.counters[1], it.sa);
        interp_sm_job<T, DIM, NS>(J, it.M, it.sp, it.out, it.in, it.g, kp, it.nftot, smem);
    });
}
// ===========================================================================
// execute: deconvolve (type 1) / amplify (type 2), one fused pass each
// ===========================================================================

// Frequency of grid index l along an axis with ms modes and nf cells, or
// INT_MIN if l holds no mode (inverse of the w(k) map of deconvolve_wrapper.cu:32-35).
CFA_HD int grid2freq(int l, int ms, int nf)
{
    if (l <= (ms - 1) / 2) return l;
    if (l >= nf - ms / 2) return l - nf;
    return INT32_MIN;
}
CFA_HD int freq2grid(int f, int nf) { return f >= 0 ? f : nf + f; }

// fk[t][k] = fw[t][w(k)] / prod_d fwkerhalf_d[|k_d - ms_d/2|]
// (Deconvolve_{2,3}d, deconvolve_wrapper.cu:26-66).  One workgroup per output
// row (fixed k2, k3, t); threads sweep k1 so both reads and writes coalesce.
template <typename T, int DIM>
__global__ void __launch_bounds__(256)
k_deconvolve(int ms, int mt, int mu, int n1, int n2, int n3, const cpx<T> *__restrict__ fw,
             cpx<T> *__restrict__ fk, const T *__restrict__ ker1, const T *__restrict__ ker2,
             const T *__restrict__ ker3, long long rows, long long nftot)
{
    for (long long r = blockIdx.x; r < rows; r += gridDim.x) {
        const int k2 = DIM > 1 ? (int)(r % mt) : 0;
        const int k3 = DIM > 2 ? (int)((r / mt) % mu) : 0;
        const long long t = r / ((long long)mt * mu);
        const int f2 = k2 - mt / 2, f3 = k3 - mu / 2;
        const T kyz = (DIM > 1 ? ker2[abs(f2)] : (T)1) * (DIM > 2 ? ker3[abs(f3)] : (T)1);
        const cpx<T> *src = fw + t * nftot +
                            ((long long)(DIM > 2 ? freq2grid(f3, n3) : 0) * n2 +
                             (DIM > 1 ? freq2grid(f2, n2) : 0)) *
                                n1;
        cpx<T> *dst = fk + r * ms;
        for (int k1 = threadIdx.x; k1 < ms; k1 += blockDim.x) {
            const int f1 = k1 - ms / 2;
            const T kv = ker1[abs(f1)] * kyz;
            const cpx<T> v = src[freq2grid(f1, n1)];
            cpx<T> o;
            o.x = v.x / kv;
            o.y = v.y / kv;
            dst[k1] = o;
        }
    }
}

// Type-2 step 1 as one pass over the fine grid: every cell is written once,
// with fk[k]/prod fwkerhalf at mode cells and 0 elsewhere (Amplify_{2,3}d plus
// the memset before it, deconvolve_wrapper.cu:81-121,181-182).  With wax >= 0
// the modes are also multiplied by i * wsign * k_wax (k_wax the frequency
// along axis wax): the type-2 transform of those modes is the derivative of
// c_j = sum_k f_k e^{i wsign k.x_j} with respect to x_j along that axis.
template <typename T, int DIM>
__global__ void __launch_bounds__(256)
k_amplify(int ms, int mt, int mu, int n1, int n2, int n3, cpx<T> *__restrict__ fw,
          const cpx<T> *__restrict__ fk, const T *__restrict__ ker1, const T *__restrict__ ker2,
          const T *__restrict__ ker3, long long rows, long long ntot, int wax, int wsign)
{
    for (long long r = blockIdx.x; r < rows; r += gridDim.x) {
        const int l2 = DIM > 1 ? (int)(r % n2) : 0;
        const int l3 = DIM > 2 ? (int)((r / n2) % n3) : 0;
        const long long t = r / ((long long)n2 * n3);
        const int f2 = DIM > 1 ? grid2freq(l2, mt, n2) : 0;
        const int f3 = DIM > 2 ? grid2freq(l3, mu, n3) : 0;
        cpx<T> *dst = fw + r * n1;
        cpx<T> zero;
        zero.x = 0;
        zero.y = 0;
        if (f2 == INT32_MIN || f3 == INT32_MIN) {
            for (int l1 = threadIdx.x; l1 < n1; l1 += blockDim.x) dst[l1] = zero;
            continue;
        }
        const T kyz = (DIM > 1 ? ker2[abs(f2)] : (T)1) * (DIM > 2 ? ker3[abs(f3)] : (T)1);
        const cpx<T> *src =
            fk + t * ntot + ((long long)(f3 + mu / 2) * mt + (f2 + mt / 2)) * (long long)ms;
        // weight i * wsign * k of a fixed row axis (y or z), or per cell (x)
        const T wrow = wax == 1 ? (T)(wsign * f2) : (wax == 2 ? (T)(wsign * f3) : (T)1);
        for (int l1 = threadIdx.x; l1 < n1; l1 += blockDim.x) {
            const int f1 = grid2freq(l1, ms, n1);
            cpx<T> o = zero;
            if (f1 != INT32_MIN) {
                const T kv = ker1[abs(f1)] * kyz;
                const cpx<T> v = src[f1 + ms / 2];
                o.x = v.x / kv;
                o.y = v.y / kv;
                if (wax >= 0) {
                    const T w = wax == 0 ? (T)(wsign * f1) : wrow;
                    const T re = -o.y * w;
                    o.y = o.x * w;
                    o.x = re;
                }
            }
            dst[l1] = o;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_zero(cpx<T> *__restrict__ p, long long n)
{
    cpx<T> zero;
    zero.x = 0;
    zero.y = 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = zero;
}

}  // namespace cfa
