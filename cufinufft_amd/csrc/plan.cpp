// Plan lifecycle and C ABI of the MI355X NUFFT library: makeplan / setpts /
// execute / destroy for single and double precision (include/cufinufft.h).
//
// Host responsibilities: spreader parameters (in the plan precision, as the
// reference's setup_spreader computes them), fine-grid sizes, the Fourier
// series of the ES kernel (FP64 Gauss-Legendre quadrature), device buffers,
// the hipFFT plans and the stage sequence of each execute.  All device work
// is asynchronous on the plan's stream.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/cufinufft.h"
#include "nufft_device.cuh"
#include "nufft_launch.h"

namespace cfa {

// Error codes of contrib/utils.h:28-36.
constexpr int kWarnEpsTooSmall = 1;
constexpr int kErrSpreadBoxSmall = 3;
constexpr int kErrUpsampTooSmall = 7;
constexpr int kErrHornerWrongBeta = 8;
constexpr int kErrNdataNotValid = 9;
constexpr int kMaxNspread = 16;

// Stage ids of the event timer (cufinufft_get_timing).
constexpr int kStageSort = 0;           // setpts: bin sort / index setup
constexpr int kStageGrid = 1;           // type 1: fine-grid zero fill
constexpr int kStageSpreadInterp = 2;   // spread (type 1) or interp (type 2)
constexpr int kStageFft = 3;            // hipFFT
constexpr int kStageDeconv = 4;         // deconvolve (type 1) or amplify (type 2)
constexpr int kNumStages = 8;
// roctx range names of the stages (shown by rocprofv3 --marker-trace).
constexpr const char *kStageName[5] = {"cufinufft sort", "cufinufft grid zero",
                                       "cufinufft spread/interp", "cufinufft fft",
                                       "cufinufft deconvolve/amplify"};

// roctx range over one API call (the reference's NVTX ranges,
// include/profile.h:14-15, src/profile.cu:10-31).  A push and pop per call;
// without a tool attached the library's dispatch is a no-op.
struct ApiRange {
    explicit ApiRange(const char *name) { roctxRangePushA(name); }
    ~ApiRange() { roctxRangePop(); }
    ApiRange(const ApiRange &) = delete;
    ApiRange &operator=(const ApiRange &) = delete;
};

// ---------------------------------------------------------------------------
// Spreader parameters, contrib/spreadinterp.cpp:6-67, evaluated in the plan
// precision T exactly as the reference's overloads resolve (the single
// precision build calls the float log10 / ceil / sqrt).
template <typename T> struct Spread {
    int ns = 0;
    T beta = 0, c = 0, hw = 0;
    T upsampfac = 2;
};

inline float m_log10(float v) { return std::log10(v); }
inline double m_log10(double v) { return std::log10(v); }
inline float m_log(float v) { return std::log(v); }
inline double m_log(double v) { return std::log(v); }
inline float m_sqrt(float v) { return std::sqrt(v); }
inline double m_sqrt(double v) { return std::sqrt(v); }
inline float m_ceil(float v) { return std::ceil(v); }
inline double m_ceil(double v) { return std::ceil(v); }

template <typename T> constexpr T eps_mach();
template <> constexpr float eps_mach<float>() { return (float)6e-08; }   // contrib/utils_fp.h:34
template <> constexpr double eps_mach<double>() { return 1.1e-16; }     // contrib/utils_fp.h:43

template <typename T>
int setup_spreader(Spread<T> &sp, T eps, T upsampfac, int kerevalmeth)
{
    const T pi = (T)M_PI;
    if (upsampfac != 2.0) {
        if (kerevalmeth == 1) return kErrHornerWrongBeta;
        if (upsampfac <= 1.0) return kErrUpsampTooSmall;
    }
    int ier = 0;
    if (eps < eps_mach<T>()) {
        eps = eps_mach<T>();
        ier = kWarnEpsTooSmall;
    }
    int ns = (int)m_ceil(-m_log10(eps / (T)10.0));
    if (upsampfac != 2.0) ns = (int)m_ceil(-m_log(eps) / (pi * m_sqrt(1 - 1 / upsampfac)));
    ns = std::max(2, ns);
    if (ns > kMaxNspread) {
        ns = kMaxNspread;
        ier = kWarnEpsTooSmall;
    }
    sp.ns = ns;
    sp.hw = (T)ns / 2;
    sp.c = (T)(4.0 / (double)(T)(ns * ns));
    T betaoverns = (T)2.30;
    if (ns == 2) betaoverns = (T)2.20;
    if (ns == 3) betaoverns = (T)2.26;
    if (ns == 4) betaoverns = (T)2.38;
    if (upsampfac != 2.0) {
        const T gamma = (T)0.97;
        betaoverns = gamma * pi * (1 - 1 / (2 * upsampfac));
    }
    sp.beta = betaoverns * (T)ns;
    sp.upsampfac = upsampfac;
    return ier;
}

static bool smooth235(long long m)
{
    if (m < 1) return false;
    while (m % 2 == 0) m /= 2;
    while (m % 3 == 0) m /= 3;
    while (m % 5 == 0) m /= 5;
    return m == 1;
}

// Smallest even 2,3,5-smooth integer >= n that is a multiple of b
// (next235beven, contrib/utils.cpp:3-22; n <= 2 gives 2 as there).  b must
// be 2,3,5-smooth itself (the reference's search never ends otherwise), so
// the candidates are the even multiples of b from n up.
long long next235beven(long long n, long long b)
{
    if (n <= 2) return 2;
    const long long step = b % 2 == 0 ? b : 2 * b;
    for (long long m = (n + step - 1) / step * step;; m += step)
        if (smooth235(m)) return m;
}

// Fine-grid size of one axis (SET_NF_TYPE12, contrib/common.cpp:24-37): b is
// the output block size for method 4, else 1.
int fine_size(int ms, double upsampfac, int ns, int b)
{
    long long nf = (long long)(upsampfac * ms);
    if (nf < 2 * ns) nf = 2 * ns;
    if (nf < 2147483647LL) nf = next235beven(nf, b);
    return nf > 2147483647LL ? 2147483647 : (int)nf;
}

// ---------------------------------------------------------------------------
// Fourier series of the ES kernel on a grid of nf cells, k = 0..nf/2
// (onedim_fseries_kernel, contrib/common.cpp:39-124): the kernel is even and
// supported on |z| < ns/2, so its transform is 2 * integral over [0, ns/2] of
// phi(z) cos(2 pi k z / nf) dz, evaluated by q = (int)(2 + 1.5 ns)-point
// Gauss-Legendre quadrature on the half support; the factor (-1)^k is the
// phase of the nf/2 grid offset at which RESCALE places x = 0.  Computed in
// FP64 with phasor recursion per node.
static void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        const double dp = n * (z * p1 - p0) / (z * z - 1.0);
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

template <typename T>
static void fseries(int nf, const Spread<T> &sp, std::vector<double> &out)
{
    const double J2 = sp.ns / 2.0;
    const int q = (int)(2 + 3.0 * J2);
    std::vector<double> z, w;
    gauss_legendre(2 * q, z, w);
    const double beta = (double)sp.beta, c = (double)sp.c;
    std::vector<double> f(q), cs(q), sn(q), cr(q), ci(q);
    for (int n = 0; n < q; ++n) {
        const double zn = J2 * z[q + n];  // positive half of the 2q-node rule
        const double a = 1.0 - c * zn * zn;
        f[n] = J2 * w[q + n] * (std::fabs(zn) < J2 ? std::exp(beta * std::sqrt(a > 0 ? a : 0)) : 0);
        const double th = 2.0 * M_PI * (nf / 2 - zn) / nf;
        cs[n] = std::cos(th);
        sn[n] = std::sin(th);
        cr[n] = 1.0;
        ci[n] = 0.0;
    }
    out.assign(nf / 2 + 1, 0.0);
    for (int k = 0; k <= nf / 2; ++k) {
        double s = 0;
        for (int n = 0; n < q; ++n) {
            s += 2.0 * f[n] * cr[n];
            const double r = cr[n] * cs[n] - ci[n] * sn[n];
            ci[n] = cr[n] * sn[n] + ci[n] * cs[n];
            cr[n] = r;
        }
        out[k] = s;
        if ((k & 63) == 63) {  // re-anchor the phasors to bound drift
            for (int n = 0; n < q; ++n) {
                const double zn = J2 * z[q + n];
                const double th = 2.0 * M_PI * (double)(k + 1) * (nf / 2 - zn) / nf;
                cr[n] = std::cos(th);
                ci[n] = std::sin(th);
            }
        }
    }
}

// Accumulator width of the FP32 fixed-point spread tiles.  A 32-bit tile
// rounds every tap to A * 2^-31 with A the strength mass of the points
// accumulated under one scale, so its error grows with their number n: CPU
// emulation of the tile on a clustered 3D tol-1e-6 case adds about
// 2-3 * n * 2^-30 to the relative error of the transform.  fx32_max is the
// largest n with 4 * n * 2^-30 within tol/5 (536 at tol 1e-5).  The SM
// spread accumulates 32-bit chunks of at most that many points, each under
// its own scale and drained to FP32 between chunks, when a workgroup-wide
// chunk fits the bound, else 64-bit cells.  Method 4 takes 32 bits for the
// whole job when its subproblems fit the bound (fx_wide = 0).
//
// The rounding is white noise on the fine grid, while the spread signal's
// spectrum is shaped by the kernel: deconvolution divides mode k by the
// kernel's Fourier factor phi^(k), so the noise reaches the modes amplified
// by rms_k phi^(0) / phi^(k) relative to the signal.  That rms is 1.1-3.2
// for every sigma = 2 plan (largest in 3D at tol 1e-6, the calibration case
// above) but ~250 (2D) and ~3000 (3D) at sigma = 1.25, whose passband reaches
// further into the kernel's decay.  The bound is therefore divided by the
// plan's amplification over kAmpRef = 3.25 when it exceeds that.
constexpr double kAmpRef = 3.25;
static double deconv_amp(int dim, const int *ms, const std::vector<double> *fser)
{
    double amp = 1.0;
    for (int d = 0; d < dim; ++d) {
        if (fser[d].empty()) continue;
        const int h = ms[d] / 2;
        double s = 0.0;
        long long n = 0;
        for (int k = -h; k <= (ms[d] - 1) / 2; ++k) {
            const double r = fser[d][0] / fser[d][k < 0 ? -k : k];
            s += r * r;
            ++n;
        }
        amp *= std::sqrt(s / (double)n);
    }
    return amp;
}
static int fx32_max(double tol, double amp)
{
    const int n = fx32_bound(tol);
    return amp > kAmpRef ? (int)((double)n * kAmpRef / amp) : n;
}
static int fx_wide(int maxsub, double tol, double amp)
{
    return maxsub > fx32_max(tol, amp) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Piecewise-polynomial ES kernel for gpu_kerevalmeth = 1.  The reference
// evaluates fixed fitted tables (contrib/ker_horner_allw_loop.c, sigma = 2
// only); this library fits its own at plan time.  For a point whose first
// tap lies at x1 = start - v in [-ns/2, -ns/2 + 1), tap i sits at offset
// x1 + i; with z = 2(x1 + ns/2) - 1 in [-1, 1), phi on each tap's unit
// interval is interpolated at deg+1 Chebyshev nodes and re-expanded in powers
// of z (coef[k*ns + i] multiplies z^k for tap i).  The degree is the
// smallest in [3, 24] whose error, sampled at 128 interior points per
// interval, is at most max(tol/20, 4e-14) of the kernel peak e^beta.
// Returns the degree and the achieved error relative to the peak.
static int horner_fit(int ns, double beta, double c, double tol, std::vector<double> &coef,
                      double *rel_err)
{
    const double hw = 0.5 * ns;
    auto phi = [&](double t) {
        const double a = 1.0 - c * t * t;
        return std::fabs(t) <= hw ? std::exp(beta * std::sqrt(a > 0 ? a : 0.0)) : 0.0;
    };
    const double peak = std::exp(beta);
    const double target = std::max(tol / 20.0, 4e-14);
    int best_deg = 0;
    double best_err = 1e300;
    std::vector<double> C, f, a, tprev, tcur, tnext;
    for (int deg = 3; deg <= 24; ++deg) {
        const int K = deg + 1;
        C.assign((size_t)K * ns, 0.0);
        f.assign(K, 0.0);
        a.assign(K, 0.0);
        for (int i = 0; i < ns; ++i) {
            for (int k = 0; k < K; ++k)
                f[k] = phi(0.5 * (std::cos(M_PI * (k + 0.5) / K) + 1.0) - hw + i);
            for (int j = 0; j < K; ++j) {
                double s = 0;
                for (int k = 0; k < K; ++k) s += f[k] * std::cos(M_PI * j * (k + 0.5) / K);
                a[j] = (j == 0 ? 1.0 : 2.0) * s / K;
            }
            // sum_j a_j T_j(z) in powers of z, T_{j+1} = 2 z T_j - T_{j-1}
            tprev.assign(K, 0.0);
            tcur.assign(K, 0.0);
            tprev[0] = 1.0;
            C[i] += a[0];
            if (K > 1) {
                tcur[1] = 1.0;
                C[(size_t)ns + i] += a[1];
            }
            for (int j = 2; j < K; ++j) {
                tnext.assign(K, 0.0);
                for (int m = 0; m < K; ++m)
                    tnext[m] = (m > 0 ? 2.0 * tcur[m - 1] : 0.0) - tprev[m];
                for (int m = 0; m <= j; ++m) C[(size_t)m * ns + i] += a[j] * tnext[m];
                tprev.swap(tcur);
                tcur.swap(tnext);
            }
        }
        double err = 0;
        for (int i = 0; i < ns; ++i) {
            for (int s = 0; s < 128; ++s) {
                const double z = -1.0 + 2.0 * (s + 0.5) / 128.0;
                double v = C[(size_t)deg * ns + i];
                for (int k = deg - 1; k >= 0; --k) v = v * z + C[(size_t)k * ns + i];
                err = std::max(err, std::fabs(v - phi(0.5 * (z + 1.0) - hw + i)));
            }
        }
        err /= peak;
        if (err < best_err) {
            best_err = err;
            best_deg = deg;
            coef = C;
        }
        if (err <= target) break;
    }
    if (rel_err) *rel_err = best_err;
    return best_deg;
}

// ---------------------------------------------------------------------------
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        if (prev >= 0) hipSetDevice(prev);
    }
};

template <typename T>
static int dmalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void **)p, n * sizeof(T)) == hipSuccess ? 0 : kErrHip;
}
template <typename T>
static void dfree(T *&p)
{
    if (p) hipFree((void *)p);
    p = nullptr;
}

template <typename T> struct FftType;
template <> struct FftType<float> {
    static constexpr hipfftType type = HIPFFT_C2C;
    static hipfftResult exec(hipfftHandle h, cpx<float> *d, int dir)
    {
        return hipfftExecC2C(h, (hipfftComplex *)d, (hipfftComplex *)d, dir);
    }
};
template <> struct FftType<double> {
    static constexpr hipfftType type = HIPFFT_Z2Z;
    static hipfftResult exec(hipfftHandle h, cpx<double> *d, int dir)
    {
        return hipfftExecZ2Z(h, (hipfftDoubleComplex *)d, (hipfftDoubleComplex *)d, dir);
    }
};

template <typename T>
struct Plan {
    cufinufft_opts opts{};
    Spread<T> sp;
    int type = 0, dim = 0;
    int ms[3] = {1, 1, 1};
    int nf[3] = {1, 1, 1};
    int ntransf = 1, maxbatch = 1, iflag = 1;
    int method = 1;
    bool eps_warning = false;
    BinGeom geom{};
    long long nbins = 1;
    long long nftot = 1, ntot = 1;

    // points
    int M = 0;
    const T *kx = nullptr, *ky = nullptr, *kz = nullptr;
    bool pts_set = false;
    int cap_M = -1;
    long long cap_sub = -1;
    long long sub_bound = 0;

    // device buffers
    cpx<T> *fw = nullptr;
    T *fwker[3] = {nullptr, nullptr, nullptr};
    std::vector<double> fser[3];
    // piecewise-polynomial kernel (gpu_kerevalmeth = 1): degree, FP64 host
    // table, device tables (plain, then divided by e^beta)
    int hdeg = 0;
    double herr = 0;
    std::vector<double> hcoef_host;
    T *hcoef = nullptr;
    SortBufs sb;

    int fxwide = 0;   // FP32 method 4: 64-bit (1) or 32-bit (0) block tiles
    int fx32max = 0;  // FP32: most points one 32-bit fixed-point scale may cover

    KerParams<T> ker_params() const
    {
        const double bl2e = (double)sp.beta * 1.4426950408889634;
        const T hi = (T)bl2e;
        return KerParams<T>{sp.beta, sp.c,    sp.hw,   hcoef, hdeg,
                            fxwide,  fx32max, hi,      (T)(bl2e - (double)hi)};
    }

    hipfftHandle fft_full = 0, fft_rem = 0;
    int fft_rem_batch = 0;
    bool fft_full_ok = false, fft_rem_ok = false;
    // Pruned 3D FFT (fft_prune.hip): one contiguous batched 1D hipFFT per
    // axis, over the lines that hold modes, with the transposes between them
    // going through fwh ([batch][ms0][nf2][nf1]); handles per batch size.
    bool pruned = false;
    cpx<T> *fwh = nullptr;
    struct Fft1d {
        hipfftHandle h[3] = {0, 0, 0};
        int batch = -1;
    } pf_full, pf_rem;
    Launch L;

    // Stage timer: HIP events recorded on the plan stream around every stage
    // while enabled; read (and cleared) by get_timing after a synchronise.
    struct TimedSpan {
        int stage;
        hipEvent_t a, b;
    };
    bool timing = false;
    std::vector<TimedSpan> spans;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    hipEvent_t ev_open = nullptr;
    static constexpr size_t kMaxSpans = 1 << 16;

    hipEvent_t next_event()
    {
        if (ev_used < ev_pool.size()) return ev_pool[ev_used++];
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ev_pool.push_back(e);
        ++ev_used;
        return e;
    }
    void tbegin()
    {
        if (!timing || spans.size() >= kMaxSpans) return;
        ev_open = next_event();
        if (ev_open) hipEventRecord(ev_open, L.stream);
    }
    void tend(int stage)
    {
        if (!timing || !ev_open) return;
        hipEvent_t b = next_event();
        if (!b) return;
        hipEventRecord(b, L.stream);
        spans.push_back({stage, ev_open, b});
        ev_open = nullptr;
    }
    // One stage: a roctx range and, while timing is on, an event span, both
    // closed when the scope ends (every return path included).
    struct Stage {
        Plan *p;
        int id;
        Stage(Plan *p_, int id_) : p(p_), id(id_)
        {
            roctxRangePushA(kStageName[id]);
            p->tbegin();
        }
        ~Stage()
        {
            p->tend(id);
            roctxRangePop();
        }
        Stage(const Stage &) = delete;
        Stage &operator=(const Stage &) = delete;
    };

    ~Plan() { release(); }

    void release()
    {
        DeviceGuard dg(opts.gpu_device_id);
        for (auto e : ev_pool) hipEventDestroy(e);
        ev_pool.clear();
        spans.clear();
        ev_used = 0;
        if (fft_full_ok) hipfftDestroy(fft_full);
        if (fft_rem_ok) hipfftDestroy(fft_rem);
        fft_full_ok = fft_rem_ok = false;
        drop_fft1d(pf_full);
        drop_fft1d(pf_rem);
        dfree(fw);
        dfree(fwh);
        for (auto &k : fwker) dfree(k);
        dfree(hcoef);
        dfree(sb.binsize);
        dfree(sb.binstart);
        dfree(sb.substart);
        dfree(sb.s2b);
        dfree(sb.skey);
        dfree(sb.skey_out);
        for (auto &r : sb.rec) {
            if (r) hipFree(r);
            r = nullptr;
        }
        dfree(sb.sort_counts);
        if (sb.sort_part) hipFree(sb.sort_part);
        sb.sort_part = nullptr;
        dfree(sb.idx);
        dfree(sb.pos);
        dfree(sb.counters);
        dfree(sb.jobq);
        if (sb.scan_part) hipFree(sb.scan_part);
        sb.scan_part = nullptr;
        dfree(sb.kcount);
        dfree(sb.kstart);
        if (sb.kscan_part) hipFree(sb.kscan_part);
        sb.kscan_part = nullptr;
        if (sb.slot) hipFree(sb.slot);
        sb.slot = nullptr;
    }

    int make_fft(hipfftHandle *h, int batch)
    {
        int n[3];
        for (int d = 0; d < dim; ++d) n[d] = nf[dim - 1 - d];  // slowest first
        if (hipfftCreate(h) != HIPFFT_SUCCESS) return kErrFft;
        if (hipfftPlanMany(h, dim, n, n, 1, (int)nftot, n, 1, (int)nftot, FftType<T>::type,
                           batch) != HIPFFT_SUCCESS) {
            hipfftDestroy(*h);
            return kErrFft;
        }
        return 0;
    }

    int fft(cpx<T> *data, int batch)
    {
        hipfftHandle h;
        if (batch == maxbatch) {
            h = fft_full;
        } else {
            if (!fft_rem_ok || fft_rem_batch != batch) {
                if (fft_rem_ok) hipfftDestroy(fft_rem);
                fft_rem_ok = false;
                int e = make_fft(&fft_rem, batch);
                if (e) return e;
                fft_rem_ok = true;
                fft_rem_batch = batch;
            }
            h = fft_rem;
        }
        if (hipfftSetStream(h, L.stream) != HIPFFT_SUCCESS) return kErrFft;
        return FftType<T>::exec(h, data, iflag) == HIPFFT_SUCCESS ? 0 : kErrFft;
    }

    static void drop_fft1d(Fft1d &f)
    {
        for (auto &h : f.h) {
            if (h) hipfftDestroy(h);
            h = 0;
        }
        f.batch = -1;
    }
    // The 1D plans of the pruned FFT for `batch` transforms: x over every
    // grid line, y over the ms0 mode columns (times nf2 planes in 3D), and in
    // 3D z over the ms0 x ms1 mode columns (each line contiguous, lines nf
    // apart).
    int make_fft1d(Fft1d &f, int batch)
    {
        drop_fft1d(f);
        const long long lines[3] = {(long long)nf[1] * nf[2], (long long)ms[0] * nf[2],
                                    (long long)ms[0] * ms[1]};
        for (int a = 0; a < dim; ++a) {
            int n = nf[a];
            if (hipfftCreate(&f.h[a]) != HIPFFT_SUCCESS) {
                f.h[a] = 0;
                drop_fft1d(f);
                return kErrFft;
            }
            if (hipfftPlanMany(&f.h[a], 1, &n, &n, 1, n, &n, 1, n, FftType<T>::type,
                               (int)(lines[a] * batch)) != HIPFFT_SUCCESS) {
                drop_fft1d(f);
                return kErrFft;
            }
        }
        f.batch = batch;
        return 0;
    }
    int fft1d(Fft1d *&f, int batch)
    {
        if (batch == pf_full.batch) {
            f = &pf_full;
            return 0;
        }
        if (batch != pf_rem.batch) {
            const int e = make_fft1d(pf_rem, batch);
            if (e) return e;
        }
        f = &pf_rem;
        return 0;
    }
    int exec1d(hipfftHandle h, cpx<T> *data)
    {
        if (hipfftSetStream(h, L.stream) != HIPFFT_SUCCESS) return kErrFft;
        return FftType<T>::exec(h, data, iflag) == HIPFFT_SUCCESS ? 0 : kErrFft;
    }

    // Type 1 after the spread: the FFT of the fine grid in fw.  Pruned, it
    // leaves the ms0 x ms1 x nf2 block [x'][y'][z] (mode columns, z
    // transformed) at the start of each transform's fw.
    int fft_type1(int blk)
    {
        if (!pruned) return fft(fw, blk);
        Fft1d *f = nullptr;
        int e = fft1d(f, blk);
        if (e) return e;
        if ((e = exec1d(f->h[0], fw))) return e;
        if (dim == 2) {  // leaves B1 = [x'][y] (y transformed) in fwh
            const auto g1 = prune2_fw_to_b1(ms, nf);
            if ((e = run_prune_gather<T>(fw, fwh, g1, blk, nullptr, nullptr, nullptr, L)))
                return e;
            return exec1d(f->h[1], fwh);
        }
        const auto g1 = prune_fw_to_b1(ms, nf);
        if ((e = run_prune_gather<T>(fw, fwh, g1, blk, nullptr, nullptr, nullptr, L))) return e;
        if ((e = exec1d(f->h[1], fwh))) return e;
        const auto g2 = prune_b1_to_b2(ms, nf);
        if ((e = run_prune_gather<T>(fwh, fw, g2, blk, nullptr, nullptr, nullptr, L))) return e;
        return exec1d(f->h[2], fw);
    }
    // Type 1, last step: fk = the modes of fw divided by the kernel's
    // Fourier series (the pruned form reads the block fft_type1 leaves).
    int deconv_type1(cpx<T> *fk, int blk)
    {
        const T *kers[3] = {fwker[0], fwker[1], fwker[2]};
        if (!pruned) return run_deconvolve<T>(dim, ms, nf, fw, fk, kers, blk, L);
        if (dim == 2)
            return run_prune_gather<T>(fwh, fk, prune2_b1_to_fk(ms, nf), blk, kers[0], nullptr,
                                       kers[1], L);
        const auto d = prune_b2_to_fk(ms, nf);
        return run_prune_gather<T>(fw, fk, d, blk, kers[0], kers[1], kers[2], L);
    }
    // Type 2, first step: the modes fk divided by the kernel's Fourier
    // series (times i * wsign * k_wax for wax >= 0), zero-padded into fw
    // (pruned: into the ms0 x ms1 x nf2 block [x'][y'][z]).
    int amplify_type2(const cpx<T> *fk, int blk, int wax, int wsign)
    {
        const T *kers[3] = {fwker[0], fwker[1], fwker[2]};
        if (!pruned) return run_amplify<T>(dim, ms, nf, fw, fk, kers, blk, L, wax, wsign);
        if (dim == 2)  // y is the scatter's c axis
            return run_prune_scatter<T>(fk, fwh, prune2_fk_to_b1(ms, nf), blk, kers[0], nullptr,
                                        kers[1], wax == 1 ? 2 : wax, wsign, L);
        const auto d = prune_fk_to_b2(ms, nf);
        return run_prune_scatter<T>(fk, fw, d, blk, kers[0], kers[1], kers[2], wax, wsign, L);
    }
    // Type 2: the FFT that turns what amplify_type2 wrote into the fine grid.
    int fft_type2(int blk)
    {
        if (!pruned) return fft(fw, blk);
        Fft1d *f = nullptr;
        int e = fft1d(f, blk);
        if (e) return e;
        if (dim == 2) {  // B1 = [x'][y] in fwh -> fw
            if ((e = exec1d(f->h[1], fwh))) return e;
            if ((e = run_prune_scatter<T>(fwh, fw, prune2_b1_to_fw(ms, nf), blk, nullptr, nullptr,
                                          nullptr, -1, 1, L)))
                return e;
            return exec1d(f->h[0], fw);
        }
        if ((e = exec1d(f->h[2], fw))) return e;
        const auto s1 = prune_b2_to_b1(ms, nf);
        if ((e = run_prune_scatter<T>(fw, fwh, s1, blk, nullptr, nullptr, nullptr, -1, 1, L)))
            return e;
        if ((e = exec1d(f->h[1], fwh))) return e;
        const auto s2 = prune_b1_to_fw(ms, nf);
        if ((e = run_prune_scatter<T>(fwh, fw, s2, blk, nullptr, nullptr, nullptr, -1, 1, L)))
            return e;
        return exec1d(f->h[0], fw);
    }

    int makeplan(int type_, int dim_, const int *nmodes, int iflag_, int ntransf_, T tol,
                 int maxbatch_, const cufinufft_opts *o, int flags = 0);
    int setpts(int M_, const T *x, const T *y, const T *z);
    int setpts_noalloc(int M_, const T *x, const T *y, const T *z, unsigned char *scr,
                       size_t scr_bytes);
    int setpts_blocks();
    int cap_copies = -1;  // capacity of idx / pos in point copies (method 4)
    long long cap_counts = -1;          // capacity of sb.sort_counts
    long long cap_rec[2] = {-1, -1};    // capacity of sb.rec[] in points
    int execute(cpx<T> *c, cpx<T> *fk);
    int execute_grad(cpx<T> *c, cpx<T> *dc, cpx<T> *fk);
};

int default_opts(int type, int dim, cufinufft_opts *opts)
{
    if (!opts) return 1;
    opts->upsampfac = 2.0;
    opts->gpu_nstreams = 0;
    opts->gpu_sort = 1;
    opts->gpu_maxsubprobsize = 1024;
    opts->gpu_obinsizex = opts->gpu_obinsizey = opts->gpu_obinsizez = -1;
    opts->gpu_binsizex = opts->gpu_binsizey = opts->gpu_binsizez = -1;
    opts->gpu_spreadinterponly = 0;
    opts->gpu_kerevalmeth = 0;
    opts->gpu_device_id = 0;
    if (dim < 1 || dim > 3) return 1;
    if (type == 1)
        opts->gpu_method = 2;
    else if (type == 2)
        opts->gpu_method = 1;
    else
        return 1;  // type 3 is not part of this library (nor of the reference)
    return 0;
}

// Default bin sizes (SETUP_BINSIZE, src/cufinufft.cu:17-73).  Method 4
// takes the reference's 3D block defaults (output blocks 8^3, bins 4^3,
// src/cufinufft.cu:53-66); 1D / 2D blocks, which the reference does not
// have, default to the SM bin sizes.
static void default_binsizes(int dim, cufinufft_opts &o)
{
    if (o.gpu_method == 4) {
        const int ob = dim == 3 ? 8 : (dim == 2 ? 32 : 1024);
        const int bs = dim == 3 ? 4 : ob;
        if (o.gpu_obinsizex < 0) o.gpu_obinsizex = ob;
        if (o.gpu_obinsizey < 0) o.gpu_obinsizey = dim > 1 ? ob : 1;
        if (o.gpu_obinsizez < 0) o.gpu_obinsizez = dim > 2 ? ob : 1;
        if (o.gpu_binsizex < 0) o.gpu_binsizex = bs;
        if (o.gpu_binsizey < 0) o.gpu_binsizey = dim > 1 ? bs : 1;
        if (o.gpu_binsizez < 0) o.gpu_binsizez = dim > 2 ? bs : 1;
        if (dim < 3) o.gpu_obinsizez = o.gpu_binsizez = 1;
        if (dim < 2) o.gpu_obinsizey = o.gpu_binsizey = 1;
        return;
    }
    if (dim == 1) {
        if (o.gpu_binsizex < 0) o.gpu_binsizex = 1024;
        o.gpu_binsizey = o.gpu_binsizez = 1;
    } else if (dim == 2) {
        if (o.gpu_binsizex < 0) o.gpu_binsizex = 32;
        if (o.gpu_binsizey < 0) o.gpu_binsizey = 32;
        o.gpu_binsizez = 1;
    } else {
        if (o.gpu_binsizex < 0) o.gpu_binsizex = 16;
        if (o.gpu_binsizey < 0) o.gpu_binsizey = 16;
        if (o.gpu_binsizez < 0) o.gpu_binsizez = 2;
    }
}

template <typename T>
int Plan<T>::makeplan(int type_, int dim_, const int *nmodes, int iflag_, int ntransf_, T tol,
                      int maxbatch_, const cufinufft_opts *o, int flags)
{
    if (type_ != 1 && type_ != 2) return 1;
    if (dim_ < 1 || dim_ > 3) return 1;
    if (flags & ~(CUFINUFFT_ACCURACY_FIRST | CUFINUFFT_TILE_SLOTS | CUFINUFFT_FULL_FFT))
        return 1;
    // Accuracy-first: the reference's width rule ns = ceil(-log10(tol/10))
    // leaves the transform's relative l2 error at 0.6-1.1, 1-2 and 1.5-3.2
    // times tol in 1, 2 and 3 dimensions (the same algorithm in FP64 against
    // direct sums, tol 1e-5 and 1e-9).  Planning for tol * 2 / (dim + 2)
    // instead (tol/2.5 in 3D: one more tap per axis at tol 1e-5 and 1e-9)
    // brings it to 0.1-0.4 times tol.
    if (flags & CUFINUFFT_ACCURACY_FIRST) tol = (T)((double)tol * 2.0 / (dim_ + 2));
    if (!nmodes) return kErrNdataNotValid;
    if (o) {
        opts = *o;
    } else {
        int e = default_opts(type_, dim_, &opts);
        if (e) return e;
    }
    type = type_;
    dim = dim_;
    for (int d = 0; d < 3; ++d) ms[d] = d < dim ? nmodes[d] : 1;
    for (int d = 0; d < dim; ++d)
        if (ms[d] < 1) return kErrNdataNotValid;
    if (ntransf_ < 1) return kErrNdataNotValid;
    // methods: 0 auto, 1 GM(-sort), 2 SM, 4 owner-computes blocks (type 1);
    // the reference's method 3 (Paul) is not built
    if (opts.gpu_method < 0 || opts.gpu_method > 4 || opts.gpu_method == 3) return 2;
    if (opts.gpu_method == 4 && type != 1) return 2;
    if (opts.gpu_maxsubprobsize < 1) return 1;
    if (opts.gpu_sort < 0 || opts.gpu_sort > 2) return 1;
    const int ier = setup_spreader<T>(sp, tol, (T)opts.upsampfac, opts.gpu_kerevalmeth);
    if (ier > 1) return ier;
    eps_warning = ier == kWarnEpsTooSmall;
    if (opts.gpu_kerevalmeth != 0 && opts.gpu_kerevalmeth != 1) return 1;

    const bool bin_default[3] = {opts.gpu_binsizex < 0, opts.gpu_binsizey < 0,
                                 opts.gpu_binsizez < 0};
    default_binsizes(dim, opts);
    if (opts.gpu_binsizex < 1 || opts.gpu_binsizey < 1 || opts.gpu_binsizez < 1) return 1;
    iflag = iflag_ >= 0 ? 1 : -1;
    ntransf = ntransf_;
    maxbatch = maxbatch_ == 0 ? std::min(ntransf, 8) : maxbatch_;
    if (maxbatch < 1) return kErrNdataNotValid;
    maxbatch = std::min(maxbatch, ntransf);

    const int obv[3] = {opts.gpu_obinsizex, opts.gpu_obinsizey, opts.gpu_obinsizez};
    if (opts.gpu_method == 4)
        for (int d = 0; d < dim; ++d)
            if (!smooth235(obv[d])) return 1;  // nf must become a multiple of the block
    for (int d = 0; d < dim; ++d) {
        if (opts.gpu_spreadinterponly) {
            nf[d] = ms[d];
            if (nf[d] < 2 * sp.ns) return kErrSpreadBoxSmall;
        } else {
            nf[d] = fine_size(ms[d], opts.upsampfac, sp.ns, opts.gpu_method == 4 ? obv[d] : 1);
        }
    }
    nftot = (long long)nf[0] * nf[1] * nf[2];
    ntot = (long long)ms[0] * ms[1] * ms[2];
    if (nftot * maxbatch > 2147483647LL && !opts.gpu_spreadinterponly) {
        // hipFFT's plan-many interface takes int distances and batch sizes
        if (nftot > 2147483647LL) return kErrNdataNotValid;
    }

    // Method: 0 picks SM when the tile fits the LDS, GM-sort otherwise.  The
    // spread tile accumulates in FP64 (FP64 plans) or fixed point (FP32
    // plans); the interp tile holds T.
    // Fourier-series factors (FP64 on the host; uploaded below), which also
    // set the fixed-point noise amplification of the type-1 deconvolution
    if (!opts.gpu_spreadinterponly)
        for (int d = 0; d < dim; ++d) fseries<T>(nf[d], sp, fser[d]);
    const double amp = deconv_amp(dim, ms, fser);
    fxwide = fx_wide(opts.gpu_maxsubprobsize, (double)tol, amp);
    fx32max = sizeof(T) == 4 ? fx32_max((double)tol, amp) : 0;
    int bsv[3] = {opts.gpu_binsizex, opts.gpu_binsizey, opts.gpu_binsizez};
    auto tile_lds = [&]() {
        return type == 1 ? spread_cfg(dim, sp.ns, bsv, sizeof(T), fx32max,
                                      opts.gpu_maxsubprobsize, false)
                               .lds
                         : sm_tile_bytes(dim, sp.ns, bsv, sizeof(T));
    };
    size_t tile_bytes = tile_lds();
    // Default bins (the reference's sizes) whose SM tile does not fit the LDS
    // with this kernel width (e.g. FP64 3D type 1 at ns 11) are halved along
    // their largest defaulted side until it fits; bins the caller chose are
    // kept as given.
    if ((opts.gpu_method == 2 || opts.gpu_method == 0) && tile_bytes > kMaxLdsBytes) {
        while (tile_bytes > kMaxLdsBytes) {
            int k = -1;
            for (int d = 0; d < dim; ++d)
                if (bin_default[d] && bsv[d] > 2 && (k < 0 || bsv[d] > bsv[k])) k = d;
            if (k < 0) break;
            bsv[k] /= 2;
            tile_bytes = tile_lds();
        }
        opts.gpu_binsizex = bsv[0];
        opts.gpu_binsizey = bsv[1];
        opts.gpu_binsizez = bsv[2];
    }
    method = opts.gpu_method;
    if (method == 0) method = tile_bytes <= kMaxLdsBytes ? 2 : 1;
    if (method == 2 && tile_bytes > kMaxLdsBytes) return kErrLds;
    if (method == 4) {
        // blocks must tile the fine grid (nf was rounded to a multiple of
        // the block above, except in spread/interp-only mode, where nf is the
        // caller's grid) and be whole numbers of bins, as in the reference
        // (spread3d_wrapper.cu:410-441)
        for (int d = 0; d < dim; ++d)
            if (obv[d] < 1 || nf[d] % obv[d] != 0 || obv[d] % bsv[d] != 0) return 1;
        if (blk_lds_bytes(dim, obv, sizeof(T), opts.gpu_maxsubprobsize, fxwide) > kMaxLdsBytes)
            return kErrLds;
    }

    for (int d = 0; d < 3; ++d) {
        geom.nf[d] = nf[d];
        geom.bs[d] = d < dim ? (method == 4 ? obv[d] : bsv[d]) : 1;
        if (d >= dim) {
            geom.nb[d] = 1;
        } else if (sizeof(T) == 4) {
            geom.nb[d] = (int)std::ceil((float)nf[d] / (float)geom.bs[d]);
        } else {
            geom.nb[d] = (int)std::ceil((double)nf[d] / (double)geom.bs[d]);
        }
    }
    nbins = (long long)geom.nb[0] * geom.nb[1] * geom.nb[2];

    DeviceGuard dg(opts.gpu_device_id);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return kErrHip;
    if (hipDeviceGetAttribute(&L.ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return kErrHip;

    // Fourier-series factors (FP64 on the host, stored in T on the device).
    if (!opts.gpu_spreadinterponly) {
        for (int d = 0; d < dim; ++d) {
            std::vector<T> h(fser[d].begin(), fser[d].end());
            if (dmalloc(&fwker[d], h.size())) return kErrHip;
            if (hipMemcpy(fwker[d], h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) !=
                hipSuccess)
                return kErrHip;
        }
        if (dmalloc(&fw, (size_t)(nftot * maxbatch))) return kErrHip;
        // 2D and 3D grids take the pruned FFT unless the caller asks for the
        // full one; the transposes' plane counts (grid y: nf2, N1, N2 in 3D)
        // and batch (grid z) and hipFFT's int batch counts bound the sizes.
        pruned = dim >= 2 && !(flags & CUFINUFFT_FULL_FFT) && maxbatch <= 65535 &&
                 nf[0] <= 65535 && nf[1] <= 65535 && nf[2] <= 65535 &&
                 (long long)nf[1] * nf[2] * maxbatch <= 2147483647LL &&
                 (long long)ms[0] * nf[2] * nf[1] * maxbatch <= 2147483647LL;
        if (pruned) {
            if (dmalloc(&fwh, (size_t)((long long)ms[0] * nf[2] * nf[1] * maxbatch)))
                return kErrHip;
            int e = make_fft1d(pf_full, maxbatch);
            if (e) return e;
        } else {
            int e = make_fft(&fft_full, maxbatch);
            if (e) return e;
            fft_full_ok = true;
        }
    }
    if (opts.gpu_kerevalmeth == 1) {
        hdeg = horner_fit(sp.ns, (double)sp.beta, (double)sp.c, (double)tol, hcoef_host, &herr);
        const size_t n = hcoef_host.size();
        const double inv_peak = std::exp(-(double)sp.beta);
        std::vector<T> h(2 * n);
        for (size_t k = 0; k < n; ++k) {
            h[k] = (T)hcoef_host[k];
            h[n + k] = (T)(hcoef_host[k] * inv_peak);
        }
        if (dmalloc(&hcoef, h.size())) return kErrHip;
        if (hipMemcpy(hcoef, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return kErrHip;
    }
    if (dmalloc(&sb.binsize, nbins) || dmalloc(&sb.binstart, nbins) ||
        dmalloc(&sb.substart, nbins + 1) || dmalloc(&sb.counters, 2) ||
        dmalloc(&sb.jobq, kJobqWords) ||
        hipMemset(sb.jobq, 0, kJobqWords * sizeof(unsigned long long)) != hipSuccess)
        return kErrHip;
    const long long nparts = (nbins + kScanChunk - 1) / kScanChunk;
    if (hipMalloc(&sb.scan_part, sizeof(int2) * (size_t)nparts) != hipSuccess) return kErrHip;
    // Tile slots of the SM spread (CUFINUFFT_TILE_SLOTS): a bin with one
    // subproblem stores its whole padded tile with plain stores and a gather
    // pass adds the tiles covering each cell, instead of flushing the tile
    // with two global float atomics per cell (on uniform 3D points, 0.4 per
    // cell, those atomics were half of the spread; the gather's pass over the
    // whole grid costs more than it saves where most bins are dense or
    // empty, hence a flag).  The gather reads the tiles of a cell's own bin and its two
    // neighbours per axis, so every axis needs three bins or more, a halo no
    // wider than a bin and a tile no wider than the grid; and the slots are
    // kept within 8 fine grids of memory.
    if (type == 1 && method == 2 && (flags & CUFINUFFT_TILE_SLOTS)) {
        const int P = (sp.ns + 1) / 2;
        bool ok = true;
        size_t cells = 1;
        for (int d = 0; d < dim; ++d) {
            const int tw = geom.bs[d] + 2 * P;
            ok = ok && geom.nb[d] >= 3 && P <= geom.bs[d] && tw <= nf[d];
            cells *= (size_t)tw;
        }
        const size_t bytes = (size_t)nbins * cells * maxbatch * sizeof(cpx<T>);
        const size_t grid = (size_t)nftot * maxbatch * sizeof(cpx<T>);
        if (ok && bytes <= 8 * grid && hipMalloc(&sb.slot, bytes) != hipSuccess) {
            sb.slot = nullptr;
            (void)hipGetLastError();  // no slots: the spread flushes with atomics
        }
    }
    if (opts.gpu_sort == 2 &&
        nbins * (long long)geom.bs[0] * geom.bs[1] * geom.bs[2] > 2147483647LL)
        return kErrNdataNotValid;  // (bin, cell) keys are int
    if (method == 4) {  // copies counted per block
        const long long nkeys = nbins;
        const long long kparts = (nkeys + kScanChunk - 1) / kScanChunk;
        if (dmalloc(&sb.kcount, (size_t)nkeys) || dmalloc(&sb.kstart, (size_t)nkeys) ||
            hipMalloc(&sb.kscan_part, sizeof(int2) * (size_t)kparts) != hipSuccess)
            return kErrHip;
    }
    return 0;
}

// Method 4 setpts: count the copies of every point per block (one copy per
// block its support touches), read the total back (the copy arrays are sized
// by it; the reference's setpts also synchronises), then scatter.
template <typename T>
int Plan<T>::setpts_blocks()
{
    const int maxsub = opts.gpu_maxsubprobsize;
    const int obv[3] = {geom.bs[0], geom.bs[1], geom.bs[2]};
    if ((long long)M * blk_max_copies(dim, sp.ns, obv) > 2147483647LL) return kErrNdataNotValid;
    Stage st(this, kStageSort);
    int e = run_blk_count<T>(dim, geom, sp.ns, M, kx, ky, kz, sb, L);
    if (e) return e;
    int ncopies = 0;
    if (hipMemcpyAsync(&ncopies, sb.counters, sizeof(int), hipMemcpyDeviceToHost, L.stream) !=
            hipSuccess ||
        hipStreamSynchronize(L.stream) != hipSuccess)
        return kErrHip;
    if (ncopies > cap_copies) {
        dfree(sb.idx);
        dfree(sb.pos);
        if (dmalloc(&sb.idx, (size_t)std::max(ncopies, 1)) ||
            dmalloc(&sb.pos, (size_t)std::max(ncopies, 1) * dim)) {
            cap_copies = -1;
            return kErrHip;
        }
        cap_copies = ncopies;
        sb.pos_stride = std::max(ncopies, 1);
    }
    sub_bound = nbins + (ncopies + (long long)maxsub - 1) / maxsub;
    if (sub_bound > cap_sub) {
        dfree(sb.s2b);
        if (dmalloc(&sb.s2b, (size_t)sub_bound)) {
            cap_sub = -1;
            return kErrHip;
        }
        cap_sub = sub_bound;
    }
    e = run_blk_scatter<T>(dim, geom, sp.ns, M, ncopies, kx, ky, kz, maxsub, sb, L);
    if (e) return e;
    pts_set = true;
    return 0;
}

template <typename T>
int Plan<T>::setpts(int M_, const T *x, const T *y, const T *z)
{
    ApiRange ar("cufinufft_setpts");
    if (M_ < 0) return kErrNdataNotValid;
    if (M_ > 0 && (!x || (dim > 1 && !y) || (dim > 2 && !z))) return 1;
    DeviceGuard dg(opts.gpu_device_id);
    M = M_;
    kx = x;
    ky = dim > 1 ? y : nullptr;
    kz = dim > 2 ? z : nullptr;
    if (M > cap_M && method != 4) {
        dfree(sb.skey);
        dfree(sb.skey_out);
        dfree(sb.idx);
        dfree(sb.pos);
        if (dmalloc(&sb.skey, M) || dmalloc(&sb.skey_out, M) || dmalloc(&sb.idx, M) ||
            dmalloc(&sb.pos, (size_t)M * dim)) {
            cap_M = -1;
            return kErrHip;
        }
        cap_M = M;
        sb.pos_stride = M;
    }
    if (method != 4 && !(method == 1 && !opts.gpu_sort) && M > 0) {
        // counting-sort scratch: (digit, tile) counts and their scan
        // partials, and the point records between passes (none for one pass)
        const long long nkeys =
            opts.gpu_sort == 2 ? nbins * (long long)geom.bs[0] * geom.bs[1] * geom.bs[2] : nbins;
        const BinSortShape sh = bin_sort_shape(M, nkeys);
        if (sh.counts > cap_counts) {
            dfree(sb.sort_counts);
            if (sb.sort_part) hipFree(sb.sort_part);
            sb.sort_part = nullptr;
            cap_counts = -1;
            const long long parts = (sh.counts + kScanChunk - 1) / kScanChunk;
            if (dmalloc(&sb.sort_counts, (size_t)sh.counts) ||
                hipMalloc(&sb.sort_part, sizeof(int2) * (size_t)parts) != hipSuccess)
                return kErrHip;
            cap_counts = sh.counts;
        }
        const size_t rb = bin_sort_rec_bytes(dim, sizeof(T));
        for (int r = 0; r < 2; ++r) {
            const long long need = sh.passes > r + 1 ? M : 0;  // pass r writes records r
            if (need > cap_rec[r]) {
                if (sb.rec[r]) hipFree(sb.rec[r]);
                sb.rec[r] = nullptr;
                cap_rec[r] = -1;
                if (hipMalloc(&sb.rec[r], rb * (size_t)need) != hipSuccess) return kErrHip;
                cap_rec[r] = need;
            }
        }
    }
    const int maxsub = opts.gpu_maxsubprobsize;
    if (method == 4) return setpts_blocks();
    sub_bound = nbins + (M + (long long)maxsub - 1) / maxsub;
    const bool need_sub = method == 2;
    if (need_sub && sub_bound > cap_sub) {
        dfree(sb.s2b);
        if (dmalloc(&sb.s2b, (size_t)sub_bound)) {
            cap_sub = -1;
            return kErrHip;
        }
        cap_sub = sub_bound;
    }
    int e;
    {
        Stage st(this, kStageSort);
        if (method == 1 && !opts.gpu_sort) {
            e = run_natural_order<T>(dim, geom, M, kx, ky, kz, sb, L);
        } else {
            e = run_bin_sort<T>(dim, geom, nbins, M, kx, ky, kz, maxsub, opts.gpu_sort == 2,
                                need_sub, sb, L);
        }
    }
    if (e) return e;
    pts_set = true;
    return 0;
}

// setpts without allocation (cufinufftf_setpts_noalloc): the same bin sort,
// with its transient buffers (the key ping-pong pair, the point records
// between passes, the per-pass (digit, tile) counts and their scan partials)
// carved from the caller-supplied device scratch instead of the plan's own.
// The plan's lasting point arrays (sorted index, positions, subproblem map)
// must already hold M points: an earlier setpts of at least M points sized
// them.  Anything larger returns kErrNoAlloc and leaves the plan as it was.
template <typename T>
int Plan<T>::setpts_noalloc(int M_, const T *x, const T *y, const T *z, unsigned char *scr,
                            size_t scr_bytes)
{
    ApiRange ar("cufinufft_setpts_noalloc");
    if (M_ < 0) return kErrNdataNotValid;
    if (M_ > 0 && (!x || (dim > 1 && !y) || (dim > 2 && !z))) return 1;
    if (method == 4) return 2;  // block copies: their count is known only after counting
    if (M_ > cap_M) return kErrNoAlloc;
    const int maxsub = opts.gpu_maxsubprobsize;
    const long long sub_need = nbins + (M_ + (long long)maxsub - 1) / maxsub;
    const bool need_sub = method == 2;
    if (need_sub && sub_need > cap_sub) return kErrNoAlloc;
    SortBufs s2 = sb;
    if (!(method == 1 && !opts.gpu_sort) && M_ > 0) {
        const long long nkeys =
            opts.gpu_sort == 2 ? nbins * (long long)geom.bs[0] * geom.bs[1] * geom.bs[2] : nbins;
        const BinSortShape sh = bin_sort_shape(M_, nkeys);
        const size_t rb = bin_sort_rec_bytes(dim, sizeof(T));
        const long long parts = (sh.counts + kScanChunk - 1) / kScanChunk;
        size_t off = 0;
        auto take = [&](size_t bytes) -> unsigned char * {
            unsigned char *p = scr + off;
            off += (bytes + 255) / 256 * 256;
            return p;
        };
        s2.skey = reinterpret_cast<int *>(take(sizeof(int) * (size_t)M_));
        s2.skey_out = reinterpret_cast<int *>(take(sizeof(int) * (size_t)M_));
        s2.rec[0] = sh.passes > 1 ? take(rb * (size_t)M_) : nullptr;
        s2.rec[1] = sh.passes > 2 ? take(rb * (size_t)M_) : nullptr;
        s2.sort_counts = reinterpret_cast<int *>(take(sizeof(int) * (size_t)sh.counts));
        s2.sort_part = take(sizeof(int2) * (size_t)parts);
        if (off > scr_bytes) return kErrNoAlloc;
    }
    DeviceGuard dg(opts.gpu_device_id);
    M = M_;
    kx = x;
    ky = dim > 1 ? y : nullptr;
    kz = dim > 2 ? z : nullptr;
    sub_bound = sub_need;
    int e;
    {
        Stage st(this, kStageSort);
        if (method == 1 && !opts.gpu_sort) {
            e = run_natural_order<T>(dim, geom, M, kx, ky, kz, s2, L);
        } else {
            e = run_bin_sort<T>(dim, geom, nbins, M, kx, ky, kz, maxsub, opts.gpu_sort == 2,
                                need_sub, s2, L);
        }
    }
    pts_set = e == 0;
    return e;
}

template <typename T>
int Plan<T>::execute(cpx<T> *c, cpx<T> *fk)
{
    ApiRange ar("cufinufft_execute");
    if (!pts_set) return 1;
    if ((M > 0 && !c) || !fk) return 1;
    DeviceGuard dg(opts.gpu_device_id);
    const KerParams<T> kp = ker_params();
    const int maxsub = opts.gpu_maxsubprobsize;
    int e = 0;
    if (opts.gpu_spreadinterponly) {
        // The fine grid is the caller's fk (nf = nmodes): spread into it or
        // interpolate from it, one batch at a time.
        for (int b0 = 0; b0 < ntransf; b0 += maxbatch) {
            const int blk = std::min(maxbatch, ntransf - b0);
            cpx<T> *cb = c + (long long)b0 * M;
            cpx<T> *fb = fk + (long long)b0 * nftot;
            if (type == 1) {
                if ((e = run_zero<T>(fb, nftot * blk, L))) return e;
                if ((e = run_spread<T>(method, dim, sp.ns, M, sb, cb, fb, geom, kp,
                                       maxsub, sub_bound, blk, nftot, L)))
                    return e;
            } else {
                if ((e = run_interp<T>(method, dim, sp.ns, M, sb, cb, fb, geom, kp,
                                       maxsub, sub_bound, blk, nftot, L)))
                    return e;
            }
        }
        return 0;
    }
    for (int b0 = 0; b0 < ntransf; b0 += maxbatch) {
        const int blk = std::min(maxbatch, ntransf - b0);
        cpx<T> *cb = c + (long long)b0 * M;
        cpx<T> *fb = fk + (long long)b0 * ntot;
        if (type == 1) {
            {
                Stage st(this, kStageGrid);
                e = run_zero<T>(fw, nftot * blk, L);
            }
            if (e) return e;
            {
                Stage st(this, kStageSpreadInterp);
                e = run_spread<T>(method, dim, sp.ns, M, sb, cb, fw, geom, kp, maxsub,
                                  sub_bound, blk, nftot, L);
            }
            if (e) return e;
            {
                Stage st(this, kStageFft);
                e = fft_type1(blk);
            }
            if (e) return e;
            {
                Stage st(this, kStageDeconv);
                e = deconv_type1(fb, blk);
            }
            if (e) return e;
        } else {
            {
                Stage st(this, kStageDeconv);
                e = amplify_type2(fb, blk, -1, 1);
            }
            if (e) return e;
            {
                Stage st(this, kStageFft);
                e = fft_type2(blk);
            }
            if (e) return e;
            {
                Stage st(this, kStageSpreadInterp);
                e = run_interp<T>(method, dim, sp.ns, M, sb, cb, fw, geom, kp, maxsub,
                                  sub_bound, blk, nftot, L);
            }
            if (e) return e;
        }
    }
    return 0;
}

// Type-2 execute that also returns the derivatives of every output with
// respect to its point's coordinates: dc[d][t*M + j] = d c[t*M + j] / d x_d(j).
// Each derivative is itself a type-2 transform, of the modes weighted by
// i * iflag * k_d (the weight is applied in the amplify pass), so it carries
// the transform's own error bound: amplify, FFT and the plan's regular
// interpolation run once for the values and once per dimension.
template <typename T>
int Plan<T>::execute_grad(cpx<T> *c, cpx<T> *dc, cpx<T> *fk)
{
    ApiRange ar("cufinufft_execute_grad");
    if (type != 2 || !pts_set || opts.gpu_spreadinterponly) return 1;
    if ((M > 0 && (!c || !dc)) || !fk) return 1;
    DeviceGuard dg(opts.gpu_device_id);
    const KerParams<T> kp = ker_params();
    const int maxsub = opts.gpu_maxsubprobsize;
    const long long plane = (long long)M * ntransf;
    int e = 0;
    for (int comp = 0; comp <= dim; ++comp) {  // 0: values, 1 + d: d/dx_d
        cpx<T> *out = comp == 0 ? c : dc + (long long)(comp - 1) * plane;
        for (int b0 = 0; b0 < ntransf; b0 += maxbatch) {
            const int blk = std::min(maxbatch, ntransf - b0);
            {
                Stage st(this, kStageDeconv);
                e = amplify_type2(fk + (long long)b0 * ntot, blk, comp - 1, iflag);
            }
            if (e) return e;
            {
                Stage st(this, kStageFft);
                e = fft_type2(blk);
            }
            if (e) return e;
            {
                Stage st(this, kStageSpreadInterp);
                e = run_interp<T>(method, dim, sp.ns, M, sb, out + (long long)b0 * M, fw, geom, kp,
                                  maxsub, sub_bound, blk, nftot, L);
            }
            if (e) return e;
        }
    }
    return 0;
}

template <typename T>
int get_timing(Plan<T> &p, double *ms, int *count)
{
    for (int s = 0; s < kNumStages; ++s) {
        if (ms) ms[s] = 0;
        if (count) count[s] = 0;
    }
    DeviceGuard dg(p.opts.gpu_device_id);
    if (hipStreamSynchronize(p.L.stream) != hipSuccess) return kErrHip;
    for (const auto &sp : p.spans) {
        float t = 0;
        if (hipEventElapsedTime(&t, sp.a, sp.b) != hipSuccess) return kErrHip;
        if (ms) ms[sp.stage] += t;
        if (count) count[sp.stage] += 1;
    }
    p.spans.clear();
    p.ev_used = 0;
    return 0;
}

template struct Plan<float>;
template struct Plan<double>;

// ---------------------------------------------------------------------------
// Plan group: many independent plans (same type, dimension, precision and
// kernel width; any N, M, bins) executed as one launch sequence whose
// spread / interp is a single persistent kernel over the subproblems of every
// member.  The members' device arrays and the caller's c / fk pointers are
// snapshotted into a device descriptor table at make time, so the group must
// be re-made after any member's setpts.  The sequence runs on the group's own
// stream, ordered after and before the caller's stream with events; with
// use_graph it is captured into a hipGraph on the first call and replayed.
template <typename T>
struct Group {
    std::vector<Plan<T> *> plans;
    std::vector<cpx<T> *> c, fk;
    int type = 0, dim = 0, ns = 0, dev = 0;
    int bs[3] = {1, 1, 1};
    int maxsub = 1;
    long long jobs_bound = 0;
    void *items = nullptr;
    long long *off = nullptr;  // n + 1 job offsets, then the job-queue words (zeroed)
    hipStream_t stream = nullptr, user = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    Launch L;
    // Branch streams for the per-member passes (zero fill, FFT, deconvolve /
    // amplify): small members leave most CUs idle, so members run
    // concurrently, round-robin over kBranches streams forked from and joined
    // back to the group stream with events (parallel branches in a captured
    // graph).  Four: the hardware queues a process gets by default.
    static constexpr int kBranches = 4;
    hipStream_t br[kBranches] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kBranches] = {};

    ~Group()
    {
        DeviceGuard dg(dev);
        if (gexec) hipGraphExecDestroy(gexec);
        if (graph) hipGraphDestroy(graph);
        if (ev_in) hipEventDestroy(ev_in);
        if (ev_out) hipEventDestroy(ev_out);
        if (ev_fork) hipEventDestroy(ev_fork);
        for (int b = 0; b < kBranches; ++b) {
            if (ev_join[b]) hipEventDestroy(ev_join[b]);
            if (br[b]) hipStreamDestroy(br[b]);
        }
        if (stream) hipStreamDestroy(stream);
        if (items) hipFree(items);
        if (off) hipFree(off);
    }

    int make(int n, Plan<T> *const *ps, void *const *cs, void *const *fks)
    {
        if (n < 1 || !ps || !cs || !fks) return 1;
        const Plan<T> &p0 = *ps[0];
        type = p0.type;
        dim = p0.dim;
        ns = p0.sp.ns;
        dev = p0.opts.gpu_device_id;
        for (int i = 0; i < n; ++i) {
            const Plan<T> *p = ps[i];
            if (!p || !p->pts_set || p->type != type || p->dim != dim || p->sp.ns != ns ||
                p->sp.beta != p0.sp.beta || p->method != 2 || p->maxbatch != p->ntransf ||
                p->opts.gpu_spreadinterponly || p->opts.gpu_device_id != dev || !cs[i] ||
                !fks[i])
                return 1;
            for (int d = 0; d < 3; ++d) bs[d] = std::max(bs[d], p->geom.bs[d]);
            jobs_bound += p->sub_bound * p->ntransf;
            maxsub = std::max(maxsub, p->opts.gpu_maxsubprobsize);
            plans.push_back(ps[i]);
            c.push_back(static_cast<cpx<T> *>(cs[i]));
            fk.push_back(static_cast<cpx<T> *>(fks[i]));
        }
        DeviceGuard dg(dev);
        L.ncu = p0.L.ncu;
        user = p0.L.stream;
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return kErrHip;
        L.stream = stream;
        if (hipEventCreateWithFlags(&ev_in, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ev_out, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming) != hipSuccess)
            return kErrHip;
        for (int b = 0; b < kBranches; ++b)
            if (hipStreamCreateWithFlags(&br[b], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&ev_join[b], hipEventDisableTiming) != hipSuccess)
                return kErrHip;
        const size_t isz = group_item_bytes(sizeof(T) == 8);
        std::vector<unsigned char> h(isz * n);
        for (int i = 0; i < n; ++i) {
            Plan<T> &p = *plans[i];
            const void *in = type == 1 ? (const void *)c[i] : (const void *)p.fw;
            void *out = type == 1 ? (void *)p.fw : (void *)c[i];
            unsigned char *dst = h.data() + isz * i;
            if (dim == 1)
                group_item_fill<T, 1>(dst, p.geom, p.sb, p.opts.gpu_maxsubprobsize, in, out, p.nftot,
                                      p.M, p.ntransf);
            else if (dim == 2)
                group_item_fill<T, 2>(dst, p.geom, p.sb, p.opts.gpu_maxsubprobsize, in, out, p.nftot,
                                      p.M, p.ntransf);
            else
                group_item_fill<T, 3>(dst, p.geom, p.sb, p.opts.gpu_maxsubprobsize, in, out, p.nftot,
                                      p.M, p.ntransf);
        }
        if (hipMalloc(&items, h.size()) != hipSuccess ||
            hipMalloc((void **)&off, sizeof(long long) * (n + 1 + kJobqWords)) != hipSuccess ||
            hipMemset(off + n + 1, 0, sizeof(long long) * kJobqWords) != hipSuccess)
            return kErrHip;
        if (hipMemcpy(items, h.data(), h.size(), hipMemcpyHostToDevice) != hipSuccess)
            return kErrHip;
        return 0;
    }

    // The launch sequence, on the group stream.
    int body()
    {
        int e = 0;
        const Plan<T> &p0 = *plans[0];
        KerParams<T> kp = p0.ker_params();
        // the strictest member's 32-bit bound applies to every job of the
        // group (members may differ in tol and maxsub)
        for (const auto *p : plans) kp.fx32max = std::min(kp.fx32max, p->fx32max);
        const int n = (int)plans.size();
        std::vector<hipStream_t> saved(n);
        for (int i = 0; i < n; ++i) {
            saved[i] = plans[i]->L.stream;
            plans[i]->L.stream = stream;  // the members' FFTs run on the group stream
        }
        auto spread_or_interp = [&]() {
            if (dim == 1)
                return group_dim<T, 1>(type, ns, items, n, off, kp, bs, maxsub, jobs_bound, L);
            if (dim == 2)
                return group_dim<T, 2>(type, ns, items, n, off, kp, bs, maxsub, jobs_bound, L);
            return group_dim<T, 3>(type, ns, items, n, off, kp, bs, maxsub, jobs_bound, L);
        };
        // per-member passes on the branch streams, between a fork and a join
        auto branches = [&](auto member) {
            if (hipEventRecord(ev_fork, stream) != hipSuccess) return kErrHip;
            for (int b = 0; b < kBranches; ++b)
                if (hipStreamWaitEvent(br[b], ev_fork, 0) != hipSuccess) return kErrHip;
            int err = 0;
            for (int i = 0; i < n && !err; ++i) {
                Launch Lb = L;
                Lb.stream = br[i % kBranches];
                plans[i]->L.stream = Lb.stream;  // the member's FFT follows its branch
                err = member(i, Lb);
            }
            for (int b = 0; b < kBranches; ++b)
                if (hipEventRecord(ev_join[b], br[b]) != hipSuccess ||
                    hipStreamWaitEvent(stream, ev_join[b], 0) != hipSuccess)
                    return kErrHip;
            return err;
        };
        if (type == 1) {
            e = branches([&](int i, const Launch &Lb) {
                return run_zero<T>(plans[i]->fw, plans[i]->nftot * plans[i]->ntransf, Lb);
            });
            if (!e) e = spread_or_interp();
            if (!e)
                e = branches([&](int i, const Launch &Lb) {
                    Plan<T> &p = *plans[i];
                    (void)Lb;  // the member's own launches follow its branch stream
                    int err = p.fft_type1(p.ntransf);
                    if (!err) err = p.deconv_type1(fk[i], p.ntransf);
                    return err;
                });
        } else {
            e = branches([&](int i, const Launch &Lb) {
                Plan<T> &p = *plans[i];
                (void)Lb;  // the member's own launches follow its branch stream
                int err = p.amplify_type2(fk[i], p.ntransf, -1, 1);
                if (!err) err = p.fft_type2(p.ntransf);
                return err;
            });
            if (!e) e = spread_or_interp();
        }
        for (int i = 0; i < n; ++i) plans[i]->L.stream = saved[i];
        return e;
    }

    int execute(int use_graph)
    {
        ApiRange ar("cufinufft_group_execute");
        DeviceGuard dg(dev);
        if (hipEventRecord(ev_in, user) != hipSuccess ||
            hipStreamWaitEvent(stream, ev_in, 0) != hipSuccess)
            return kErrHip;
        int e = 0;
        if (!use_graph) {
            e = body();
        } else {
            if (!gexec) {
                if (hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed) != hipSuccess)
                    return kErrHip;
                e = body();
                hipGraph_t g = nullptr;
                const hipError_t ce = hipStreamEndCapture(stream, &g);
                if (e) {
                    if (g) hipGraphDestroy(g);
                    return e;
                }
                if (ce != hipSuccess || !g) return kErrHip;
                graph = g;
                if (hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
                    gexec = nullptr;
                    return kErrHip;
                }
            }
            if (hipGraphLaunch(gexec, stream) != hipSuccess) return kErrHip;
        }
        if (e) return e;
        if (hipEventRecord(ev_out, stream) != hipSuccess ||
            hipStreamWaitEvent(user, ev_out, 0) != hipSuccess)
            return kErrHip;
        return 0;
    }
};

template struct Group<float>;
template struct Group<double>;

}  // namespace cfa

struct cufinufft_group_s : cfa::Group<double> {};
struct cufinufftf_group_s : cfa::Group<float> {};

struct cufinufft_plan_s : cfa::Plan<double> {};
struct cufinufftf_plan_s : cfa::Plan<float> {};

namespace {

template <typename P, typename T>
int makeplan_impl(int type, int dim, int *nmodes, int iflag, int ntransf, T tol, int maxbatch,
                  P **out, cufinufft_opts *opts, int flags = 0)
{
    cfa::ApiRange ar("cufinufft_makeplan");
    if (!out) return 1;
    *out = nullptr;
    P *p = new (std::nothrow) P();
    if (!p) return 1;
    // Any failure frees the plan; a successful plan may still carry the
    // reference's WARN_EPS_TOO_SMALL (tol clamped to machine precision).
    const int ier = p->makeplan(type, dim, nmodes, iflag, ntransf, tol, maxbatch, opts, flags);
    if (ier != 0) {
        delete p;
        return ier;
    }
    *out = p;
    return p->eps_warning ? 1 : 0;
}

template <typename P>
int get_info_impl(P *p, int *out)
{
    if (!p || !out) return 1;
    int total = 0;
    if (p->pts_set && p->sb.counters) {
        cfa::DeviceGuard dg(p->opts.gpu_device_id);
        if (hipStreamSynchronize(p->L.stream) != hipSuccess) return cfa::kErrHip;
        if (hipMemcpy(&total, p->sb.counters + 1, sizeof(int), hipMemcpyDeviceToHost) !=
            hipSuccess)
            return cfa::kErrHip;
    }
    const int v[16] = {p->sp.ns,        p->nf[0],         p->nf[1],         p->nf[2],
                       p->geom.nb[0],   p->geom.nb[1],    p->geom.nb[2],    p->method,
                       p->geom.bs[0],   p->geom.bs[1],    p->geom.bs[2],    p->opts.gpu_maxsubprobsize,
                       p->M,            total,            p->maxbatch,      p->opts.gpu_spreadinterponly};
    std::memcpy(out, v, sizeof(v));
    return 0;
}

template <typename P>
int get_sort_impl(P *p, int *binsize, int *binstart, int *idx, int *substart, int *s2b)
{
    if (!p || !p->pts_set) return 1;
    cfa::DeviceGuard dg(p->opts.gpu_device_id);
    if (hipStreamSynchronize(p->L.stream) != hipSuccess) return cfa::kErrHip;
    const size_t nb = (size_t)p->nbins;
    auto cp = [](int *dst, const int *src, size_t n) {
        return !dst || !src || n == 0 ||
               hipMemcpy(dst, src, n * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    };
    int total = 0;
    if (hipMemcpy(&total, p->sb.counters + 1, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return cfa::kErrHip;
    bool ok = cp(binsize, p->sb.binsize, nb) && cp(binstart, p->sb.binstart, nb) &&
              cp(idx, p->sb.idx, (size_t)p->M) && cp(substart, p->sb.substart, nb + 1) &&
              (p->method != 2 || cp(s2b, p->sb.s2b, (size_t)total));
    return ok ? 0 : cfa::kErrHip;
}

template <typename P>
int get_fseries_impl(P *p, double *k1, double *k2, double *k3)
{
    if (!p) return 1;
    double *o[3] = {k1, k2, k3};
    for (int d = 0; d < p->dim; ++d)
        if (o[d] && !p->fser[d].empty())
            std::memcpy(o[d], p->fser[d].data(), p->fser[d].size() * sizeof(double));
    return 0;
}

template <typename P>
int get_kernel_poly_impl(P *p, int *deg, double *rel_err, double *coef)
{
    if (!p) return 1;
    if (deg) *deg = p->hdeg;
    if (rel_err) *rel_err = p->herr;
    if (coef && p->hdeg > 0)
        std::memcpy(coef, p->hcoef_host.data(), p->hcoef_host.size() * sizeof(double));
    return 0;
}

}  // namespace

extern "C" {

int cufinufft_default_opts(int type, int dim, cufinufft_opts *opts)
{
    return cfa::default_opts(type, dim, opts);
}
int cufinufftf_default_opts(int type, int dim, cufinufft_opts *opts)
{
    return cfa::default_opts(type, dim, opts);
}

int cufinufft_makeplan(int type, int dim, int *nmodes, int iflag, int ntransf, double tol,
                       int maxbatchsize, cufinufft_plan *plan, cufinufft_opts *opts)
{
    return makeplan_impl<cufinufft_plan_s, double>(type, dim, nmodes, iflag, ntransf, tol,
                                                   maxbatchsize, plan, opts);
}
int cufinufftf_makeplan(int type, int dim, int *nmodes, int iflag, int ntransf, float tol,
                        int maxbatchsize, cufinufftf_plan *plan, cufinufft_opts *opts)
{
    return makeplan_impl<cufinufftf_plan_s, float>(type, dim, nmodes, iflag, ntransf, tol,
                                                   maxbatchsize, plan, opts);
}

int cufinufft_makeplan_ext(int type, int dim, int *nmodes, int iflag, int ntransf, double tol,
                           int maxbatchsize, cufinufft_plan *plan, cufinufft_opts *opts, int flags)
{
    return makeplan_impl<cufinufft_plan_s, double>(type, dim, nmodes, iflag, ntransf, tol,
                                                   maxbatchsize, plan, opts, flags);
}
int cufinufftf_makeplan_ext(int type, int dim, int *nmodes, int iflag, int ntransf, float tol,
                            int maxbatchsize, cufinufftf_plan *plan, cufinufft_opts *opts,
                            int flags)
{
    return makeplan_impl<cufinufftf_plan_s, float>(type, dim, nmodes, iflag, ntransf, tol,
                                                   maxbatchsize, plan, opts, flags);
}

int cufinufft_setpts(int M, double *x, double *y, double *z, int, double *, double *, double *,
                     cufinufft_plan plan)
{
    return plan ? plan->setpts(M, x, y, z) : 1;
}
int cufinufftf_setpts(int M, float *x, float *y, float *z, int, float *, float *, float *,
                      cufinufftf_plan plan)
{
    return plan ? plan->setpts(M, x, y, z) : 1;
}

}  // extern "C"

// Device scratch of cufinufftf_setpts_noalloc: one fixed array per GPU, held
// until the process ends.  Host code never uses the array's name as an
// address (on the host it names a stand-in, not device memory): each device's
// address is looked up once, with that device current, and cached.
namespace {
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of the scratch on the current device (resolved on first use).
int scratch_address(unsigned char **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxScratchDevices)
        return cfa::kErrHip;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return cfa::kErrHip;
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" {

int cufinufftf_setpts_noalloc(int M, float *x, float *y, float *z, cufinufftf_plan plan)
{
    if (!plan) return 1;
    cfa::DeviceGuard dg(plan->opts.gpu_device_id);
    unsigned char *scr = nullptr;
    const int e = scratch_address(&scr);
    if (e) return e;
    return plan->setpts_noalloc(M, x, y, z, scr, kNoAllocScratchBytes);
}

int cufinufft_execute(void *c, void *fk, cufinufft_plan plan)
{
    return plan ? plan->execute((cfa::cpx<double> *)c, (cfa::cpx<double> *)fk) : 1;
}
int cufinufftf_execute(void *c, void *fk, cufinufftf_plan plan)
{
    return plan ? plan->execute((cfa::cpx<float> *)c, (cfa::cpx<float> *)fk) : 1;
}

int cufinufft_group_make(int n, cufinufft_plan *plans, void **c, void **fk,
                         cufinufft_group *out)
{
    if (!out || !plans || n < 1) return 1;
    *out = nullptr;
    std::vector<cfa::Plan<double> *> ps(plans, plans + n);
    auto *g = new (std::nothrow) cufinufft_group_s();
    if (!g) return 1;
    const int e = g->make(n, ps.data(), c, fk);
    if (e) {
        delete g;
        return e;
    }
    *out = g;
    return 0;
}
int cufinufftf_group_make(int n, cufinufftf_plan *plans, void **c, void **fk,
                          cufinufftf_group *out)
{
    if (!out || !plans || n < 1) return 1;
    *out = nullptr;
    std::vector<cfa::Plan<float> *> ps(plans, plans + n);
    auto *g = new (std::nothrow) cufinufftf_group_s();
    if (!g) return 1;
    const int e = g->make(n, ps.data(), c, fk);
    if (e) {
        delete g;
        return e;
    }
    *out = g;
    return 0;
}
int cufinufft_group_execute(cufinufft_group g, int use_graph)
{
    return g ? g->execute(use_graph) : 1;
}
int cufinufftf_group_execute(cufinufftf_group g, int use_graph)
{
    return g ? g->execute(use_graph) : 1;
}
int cufinufft_group_setstream(cufinufft_group g, void *stream)
{
    if (!g) return 1;
    g->user = (hipStream_t)stream;
    return 0;
}
int cufinufftf_group_setstream(cufinufftf_group g, void *stream)
{
    if (!g) return 1;
    g->user = (hipStream_t)stream;
    return 0;
}
int cufinufft_group_destroy(cufinufft_group g)
{
    if (!g) return 1;
    delete g;
    return 0;
}
int cufinufftf_group_destroy(cufinufftf_group g)
{
    if (!g) return 1;
    delete g;
    return 0;
}

int cufinufft_execute_grad(void *c, void *dc, void *fk, cufinufft_plan plan)
{
    return plan ? plan->execute_grad((cfa::cpx<double> *)c, (cfa::cpx<double> *)dc,
                                     (cfa::cpx<double> *)fk)
                : 1;
}
int cufinufftf_execute_grad(void *c, void *dc, void *fk, cufinufftf_plan plan)
{
    return plan ? plan->execute_grad((cfa::cpx<float> *)c, (cfa::cpx<float> *)dc,
                                     (cfa::cpx<float> *)fk)
                : 1;
}

int cufinufft_destroy(cufinufft_plan plan)
{
    if (!plan) return 1;
    delete plan;
    return 0;
}
int cufinufftf_destroy(cufinufftf_plan plan)
{
    if (!plan) return 1;
    delete plan;
    return 0;
}

int cufinufft_setstream(cufinufft_plan plan, void *stream)
{
    if (!plan) return 1;
    plan->L.stream = (hipStream_t)stream;
    return 0;
}
int cufinufftf_setstream(cufinufftf_plan plan, void *stream)
{
    if (!plan) return 1;
    plan->L.stream = (hipStream_t)stream;
    return 0;
}

int cufinufft_get_info(cufinufft_plan plan, int *out) { return get_info_impl(plan, out); }
int cufinufftf_get_info(cufinufftf_plan plan, int *out) { return get_info_impl(plan, out); }

int cufinufft_get_sort(cufinufft_plan plan, int *binsize, int *binstart, int *idxnupts,
                       int *subprobstart, int *subprob_to_bin)
{
    return get_sort_impl(plan, binsize, binstart, idxnupts, subprobstart, subprob_to_bin);
}
int cufinufftf_get_sort(cufinufftf_plan plan, int *binsize, int *binstart, int *idxnupts,
                        int *subprobstart, int *subprob_to_bin)
{
    return get_sort_impl(plan, binsize, binstart, idxnupts, subprobstart, subprob_to_bin);
}

int cufinufft_set_timing(cufinufft_plan plan, int on)
{
    if (!plan) return 1;
    plan->timing = on != 0;
    return 0;
}
int cufinufftf_set_timing(cufinufftf_plan plan, int on)
{
    if (!plan) return 1;
    plan->timing = on != 0;
    return 0;
}
int cufinufft_get_timing(cufinufft_plan plan, double *ms, int *count)
{
    return plan ? cfa::get_timing<double>(*plan, ms, count) : 1;
}
int cufinufftf_get_timing(cufinufftf_plan plan, double *ms, int *count)
{
    return plan ? cfa::get_timing<float>(*plan, ms, count) : 1;
}

int cufinufft_get_fseries(cufinufft_plan plan, double *k1, double *k2, double *k3)
{
    return get_fseries_impl(plan, k1, k2, k3);
}
int cufinufftf_get_fseries(cufinufftf_plan plan, double *k1, double *k2, double *k3)
{
    return get_fseries_impl(plan, k1, k2, k3);
}

int cufinufft_get_kernel_poly(cufinufft_plan plan, int *deg, double *rel_err, double *coef)
{
    return get_kernel_poly_impl(plan, deg, rel_err, coef);
}
int cufinufftf_get_kernel_poly(cufinufftf_plan plan, int *deg, double *rel_err, double *coef)
{
    return get_kernel_poly_impl(plan, deg, rel_err, coef);
}

}  // extern "C"
