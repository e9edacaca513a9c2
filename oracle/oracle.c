/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C CPU restatement of cuFINUFFT v1.3's type-1/type-2 NUFFT path, used
 * only as the checker: by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  It is never linked into, loaded by or called from the
 * product library (cufinufft_amd/libcufinufft.so).
 *
 * Each function cites the reference file:line it restates.  Floating-point
 * results are computed in FP64 (the oracle's job is to be more accurate than
 * the path under test); the integer chain that fixes the bin sort (RESCALE,
 * bin index, counts, scans) is restated bit-faithfully in the plan's precision,
 * because bin indices and per-bin counts must match exactly.
 *
 * Pinning (what checks each function independently of its own formula) is
 * recorded in the test files that exercise it: tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_M_1_2PI 0.159154943091895336 /* contrib/utils.h:18 */
#define ORC_MAX_NSPREAD 16               /* contrib/spreadinterp.h:10 */

/* ------------------------------------------------------------------------ */
/* next235beven: contrib/utils.cpp:3-22.  Smallest even n' >= n whose prime
   factors are in {2,3,5} and which is a multiple of b.                     */
int orc_next235beven(int n, int b)
{
    if (n <= 2) return 2;
    if (n % 2 == 1) n += 1;
    int nplus = n - 2;
    int numdiv = 2;
    while (numdiv > 1 || nplus % b != 0) {
        nplus += 2;
        numdiv = nplus;
        while (numdiv % 2 == 0) numdiv /= 2;
        while (numdiv % 3 == 0) numdiv /= 3;
        while (numdiv % 5 == 0) numdiv /= 5;
    }
    return nplus;
}

/* ------------------------------------------------------------------------ */
/* setup_spreader: contrib/spreadinterp.cpp:6-67, evaluated in the plan's
   precision exactly as the reference's C++ overloads resolve (float log10,
   ceil and sqrt for single precision).  Returns the reference's ier codes:
   0 ok, 1 WARN_EPS_TOO_SMALL, 7 ERR_UPSAMPFAC_TOO_SMALL, 8 HORNER_WRONG_BETA
   (contrib/utils.h:28-36).                                                   */
int orc_setup_spreader_f(float eps, float upsampfac, int kerevalmeth,
                         int *ns_out, float *beta_out, float *c_out,
                         float *halfwidth_out)
{
    const float PIf = (float)M_PI;
    if (upsampfac != 2.0) {
        if (kerevalmeth == 1) return 8;
        if (upsampfac <= 1.0) return 7;
    }
    int ier = 0;
    const float EPS = (float)6e-08; /* contrib/utils_fp.h:34 */
    if (eps < EPS) { eps = EPS; ier = 1; }
    int ns = (int)ceilf(-log10f(eps / (float)10.0));
    if (upsampfac != 2.0)
        ns = (int)ceilf(-logf(eps) / (PIf * sqrtf(1 - 1 / upsampfac)));
    if (ns < 2) ns = 2;
    if (ns > ORC_MAX_NSPREAD) { ns = ORC_MAX_NSPREAD; ier = 1; }
    *ns_out = ns;
    *halfwidth_out = (float)ns / 2;
    *c_out = (float)(4.0 / (double)(float)(ns * ns));
    float betaoverns = 2.30f;
    if (ns == 2) betaoverns = 2.20f;
    if (ns == 3) betaoverns = 2.26f;
    if (ns == 4) betaoverns = 2.38f;
    if (upsampfac != 2.0) {
        float gamma = 0.97f;
        betaoverns = gamma * PIf * (1 - 1 / (2 * upsampfac));
    }
    *beta_out = betaoverns * (float)ns;
    return ier;
}

int orc_setup_spreader_d(double eps, double upsampfac, int kerevalmeth,
                         int *ns_out, double *beta_out, double *c_out,
                         double *halfwidth_out)
{
    const double PId = M_PI;
    if (upsampfac != 2.0) {
        if (kerevalmeth == 1) return 8;
        if (upsampfac <= 1.0) return 7;
    }
    int ier = 0;
    const double EPS = 1.1e-16; /* contrib/utils_fp.h:43 */
    if (eps < EPS) { eps = EPS; ier = 1; }
    int ns = (int)ceil(-log10(eps / 10.0));
    if (upsampfac != 2.0)
        ns = (int)ceil(-log(eps) / (PId * sqrt(1 - 1 / upsampfac)));
    if (ns < 2) ns = 2;
    if (ns > ORC_MAX_NSPREAD) { ns = ORC_MAX_NSPREAD; ier = 1; }
    *ns_out = ns;
    *halfwidth_out = (double)ns / 2;
    *c_out = 4.0 / (double)(ns * ns);
    double betaoverns = 2.30;
    if (ns == 2) betaoverns = 2.20;
    if (ns == 3) betaoverns = 2.26;
    if (ns == 4) betaoverns = 2.38;
    if (upsampfac != 2.0) {
        double gamma = 0.97;
        betaoverns = gamma * PId * (1 - 1 / (2 * upsampfac));
    }
    *beta_out = betaoverns * (double)ns;
    return ier;
}

/* SET_NF_TYPE12: contrib/common.cpp:24-37.  b is the output block size for
   gpu_method 4 (nf becomes a multiple of it), 1 for the other methods.     */
int orc_set_nf_type12(int ms, double upsampfac, int ns, int b)
{
    int nf = (int)(upsampfac * ms);
    if (nf < 2 * ns) nf = 2 * ns;
    if (nf < 2147483647) nf = orc_next235beven(nf, b);
    return nf;
}

/* ------------------------------------------------------------------------ */
/* RESCALE with pirange=1: contrib/spreadinterp.h:36-38.  The reference
   evaluates x*M_1_2PI + offset in double (M_1_2PI is a double literal), which
   nvcc's default --fmad=true contracts to one fused multiply-add; the product
   with N is a second double multiply; the result is rounded to FLT.         */
static inline float rescale_f(float x, int n)
{
    const float PIf = (float)M_PI;
    double off = x < -PIf ? 1.5 : (x >= PIf ? -0.5 : 0.5);
    return (float)(fma((double)x, ORC_M_1_2PI, off) * (double)n);
}
static inline double rescale_d(double x, int n)
{
    double off = x < -M_PI ? 1.5 : (x >= M_PI ? -0.5 : 0.5);
    return fma(x, ORC_M_1_2PI, off) * (double)n;
}

void orc_rescale_f(int M, const float *x, int n, float *xr)
{
    for (int i = 0; i < M; ++i) xr[i] = rescale_f(x[i], n);
}
void orc_rescale_d(int M, const double *x, int n, double *xr)
{
    for (int i = 0; i < M; ++i) xr[i] = rescale_d(x[i], n);
}

/* The same RESCALE of single-precision coordinates, but returning the double
   value before its final rounding to float.  The float-rounded value fixes
   the bin (above); this unrounded one is the position that the library's
   single-precision spreader takes its tap offsets from.  The reference takes
   them from the rounded value, which costs up to half an ulp of nf in grid
   units (an error floor near 2e-5 at nf=1024).                             */
void orc_rescale_f_exact(int M, const float *x, int n, double *xr)
{
    const float PIf = (float)M_PI;
    for (int i = 0; i < M; ++i) {
        double off = x[i] < -PIf ? 1.5 : (x[i] >= PIf ? -0.5 : 0.5);
        xr[i] = fma((double)x[i], ORC_M_1_2PI, off) * (double)n;
    }
}

/* Bin index of one coordinate: spreadinterp2d.cu:113-118,
   spreadinterp3d.cu:27-37 -- floor in FLT, then ">= nbin -> -1", "< 0 -> 0". */
static inline int binof_f(float xr, int bs, int nb)
{
    int b = (int)floorf(xr / (float)bs);
    b = b >= nb ? b - 1 : b;
    return b < 0 ? 0 : b;
}
static inline int binof_d(double xr, int bs, int nb)
{
    int b = (int)floor(xr / (double)bs);
    b = b >= nb ? b - 1 : b;
    return b < 0 ? 0 : b;
}

/* Number of bins per axis: spread2d_wrapper.cu:182-183 (ceil in FLT).     */
int orc_numbins(int single, int nf, int bs)
{
    if (single) return (int)ceilf((float)nf / (float)bs);
    return (int)ceil((double)nf / (double)bs);
}

/* Counting bin sort with prefix scans, the setpts stage of the SM / GM-sort
   methods: CalcBinSize_noghost_{2,3}d (spreadinterp2d.cu:102-126,
   spreadinterp3d.cu:15-42), exclusive_scan (spread2d_wrapper.cu:270-272),
   CalcInvertofGlobalSortIdx (spreadinterp2d.cu:128-149), CalcSubProb_*
   (precision_independent.cu:65-73,107-115), inclusive_scan with [0]=0
   (spread2d_wrapper.cu:552-555), MapBintoSubProb (precision_independent.cu:
   75-85,117-127).  Bin index = bx + nbx*(by + nby*bz).  The in-bin order here
   is ascending point index (the reference's is atomic arrival order, so only
   the per-bin sets are comparable).  Outputs: binof[M], binsize[nbins],
   binstart[nbins], idx[M], numsub[nbins], substart[nbins+1]; returns the
   total number of subproblems.                                            */
int orc_bin_sort(int single, int dim, int M, const void *x, const void *y,
                 const void *z, const int *nf, const int *bs, int maxsub,
                 int *binof, int *binsize, int *binstart, int *idx,
                 int *numsub, int *substart)
{
    int nb[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d) nb[d] = orc_numbins(single, nf[d], bs[d]);
    long nbins = (long)nb[0] * nb[1] * nb[2];
    memset(binsize, 0, sizeof(int) * nbins);
    const void *co[3] = {x, y, z};
    for (int i = 0; i < M; ++i) {
        int b[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) {
            if (single) {
                float xr = rescale_f(((const float *)co[d])[i], nf[d]);
                b[d] = binof_f(xr, bs[d], nb[d]);
            } else {
                double xr = rescale_d(((const double *)co[d])[i], nf[d]);
                b[d] = binof_d(xr, bs[d], nb[d]);
            }
        }
        int bi = b[0] + nb[0] * (b[1] + nb[1] * b[2]);
        binof[i] = bi;
        binsize[bi]++;
    }
    int acc = 0;
    for (long b = 0; b < nbins; ++b) { binstart[b] = acc; acc += binsize[b]; }
    int *fill = (int *)calloc(nbins, sizeof(int));
    for (int i = 0; i < M; ++i) {
        int bi = binof[i];
        idx[binstart[bi] + fill[bi]++] = i;
    }
    free(fill);
    int tot = 0;
    substart[0] = 0;
    for (long b = 0; b < nbins; ++b) {
        numsub[b] = (binsize[b] + maxsub - 1) / maxsub;
        tot += numsub[b];
        substart[b + 1] = tot;
    }
    return tot;
}

/* ------------------------------------------------------------------------ */
/* ES kernel: contrib/spreadinterp.cpp:69-81 and src/cuspreadinterp.h:6-16.
   phi(x) = exp(beta*sqrt(1 - c*x^2)) for |x| < ns/2, else 0.               */
double orc_es_kernel(double x, int ns, double beta, double c)
{
    if (fabs(x) >= ns / 2.0) return 0.0;
    return exp(beta * sqrt(1.0 - c * x * x));
}

/* Kernel values at the ns grid points that one coordinate touches:
   start = ceil(xr - ns/2), ker[i] = phi(start + i - xr), i < ns
   (spreadinterp2d.cu:33-42, cuspreadinterp.h:33-40).                       */
static inline int kervec(double xr, int ns, double beta, double c, double *ker)
{
    int start = (int)ceil(xr - ns / 2.0);
    double x1 = (double)start - xr;
    for (int i = 0; i < ns; ++i) ker[i] = orc_es_kernel(x1 + i, ns, beta, c);
    return start;
}

static inline int wrapi(int i, int n)
{
    i %= n;
    return i < 0 ? i + n : i;
}

/* Spread (type-1 step 1): fw[l] += c_j * prod_d phi(l_d - xr_jd), periodic
   wrap; the GM / SM kernels of spreadinterp2d.cu:16-58,152-231 and
   spreadinterp3d.cu:124-175,285-389 all compute this sum.  xr are rescaled
   grid coordinates (FP64 copies of the plan-precision values); c and fw are
   interleaved complex128; fw is x-fastest with sizes nf[0..2].              */
void orc_spread(int dim, int M, const double *xr, const double *yr,
                const double *zr, const double *c, const int *nf, int ns,
                double beta, double es_c, double *fw)
{
    const int n1 = nf[0], n2 = dim > 1 ? nf[1] : 1, n3 = dim > 2 ? nf[2] : 1;
    memset(fw, 0, sizeof(double) * 2 * (size_t)n1 * n2 * n3);
    double k1[ORC_MAX_NSPREAD], k2[ORC_MAX_NSPREAD], k3[ORC_MAX_NSPREAD];
    for (int j = 0; j < M; ++j) {
        int s1 = kervec(xr[j], ns, beta, es_c, k1), s2 = 0, s3 = 0;
        int w2 = 1, w3 = 1;
        k2[0] = k3[0] = 1.0;
        if (dim > 1) { s2 = kervec(yr[j], ns, beta, es_c, k2); w2 = ns; }
        if (dim > 2) { s3 = kervec(zr[j], ns, beta, es_c, k3); w3 = ns; }
        double cr = c[2 * j], ci = c[2 * j + 1];
        for (int c3 = 0; c3 < w3; ++c3) {
            size_t o3 = (size_t)wrapi(s3 + c3, n3) * n2;
            for (int c2 = 0; c2 < w2; ++c2) {
                size_t o2 = (o3 + wrapi(s2 + c2, n2)) * n1;
                double k23 = k2[c2] * k3[c3];
                for (int c1 = 0; c1 < ns; ++c1) {
                    size_t o = o2 + wrapi(s1 + c1, n1);
                    double kv = k1[c1] * k23;
                    fw[2 * o] += cr * kv;
                    fw[2 * o + 1] += ci * kv;
                }
            }
        }
    }
}

/* The same spread on all OpenMP threads, for the CPU baseline (the bench
   times it; the checks use orc_spread): the last axis of the grid is cut
   into one slab of planes per thread, points are bucketed by their first
   tap plane on that axis (counting sort), and each thread visits the
   buckets whose taps can reach its slab and adds only the taps that land in
   it -- every grid cell has one writer, so no atomics.  Points near a slab
   boundary have their kernel values computed by two threads.  The sum per
   cell is orc_spread's, in another order.                                  */
void orc_spread_omp(int dim, int M, const double *xr, const double *yr,
                    const double *zr, const double *c, const int *nf, int ns,
                    double beta, double es_c, double *fw)
{
    const int n1 = nf[0], n2 = dim > 1 ? nf[1] : 1, n3 = dim > 2 ? nf[2] : 1;
    const double *ax = dim == 3 ? zr : (dim == 2 ? yr : xr);  /* slab axis */
    const int n = dim == 3 ? n3 : (dim == 2 ? n2 : n1);
    memset(fw, 0, sizeof(double) * 2 * (size_t)n1 * n2 * n3);
    int *start = (int *)calloc((size_t)n + 1, sizeof(int));
    int *fill = (int *)malloc(sizeof(int) * (size_t)n);
    int *order = (int *)malloc(sizeof(int) * (size_t)(M > 0 ? M : 1));
    int *pl = (int *)malloc(sizeof(int) * (size_t)(M > 0 ? M : 1));
    if (!start || !fill || !order || !pl) {
        free(start);
        free(fill);
        free(order);
        free(pl);
        orc_spread(dim, M, xr, yr, zr, c, nf, ns, beta, es_c, fw);
        return;
    }
    double kdummy[ORC_MAX_NSPREAD];
    for (int j = 0; j < M; ++j) {
        pl[j] = wrapi(kervec(ax[j], ns, beta, es_c, kdummy), n);
        start[pl[j] + 1]++;
    }
    for (int p = 0; p < n; ++p) start[p + 1] += start[p];
    memcpy(fill, start, sizeof(int) * (size_t)n);
    for (int j = 0; j < M; ++j) order[fill[pl[j]]++] = j;
    free(fill);
#pragma omp parallel
    {
        const int nt = omp_get_num_threads(), t = omp_get_thread_num();
        const int a = (int)((long long)n * t / nt), b = (int)((long long)n * (t + 1) / nt);
        double k1[ORC_MAX_NSPREAD], k2[ORC_MAX_NSPREAD], k3[ORC_MAX_NSPREAD];
        /* buckets whose taps reach [a, b): first planes a-ns+1 .. b-1, each once */
        int span = b - a + ns - 1;
        if (span > n) span = n;
        for (int q = 0; q < (b > a ? span : 0); ++q) {
            const int p = wrapi(a - ns + 1 + q, n);
            for (int e = start[p]; e < start[p + 1]; ++e) {
                const int j = order[e];
                int s1 = kervec(xr[j], ns, beta, es_c, k1), s2 = 0, s3 = 0;
                int w2 = 1, w3 = 1;
                k2[0] = k3[0] = 1.0;
                if (dim > 1) { s2 = kervec(yr[j], ns, beta, es_c, k2); w2 = ns; }
                if (dim > 2) { s3 = kervec(zr[j], ns, beta, es_c, k3); w3 = ns; }
                const double cr = c[2 * j], ci = c[2 * j + 1];
                for (int c3 = 0; c3 < w3; ++c3) {
                    const int l3 = wrapi(s3 + c3, n3);
                    if (dim == 3 && (l3 < a || l3 >= b)) continue;
                    const size_t o3 = (size_t)l3 * n2;
                    for (int c2 = 0; c2 < w2; ++c2) {
                        const int l2 = wrapi(s2 + c2, n2);
                        if (dim == 2 && (l2 < a || l2 >= b)) continue;
                        const size_t o2 = (o3 + l2) * n1;
                        const double k23 = k2[c2] * k3[c3];
                        for (int c1 = 0; c1 < ns; ++c1) {
                            const int l1 = wrapi(s1 + c1, n1);
                            if (dim == 1 && (l1 < a || l1 >= b)) continue;
                            const size_t o = o2 + l1;
                            const double kv = k1[c1] * k23;
                            fw[2 * o] += cr * kv;
                            fw[2 * o + 1] += ci * kv;
                        }
                    }
                }
            }
        }
    }
    free(start);
    free(order);
    free(pl);
}

/* Interp (type-2 step 3): c_j = sum_l fw[l] prod_d phi(l_d - xr_jd);
   Interp_{2,3}d_NUptsdriven (spreadinterp2d.cu:516-554,
   spreadinterp3d.cu:654-705).  Points are independent, so OpenMP splits
   them across threads.                                                     */
void orc_interp(int dim, int M, const double *xr, const double *yr,
                const double *zr, double *c, const int *nf, int ns,
                double beta, double es_c, const double *fw)
{
    const int n1 = nf[0], n2 = dim > 1 ? nf[1] : 1, n3 = dim > 2 ? nf[2] : 1;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < M; ++j) {
        double k1[ORC_MAX_NSPREAD], k2[ORC_MAX_NSPREAD], k3[ORC_MAX_NSPREAD];
        int s1 = kervec(xr[j], ns, beta, es_c, k1), s2 = 0, s3 = 0;
        int w2 = 1, w3 = 1;
        k2[0] = k3[0] = 1.0;
        if (dim > 1) { s2 = kervec(yr[j], ns, beta, es_c, k2); w2 = ns; }
        if (dim > 2) { s3 = kervec(zr[j], ns, beta, es_c, k3); w3 = ns; }
        double ar = 0, ai = 0;
        for (int c3 = 0; c3 < w3; ++c3) {
            size_t o3 = (size_t)wrapi(s3 + c3, n3) * n2;
            for (int c2 = 0; c2 < w2; ++c2) {
                size_t o2 = (o3 + wrapi(s2 + c2, n2)) * n1;
                double k23 = k2[c2] * k3[c3];
                for (int c1 = 0; c1 < ns; ++c1) {
                    size_t o = o2 + wrapi(s1 + c1, n1);
                    double kv = k1[c1] * k23;
                    ar += fw[2 * o] * kv;
                    ai += fw[2 * o + 1] * kv;
                }
            }
        }
        c[2 * j] = ar;
        c[2 * j + 1] = ai;
    }
}

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre nodes and weights on [-1,1] by Newton's method on the
   three-term recurrence of P_n (the textbook method; the reference uses the
   Glaser-Liu-Rokhlin algorithm, contrib/legendre_rule_fast.c:15-93, for the
   same nodes).  Nodes ascending.                                            */
void orc_gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5));
        double pp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0; }
            pp = n * (z * p1 - p0) / (z * z - 1.0);
            double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        /* recompute derivative at the converged root */
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
            double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        pp = n * (z * p1 - p0) / (z * z - 1.0);
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

/* Fourier-series coefficients of the ES kernel on the fine grid:
   onedim_fseries_kernel_precomp + _compute (contrib/common.cpp:39-124) and
   FseriesKernelCompute (src/common.cu:16-45).  With q = (int)(2 + 1.5 ns)
   Gauss-Legendre nodes on the half-support [0, ns/2] (the positive half of a
   2q-point rule), fwkerhalf[k] = sum_n (ns/2) w_n phi(z_n) 2 cos(2 pi k
   (nf/2 - z_n)/nf) for k = 0..nf/2; the nf/2 phase shift is the grid offset
   that RESCALE puts on x = 0.                                               */
void orc_fseries(int nf, int ns, double beta, double es_c, double *fwkerhalf)
{
    double J2 = ns / 2.0;
    int q = (int)(2 + 3.0 * J2);
    double z[200], w[200];
    orc_gauss_legendre(2 * q, z, w);
    double f[100], zz[100];
    for (int n = 0; n < q; ++n) {
        zz[n] = z[q + n] * J2; /* positive half of the 2q-node rule */
        f[n] = J2 * w[q + n] * orc_es_kernel(zz[n], ns, beta, es_c);
    }
    for (int k = 0; k <= nf / 2; ++k) {
        double s = 0;
        for (int n = 0; n < q; ++n)
            s += f[n] * 2.0 * cos(2.0 * M_PI * (double)k * (nf / 2 - zz[n]) / nf);
        fwkerhalf[k] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* Mode index -> fine-grid index along one axis: deconvolve_wrapper.cu:32-35.
   Modes are k = 0..ms-1 in CMCL order (frequency k - ms/2).                */
static inline int mode2grid(int k, int ms, int nf)
{
    int f = k - ms / 2;
    return f >= 0 ? f : nf + f;
}

/* Deconvolve (type-1 step 3), Deconvolve_{2,3}d: deconvolve_wrapper.cu:26-66.
   fk[k] = fw[w(k)] / prod_d fwkerhalf_d[|k_d - ms_d/2|].                   */
void orc_deconvolve(int dim, const int *ms, const int *nf, const double *fw,
                    const double *ker1, const double *ker2, const double *ker3,
                    double *fk)
{
    int m1 = ms[0], m2 = dim > 1 ? ms[1] : 1, m3 = dim > 2 ? ms[2] : 1;
    int n1 = nf[0], n2 = dim > 1 ? nf[1] : 1;
    for (int k3 = 0; k3 < m3; ++k3)
        for (int k2 = 0; k2 < m2; ++k2)
            for (int k1 = 0; k1 < m1; ++k1) {
                size_t out = k1 + (size_t)m1 * (k2 + (size_t)m2 * k3);
                int w1 = mode2grid(k1, m1, n1);
                int w2 = dim > 1 ? mode2grid(k2, m2, n2) : 0;
                int w3 = dim > 2 ? mode2grid(k3, m3, nf[2]) : 0;
                size_t in = w1 + (size_t)n1 * (w2 + (size_t)n2 * w3);
                double kv = ker1[abs(k1 - m1 / 2)];
                if (dim > 1) kv *= ker2[abs(k2 - m2 / 2)];
                if (dim > 2) kv *= ker3[abs(k3 - m3 / 2)];
                fk[2 * out] = fw[2 * in] / kv;
                fk[2 * out + 1] = fw[2 * in + 1] / kv;
            }
}

/* Amplify (type-2 step 1), Amplify_{2,3}d after the memset of fw:
   deconvolve_wrapper.cu:81-121,181-182.                                    */
void orc_amplify(int dim, const int *ms, const int *nf, const double *fk,
                 const double *ker1, const double *ker2, const double *ker3,
                 double *fw)
{
    int m1 = ms[0], m2 = dim > 1 ? ms[1] : 1, m3 = dim > 2 ? ms[2] : 1;
    int n1 = nf[0], n2 = dim > 1 ? nf[1] : 1, n3 = dim > 2 ? nf[2] : 1;
    memset(fw, 0, sizeof(double) * 2 * (size_t)n1 * n2 * n3);
    for (int k3 = 0; k3 < m3; ++k3)
        for (int k2 = 0; k2 < m2; ++k2)
            for (int k1 = 0; k1 < m1; ++k1) {
                size_t in = k1 + (size_t)m1 * (k2 + (size_t)m2 * k3);
                int w1 = mode2grid(k1, m1, n1);
                int w2 = dim > 1 ? mode2grid(k2, m2, n2) : 0;
                int w3 = dim > 2 ? mode2grid(k3, m3, n3) : 0;
                size_t out = w1 + (size_t)n1 * (w2 + (size_t)n2 * w3);
                double kv = ker1[abs(k1 - m1 / 2)];
                if (dim > 1) kv *= ker2[abs(k2 - m2 / 2)];
                if (dim > 2) kv *= ker3[abs(k3 - m3 / 2)];
                fw[2 * out] = fk[2 * in] / kv;
                fw[2 * out + 1] = fk[2 * in + 1] / kv;
            }
}

/* ------------------------------------------------------------------------ */
/* Direct sums, the FP64 accuracy oracle.  dirft2d1 / dirft2d2
   (contrib/dirft2d.cpp:7-74) extended to 1D and 3D from the same definition
   (tests/utils.py:56-71 of the python package is the same sum):
     type 1: f[k] = sum_j c_j exp(i*sgn*(k1 x_j + k2 y_j + k3 z_j))
     type 2: c_j = sum_k f[k] exp(i*sgn*(k . x_j))
   with k_d in [-ms_d/2, (ms_d-1)/2], x fastest.  Only the requested mode
   indices (type 1) or point indices (type 2) are evaluated, so the oracle
   checks a sample of 1000 entries in O(1000*M) or O(1000*N) work.
   Phases are evaluated directly with sincos per term in FP64.              */
void orc_dirft_type1(int dim, int M, const double *x, const double *y,
                     const double *z, const double *c, int iflag,
                     const int *ms, int nk, const long *kidx, double *out)
{
    int m1 = ms[0], m2 = dim > 1 ? ms[1] : 1;
    double sgn = iflag >= 0 ? 1.0 : -1.0;
#pragma omp parallel for schedule(dynamic, 4)
    for (int t = 0; t < nk; ++t) {
        long k = kidx[t];
        double f1 = (double)(k % m1) - m1 / 2;
        double f2 = dim > 1 ? (double)((k / m1) % m2) - m2 / 2 : 0;
        double f3 = dim > 2 ? (double)(k / ((long)m1 * m2)) - ms[2] / 2 : 0;
        double ar = 0, ai = 0;
        for (int j = 0; j < M; ++j) {
            double ph = f1 * x[j];
            if (dim > 1) ph += f2 * y[j];
            if (dim > 2) ph += f3 * z[j];
            ph *= sgn;
            double s = sin(ph), co = cos(ph);
            ar += c[2 * j] * co - c[2 * j + 1] * s;
            ai += c[2 * j] * s + c[2 * j + 1] * co;
        }
        out[2 * t] = ar;
        out[2 * t + 1] = ai;
    }
}

void orc_dirft_type2(int dim, const double *x, const double *y,
                     const double *z, int np, const long *pidx,
                     const double *f, int iflag, const int *ms, double *out)
{
    int m1 = ms[0], m2 = dim > 1 ? ms[1] : 1, m3 = dim > 2 ? ms[2] : 1;
    double sgn = iflag >= 0 ? 1.0 : -1.0;
#pragma omp parallel for schedule(dynamic, 4)
    for (int t = 0; t < np; ++t) {
        long j = pidx[t];
        double ar = 0, ai = 0;
        for (int k3 = 0; k3 < m3; ++k3)
            for (int k2 = 0; k2 < m2; ++k2)
                for (int k1 = 0; k1 < m1; ++k1) {
                    double ph = (double)(k1 - m1 / 2) * x[j];
                    if (dim > 1) ph += (double)(k2 - m2 / 2) * y[j];
                    if (dim > 2) ph += (double)(k3 - m3 / 2) * z[j];
                    ph *= sgn;
                    size_t k = k1 + (size_t)m1 * (k2 + (size_t)m2 * k3);
                    double s = sin(ph), co = cos(ph);
                    ar += f[2 * k] * co - f[2 * k + 1] * s;
                    ai += f[2 * k] * s + f[2 * k + 1] * co;
                }
        out[2 * t] = ar;
        out[2 * t + 1] = ai;
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
