/* C ABI of the MI355X NUFFT library (libcufinufft.so).
 *
 * A drop-in for the type-1 / type-2 entry points of cuFINUFFT v1.3: same
 * symbol names, argument order, meaning and integer return codes.  Every
 * pointer to coordinates, strengths or modes is a caller-owned DEVICE pointer
 * (HIP memory on the plan's device); complex arrays are interleaved {re, im}
 * (std::complex / hipFloatComplex / hipDoubleComplex layout).
 *
 * Return codes (contrib/utils.h:28-36 of the reference, plus this library's
 * device-side codes, which the reference turned into exit() instead):
 *   0 ok, 1 generic error or WARN_EPS_TOO_SMALL (makeplan only), 2 bad method,
 *   3 ERR_SPREAD_BOX_SMALL, 4 ERR_SPREAD_PTS_OUT_RANGE, 5 ERR_SPREAD_ALLOC,
 *   6 ERR_SPREAD_DIR, 7 ERR_UPSAMPFAC_TOO_SMALL, 8 HORNER_WRONG_BETA,
 *   9 ERR_NDATA_NOTVALID, 10 HIP runtime error, 11 FFT library error,
 *   12 tile exceeds the 160 KiB LDS of one workgroup,
 *   13 input beyond what is allocated (cufinufftf_setpts_noalloc only).
 */
#ifndef CUFINUFFT_H
#define CUFINUFFT_H

#include "cufinufft_opts.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque plans (reference: include/cufinufft_eitherprec.h:299-302). */
typedef struct cufinufft_plan_s *cufinufft_plan;   /* double precision */
typedef struct cufinufftf_plan_s *cufinufftf_plan; /* single precision */

/* Defaults for a (type, dim): upsampfac 2, method 2 (type 1) or 1 (type 2),
 * sort on, maxsubprobsize 1024, bins auto, exp(sqrt()) kernel, device 0.
 * Replaces src/cufinufft.cu:639-730 (cufinufft_default_opts; the single
 * precision name maps to the same function in the reference macro layer,
 * include/cufinufft_eitherprec.h:103). */
int cufinufft_default_opts(int type, int dim, cufinufft_opts *opts);
int cufinufftf_default_opts(int type, int dim, cufinufft_opts *opts);

/* Plan a transform: include/cufinufft_eitherprec.h:314-316,
 * src/cufinufft.cu:78-273.  nmodes has 3 entries (unused ones ignored),
 * maxbatchsize 0 picks min(ntransf, 8), opts may be NULL. */
int cufinufft_makeplan(int type, int dim, int *nmodes, int iflag, int ntransf,
                       double tol, int maxbatchsize, cufinufft_plan *plan,
                       cufinufft_opts *opts);
int cufinufftf_makeplan(int type, int dim, int *nmodes, int iflag, int ntransf,
                        float tol, int maxbatchsize, cufinufftf_plan *plan,
                        cufinufft_opts *opts);

/* Set nonuniform points (device arrays, kept by pointer, not copied, and
 * re-read by execute) and bin-sort them: include/cufinufft_eitherprec.h:
 * 317-318, src/cufinufft.cu:275-493.  N, s, t, u are reserved for type 3
 * and ignored. */
int cufinufft_setpts(int M, double *x, double *y, double *z, int N, double *s,
                     double *t, double *u, cufinufft_plan plan);
int cufinufftf_setpts(int M, float *x, float *y, float *z, int N, float *s,
                      float *t, float *u, cufinufftf_plan plan);

/* Execute ntransf transforms: c has M*ntransf entries, fk N1*N2(*N3)*ntransf,
 * transform index slowest, x fastest: include/cufinufft_eitherprec.h:319,
 * src/cufinufft.cu:495-569.  The pointers are void* here so that C callers
 * can pass any interleaved complex type. */
int cufinufft_execute(void *c, void *fk, cufinufft_plan plan);
int cufinufftf_execute(void *c, void *fk, cufinufftf_plan plan);

/* Free all plan memory and the plan itself: include/cufinufft_eitherprec.h:
 * 320, src/cufinufft.cu:571-637.  NULL returns 1, as in the reference. */
int cufinufft_destroy(cufinufft_plan plan);
int cufinufftf_destroy(cufinufftf_plan plan);

/* ---- Extensions (no counterpart in the reference ABI) ------------------ */

/* makeplan with planning flags (0 gives makeplan exactly).
 * CUFINUFFT_ACCURACY_FIRST: the reference's kernel width rule
 * ns = ceil(-log10(tol/10)) (contrib/spreadinterp.cpp:40-50) leaves the
 * transform's relative l2 error at 0.6-1.1, 1-2 and 1.5-3.2 times tol in 1,
 * 2 and 3 dimensions (FP64, against direct sums, tol 1e-5 and 1e-9); this
 * flag plans for tol * 2 / (dim + 2) instead (tol/2.5 in 3D: one more tap
 * per axis at those tolerances), which measures 0.1-0.4 times tol.  For
 * callers that need the requested tolerance met, at the cost of the wider
 * kernel. */
#define CUFINUFFT_ACCURACY_FIRST 1
/* CUFINUFFT_TILE_SLOTS (type 1, gpu_method 2): a bin whose points form one
 * subproblem stores its padded spread tile whole (plain stores into plan
 * memory, at most 8 fine grids of it) and a gather pass adds the stored
 * tiles covering each grid cell, instead of the tile being flushed with two
 * global float atomics per cell.  Faster where most bins hold one
 * subproblem (uniform 3D points: 10.5 -> 7.8 ms for 5e7 points on 512^3),
 * slower where few do (the gather passes over the whole grid).  Applies
 * when every axis has 3 bins or more, the halo ceil(ns/2) is no wider than
 * a bin and the tile no wider than the grid; ignored otherwise. */
#define CUFINUFFT_TILE_SLOTS 2
/* CUFINUFFT_FULL_FFT (2D, 3D): transform the whole fine grid with one 2D / 3D
 * hipFFT plan, as the reference does with cuFFT.  Without it such a plan runs
 * the FFT as contiguous batched 1D passes over only the grid lines that hold
 * modes (type 1: the mode columns after the x pass, in 3D the mode rows
 * after the y pass; type 2 in reverse), with its own transposes between the
 * passes and the deconvolve / amplify folded into the first or last of them;
 * the plan then holds a further ms0 * nf1 * nf2 complex values per batch
 * member.  The modes agree with the full transform's to rounding. */
#define CUFINUFFT_FULL_FFT 4
int cufinufft_makeplan_ext(int type, int dim, int *nmodes, int iflag, int ntransf,
                           double tol, int maxbatchsize, cufinufft_plan *plan,
                           cufinufft_opts *opts, int flags);
int cufinufftf_makeplan_ext(int type, int dim, int *nmodes, int iflag, int ntransf,
                            float tol, int maxbatchsize, cufinufftf_plan *plan,
                            cufinufft_opts *opts, int flags);

/* Stream for every launch and FFT of this plan (a hipStream_t; NULL is the
 * null stream, the default).  For callers that order the library with their
 * own streams, e.g. PyTorch's current stream or a HIP graph capture. */
int cufinufft_setstream(cufinufft_plan plan, void *stream);
int cufinufftf_setstream(cufinufftf_plan plan, void *stream);

/* Plan groups: many independent plans of one type, dimension, precision and
 * tolerance (any modes, points and bins; method 2; maxbatchsize = ntransf)
 * executed together, with the spread (type 1) or interpolation (type 2) of
 * every member in ONE persistent kernel launch over a global subproblem
 * queue.  make snapshots the members' setpts arrays and the c / fk device
 * pointers (plans[i] with c[i], fk[i]); re-make after any member's setpts.
 * execute runs the group on its own stream, ordered after and before the
 * caller's stream (setstream; null stream by default); use_graph = 1
 * captures the launch sequence into a hipGraph on the first call and replays
 * it afterwards.  For callers batching many small transforms (no counterpart
 * in the reference). */
typedef struct cufinufft_group_s *cufinufft_group;
typedef struct cufinufftf_group_s *cufinufftf_group;
int cufinufft_group_make(int n, cufinufft_plan *plans, void **c, void **fk,
                         cufinufft_group *group);
int cufinufftf_group_make(int n, cufinufftf_plan *plans, void **c, void **fk,
                          cufinufftf_group *group);
int cufinufft_group_execute(cufinufft_group group, int use_graph);
int cufinufftf_group_execute(cufinufftf_group group, int use_graph);
int cufinufft_group_setstream(cufinufft_group group, void *stream);
int cufinufftf_group_setstream(cufinufftf_group group, void *stream);
int cufinufft_group_destroy(cufinufft_group group);
int cufinufftf_group_destroy(cufinufftf_group group);

/* Type-2 execute that also returns each output's derivatives with respect to
 * its point's coordinates (the differentiable form): c as in execute, and dc
 * with dim planes of M*ntransf complex values, dc[d*M*ntransf + t*M + j] =
 * d c[t*M + j] / d x_d[j] (x_0 = x, x_1 = y, x_2 = z).  Each derivative is
 * the type-2 transform of the modes weighted by i * iflag * k_d (the weight
 * is applied in the amplify pass), so it meets the plan's tolerance like the
 * values; the cost is dim + 1 type-2 executes.  For autograd callers
 * (gradients of losses with respect to point coordinates); returns 1 for
 * type-1 and spread/interp-only plans. */
int cufinufft_execute_grad(void *c, void *dc, void *fk, cufinufft_plan plan);
int cufinufftf_execute_grad(void *c, void *dc, void *fk, cufinufftf_plan plan);

/* Plan facts, for tests and tuners: out[0..15] = ns, nf1, nf2, nf3,
 * numbins x, y, z, method, binsize x, y, z, maxsubprobsize, M,
 * total subproblems of the last setpts (synchronises), maxbatchsize,
 * spreadinterponly. */
int cufinufft_get_info(cufinufft_plan plan, int *out);
int cufinufftf_get_info(cufinufftf_plan plan, int *out);

/* Copy the bin-sort arrays of the last setpts to host memory (synchronous;
 * any pointer may be NULL): binsize[nbins], binstart[nbins], idxnupts[M],
 * subprobstart[nbins+1], subprob_to_bin[total subproblems].  The test
 * surface of the reference's internal arrays (spread*_wrapper.cu). */
int cufinufft_get_sort(cufinufft_plan plan, int *binsize, int *binstart,
                       int *idxnupts, int *subprobstart, int *subprob_to_bin);
int cufinufftf_get_sort(cufinufftf_plan plan, int *binsize, int *binstart,
                        int *idxnupts, int *subprobstart, int *subprob_to_bin);

/* Stage timer for profiling and tuning: while on, every stage of setpts and
 * execute is bracketed by HIP events on the plan's stream (no host
 * synchronisation).  get_timing synchronises the stream, returns per-stage
 * total milliseconds ms[8] and launch counts count[8] since the last read,
 * and clears them.  Stages: 0 bin sort, 1 fine-grid zero fill, 2 spread or
 * interp, 3 FFT, 4 deconvolve or amplify; 5-7 reserved. */
int cufinufft_set_timing(cufinufft_plan plan, int on);
int cufinufftf_set_timing(cufinufftf_plan plan, int on);
int cufinufft_get_timing(cufinufft_plan plan, double *ms, int *count);
int cufinufftf_get_timing(cufinufftf_plan plan, double *ms, int *count);

/* Fourier-series factors of the plan (host copies of fwkerhalf_d, length
 * nf_d/2+1, FP64 as computed at plan time; NULL skips an axis). */
int cufinufft_get_fseries(cufinufft_plan plan, double *k1, double *k2, double *k3);
int cufinufftf_get_fseries(cufinufftf_plan plan, double *k1, double *k2, double *k3);

/* Piecewise-polynomial kernel of a plan made with gpu_kerevalmeth = 1 (no
 * counterpart in the reference, whose Horner tables are compiled in,
 * contrib/ker_horner_allw_loop.c): *deg receives the degree (0 when the plan
 * evaluates exp(sqrt())), *rel_err the fit's sampled maximum error relative
 * to the kernel peak e^beta, and coef (if not NULL, (deg+1)*ns doubles) the
 * FP64 coefficients, coef[k*ns + i] multiplying z^k for tap i, where
 * z = 2(start - x) + ns - 1 and start = ceil(x - ns/2). */
int cufinufft_get_kernel_poly(cufinufft_plan plan, int *deg, double *rel_err, double *coef);
int cufinufftf_get_kernel_poly(cufinufftf_plan plan, int *deg, double *rel_err, double *coef);

/* setpts that allocates nothing (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same bin sort
 * as cufinufftf_setpts, with its transient buffers (sort keys, point records
 * between passes, per-pass digit counts) taken from one fixed 64 MiB device
 * array of the library, held on each GPU it runs on until the process ends.
 * The plan's own point arrays must already hold M points (an earlier
 * cufinufftf_setpts of at least M points on this plan sized them); M beyond
 * that, or a sort needing more than the 64 MiB (about 8 bytes per point plus
 * 2 x (3 coordinates + index) bytes per point when the keys need more than
 * one sort pass: ~2.6 million 3D points), returns 13 and changes nothing.
 * Calls that overlap on one GPU share that array and race: serialise them
 * (one stream, or synchronise between them). */
int cufinufftf_setpts_noalloc(int M, float *x, float *y, float *z,
                              cufinufftf_plan plan);

#ifdef __cplusplus
}
#endif

#endif
