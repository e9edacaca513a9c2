/* Options struct of the cufinufft C ABI.
 *
 * Field order and types are the ABI: one double followed by 13 ints (64 bytes
 * with tail padding), exactly as include/cufinufft_opts.h:4-26 of cuFINUFFT
 * v1.3 and as its ctypes mirror python/cufinufft/_cufinufft.py:77-93.
 * New knobs of this library are expressed through values of existing fields
 * (gpu_method = 0, gpu_sort = 2) or through new entry points, never through
 * new fields.
 */
#ifndef CUFINUFFT_OPTS_H
#define CUFINUFFT_OPTS_H

typedef struct cufinufft_opts {
    double upsampfac;        /* sigma; 2.0 is the standard choice */
    int gpu_method;          /* 1 GM / GM-sort, 2 SM subproblems, 4 owner-computes
                                blocks (type 1 only); 0 picks 2 when its LDS tile
                                fits, else 1 (the autotuner lives in Python:
                                cufinufft_amd/autotune.py) */
    int gpu_sort;            /* 0 natural order (method 1 only), 1 bin-sorted order,
                                2 bin-sorted and by fine-grid cell inside each bin */
    int gpu_binsizex;        /* bin (subproblem tile) size, -1 = default */
    int gpu_binsizey;
    int gpu_binsizez;
    int gpu_obinsizex;       /* method 4: block size (must hold whole bins) */
    int gpu_obinsizey;
    int gpu_obinsizez;
    int gpu_maxsubprobsize;  /* max points per subproblem */
    int gpu_nstreams;        /* kept for ABI; unused (as in the reference) */
    int gpu_kerevalmeth;     /* 0 exp(sqrt()); 1 piecewise polynomial fitted at
                                plan time (upsampfac 2.0, as in the reference) */
    int gpu_spreadinterponly;/* 1: execute only spreads (type 1) or interpolates (type 2) */
    int gpu_device_id;       /* HIP device the plan lives on */
} cufinufft_opts;

#endif
