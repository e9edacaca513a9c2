"""Autotuner for the spread / interp method and its geometry.

For a transform shape (type, modes, precision, tol, ntransf) and the caller's
actual points, time every candidate option set -- method (1 GM-sort, 2 SM
subproblems), bin size and maximum subproblem size -- on the device and
return the fastest.  Timing uses the library's HIP-event stage timer around
the spread / interp kernel plus the whole execute, so it measures exactly the
kernels the plan will run.  Results are cached in memory per shape key and,
if ``CUFINUFFT_AMD_TUNE_CACHE`` names a JSON file, across processes.

    from cufinufft_amd import autotune
    best, table = autotune.tune(1, (256, 256, 256), pts, eps=1e-5, dtype=np.float32)
    plan = cufinufft(1, (256, 256, 256), eps=1e-5, dtype=np.float32, **best)

The candidate list follows the hardware: spread tiles accumulate in 64-bit
fixed point (16 bytes per cell) for FP64 plans and in 32-bit (chunked, with
FP32 drain planes) or 64-bit fixed point for FP32 plans, interp tiles hold
one complex per cell, all capped at the 160 KiB LDS of a gfx950 workgroup;
plans whose tile does not fit are skipped (makeplan reports it).  The
winner is also tried with the polynomial kernel and on the sigma = 1.25 grid
(with bins shrunk until its wider kernel's tile fits).
"""
import json
import os
import time

import numpy as np

_CACHE = {}


def _key(nufft_type, modes, dtype, eps, n_trans, M, upsampfac, accuracy_first=False):
    return json.dumps([nufft_type, list(modes), np.dtype(dtype).name, float(eps), int(n_trans),
                       int(M), float(upsampfac)] + (["accuracy_first"] if accuracy_first else []))


def candidates(nufft_type, dim, M=None):
    """Option sets to try (cufinufft_opts fields) for a transform with M points."""
    if dim == 3:
        bins = [(16, 16, 2), (8, 8, 8), (8, 8, 4), (16, 8, 4), (4, 4, 4), (16, 16, 4)]
        if nufft_type == 1:
            # larger tiles: less halo per interior cell in the flush
            bins = [(16, 16, 2), (8, 8, 8), (16, 8, 4), (16, 8, 8), (16, 16, 4), (16, 16, 8),
                    (32, 16, 8)]
    elif dim == 2:
        bins = [(32, 32, 1), (16, 16, 1), (64, 32, 1), (64, 64, 1), (32, 16, 1)]
    else:
        bins = [(1024, 1, 1), (512, 1, 1), (2048, 1, 1)]
    # type 1: subproblems are a load-balancing unit only (the spread
    # accumulates long jobs in chunks), so large ones cut the flushes of dense
    # bins; type 2: subproblems of >= 1024 points run 1024-thread workgroups
    # (512-thread ones from 512 points)
    subs = (1024, 4096, 16384) if nufft_type == 1 else (256, 512, 1024, 2048, 4096)
    out = []
    for b in bins:
        for msub in subs:
            out.append(dict(gpu_method=2, gpu_binsizex=b[0], gpu_binsizey=b[1],
                            gpu_binsizez=b[2], gpu_maxsubprobsize=msub))
            # cell-ordered points inside bins (gpu_sort=2): fewer LDS bank
            # conflicts for the interpolation gathers
            if nufft_type == 2:
                out.append(dict(gpu_method=2, gpu_sort=2, gpu_binsizex=b[0], gpu_binsizey=b[1],
                                gpu_binsizez=b[2], gpu_maxsubprobsize=msub))
    # GM-sort: the bin size only sets the sort order; its global atomics (type
    # 1) and gathers (type 2) lose by an order of magnitude at large 3D point
    # counts, where it is not tried
    if not (dim == 3 and M is not None and M > 20_000_000):
        for b in bins[:2]:
            out.append(dict(gpu_method=1, gpu_binsizex=b[0], gpu_binsizey=b[1],
                            gpu_binsizez=b[2], gpu_maxsubprobsize=1024))
    # owner-computes blocks (type 1): halo-free tiles, so blocks can be large;
    # blocks that do not tile the fine grid are rejected by makeplan
    if nufft_type == 1:
        if dim == 3:
            blocks = [(16, 16, 8), (32, 16, 8), (32, 32, 8)]
        elif dim == 2:
            blocks = [(64, 64, 1), (128, 64, 1), (128, 128, 1), (256, 64, 1)]
        else:
            blocks = [(2048, 1, 1), (4096, 1, 1), (8192, 1, 1)]
        for b in blocks:
            for msub in (1024, 4096):
                out.append(dict(gpu_method=4, gpu_obinsizex=b[0], gpu_obinsizey=b[1],
                                gpu_obinsizez=b[2], gpu_binsizex=b[0], gpu_binsizey=b[1],
                                gpu_binsizez=b[2], gpu_maxsubprobsize=msub))
    return out


def sigma125_candidates(opts, dim):
    """The option set on the sigma = 1.25 grid, with the bins halved along
    their largest side (down to 4 cells) until the wider kernel's tile fits:
    the sequence of sets to try, largest tile first (method 2 only)."""
    base = dict(opts, upsampfac=1.25)
    # the polynomial kernel is planned for sigma = 2 only (makeplan returns
    # 8 otherwise, as the reference does)
    base.pop("gpu_kerevalmeth", None)
    if base.get("gpu_method", 2) != 2:
        return [base]
    b = [base.get("gpu_binsizex", 16 if dim > 1 else 1024),
         base.get("gpu_binsizey", 16 if dim > 1 else 1),
         base.get("gpu_binsizez", 2 if dim == 3 else 1)]
    b = [x if x > 0 else 8 for x in b]
    out = []
    while True:
        out.append(dict(base, gpu_binsizex=b[0], gpu_binsizey=b[1], gpu_binsizez=b[2]))
        k = max(range(dim), key=lambda d: b[d])
        if b[k] <= 4:
            return out
        b[k] //= 2


def _load_disk():
    path = os.environ.get("CUFINUFFT_AMD_TUNE_CACHE")
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                _CACHE.update(json.load(f))
        except (OSError, ValueError):
            pass


def _save_disk():
    path = os.environ.get("CUFINUFFT_AMD_TUNE_CACHE")
    if path:
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(_CACHE, f)
        os.replace(tmp, path)


def tune(nufft_type, modes, pts, eps, dtype=np.float32, n_trans=1, maxbatchsize=1,
         upsampfac=2.0, reps=2, cands=None, use_cache=True, log=None, accuracy_first=False):
    """Time candidate options on the given device points.

    modes: numpy-order shape (nZ, nY, nX) as for the ``cufinufft`` class;
    pts: device arrays in numpy order (as for ``set_pts``).  accuracy_first
    plans every candidate with the accuracy-first flag (the plan meets eps;
    the caller must plan the winner the same way).
    Returns (best_opts, [(opts, exec_ms) ...]).
    """
    import torch

    from cufinufft_amd.cufinufft import cufinufft
    dim = len(modes)
    M = int(pts[0].numel()) if hasattr(pts[0], "numel") else int(pts[0].size)
    key = _key(nufft_type, modes, dtype, eps, n_trans, M, upsampfac, accuracy_first)
    if use_cache:
        if not _CACHE:
            _load_disk()
        if key in _CACHE:
            return dict(_CACHE[key]), []
    dev = pts[0].device
    cd = torch.complex64 if np.dtype(dtype) == np.float32 else torch.complex128
    shape = tuple(modes)
    g = torch.Generator(device=dev).manual_seed(0)
    if nufft_type == 1:
        inp = torch.randn(M * n_trans, dtype=cd, device=dev, generator=g)
        out = torch.empty(((n_trans,) if n_trans > 1 else ()) + shape, dtype=cd, device=dev)
        args = (inp, out)
    else:
        inp = torch.randn(((n_trans,) if n_trans > 1 else ()) + shape, dtype=cd, device=dev,
                          generator=g)
        out = torch.empty(M * n_trans, dtype=cd, device=dev)
        args = (out, inp)
    table = []
    best, best_t = None, float("inf")

    def trial(opts, n=reps):
        kw = dict(upsampfac=upsampfac)
        kw.update(opts)
        plan = None
        try:
            plan = cufinufft(nufft_type, shape, n_trans=n_trans, eps=eps, dtype=dtype,
                             maxbatchsize=maxbatchsize, accuracy_first=accuracy_first, **kw)
            plan.set_pts(*pts)
            plan.execute(*args)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(n):
                plan.execute(*args)
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / n * 1e3
        except RuntimeError as e:
            if log:
                log(f"  skip {opts}: {e}")
            return None
        finally:
            plan = None
        table.append((opts, t))
        if log:
            log(f"  {opts}: {t:.3f} ms")
        return t

    for opts in (cands or candidates(nufft_type, dim, M)):
        t = trial(opts)
        if t is not None and t < best_t:
            best, best_t = opts, t
    if best is None:
        raise RuntimeError("no candidate option set could be planned")
    if cands is None:
        # The short first-stage timings are noisy at the few-percent level:
        # re-time the three fastest with more executes and, on the new
        # winner, the polynomial kernel (sigma = 2 only) and the smaller
        # sigma = 1.25 grid (wider kernel, smaller FFT).
        long_reps = max(3 * reps, 5)
        top = sorted(table, key=lambda r: r[1])[:3]
        best, best_t = None, float("inf")
        for opts, _ in top:
            t = trial(opts, long_reps)
            if t is not None and t < best_t:
                best, best_t = opts, t
        if upsampfac == 2.0:
            t = trial(dict(best, gpu_kerevalmeth=1), long_reps)
            if t is not None and t < best_t:
                best, best_t = dict(best, gpu_kerevalmeth=1), t
        # type-1 SM: stored tiles and a gather pass instead of atomic flushes
        # (pays where most bins hold one subproblem)
        if nufft_type == 1 and best.get("gpu_method", 2) == 2:
            t = trial(dict(best, tile_slots=True), long_reps)
            if t is not None and t < best_t:
                best, best_t = dict(best, tile_slots=True), t
        if upsampfac == 2.0:
            # sigma = 1.25 widens the kernel (ns 9 instead of 6 at tol 1e-5,
            # 15 instead of 10 at 1e-9), so the winner's bins may not fit the
            # LDS with it: halve the largest bin side until a plan is made
            for opts in sigma125_candidates(best, dim):
                t = trial(opts, long_reps)
                if t is None:
                    continue
                if t < best_t:
                    best, best_t = opts, t
                break
    _CACHE[key] = best
    if use_cache:
        _save_disk()
    return dict(best), table


def group_candidates(dim):
    """Option sets to try for a plan group (one SM launch over every member,
    so one method -- 2 -- and one set of bins and subproblem size, whose
    largest tile sizes the launch's LDS)."""
    if dim == 3:
        bins = [(16, 8, 8), (8, 8, 8), (16, 16, 4)]
    elif dim == 2:
        bins = [(32, 32, 1), (16, 16, 1), (64, 32, 1), (32, 16, 1), (64, 64, 1)]
    else:
        bins = [(1024, 1, 1), (512, 1, 1)]
    return [dict(gpu_method=2, gpu_binsizex=b[0], gpu_binsizey=b[1], gpu_binsizez=b[2],
                 gpu_maxsubprobsize=msub) for b in bins for msub in (512, 1024, 4096)]


def tune_group(build, cands, reps=5, log=None):
    """Time a plan group under each candidate option set.

    build(opts) makes the members with opts (points set), forms the group and
    returns a callable that executes it once (or raises RuntimeError when
    the options cannot be planned).  Returns (best_opts, [(opts, ms) ...]).
    """
    import torch
    table = []
    best, best_t = None, float("inf")
    for opts in cands:
        try:
            run = build(opts)
            run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                run()
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / reps * 1e3
        except RuntimeError as e:
            if log:
                log(f"  skip {opts}: {e}")
            continue
        finally:
            run = None
        table.append((opts, t))
        if log:
            log(f"  group {opts}: {t:.3f} ms")
        if t < best_t:
            best, best_t = opts, t
    if best is None:
        raise RuntimeError("no candidate option set could be planned for the group")
    return dict(best), table
