"""ORACLE -- test infrastructure only.

CPU restatement of cuFINUFFT v1.3's type-1/type-2 NUFFT path (the C core is
``oracle/oracle.c``; this module binds it with ctypes and adds the FFT step with
numpy / scipy).  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this module, and only as the checker.  The
product library never calls it.

Functions and the reference code they restate:

* ``setup_spreader``   contrib/spreadinterp.cpp:6-67
* ``set_nf_type12``    contrib/common.cpp:24-37, next235beven contrib/utils.cpp:3-22
* ``rescale``          contrib/spreadinterp.h:36-38 (RESCALE, pirange=1)
* ``bin_sort``         spreadinterp{2,3}d.cu CalcBinSize_noghost / CalcInvertofGlobalSortIdx,
                       precision_independent.cu CalcSubProb / MapBintoSubProb,
                       spread2d_wrapper.cu:268-300,528-592 (scans)
* ``fseries``          contrib/common.cpp:39-124, src/common.cu:16-45
* ``spread``/``interp`` spreadinterp{2,3}d.cu (every method computes the same sum)
* ``deconvolve``/``amplify`` src/deconvolve_wrapper.cu:26-121
* ``type1``/``type2``  CUFINUFFT{2,3}D{1,2}_EXEC (src/2d/cufinufft2d.cu:15-165)
* ``dirft_type1``/``dirft_type2`` contrib/dirft2d.cpp:7-74 (and its 3D analogue)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    """Load (building on first use if needed) liboracle.so."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "liboracle.so"], cwd=_HERE)
        _LIB = ctypes.CDLL(path)
        _declare(_LIB)
    return _LIB


def _declare(L):
    i, d, f, p = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)
    L.orc_next235beven.argtypes = [i, i]
    L.orc_next235beven.restype = i
    L.orc_setup_spreader_f.argtypes = [f, f, i, ip, ctypes.POINTER(f), ctypes.POINTER(f), ctypes.POINTER(f)]
    L.orc_setup_spreader_f.restype = i
    L.orc_setup_spreader_d.argtypes = [d, d, i, ip, ctypes.POINTER(d), ctypes.POINTER(d), ctypes.POINTER(d)]
    L.orc_setup_spreader_d.restype = i
    L.orc_set_nf_type12.argtypes = [i, d, i, i]
    L.orc_set_nf_type12.restype = i
    L.orc_rescale_f.argtypes = [i, p, i, p]
    L.orc_rescale_d.argtypes = [i, p, i, p]
    L.orc_rescale_f_exact.argtypes = [i, p, i, p]
    L.orc_numbins.argtypes = [i, i, i]
    L.orc_numbins.restype = i
    L.orc_bin_sort.argtypes = [i, i, i, p, p, p, p, p, i, p, p, p, p, p, p]
    L.orc_bin_sort.restype = i
    L.orc_es_kernel.argtypes = [d, i, d, d]
    L.orc_es_kernel.restype = d
    L.orc_spread.argtypes = [i, i, p, p, p, p, p, i, d, d, p]
    L.orc_spread_omp.argtypes = [i, i, p, p, p, p, p, i, d, d, p]
    L.orc_interp.argtypes = [i, i, p, p, p, p, p, i, d, d, p]
    L.orc_gauss_legendre.argtypes = [i, p, p]
    L.orc_fseries.argtypes = [i, i, d, d, p]
    L.orc_deconvolve.argtypes = [i, p, p, p, p, p, p, p]
    L.orc_amplify.argtypes = [i, p, p, p, p, p, p, p]
    L.orc_dirft_type1.argtypes = [i, i, p, p, p, p, i, p, i, p, p]
    L.orc_dirft_type2.argtypes = [i, p, p, p, i, p, p, i, p, p]
    L.orc_num_threads.restype = i


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def num_threads():
    return lib().orc_num_threads()


# --------------------------------------------------------------------------
class SpreadParams:
    """Spreader parameters as set up in the plan's precision."""

    def __init__(self, ns, beta, c, halfwidth, ier):
        self.ns, self.beta, self.c, self.halfwidth, self.ier = ns, beta, c, halfwidth, ier

    def __repr__(self):
        return f"SpreadParams(ns={self.ns}, beta={self.beta!r}, c={self.c!r}, ier={self.ier})"


def setup_spreader(tol, single, upsampfac=2.0, kerevalmeth=0):
    L = lib()
    ns = ctypes.c_int(0)
    if single:
        b, c, h = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        ier = L.orc_setup_spreader_f(float(np.float32(tol)), float(np.float32(upsampfac)),
                                     kerevalmeth, ctypes.byref(ns), ctypes.byref(b),
                                     ctypes.byref(c), ctypes.byref(h))
    else:
        b, c, h = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ier = L.orc_setup_spreader_d(float(tol), float(upsampfac), kerevalmeth,
                                     ctypes.byref(ns), ctypes.byref(b), ctypes.byref(c),
                                     ctypes.byref(h))
    return SpreadParams(ns.value, b.value, c.value, h.value, ier)


def next235beven(n, b=1):
    return lib().orc_next235beven(int(n), int(b))


def set_nf_type12(ms, ns, upsampfac=2.0, b=1):
    return lib().orc_set_nf_type12(int(ms), float(upsampfac), int(ns), int(b))


def numbins(nf, bs, single):
    return lib().orc_numbins(1 if single else 0, int(nf), int(bs))


def rescale(x, nf):
    """RESCALE in the dtype of x (float32 or float64)."""
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    if x.dtype == np.float32:
        lib().orc_rescale_f(x.size, _ptr(x), int(nf), _ptr(out))
    elif x.dtype == np.float64:
        lib().orc_rescale_d(x.size, _ptr(x), int(nf), _ptr(out))
    else:
        raise TypeError(x.dtype)
    return out


def bin_sort(coords, nf, bs, maxsub=1024):
    """Counting bin sort.  coords: list of dim arrays of one float dtype.

    Returns dict with binof, binsize, binstart, idx (ascending in-bin order),
    numsub, substart (length nbins+1), total subproblems and numbins.
    """
    dim = len(coords)
    single = coords[0].dtype == np.float32
    cs = [np.ascontiguousarray(c) for c in coords] + [None] * (3 - dim)
    M = cs[0].size
    nfa = np.array(list(nf) + [1] * (3 - dim), dtype=np.int32)
    bsa = np.array(list(bs) + [1] * (3 - dim), dtype=np.int32)
    nb = [numbins(nfa[d], bsa[d], single) for d in range(dim)] + [1] * (3 - dim)
    nbins = nb[0] * nb[1] * nb[2]
    binof = np.empty(M, np.int32)
    binsize = np.empty(nbins, np.int32)
    binstart = np.empty(nbins, np.int32)
    idx = np.empty(M, np.int32)
    numsub = np.empty(nbins, np.int32)
    substart = np.empty(nbins + 1, np.int32)
    tot = lib().orc_bin_sort(1 if single else 0, dim, M, _ptr(cs[0]), _ptr(cs[1]), _ptr(cs[2]),
                             _ptr(nfa), _ptr(bsa), int(maxsub), _ptr(binof), _ptr(binsize),
                             _ptr(binstart), _ptr(idx), _ptr(numsub), _ptr(substart))
    return dict(binof=binof, binsize=binsize, binstart=binstart, idx=idx, numsub=numsub,
                substart=substart, total=tot, numbins=nb[:dim])


def es_kernel(x, ns, beta, c):
    return lib().orc_es_kernel(float(x), int(ns), float(beta), float(c))


def gauss_legendre(n):
    x = np.empty(n)
    w = np.empty(n)
    lib().orc_gauss_legendre(int(n), _ptr(x), _ptr(w))
    return x, w


def fseries(nf, sp):
    out = np.empty(nf // 2 + 1)
    lib().orc_fseries(int(nf), int(sp.ns), float(sp.beta), float(sp.c), _ptr(out))
    return out


def rescale_exact(x, nf):
    """RESCALE of float32 x evaluated in double, before rounding to float."""
    x = np.ascontiguousarray(x)
    if x.dtype == np.float64:
        return rescale(x, nf)
    out = np.empty(x.size, np.float64)
    lib().orc_rescale_f_exact(x.size, _ptr(x), int(nf), _ptr(out))
    return out


def _rescaled(coords, nf, exact=True):
    """Grid coordinates used for tap offsets.  exact=True: the library's
    chain (unrounded double rescale); exact=False: the reference's chain
    (rescale rounded to the coordinate dtype)."""
    if exact:
        return [rescale_exact(c, n) for c, n in zip(coords, nf)]
    return [rescale(c, n).astype(np.float64) for c, n in zip(coords, nf)]


def spread(coords, c, nf, sp, exact=True, threads=False):
    """Spread strengths c (complex, length M) to the fine grid (shape nf[::-1]).
    threads=True: the OpenMP slab spread (the CPU baseline's; same sums)."""
    dim = len(coords)
    xr = _rescaled(coords, nf, exact) + [None] * (3 - dim)
    cc = _c(c, np.complex128)
    nfa = np.array(list(nf) + [1] * (3 - dim), dtype=np.int32)
    fw = np.empty(tuple(nf[::-1]), np.complex128)
    fn = lib().orc_spread_omp if threads else lib().orc_spread
    fn(dim, cc.size, _ptr(xr[0]), _ptr(xr[1]), _ptr(xr[2]), _ptr(cc), _ptr(nfa),
       int(sp.ns), float(sp.beta), float(sp.c), _ptr(fw))
    return fw


def interp(coords, fw, nf, sp, exact=True):
    dim = len(coords)
    xr = _rescaled(coords, nf, exact) + [None] * (3 - dim)
    fwc = _c(fw, np.complex128)
    nfa = np.array(list(nf) + [1] * (3 - dim), dtype=np.int32)
    M = xr[0].size
    out = np.empty(M, np.complex128)
    lib().orc_interp(dim, M, _ptr(xr[0]), _ptr(xr[1]), _ptr(xr[2]), _ptr(out), _ptr(nfa),
                     int(sp.ns), float(sp.beta), float(sp.c), _ptr(fwc))
    return out


def deconvolve(fw, ms, nf, kers):
    dim = len(ms)
    k = [np.ascontiguousarray(a, np.float64) for a in kers] + [None] * (3 - dim)
    fk = np.empty(tuple(ms[::-1]), np.complex128)
    msa = np.array(list(ms) + [1] * (3 - dim), np.int32)
    nfa = np.array(list(nf) + [1] * (3 - dim), np.int32)
    lib().orc_deconvolve(dim, _ptr(msa), _ptr(nfa), _ptr(_c(fw, np.complex128)), _ptr(k[0]),
                         _ptr(k[1]), _ptr(k[2]), _ptr(fk))
    return fk


def amplify(fk, ms, nf, kers):
    dim = len(ms)
    k = [np.ascontiguousarray(a, np.float64) for a in kers] + [None] * (3 - dim)
    fw = np.empty(tuple(nf[::-1]), np.complex128)
    msa = np.array(list(ms) + [1] * (3 - dim), np.int32)
    nfa = np.array(list(nf) + [1] * (3 - dim), np.int32)
    lib().orc_amplify(dim, _ptr(msa), _ptr(nfa), _ptr(_c(fk, np.complex128)), _ptr(k[0]),
                      _ptr(k[1]), _ptr(k[2]), _ptr(fw))
    return fw


def _fft(fw, iflag, workers):
    """Unnormalised DFT of the whole grid with sign iflag (cuFFT/hipFFT
    convention: iflag=+1 is the backward transform without 1/n)."""
    try:
        import scipy.fft as sfft
        if iflag >= 0:
            return sfft.ifftn(fw, norm="forward", workers=workers)
        return sfft.fftn(fw, workers=workers)
    except ImportError:  # pragma: no cover
        if iflag >= 0:
            return np.fft.ifftn(fw) * fw.size
        return np.fft.fftn(fw)


class Plan:
    """Oracle plan: same parameters as the library's makeplan (types 1/2)."""

    def __init__(self, nufft_type, n_modes, tol, single, iflag=None, upsampfac=2.0,
                 exact=True, blocks=None):
        self.type = nufft_type
        self.ms = list(n_modes)
        self.dim = len(self.ms)
        self.single = single
        self.iflag = (1 if nufft_type == 1 else -1) if iflag is None else iflag
        self.sp = setup_spreader(tol, single, upsampfac)
        # blocks: the gpu_method 4 output block sizes (nf rounded to multiples)
        bl = list(blocks) if blocks else [1] * self.dim
        self.nf = [set_nf_type12(m, self.sp.ns, upsampfac, b) for m, b in zip(self.ms, bl)]
        self.kers = [fseries(n, self.sp) for n in self.nf]
        self.exact = exact

    def type1(self, coords, c, workers=1):
        fw = spread(coords, c, self.nf, self.sp, self.exact)
        fw = _fft(fw, self.iflag, workers)
        return deconvolve(fw, self.ms, self.nf, self.kers)

    def type2(self, coords, fk, workers=1):
        fw = amplify(fk, self.ms, self.nf, self.kers)
        fw = _fft(fw, self.iflag, workers)
        return interp(coords, fw, self.nf, self.sp, self.exact)


def dirft_type1(coords, c, ms, iflag, kidx):
    """Direct FP64 type-1 sum at flat mode indices kidx (x fastest)."""
    dim = len(coords)
    xs = [_c(a, np.float64) for a in coords] + [None] * (3 - dim)
    cc = _c(c, np.complex128)
    kidx = _c(kidx, np.int64)
    msa = np.array(list(ms) + [1] * (3 - dim), np.int32)
    out = np.empty(kidx.size, np.complex128)
    lib().orc_dirft_type1(dim, cc.size, _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2]), _ptr(cc),
                          int(iflag), _ptr(msa), kidx.size, _ptr(kidx), _ptr(out))
    return out


def dirft_type2(coords, fk, ms, iflag, pidx):
    """Direct FP64 type-2 sum at point indices pidx; fk flat, x fastest."""
    dim = len(coords)
    xs = [_c(a, np.float64) for a in coords] + [None] * (3 - dim)
    f = _c(fk, np.complex128).ravel()
    pidx = _c(pidx, np.int64)
    msa = np.array(list(ms) + [1] * (3 - dim), np.int32)
    out = np.empty(pidx.size, np.complex128)
    lib().orc_dirft_type2(dim, _ptr(xs[0]), _ptr(xs[1]), _ptr(xs[2]), pidx.size, _ptr(pidx),
                          _ptr(f), int(iflag), _ptr(msa), _ptr(out))
    return out


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
