"""Upsampling factor sigma = 1.25 (contrib/spreadinterp.cpp:14-26,43,60-62 of
the reference: ns = ceil(-ln(tol) / (pi sqrt(1 - 1/sigma))), beta =
0.97 pi (1 - 1/(2 sigma)) ns, nf = next235beven(floor(sigma N))) against the
oracle's sigma path and the direct sums, in 2D and 3D, FP32 and FP64, type 1
and type 2.  The wider kernel (ns 9 at tol 1e-5, 15 at 1e-9) needs bins small
enough for its tile to fit the LDS.  Bars: within tol/5 of the oracle
(2e-6 for FP32, whose kernel values cost ~5e-7 in 3D) and rel-l2 against the
direct sum within max(2 tol, 1.25 x the oracle's own error).

At sigma = 1.25 the passband reaches far into the kernel's decay, so the
deconvolution factors 1/phi^(k) grow to ~1e2 (2D) and ~1e3 (3D) at the band
edge: the fine grid holds the modes amplified by them, and every FP32
rounding on that grid (the FFT, the interp sums) reaches the result
amplified alike.  The oracle computes in FP64, so an FP32 plan's distance
from it is bounded by the same rounding, measured on the CPU: the oracle
pipeline with its FFT done in complex64 (``_fft32_dev``) deviates by 1.4e-5
(3D type 2), against 1e-7 at sigma = 2.  The FP32 bar is three times that
deviation (the device also rounds its spread / interp sums in FP32), or
2e-6, whichever is larger; the direct-sum bar is unchanged."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import gpu_util as G
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = [
    # dim, dtype, modes (x, y[, z]), M, tol, bins
    (2, np.float32, (64, 48), 20000, 1e-5, (32, 32, 1)),
    (2, np.float64, (40, 36), 10000, 1e-9, (16, 16, 1)),
    (3, np.float32, (24, 20, 16), 20000, 1e-5, (8, 8, 8)),
    (3, np.float64, (16, 16, 12), 8000, 1e-9, (4, 4, 4)),
]


def _fft32_dev(P, kind, pts, data):
    """rel-l2 distance between the oracle and the oracle with a complex64
    FFT (the FP32 rounding of the amplified fine grid alone)."""
    import scipy.fft as sfft

    def fft32(fw):
        a = fw.astype(np.complex64)
        r = sfft.ifftn(a, norm="forward") if P.iflag >= 0 else sfft.fftn(a)
        return r.astype(np.complex128)

    if kind == 1:
        fw = O.spread(pts, data, P.nf, P.sp, P.exact)
        ref = O.deconvolve(O._fft(fw, P.iflag, 1), P.ms, P.nf, P.kers)
        got = O.deconvolve(fft32(fw), P.ms, P.nf, P.kers)
    else:
        fw = O.amplify(data, P.ms, P.nf, P.kers)
        ref = O.interp(pts, O._fft(fw, P.iflag, 1), P.nf, P.sp, P.exact)
        got = O.interp(pts, fft32(fw), P.nf, P.sp, P.exact)
    return rel_l2(got, ref)


def _opts(bins):
    return dict(upsampfac=1.25, gpu_method=2, gpu_binsizex=bins[0], gpu_binsizey=bins[1],
                gpu_binsizez=bins[2])


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-{np.dtype(c[1]).name}")
def test_sigma125_type1(torch_dev, case):
    dim, dtype, nm, M, tol, bins = case
    print(f"sigma 1.25 type 1 {case}", flush=True)
    single = dtype == np.float32
    pts = G.clustered_pts(M, dim, dtype, 61)
    c = G.strengths(M, dtype, 61)
    fk, plan = G.run_type1(torch_dev, pts, c, nm, tol, dtype, **_opts(bins))
    P = O.Plan(1, list(nm), tol, single=single, upsampfac=1.25)
    info = plan.info()
    assert info["ns"] == P.sp.ns and tuple(info["nf"][:dim]) == tuple(P.nf)
    assert P.sp.ns == (9 if single else 15)
    ofk = P.type1(pts, c)
    kidx = np.random.default_rng(3).choice(int(np.prod(nm)), 600, replace=False)
    direct = O.dirft_type1([p.astype(np.float64) for p in pts], c, list(nm), +1, kidx)
    oerr = rel_l2(ofk.ravel()[kidx], direct)
    err = rel_l2(fk.ravel()[kidx], direct)
    print(f"   gpu err {err:.3e} oracle err {oerr:.3e}", flush=True)
    assert err <= max(2 * tol, 1.25 * oerr)
    bar = max(tol / 5, 2e-6, 3 * _fft32_dev(P, 1, pts, c)) if single else tol / 5
    dev = rel_l2(fk, ofk)
    print(f"   vs oracle {dev:.3e} (bar {bar:.3e})", flush=True)
    assert dev <= bar


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-{np.dtype(c[1]).name}")
def test_sigma125_type2(torch_dev, case):
    dim, dtype, nm, M, tol, bins = case
    print(f"sigma 1.25 type 2 {case}", flush=True)
    single = dtype == np.float32
    pts = G.clustered_pts(M, dim, dtype, 62)
    fk = G.modes(nm[::-1], dtype, 62)
    c, plan = G.run_type2(torch_dev, pts, fk, tol, dtype, **_opts(bins))
    P = O.Plan(2, list(nm), tol, single=single, upsampfac=1.25)
    oc = P.type2(pts, fk)
    pidx = np.random.default_rng(4).choice(M, 600, replace=False)
    direct = O.dirft_type2([p.astype(np.float64) for p in pts], fk.ravel(), list(nm), -1, pidx)
    oerr = rel_l2(oc[pidx], direct)
    err = rel_l2(c[pidx], direct)
    print(f"   gpu err {err:.3e} oracle err {oerr:.3e}", flush=True)
    assert err <= max(2 * tol, 1.25 * oerr)
    bar = max(tol / 5, 2e-6, 3 * _fft32_dev(P, 2, pts, fk)) if single else tol / 5
    dev = rel_l2(c, oc)
    print(f"   vs oracle {dev:.3e} (bar {bar:.3e})", flush=True)
    assert dev <= bar
