"""Pinning the oracle (CPU).

Every oracle function is checked against something other than its own
formula: brute force, exact rational arithmetic, an independent library
routine (numpy's Gauss-Legendre, scipy's adaptive quadrature, numpy FFTs),
the reference's own Gauss-Legendre code compiled from its source
(oracle/_ref), closed forms, or invariants that a dropped term, wrong sign,
wrong index or transposed axis breaks.
"""
import ctypes
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
REF_GL = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "liblegendre_ref.so")


# ---------------------------------------------------------------- sizes
def _smooth_even_bruteforce(n):
    m = max(n, 2)
    while True:
        if m % 2 == 0:
            r = m
            for p in (2, 3, 5):
                while r % p == 0:
                    r //= p
            if r == 1:
                return m
        m += 1


def test_next235beven_bruteforce():
    for n in list(range(0, 3000)) + [4097, 10001, 65537, 999_999]:
        assert O.next235beven(n) == _smooth_even_bruteforce(n), n


def test_next235beven_block_multiple_bruteforce():
    # method 4 rounds nf up to an even smooth multiple of the output block
    # (contrib/common.cpp:31-32): brute force over n and smooth block sizes
    for b in (1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 24, 32):
        for n in list(range(0, 400)) + [1001, 4097]:
            m = max(n, 2)
            m += m % 2
            while True:
                r = m
                for p in (2, 3, 5):
                    while r % p == 0:
                        r //= p
                if r == 1 and m % b == 0:
                    break
                m += 2
            want = 2 if n <= 2 else m
            assert O.next235beven(n, b) == want, (n, b)
    # the example the reference's rule changes: N=50, ns=6, 8^3 blocks
    assert O.set_nf_type12(50, 6) == 100
    assert O.set_nf_type12(50, 6, 2.0, 8) == 120


def test_nf_table():
    # fine-grid sizes of the workload shapes (sigma = 2): 2N is already smooth
    for N, nf in ((256, 512), (128, 256), (512, 1024), (320, 640), (100, 200)):
        assert O.set_nf_type12(N, 6) == nf
    # small N: nf >= 2 ns, then smoothed
    assert O.set_nf_type12(3, 7) == 16  # 2*ns = 14 = 2*7 is not 5-smooth
    assert O.set_nf_type12(5, 8) == 16
    # non-standard sigma: nf = next235beven(floor(sigma N))
    assert O.set_nf_type12(100, 6, 1.25) == _smooth_even_bruteforce(125)


# --------------------------------------------------------------- spreader
@pytest.mark.parametrize("single", [True, False])
def test_setup_spreader_table(single):
    # ns = ceil(-log10(tol/10)) clipped to [2, 16]; beta/ns = 2.30 except
    # 2.20, 2.26, 2.38 at ns = 2, 3, 4; c = 4/ns^2
    cases = [(1e-1, 2), (1e-2, 3), (1e-3, 4), (1e-4, 5), (1e-5, 6), (1e-6, 7)]
    if not single:
        cases += [(1e-9, 10), (1e-12, 13), (1e-14, 15), (1e-15, 16)]
    bon = {2: 2.20, 3: 2.26, 4: 2.38}
    for tol, ns in cases:
        sp = O.setup_spreader(tol, single)
        assert sp.ns == ns and sp.ier == 0, (tol, sp)
        f = np.float32 if single else np.float64
        assert sp.beta == pytest.approx(float(f(bon.get(ns, 2.30)) * f(ns)), rel=1e-7)
        assert sp.c == pytest.approx(4.0 / ns ** 2, rel=1e-7)


def test_setup_spreader_codes():
    sp = O.setup_spreader(1e-20, False)           # below eps_mach: warn, clamp
    assert sp.ier == 1 and sp.ns == 16
    sp = O.setup_spreader(1e-9, True)             # float: 6e-8 floor -> ns 9
    assert sp.ier == 1 and sp.ns == 9
    assert O.setup_spreader(1e-6, False, upsampfac=0.9).ier == 7
    assert O.setup_spreader(1e-6, False, upsampfac=1.25, kerevalmeth=1).ier == 8
    sp = O.setup_spreader(1e-6, False, upsampfac=1.25)
    ns = int(np.ceil(-np.log(1e-6) / (np.pi * np.sqrt(1 - 1 / 1.25))))
    assert sp.ns == ns
    assert sp.beta == pytest.approx(0.97 * np.pi * (1 - 1 / 2.5) * ns, rel=1e-12)


# ------------------------------------------------------- rescale and bins
def _rescale_exact_rational(x32, n):
    """RESCALE of one float32 by exact rational arithmetic: fma(x, 1/2pi, off)
    (one rounding to double), times n (a second rounding), then to float32."""
    c = Fraction(0.159154943091895336)
    pi32 = np.float32(np.pi)
    off = 1.5 if x32 < -pi32 else (-0.5 if x32 >= pi32 else 0.5)
    t = float(Fraction(float(x32)) * c + Fraction(off))  # correctly rounded
    v = float(Fraction(t) * n)
    return np.float32(v), v


def test_rescale_rounding_chain():
    rng = np.random.default_rng(0)
    pi32 = np.float32(np.pi)
    xs = list(rng.uniform(-3 * np.pi, 3 * np.pi, 2000).astype(np.float32))
    xs += [pi32, -pi32, np.nextafter(pi32, np.float32(0)), np.nextafter(-pi32, np.float32(0)),
           np.float32(0), np.float32(-0.0), np.float32(1e-30), np.float32(3 * np.pi - 1e-6)]
    xs = np.array(xs, np.float32)
    for n in (512, 1000, 640):
        got = O.rescale(xs, n)
        exact = O.rescale_exact(xs, n)
        for i, x in enumerate(xs):
            r32, r64 = _rescale_exact_rational(x, n)
            assert got[i] == r32, (x, n)
            assert exact[i] == r64, (x, n)


def test_rescale_double_and_fold():
    rng = np.random.default_rng(1)
    x = rng.uniform(-np.pi, np.pi, 1000)
    got = O.rescale(x, 256)
    ref = np.array([float(Fraction(float(Fraction(v) * Fraction(0.159154943091895336)
                                         + Fraction(0.5))) * 256) for v in x])
    np.testing.assert_array_equal(got, ref)
    # one period below / above folds onto the same grid position
    a = O.rescale(np.array([-1.0 - 2 * np.pi, -1.0, -1.0 + 2 * np.pi]), 256)
    assert abs(a[0] - a[1]) < 1e-9 and abs(a[2] - a[1]) < 1e-9


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("bs", [(32, 32), (12, 7)])
def test_bin_sort_invariants(dtype, bs):
    rng = np.random.default_rng(2)
    M = 20000
    pts = [rng.uniform(-np.pi, np.pi, M).astype(dtype) for _ in range(2)]
    # a few points exactly at the period edges
    pts[0][:4] = np.array([-np.pi, np.pi, 0, np.nextafter(np.pi, 0)], dtype)
    nf = (200, 96)
    r = O.bin_sort(pts, nf, bs, maxsub=37)
    nb = [int(np.ceil(np.float32(n) / np.float32(b))) if dtype == np.float32 else
          int(np.ceil(n / b)) for n, b in zip(nf, bs)]
    assert list(r["numbins"]) == nb
    # independent bin computation: numpy floor of the dtype division, clamps
    bins = np.zeros(M, np.int64)
    stride = 1
    for d in range(2):
        xr = O.rescale(pts[d], nf[d])
        q = np.floor(xr / dtype(bs[d])).astype(np.int64)
        q = np.where(q >= nb[d], q - 1, q)
        q = np.maximum(q, 0)
        bins += q * stride
        stride *= nb[d]
    np.testing.assert_array_equal(r["binof"], bins)
    cnt = np.bincount(bins, minlength=nb[0] * nb[1])
    np.testing.assert_array_equal(r["binsize"], cnt)
    np.testing.assert_array_equal(r["binstart"], np.concatenate([[0], np.cumsum(cnt)[:-1]]))
    ns = -(-cnt // 37)
    np.testing.assert_array_equal(r["numsub"], ns)
    np.testing.assert_array_equal(r["substart"], np.concatenate([[0], np.cumsum(ns)]))
    assert r["total"] == ns.sum()
    assert np.array_equal(np.sort(r["idx"]), np.arange(M))
    np.testing.assert_array_equal(bins[r["idx"]], np.repeat(np.arange(cnt.size), cnt))


# ------------------------------------------------------------ ES kernel
def test_es_kernel_closed_form():
    for ns, beta in ((6, 13.8), (10, 23.0)):
        c = 4.0 / ns ** 2
        assert O.es_kernel(0.0, ns, beta, c) == pytest.approx(np.exp(beta), rel=1e-15)
        assert O.es_kernel(ns / 2, ns, beta, c) == 0.0
        assert O.es_kernel(-ns / 2 - 0.1, ns, beta, c) == 0.0
        x = 1.234
        assert O.es_kernel(x, ns, beta, c) == pytest.approx(np.exp(beta * np.sqrt(1 - c * x * x)))
        assert O.es_kernel(-x, ns, beta, c) == O.es_kernel(x, ns, beta, c)


# --------------------------------------------------- quadrature, fseries
@pytest.mark.parametrize("n", [4, 10, 17, 34, 50])
def test_gauss_legendre_numpy(n):
    x, w = O.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(x, xr, atol=1e-14)
    np.testing.assert_allclose(w, wr, rtol=1e-12, atol=1e-15)


@pytest.mark.skipif(not os.path.exists(REF_GL), reason="oracle/_ref not built")
@pytest.mark.parametrize("n", [8, 20, 34, 50])
def test_gauss_legendre_reference_code(n):
    """Against the reference's own Glaser-Liu-Rokhlin routine
    (contrib/legendre_rule_fast.c, compiled unmodified into oracle/_ref)."""
    lib = ctypes.CDLL(REF_GL)
    xr = np.zeros(n)
    wr = np.zeros(n)
    lib.legendre_compute_glr(ctypes.c_int(n), xr.ctypes.data_as(ctypes.c_void_p),
                             wr.ctypes.data_as(ctypes.c_void_p))
    order = np.argsort(xr)
    x, w = O.gauss_legendre(n)
    np.testing.assert_allclose(x, xr[order], atol=1e-13)
    np.testing.assert_allclose(w, wr[order], rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("tol,single,nf", [(1e-5, True, 512), (1e-9, False, 256), (1e-3, True, 64),
                                           (1e-12, False, 128)])
def test_fseries_vs_adaptive_quadrature(tol, single, nf):
    """fwkerhalf[k] = (-1)^k * integral of phi(z) cos(2 pi k z / nf) over the
    support.  The reference's q = 2 + 1.5 ns node rule is limited by the
    square-root cusp of phi at the support edge (1e-7 of the peak at ns=4,
    1e-9 at ns=6, 1e-13 at ns=10), always far below tol: bar tol/100."""
    import warnings
    from scipy.integrate import IntegrationWarning, quad
    sp = O.setup_spreader(tol, single)
    got = O.fseries(nf, sp)
    hw = sp.ns / 2
    phi = lambda z: O.es_kernel(z, sp.ns, sp.beta, sp.c)
    for k in (0, 1, 7, nf // 4, nf // 2):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", IntegrationWarning)
            val, _ = quad(lambda z: phi(z) * np.cos(2 * np.pi * k * z / nf), -hw, hw,
                          limit=400, epsabs=0, epsrel=1e-13, points=[0])
        assert abs(got[k] - (-1) ** k * val) <= tol / 100 * abs(got[0]), k


# ------------------------------------------------- deconvolve / amplify
def test_amplify_deconvolve_maps():
    ms, nf = [5, 6, 4], [12, 16, 10]
    rng = np.random.default_rng(3)
    fk = rng.standard_normal(ms[::-1]) + 1j * rng.standard_normal(ms[::-1])
    ones = [np.ones(n // 2 + 1) for n in nf]
    fw = O.amplify(fk, ms, nf, ones)
    # every mode lands at its wrapped frequency index, nothing else is set
    assert np.count_nonzero(fw) == fk.size
    for (k3, k2, k1), v in np.ndenumerate(fk):
        f = [k1 - ms[0] // 2, k2 - ms[1] // 2, k3 - ms[2] // 2]
        assert fw[f[2] % nf[2], f[1] % nf[1], f[0] % nf[0]] == v
    np.testing.assert_array_equal(O.deconvolve(fw, ms, nf, ones), fk)
    # the factors divide by prod_d ker_d[|k_d - ms_d/2|]
    kers = [1.0 + np.arange(n // 2 + 1) for n in nf]
    fk2 = O.deconvolve(fw, ms, nf, kers)
    K = np.einsum("i,j,k->ijk", *[kers[2][np.abs(np.arange(ms[2]) - ms[2] // 2)],
                                  kers[1][np.abs(np.arange(ms[1]) - ms[1] // 2)],
                                  kers[0][np.abs(np.arange(ms[0]) - ms[0] // 2)]])
    np.testing.assert_allclose(fk2, fk / K)


# --------------------------------------------------------- direct sums
def test_dirft_bruteforce_and_closed_form():
    rng = np.random.default_rng(4)
    M, N = 50, (6, 5, 4)
    x = [rng.uniform(-np.pi, np.pi, M) for _ in range(3)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    k = np.meshgrid(*[np.arange(-(n // 2), (n + 1) // 2) for n in N[::-1]], indexing="ij")
    ph = np.einsum("dk,dj->kj", np.stack([k[2].ravel(), k[1].ravel(), k[0].ravel()]),
                   np.stack(x))
    ref1 = np.exp(1j * ph) @ c
    got1 = O.dirft_type1(x, c, list(N), +1, np.arange(np.prod(N)))
    np.testing.assert_allclose(got1, ref1, rtol=1e-12, atol=1e-12)
    fk = rng.standard_normal(np.prod(N)) + 1j * rng.standard_normal(np.prod(N))
    ref2 = np.exp(-1j * ph).T @ fk
    got2 = O.dirft_type2(x, fk, list(N), -1, np.arange(M))
    np.testing.assert_allclose(got2, ref2, rtol=1e-12, atol=1e-12)
    # one unit source at the origin: every mode is 1
    z = [np.zeros(1)] * 2
    np.testing.assert_allclose(O.dirft_type1(z, np.ones(1, complex), [7, 4], 1, np.arange(28)),
                               np.ones(28))


# --------------------------------------------------- full oracle NUFFT
@pytest.mark.parametrize("dim,N,tol", [(2, (40, 36), 1e-6), (3, (12, 10, 14), 1e-8),
                                       (1, (64,), 1e-10)])
def test_oracle_nufft_vs_direct(dim, N, tol):
    rng = np.random.default_rng(5)
    M = 3000
    pts = [rng.uniform(-np.pi, np.pi, M) for _ in range(dim)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    P1 = O.Plan(1, list(N), tol, single=False)
    f = P1.type1(pts, c)
    kidx = np.arange(int(np.prod(N)))
    assert rel_l2(f.ravel(), O.dirft_type1(pts, c, list(N), 1, kidx)) < 3 * tol
    fk = rng.standard_normal(N[::-1]) + 1j * rng.standard_normal(N[::-1])
    P2 = O.Plan(2, list(N), tol, single=False)
    cc = P2.type2(pts, fk)
    assert rel_l2(cc, O.dirft_type2(pts, fk, list(N), -1, np.arange(M))) < 3 * tol
    # adjointness: <type1(c), g> == <c, type2(g)> with opposite signs
    g = rng.standard_normal(N[::-1]) + 1j * rng.standard_normal(N[::-1])
    lhs = np.vdot(g.ravel(), f.ravel())
    rhs = np.vdot(O.Plan(2, list(N), tol, single=False, iflag=-1).type2(pts, g), c)
    assert abs(lhs - rhs) < 10 * tol * abs(lhs)


def test_reference_fp32_coordinate_floor():
    """Documents why the library takes FP32 tap offsets from the unrounded
    rescale: with the reference's float-rounded coordinate the error at
    nf = 1024 has a floor near 2e-5, far above tol = 1e-6."""
    rng = np.random.default_rng(6)
    M, N = 20000, (512, 512)
    pts = [rng.uniform(-np.pi, np.pi, M).astype(np.float32) for _ in range(2)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    kidx = rng.choice(N[0] * N[1], 200, replace=False)
    ref = O.dirft_type1([p.astype(np.float64) for p in pts], c, list(N), 1, kidx)
    e_ref = rel_l2(O.Plan(1, list(N), 1e-6, True, exact=False).type1(pts, c).ravel()[kidx], ref)
    e_lib = rel_l2(O.Plan(1, list(N), 1e-6, True, exact=True).type1(pts, c).ravel()[kidx], ref)
    assert e_ref > 5e-6 and e_lib < 2e-6


@pytest.mark.parametrize("dim,nm,M", [(1, (200,), 3000), (2, (40, 36), 8000),
                                      (3, (20, 18, 16), 8000), (3, (6, 6, 6), 500),
                                      (2, (24, 8), 0)])
def test_spread_threads_matches_sequential(dim, nm, M):
    """The OpenMP slab spread (the CPU baseline's) against the sequential
    spread the checks use: the same per-cell sums in another order, so equal
    to rounding -- including grids with fewer planes per thread than taps
    (6^3 modes: 12 planes for the threads) and no points."""
    rng = np.random.default_rng(dim + M)
    P = O.Plan(1, list(nm), 1e-6, single=False)
    pts = [rng.uniform(-np.pi, np.pi, M) for _ in range(dim)]
    c = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    a = O.spread(pts, c, P.nf, P.sp)
    b = O.spread(pts, c, P.nf, P.sp, threads=True)
    if M == 0:
        assert not a.any() and not b.any()
    else:
        assert rel_l2(b, a) < 1e-13
