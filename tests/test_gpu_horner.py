"""Piecewise-polynomial kernel evaluation (gpu_kerevalmeth = 1).

The library fits its own per-tap polynomials at plan time
(contrib/ker_horner_allw_loop.c is the reference's compiled-in table for the
same option).  Bars:
* the table itself, evaluated here in FP64 at random z against the oracle's
  closed-form ES kernel: max error <= 1.5 * max(tol/20, 4e-14) of the peak
  e^beta (independent of the library's own reported fit error);
* spread / interp alone against the oracle's exp(sqrt()) sum: rel-l2 within
  max(tol/10, the FP32 floor of the exp path's own bars);
* full transforms against the FP64 direct sum: the same bar as the
  exp(sqrt()) path, for every tap path (GM, SM tile, FP32 fixed-point tile).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import gpu_util as G
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-2), (np.float32, 1e-4), (np.float32, 1e-6),
                                       (np.float64, 1e-3), (np.float64, 1e-9),
                                       (np.float64, 1e-12), (np.float64, 1e-15)])
def test_table_against_closed_form(torch_dev, dtype, tol):
    from cufinufft_amd import cufinufft
    _flush(f"horner table {np.dtype(dtype).name} tol {tol:g}")
    plan = cufinufft(1, (64, 64), eps=tol, dtype=dtype, gpu_kerevalmeth=1)
    deg, err, coef = plan.kernel_poly()
    sp = O.setup_spreader(tol, dtype == np.float32)
    ns = sp.ns
    assert deg >= 3 and coef.shape == (deg + 1, ns)
    bound = max(tol / 20, 4e-14)
    assert err <= bound
    z = np.random.default_rng(1).uniform(-1, 1, 600)
    peak = np.exp(sp.beta)
    worst = 0.0
    for i in range(ns):
        p = np.polynomial.polynomial.polyval(z, coef[:, i])
        t = (z + 1) / 2 - ns / 2 + i
        ref = np.array([O.es_kernel(v, ns, sp.beta, sp.c) for v in t])
        worst = max(worst, np.max(np.abs(p - ref)) / peak)
    assert worst <= 1.5 * bound, worst
    # the exp(sqrt()) plan reports degree 0
    assert cufinufft(1, (64, 64), eps=tol, dtype=dtype).kernel_poly()[0] == 0


SI_CASES = [
    (2, np.float32, (64, 48), 5000, 1e-5),
    (3, np.float32, (24, 20, 28), 4000, 1e-6),
    (3, np.float64, (24, 24, 24), 3000, 1e-9),
    (1, np.float64, (100,), 2000, 1e-12),
]


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("case", SI_CASES, ids=lambda c: f"{c[0]}d-{np.dtype(c[1]).name}")
def test_spread_interp_only(torch_dev, case, method):
    dim, dtype, nm, M, tol = case
    _flush(f"horner spread/interp-only {case} method {method}")
    pts = G.uniform_pts(M, dim, dtype, seed=31)
    c = G.strengths(M, dtype, seed=31)
    bins = dict(gpu_binsizex=8, gpu_binsizey=8 if dim > 1 else 1, gpu_binsizez=4 if dim > 2 else 1)
    sp = O.setup_spreader(tol, dtype == np.float32)
    bar = max(tol / 10, 8e-7 if method == 2 else 1.5e-6) if dtype == np.float32 else tol / 10
    fw, _ = G.run_type1(torch_dev, pts, c, nm, tol, dtype, gpu_method=method,
                        gpu_spreadinterponly=1, gpu_kerevalmeth=1, **bins)
    assert rel_l2(fw, O.spread(pts, c, list(nm), sp)) < bar
    f = G.modes(tuple(nm[::-1]), dtype, seed=32)
    cc, _ = G.run_type2(torch_dev, pts, f, tol, dtype, gpu_method=method,
                        gpu_spreadinterponly=1, gpu_kerevalmeth=1, **bins)
    assert rel_l2(cc, O.interp(pts, f, list(nm), sp)) < bar


NU_CASES = [
    (2, np.float32, (64, 80), 20_000, 1e-5, "uniform"),
    (3, np.float32, (32, 32, 32), 50_000, 1e-6, "clustered"),
    (3, np.float64, (24, 24, 24), 20_000, 1e-9, "clustered"),
    (1, np.float64, (400,), 5000, 1e-12, "uniform"),
]


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("case", NU_CASES, ids=lambda c: f"{c[0]}d-{np.dtype(c[1]).name}-{c[4]:g}")
def test_transforms(torch_dev, case, method):
    dim, dtype, nm, M, tol, dist = case
    _flush(f"horner transforms {case} method {method}")
    pts = (G.uniform_pts if dist == "uniform" else G.clustered_pts)(M, dim, dtype, 41)
    c = G.strengths(M, dtype, seed=41)
    fk, _ = G.run_type1(torch_dev, pts, c, nm, tol, dtype, gpu_method=method, gpu_kerevalmeth=1)
    rng = np.random.default_rng(7)
    ntot = int(np.prod(nm))
    kidx = rng.choice(ntot, min(800, ntot), replace=False)
    p64 = [p.astype(np.float64) for p in pts]
    direct = O.dirft_type1(p64, c, list(nm), +1, kidx)
    oerr = rel_l2(O.Plan(1, list(nm), tol, single=dtype == np.float32).type1(pts, c).ravel()[kidx],
                  direct)
    err = rel_l2(fk.ravel()[kidx], direct)
    _flush(f"   type 1 err {err:.3e} oracle err {oerr:.3e}")
    assert err <= max(2 * tol, 1.25 * oerr)
    f = G.modes(tuple(nm[::-1]), dtype, seed=42)
    cc, _ = G.run_type2(torch_dev, pts, f, tol, dtype, gpu_method=method, gpu_kerevalmeth=1)
    pidx = rng.choice(M, 800, replace=False)
    direct2 = O.dirft_type2(p64, f, list(nm), -1, pidx)
    oerr2 = rel_l2(O.Plan(2, list(nm), tol, single=dtype == np.float32).type2(pts, f)[pidx],
                   direct2)
    err2 = rel_l2(cc[pidx], direct2)
    _flush(f"   type 2 err {err2:.3e} oracle err {oerr2:.3e}")
    assert err2 <= max(2 * tol, 1.25 * oerr2)
