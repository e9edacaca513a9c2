"""The reference's Python-level tests, run against this library's ``cufinufft``
class with torch device tensors in place of pycuda GPUArrays:

* argument checks of python/cufinufft/tests/test_error_checks.py:12-101
  (points dtype, points sizes, unknown option, execute dtypes);
* the single-entry accuracy checks of python/cufinufft/tests/test_basic.py
  (types 1 and 2, FP32 and FP64, tol 1e-3 -> relative error < 0.01, and the
  options case with gpu_sort off and 10-point subproblems);
* the two examples examples/example2d1many.py and example2d2many.py, which
  assert relative error < 10 eps at one mode or point per transform;
* the C-ABI link path of test/cufinufft2d2api_test.cu:72-110, which makes a
  plan with opts = NULL (reference defaults).
"""
import ctypes

import numpy as np
import pytest

from tests import gpu_util as G

pytestmark = pytest.mark.gpu


def _pts(M, dim, dtype, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-np.pi, np.pi, (dim, M)).astype(dtype)


# ---------------------------------------------------------- error checks
def test_set_pts_raises_on_dtype(torch_dev):
    from cufinufft_amd import cufinufft
    k = _pts(4096, 3, np.float32, 1)
    ok = [G.to_dev(a, torch_dev) for a in k]
    bad = [G.to_dev(a.astype(np.float64), torch_dev) for a in k]
    plan = cufinufft(1, (16, 16, 16), eps=1e-3, dtype=np.float32)
    for args in ((bad[0], ok[1], ok[2]), (ok[0], bad[1], ok[2]), (ok[0], ok[1], bad[2]),
                 (bad[0], bad[1], bad[2])):
        with pytest.raises(TypeError):
            plan.set_pts(*args)


def test_set_pts_raises_on_size(torch_dev):
    from cufinufft_amd import cufinufft
    k = [G.to_dev(a, torch_dev) for a in _pts(8, 3, np.float32, 2)]
    plan = cufinufft(1, (16, 16, 16), eps=1e-3, dtype=np.float32)
    with pytest.raises(TypeError) as err:
        plan.set_pts(k[0], k[1][:4])
    assert 'kx and ky must be equal' in err.value.args[0]
    with pytest.raises(TypeError) as err:
        plan.set_pts(k[0], k[1], k[2][:4])
    assert 'kx and kz must be equal' in err.value.args[0]


def test_wrong_field_names():
    from cufinufft_amd import cufinufft
    with pytest.raises(TypeError) as err:
        cufinufft(1, (8, 8), foo="bar")
    assert "Invalid option 'foo'" in err.value.args[0]


def test_exec_raises_on_dtype(torch_dev):
    import torch
    from cufinufft_amd import cufinufft
    M, shape = 4096, (16, 16, 16)
    k = [G.to_dev(a, torch_dev) for a in _pts(M, 3, np.float32, 3)]
    c = G.strengths(M, np.float32, 3)
    c_dev = G.to_dev(c, torch_dev)
    c_wrong = G.to_dev(c.real.copy(), torch_dev)
    fk = torch.zeros(shape, dtype=torch.complex64, device=torch_dev)
    fk_wrong = torch.zeros(shape, dtype=torch.complex128, device=torch_dev)
    plan = cufinufft(1, shape, eps=1e-3, dtype=np.float32)
    plan.set_pts(*k)
    with pytest.raises(TypeError):
        plan.execute(c_dev, fk_wrong)
    with pytest.raises(TypeError):
        plan.execute(c_wrong, fk)


# ------------------------------------------------------------ basic checks
def _direct1(c, k, shape, ind):
    """One mode of the type-1 sum in numpy axis order (as tests/utils.py:56-63
    of the reference: axis d of the mode array pairs with coordinate k[d])."""
    idx = np.unravel_index(ind, shape)
    ph = sum((idx[d] - shape[d] // 2) * k[d].astype(np.float64) for d in range(len(shape)))
    return np.sum(c.astype(np.complex128) * np.exp(1j * ph))


def _direct2(fk, kj):
    grids = np.meshgrid(*[np.arange(-(n // 2), (n + 1) // 2) for n in fk.shape], indexing="ij")
    ph = sum(g * float(x) for g, x in zip(grids, kj))
    return np.sum(fk.astype(np.complex128) * np.exp(-1j * ph))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_type1_basic(torch_dev, dtype):
    import torch
    from cufinufft_amd import cufinufft
    M, shape = 4096, (16, 16, 16)
    k = _pts(M, 3, dtype, 4)
    c = G.strengths(M, dtype, 4)
    plan = cufinufft(1, shape, eps=1e-3, dtype=dtype)
    plan.set_pts(*[G.to_dev(a, torch_dev) for a in k])
    fk = torch.zeros(shape, dtype=torch.complex64 if dtype == np.float32 else torch.complex128,
                     device=torch_dev)
    plan.execute(G.to_dev(c, torch_dev), fk)
    ind = int(0.1789 * np.prod(shape))
    est = fk.cpu().numpy().ravel()[ind]
    tgt = _direct1(c, k, shape, ind)
    assert abs(tgt - est) / abs(tgt) < 0.01


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_type2_basic(torch_dev, dtype):
    import torch
    from cufinufft_amd import cufinufft
    M, shape = 4096, (16, 16, 16)
    k = _pts(M, 3, dtype, 5)
    fk = G.modes(shape, dtype, 5)
    plan = cufinufft(2, shape, eps=1e-3, dtype=dtype)
    plan.set_pts(*[G.to_dev(a, torch_dev) for a in k])
    c = torch.zeros(M, dtype=torch.complex64 if dtype == np.float32 else torch.complex128,
                    device=torch_dev)
    plan.execute(c, G.to_dev(fk, torch_dev))
    ind = M // 2
    tgt = _direct2(fk, k[:, ind])
    assert abs(tgt - c.cpu().numpy()[ind]) / abs(tgt) < 0.01


def test_opts_basic(torch_dev):
    import torch
    from cufinufft_amd import cufinufft
    M, shape = 32, (8, 8, 8)
    k = _pts(M, 3, np.float32, 6)
    c = G.strengths(M, np.float32, 6)
    plan = cufinufft(1, shape, eps=1e-3, dtype=np.float32, gpu_sort=False,
                     gpu_maxsubprobsize=10)
    plan.set_pts(*[G.to_dev(a, torch_dev) for a in k])
    fk = torch.zeros(shape, dtype=torch.complex64, device=torch_dev)
    plan.execute(G.to_dev(c, torch_dev), fk)
    ind = int(0.1789 * np.prod(shape))
    tgt = _direct1(c, k, shape, ind)
    assert abs(tgt - fk.cpu().numpy().ravel()[ind]) / abs(tgt) < 0.01


# ---------------------------------------------------------------- examples
def test_example2d1many(torch_dev):
    """examples/example2d1many.py: N = 59 x 61, M = 100, two transforms,
    eps 1e-6, FP32; relative error at mode (0.37 N1, 0.26 N2) < 10 eps."""
    import torch
    from cufinufft_amd import cufinufft
    N1, N2, M, ntr, eps = 59, 61, 100, 2, 1e-6
    rng = np.random.default_rng(7)
    kx = rng.uniform(-np.pi, np.pi, M).astype(np.float32)
    ky = rng.uniform(-np.pi, np.pi, M).astype(np.float32)
    c = (rng.standard_normal((ntr, M)) + 1j * rng.standard_normal((ntr, M))).astype(np.complex64)
    fk_dev = torch.zeros((ntr, N1, N2), dtype=torch.complex64, device=torch_dev)
    plan = cufinufft(1, (N1, N2), ntr, eps=eps, dtype=np.float32)
    plan.set_pts(G.to_dev(kx, torch_dev), G.to_dev(ky, torch_dev))
    plan.execute(G.to_dev(c, torch_dev), fk_dev)
    fk = fk_dev.cpu().numpy()
    nt1, nt2 = int(0.37 * N1), int(0.26 * N2)
    x, y = nt1 - N1 // 2, nt2 - N2 // 2
    for i in range(ntr):
        true = np.sum(c[i].astype(np.complex128) *
                      np.exp(1j * (x * kx.astype(np.float64) + y * ky.astype(np.float64))))
        rel = abs(fk[i, nt1, nt2] - true) / np.max(np.abs(fk[i]))
        assert rel < 10 * eps, (i, rel)


def test_example2d2many(torch_dev):
    """examples/example2d2many.py: N = 37 x 41, M = 17, two transforms,
    eps 1e-6, FP32; relative error at point M // 2 < 10 eps."""
    import torch
    from cufinufft_amd import cufinufft
    N1, N2, M, ntr, eps = 37, 41, 17, 2, 1e-6
    rng = np.random.default_rng(8)
    kx = rng.uniform(-np.pi, np.pi, M).astype(np.float32)
    ky = rng.uniform(-np.pi, np.pi, M).astype(np.float32)
    fk = (rng.standard_normal((ntr, N1, N2)) +
          1j * rng.standard_normal((ntr, N1, N2))).astype(np.complex64)
    c_dev = torch.zeros((ntr, M), dtype=torch.complex64, device=torch_dev)
    plan = cufinufft(2, (N1, N2), ntr, eps=eps, dtype=np.float32)
    plan.set_pts(G.to_dev(kx, torch_dev), G.to_dev(ky, torch_dev))
    plan.execute(c_dev, G.to_dev(fk, torch_dev))
    c = c_dev.cpu().numpy()
    jt = M // 2
    x, y = np.mgrid[-(N1 // 2):(N1 + 1) // 2, -(N2 // 2):(N2 + 1) // 2]
    for i in range(ntr):
        true = np.sum(fk[i].astype(np.complex128) *
                      np.exp(-1j * (x * float(kx[jt]) + y * float(ky[jt]))))
        rel = abs(c[i, jt] - true) / np.max(np.abs(c[i]))
        assert rel < 10 * eps, (i, rel)


# ------------------------------------------------------------- C ABI path
@pytest.mark.parametrize("single", [True, False])
def test_makeplan_null_opts(torch_dev, single):
    """test/cufinufft2d2api_test.cu:72-110: makeplan with opts = NULL takes
    the reference defaults (type 2: GM-sort), then setpts / execute / destroy
    through the plain C entry points; checked against the direct sum."""
    import torch
    from cufinufft_amd import _cufinufft as L
    dtype = np.float32 if single else np.float64
    N1, N2, M, tol = 100, 50, 3000, (1e-5 if single else 1e-9)
    rng = np.random.default_rng(9)
    x = rng.uniform(-np.pi, np.pi, M).astype(dtype)
    y = rng.uniform(-np.pi, np.pi, M).astype(dtype)
    fk = G.modes((N2, N1), dtype, 9)
    dx, dy, dfk = (G.to_dev(a, torch_dev) for a in (x, y, fk))
    c = torch.zeros(M, dtype=dfk.dtype, device=torch_dev)
    torch.cuda.synchronize()
    plan = ctypes.c_void_p()
    modes = (ctypes.c_int * 3)(N1, N2, 1)
    mk, sp, ex, de = ((L._make_planf, L._set_ptsf, L._exec_planf, L._destroy_planf) if single
                      else (L._make_plan, L._set_pts, L._exec_plan, L._destroy_plan))
    assert mk(2, 2, modes, -1, 1, tol, 1, ctypes.byref(plan), None) == 0
    assert plan.value
    info = (ctypes.c_int * 16)()
    get_info = L.lib.cufinufftf_get_info if single else L.lib.cufinufft_get_info
    assert get_info(plan, info) == 0
    assert info[7] == 1  # the type-2 default method
    assert sp(M, ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()), None, 0,
              None, None, None, plan) == 0
    assert ex(ctypes.c_void_p(c.data_ptr()), ctypes.c_void_p(dfk.data_ptr()), plan) == 0
    assert de(plan) == 0
    torch.cuda.synchronize()
    from oracle import oracle as O
    pidx = np.arange(0, M, 7)
    ref = O.dirft_type2([x.astype(np.float64), y.astype(np.float64)], fk.ravel(), [N1, N2], -1,
                        pidx)
    got = c.cpu().numpy()[pidx]
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    assert err < 2 * tol
