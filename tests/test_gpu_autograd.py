"""Differentiable form: gradients of NUFFT losses with respect to strengths /
modes and point coordinates.

Checks (radial-spoke points, loss = ||out - target||^2):
* against the FP64 direct-sum gradient: the same loss written as a dense FP64
  torch model (A = exp(s i k.x)) and differentiated by torch;
* against central finite differences with step 1e-3 on point coordinates.
"""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def spokes(nspk, nsamp, jitter, seed, dtype):
    rng = np.random.default_rng(seed)
    phi = (1 + 5 ** 0.5) / 2
    ang = np.mod(np.arange(nspk) * np.pi / phi, np.pi)
    rad = -np.pi + 2 * np.pi * np.arange(nsamp) / nsamp
    x = np.outer(np.cos(ang), rad).ravel() + jitter * rng.standard_normal(nspk * nsamp)
    y = np.outer(np.sin(ang), rad).ravel() + jitter * rng.standard_normal(nspk * nsamp)
    x = np.mod(x + np.pi, 2 * np.pi) - np.pi
    y = np.mod(y + np.pi, 2 * np.pi) - np.pi
    return np.stack([x, y]).astype(dtype)


def dense(points, n_modes, isign):
    """A[j, k] = exp(isign i k.x_j), k flattened x fastest (FP64)."""
    import torch
    ks = [torch.arange(-(n // 2), (n + 1) // 2, dtype=torch.float64, device=points.device)
          for n in n_modes]
    grids = torch.meshgrid(*ks[::-1], indexing="ij")  # (.., ky, kx)
    kf = [g.reshape(-1) for g in grids[::-1]]  # kx, ky, (kz) flattened x fastest
    ph = sum(points[d].double()[:, None] * kf[d][None, :] for d in range(len(n_modes)))
    return torch.exp(1j * isign * ph)


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-6), (np.float64, 1e-10)])
def test_type2_grads_vs_dense(torch_dev, dtype, tol):
    import torch
    from cufinufft_amd.torch_nufft import nufft_type2
    print(f"autograd type2 {np.dtype(dtype).name}", flush=True)
    N = (40, 36)
    P = spokes(48, 80, 1e-3, 5, dtype)
    cd = torch.complex64 if dtype == np.float32 else torch.complex128
    rng = np.random.default_rng(1)
    fk0 = (rng.standard_normal(N[::-1]) + 1j * rng.standard_normal(N[::-1]))
    tgt = rng.standard_normal(P.shape[1]) + 1j * rng.standard_normal(P.shape[1])
    pts = torch.tensor(P, device=torch_dev, requires_grad=True)
    fk = torch.tensor(fk0, dtype=cd, device=torch_dev, requires_grad=True)
    target = torch.tensor(tgt, dtype=cd, device=torch_dev)
    c = nufft_type2(pts, fk, isign=-1, eps=tol)
    loss = (c - target).abs().pow(2).sum()
    loss.backward()
    # FP64 dense model of the same loss
    p64 = torch.tensor(P.astype(np.float64), device=torch_dev, requires_grad=True)
    f64 = torch.tensor(fk0, dtype=torch.complex128, device=torch_dev, requires_grad=True)
    c64 = dense(p64, N, -1) @ f64.reshape(-1)
    l64 = (c64 - target.to(torch.complex128)).abs().pow(2).sum()
    l64.backward()
    assert rel_l2(c.detach().cpu().numpy(), c64.detach().cpu().numpy()) < 3 * max(tol, 1e-7)
    e_f = rel_l2(fk.grad.cpu().numpy(), f64.grad.cpu().numpy())
    e_x = rel_l2(pts.grad.cpu().numpy(), p64.grad.cpu().numpy())
    print(f"   grad_fk err {e_f:.2e}  grad_x err {e_x:.2e}", flush=True)
    assert e_f < 10 * max(tol, 1e-7)
    assert e_x < 30 * max(tol, 1e-7)


@pytest.mark.parametrize("dim,dtype,tol", [(2, np.float64, 1e-10), (3, np.float64, 1e-9),
                                           (2, np.float32, 1e-6)])
def test_type1_grads_vs_dense(torch_dev, dim, dtype, tol):
    import torch
    from cufinufft_amd.torch_nufft import nufft_type1
    print(f"autograd type1 {dim}d {np.dtype(dtype).name}", flush=True)
    N = (24, 20, 16)[:dim]
    rng = np.random.default_rng(2)
    M = 1500
    P = rng.uniform(-np.pi, np.pi, (dim, M)).astype(dtype)
    cd = torch.complex64 if dtype == np.float32 else torch.complex128
    c0 = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    tgt = rng.standard_normal(N[::-1]) + 1j * rng.standard_normal(N[::-1])
    pts = torch.tensor(P, device=torch_dev, requires_grad=True)
    c = torch.tensor(c0, dtype=cd, device=torch_dev, requires_grad=True)
    target = torch.tensor(tgt, dtype=cd, device=torch_dev)
    fk = nufft_type1(pts, c, N, isign=1, eps=tol)
    loss = (fk - target).abs().pow(2).sum()
    loss.backward()
    p64 = torch.tensor(P.astype(np.float64), device=torch_dev, requires_grad=True)
    c64 = torch.tensor(c0, dtype=torch.complex128, device=torch_dev, requires_grad=True)
    f64 = dense(p64, N, 1).T @ c64
    l64 = (f64 - target.to(torch.complex128).reshape(-1)).abs().pow(2).sum()
    l64.backward()
    e_c = rel_l2(c.grad.cpu().numpy(), c64.grad.cpu().numpy())
    e_x = rel_l2(pts.grad.cpu().numpy(), p64.grad.cpu().numpy())
    print(f"   grad_c err {e_c:.2e}  grad_x err {e_x:.2e}", flush=True)
    assert e_c < 10 * max(tol, 1e-7)
    assert e_x < 30 * max(tol, 1e-7)


def test_type2_points_grad_finite_difference(torch_dev):
    """Central differences, step h = 1e-3, on coordinates of a FP64 plan.

    The loss holds frequencies up to 2 N_max = 32, so the O(h^2) truncation
    error of a step-1e-3 central difference is ~1e-4 relative by itself.  The
    gradient must agree with fd(h) within fd's own error estimate
    (4/3 |fd(h) - fd(h/2)|) and with the Richardson extrapolation
    (4 fd(h/2) - fd(h)) / 3, whose error is O(h^4), to 1e-6 relative."""
    import torch
    from cufinufft_amd.torch_nufft import nufft_type2
    print("autograd finite differences", flush=True)
    N = (32, 32)
    P = spokes(32, 64, 1e-3, 7, np.float64)
    rng = np.random.default_rng(3)
    fk = torch.tensor(rng.standard_normal(N[::-1]) + 1j * rng.standard_normal(N[::-1]),
                      device=torch_dev)
    target = torch.tensor(rng.standard_normal(P.shape[1]) + 1j * rng.standard_normal(P.shape[1]),
                          device=torch_dev)

    def loss_of(p):
        c = nufft_type2(p, fk, isign=-1, eps=1e-12)
        return (c - target).abs().pow(2).sum()

    pts = torch.tensor(P, device=torch_dev, requires_grad=True)
    loss_of(pts).backward()
    g = pts.grad.cpu().numpy()
    def fd(d, j, h):
        pp = P.copy()
        pm = P.copy()
        pp[d, j] += h
        pm[d, j] -= h
        with torch.no_grad():
            lp = loss_of(torch.tensor(pp, device=torch_dev)).item()
            lm = loss_of(torch.tensor(pm, device=torch_dev)).item()
        return (lp - lm) / (2 * h)

    h = 1e-3
    for (d, j) in [(0, 5), (1, 100), (0, 777), (1, 1500), (0, 2047)]:
        f1, f2 = fd(d, j, h), fd(d, j, h / 2)
        scale = max(1.0, abs(g[d, j]))
        assert abs(f1 - g[d, j]) <= 4 / 3 * abs(f1 - f2) * 1.5 + 1e-6 * scale, (d, j, f1, g[d, j])
        rich = (4 * f2 - f1) / 3
        assert abs(rich - g[d, j]) <= 1e-6 * scale, (d, j, rich, g[d, j])


def test_points_cache_follows_inplace_writes(torch_dev):
    """A plan re-sorts when the points tensor is written in place (autograd
    version counter) and when a different tensor is passed; the forward on
    updated points matches the dense FP64 model."""
    import torch
    from cufinufft_amd.torch_nufft import nufft_type2
    n_modes = (24, 20)
    pts = torch.tensor(spokes(16, 32, 1e-3, 3, np.float64), device=torch_dev)
    rng = np.random.default_rng(4)
    fk = torch.tensor(rng.standard_normal(n_modes[::-1]) + 1j * rng.standard_normal(n_modes[::-1]),
                      device=torch_dev)
    c1 = nufft_type2(pts, fk, isign=-1, eps=1e-10)
    with torch.no_grad():
        pts.mul_(0.9)  # in place: same storage, new version
    c2 = nufft_type2(pts, fk, isign=-1, eps=1e-10)
    ref2 = dense(pts, n_modes, -1) @ fk.reshape(-1)
    assert rel_l2(c2.cpu().numpy(), ref2.cpu().numpy()) < 1e-9
    assert rel_l2(c1.cpu().numpy(), c2.cpu().numpy()) > 1e-3
    c3 = nufft_type2(pts.clone() * 1.05, fk, isign=-1, eps=1e-10)  # a different tensor
    ref3 = dense(pts * 1.05, n_modes, -1) @ fk.reshape(-1)
    assert rel_l2(c3.cpu().numpy(), ref3.cpu().numpy()) < 1e-9
