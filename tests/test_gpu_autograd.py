"""Differentiable form: gradients of NUFFT losses with respect to strengths /
modes and point coordinates.

Checks (radial-spoke points, loss = ||out - target||^2):
* against the FP64 direct-sum gradient: the same loss written as a dense FP64
  torch model (A = exp(s i k.x)) and differentiated by torch;
* against central finite differences with step 1e-3 on point coordinates;
* at the full differentiable-form size (N = 320^2, 320 golden-angle spokes x
  640 samples, tol 1e-6 FP32): the coordinate and mode gradients at sampled
  entries against FP64 direct sums of the same expressions.
Bars: 3 tol (relative l2) for values and gradients.
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
    assert e_f < 3 * max(tol, 1e-7)
    assert e_x < 3 * max(tol, 1e-7)


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
    assert e_c < 3 * max(tol, 1e-7)
    assert e_x < 3 * max(tol, 1e-7)


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


def test_type2_grads_full_size(torch_dev):
    """The differentiable-form workload: 2D type 2 FP32 at tol 1e-6, N = 320^2,
    320 spokes x 640 samples jittered by 1e-3, loss ||c - target||^2.  The
    coordinate gradient at 1000 sampled points and the mode gradient at 1000
    sampled modes are compared with FP64 direct sums:
    grad_x_d[j] = 2 Re(conj(r_j) sum_k (i s k_d) f_k e^{i s k.x_j}),
    grad_f[k] = 2 sum_j r_j e^{-i s k.x_j}, r = c - target (torch's complex
    gradient convention, conj-linear: d loss / d conj(z))."""
    import torch
    from oracle import oracle as O
    from cufinufft_amd.torch_nufft import nufft_type2
    print("autograd full size 320^2, 320x640 spokes", flush=True)
    N, tol = (320, 320), 1e-6
    P = spokes(320, 640, 1e-3, 5, np.float32)
    M = P.shape[1]
    rng = np.random.default_rng(5)
    fk0 = (rng.standard_normal(N[::-1]) + 1j * rng.standard_normal(N[::-1])).astype(np.complex64)
    tgt = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(np.complex64)
    pts = torch.tensor(P, device=torch_dev, requires_grad=True)
    fk = torch.tensor(fk0, device=torch_dev, requires_grad=True)
    target = torch.tensor(tgt, device=torch_dev)
    c = nufft_type2(pts, fk, isign=-1, eps=tol)
    loss = (c - target).abs().pow(2).sum()
    loss.backward()
    P64 = [P[0].astype(np.float64), P[1].astype(np.float64)]
    sel = np.random.default_rng(6).choice(M, 1000, replace=False)
    c_ref = O.dirft_type2(P64, fk0.ravel(), list(N), -1, sel)
    c_got = c.detach().cpu().numpy()[sel]
    assert rel_l2(c_got, c_ref) < 3 * tol
    # residual from the exact values at the sampled points
    r = c_ref - tgt[sel].astype(np.complex128)
    kx = np.arange(-(N[0] // 2), (N[0] + 1) // 2)
    ky = np.arange(-(N[1] // 2), (N[1] + 1) // 2)
    KX, KY = np.meshgrid(kx, ky)  # (N2, N1): x fastest
    gx = pts.grad.cpu().numpy()[:, sel]
    for d, K in enumerate((KX, KY)):
        w = (1j * -1 * K * fk0.astype(np.complex128)).ravel()
        dcd = O.dirft_type2(P64, w, list(N), -1, sel)
        ref = 2 * np.real(np.conj(r) * dcd)
        e = rel_l2(gx[d], ref)
        print(f"   grad_x[{d}] err {e:.2e}", flush=True)
        assert e < 3 * tol
    # mode gradient: 2 * type 1 (sign +1) of the full residual
    r_all = (c.detach().cpu().numpy().astype(np.complex128) - tgt.astype(np.complex128))
    kidx = np.random.default_rng(7).choice(N[0] * N[1], 1000, replace=False)
    gf_ref = 2 * O.dirft_type1(P64, r_all, list(N), +1, kidx)
    gf = fk.grad.cpu().numpy().ravel()[kidx]
    e = rel_l2(gf, gf_ref)
    print(f"   grad_fk err {e:.2e}", flush=True)
    assert e < 3 * tol
