"""Shared helpers of the GPU parity tests: seeded inputs of the reference's
test distributions, device transfers, and one-call wrappers around the
library's Python class."""
import numpy as np


def cdtype(dtype):
    return np.complex64 if np.dtype(dtype) == np.float32 else np.complex128


def uniform_pts(M, dim, dtype, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-np.pi, np.pi, M).astype(dtype) for _ in range(dim)]


def clustered_pts(M, dim, dtype, seed, ncl=8, sigma=0.3):
    """Mixture of isotropic Gaussian clusters wrapped into [-pi, pi)."""
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-np.pi, np.pi, (ncl, dim))
    lab = rng.integers(0, ncl, M)
    pts = cen[lab] + sigma * rng.standard_normal((M, dim))
    pts = np.mod(pts + np.pi, 2 * np.pi) - np.pi
    return [np.ascontiguousarray(pts[:, d]).astype(dtype) for d in range(dim)]


def strengths(M, dtype, seed, ntr=1):
    rng = np.random.default_rng(seed + 1000)
    n = M * ntr
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(cdtype(dtype))


def modes(shape, dtype, seed, ntr=1):
    """Mode array in numpy order (..., nY, nX), x fastest."""
    rng = np.random.default_rng(seed + 2000)
    sh = ((ntr,) if ntr > 1 else ()) + tuple(shape)
    return (rng.standard_normal(sh) + 1j * rng.standard_normal(sh)).astype(cdtype(dtype))


def to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def run_type1(dev, pts, c, shape_xyz, tol, dtype, ntr=1, **opts):
    """Type 1 through the library.  pts are (x, y[, z]); shape_xyz = (N1, N2[, N3]).
    Returns fk as numpy with shape (ntr,) + (N3,) N2, N1 and the plan."""
    import torch
    from cufinufft_amd import cufinufft
    shape_np = tuple(shape_xyz[::-1])
    plan = cufinufft(1, shape_np, n_trans=ntr, eps=tol, dtype=dtype, **opts)
    # the class reverses axes like the reference, so pass numpy-order arrays
    dpts = [to_dev(p, dev) for p in pts[::-1]]
    plan.set_pts(*dpts)
    dc = to_dev(c, dev)
    fk = torch.zeros(((ntr,) if ntr > 1 else ()) + shape_np, dtype=dc.dtype, device=dev)
    plan.execute(dc, fk)
    torch.cuda.synchronize()
    return fk.cpu().numpy(), plan


def run_type2(dev, pts, fk, tol, dtype, ntr=1, **opts):
    import torch
    from cufinufft_amd import cufinufft
    shape_np = fk.shape[1:] if ntr > 1 else fk.shape
    plan = cufinufft(2, tuple(shape_np), n_trans=ntr, eps=tol, dtype=dtype, **opts)
    dpts = [to_dev(p, dev) for p in pts[::-1]]
    plan.set_pts(*dpts)
    dfk = to_dev(fk, dev)
    M = pts[0].size
    c = torch.zeros(M * ntr, dtype=dfk.dtype, device=dev)
    plan.execute(c, dfk)
    torch.cuda.synchronize()
    return c.cpu().numpy(), plan
