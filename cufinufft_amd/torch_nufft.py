"""Differentiable NUFFTs for PyTorch-ROCm (torch.autograd.Function).

    fk = nufft_type1(points, c, n_modes, isign=+1, eps=1e-6)
    c  = nufft_type2(points, fk, isign=-1, eps=1e-6)

points: real tensor of shape (dim, M) on the GPU, rows x, y[, z] (library axis
order), in [-pi, pi); float32 or float64 selects the plan precision.
c: complex tensor of shape (M,) or (ntransf, M); fk: complex tensor of shape
(N3, N2, N1) / (N2, N1) / (N1,) (x fastest), with a leading ntransf axis when
batched.  n_modes is (N1, N2[, N3]).

Backward, with the adjoint identities of the two transforms
(f_k = sum_j c_j e^{s i k.x_j} for type 1, c_j = sum_k f_k e^{s i k.x_j} for
type 2, s = isign):

* type 2, strengths: grad_fk = type-1 transform with sign -s of grad_c;
* type 2, points:    grad_x_d[j] = Re(conj(grad_c[j]) * dc_j/dx_d), with dc
  from execute_grad of the forward plan: dc_j/dx_d is the type-2 transform
  of the modes weighted by i s k_d (weight applied in the amplify pass), so
  the gradient carries the transform's own tolerance;
* type 1, strengths: grad_c = h = type-2 transform with sign -s of grad_fk;
* type 1, points:    grad_x_d[j] = Re(c_j * conj(dh_j/dx_d)), with h and dh
  from one execute_grad of the sign -s type-2 plan.

Gradients sum over the ntransf transforms that share the points.  Plans are
cached per (type, modes, precision, eps, sign, ntransf); set_pts runs when the
points differ from the ones the plan last sorted (a forward and its backward
on unchanged points sort once per plan).
"""
import numpy as np
import torch

from cufinufft_amd.cufinufft import cufinufft

_PLANS = {}


def _plan(nufft_type, n_modes, dtype, eps, isign, ntr):
    key = (nufft_type, tuple(n_modes), np.dtype(dtype).name, float(eps), int(isign), int(ntr))
    p = _PLANS.get(key)
    if p is None:
        p = cufinufft(nufft_type, tuple(n_modes)[::-1], n_trans=ntr, eps=eps, isign=isign,
                      dtype=dtype, maxbatchsize=0)
        _PLANS[key] = p
    return p


def _np_real(t):
    return np.float32 if t.dtype == torch.float32 else np.float64


def _set_pts(plan, points):
    """set_pts unless this plan already holds these points: same storage
    address, shape, strides and dtype, and the same autograd version counter
    (bumped by every in-place write, shared by views).  The plan keeps a
    reference to the tensor it last sorted, so that memory cannot be reused by
    another tensor while the key is cached.  (Writes through ``.data`` bypass
    the version counter and are not detected.)"""
    key = (points.data_ptr(), tuple(points.shape), tuple(points.stride()), points.dtype,
           points._version)
    held = getattr(plan, "_pts_held", None)
    if held is not None and held[1] == key:
        return
    # the class takes numpy axis order (..., y, x) and reverses it
    rows = [points[d].contiguous() for d in range(points.shape[0])]
    plan.set_pts(*rows[::-1])
    plan._pts_held = (points, key)


def _cdtype(rdtype):
    return torch.complex64 if rdtype == torch.float32 else torch.complex128


def _type2_grad(points, fk, isign, eps, ntr, n_modes):
    """Value and coordinate derivatives of a type-2 transform."""
    dim, M = points.shape
    plan = _plan(2, n_modes, _np_real(points), eps, isign, ntr)
    _set_pts(plan, points)
    c = torch.empty(ntr * M, dtype=fk.dtype, device=fk.device)
    dc = torch.empty(dim, ntr * M, dtype=fk.dtype, device=fk.device)
    plan.execute_grad(c, dc, fk.contiguous())
    return c, dc


class _Type1(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, c, n_modes, isign, eps):
        dim, M = points.shape
        ntr = c.numel() // M
        plan = _plan(1, n_modes, _np_real(points), eps, isign, ntr)
        _set_pts(plan, points)
        shape = ((ntr,) if ntr > 1 else ()) + tuple(n_modes)[::-1]
        fk = torch.empty(shape, dtype=c.dtype, device=c.device)
        plan.execute(c.contiguous(), fk)
        ctx.save_for_backward(points, c)
        ctx.cfg = (tuple(n_modes), isign, eps, ntr)
        return fk

    @staticmethod
    def backward(ctx, g):
        points, c = ctx.saved_tensors
        n_modes, isign, eps, ntr = ctx.cfg
        dim, M = points.shape
        need_x, need_c = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        grad_x = grad_c = None
        if need_x:
            h, dh = _type2_grad(points, g, -isign, eps, ntr, n_modes)
            cc = c.reshape(ntr, M)
            dh = dh.reshape(dim, ntr, M)
            grad_x = (cc.unsqueeze(0) * dh.conj()).real.sum(dim=1).to(points.dtype)
            if need_c:
                grad_c = h.reshape(c.shape)
        elif need_c:
            plan = _plan(2, n_modes, _np_real(points), eps, -isign, ntr)
            _set_pts(plan, points)
            h = torch.empty(ntr * M, dtype=c.dtype, device=c.device)
            plan.execute(h, g.contiguous())
            grad_c = h.reshape(c.shape)
        return grad_x, grad_c, None, None, None


class _Type2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, points, fk, isign, eps):
        dim, M = points.shape
        n_modes = tuple(fk.shape[-dim:])[::-1]
        ntr = fk.numel() // int(np.prod(n_modes))
        plan = _plan(2, n_modes, _np_real(points), eps, isign, ntr)
        _set_pts(plan, points)
        c = torch.empty(((ntr,) if ntr > 1 else ()) + (M,), dtype=fk.dtype, device=fk.device)
        plan.execute(c, fk.contiguous())
        ctx.save_for_backward(points, fk)
        ctx.cfg = (n_modes, isign, eps, ntr)
        return c

    @staticmethod
    def backward(ctx, g):
        points, fk = ctx.saved_tensors
        n_modes, isign, eps, ntr = ctx.cfg
        dim, M = points.shape
        grad_x = grad_fk = None
        g = g.contiguous()
        if ctx.needs_input_grad[0]:
            _, dc = _type2_grad(points, fk, isign, eps, ntr, n_modes)
            gg = g.reshape(ntr, M)
            dc = dc.reshape(dim, ntr, M)
            grad_x = (gg.conj().unsqueeze(0) * dc).real.sum(dim=1).to(points.dtype)
        if ctx.needs_input_grad[1]:
            plan = _plan(1, n_modes, _np_real(points), eps, -isign, ntr)
            _set_pts(plan, points)
            grad_fk = torch.empty_like(fk)
            plan.execute(g, grad_fk)
        return grad_x, grad_fk, None, None


def nufft_type1(points, c, n_modes, isign=1, eps=1e-6):
    """Type-1 NUFFT f_k = sum_j c_j exp(isign i k.x_j), differentiable in c and points."""
    if c.dtype != _cdtype(points.dtype):
        raise TypeError("c must be complex64 for float32 points, complex128 for float64")
    return _Type1.apply(points, c, tuple(n_modes), int(isign), float(eps))


def nufft_type2(points, fk, isign=-1, eps=1e-6):
    """Type-2 NUFFT c_j = sum_k f_k exp(isign i k.x_j), differentiable in fk and points."""
    if fk.dtype != _cdtype(points.dtype):
        raise TypeError("fk must be complex64 for float32 points, complex128 for float64")
    return _Type2.apply(points, fk, int(isign), float(eps))
