"""The pruned 2D / 3D FFT against the one-plan FFT (full_fft=True, the
reference's cuFFT structure) on the same points and inputs.

Both run the same spread / interp; only the FFT stage differs (three
contiguous 1D passes over the grid lines that hold modes, with the
deconvolve / amplify folded into the transposes, against one 3D hipFFT plus
the separate deconvolve / amplify).  Their outputs agree to rounding: bars 1e-12 (FP64)
relative l2, and for FP32 1e-6 or, where larger, twice the full-FFT FP32
result's own distance from the FP64 pipeline on the same inputs (same kernel
width): both are FP32 FFTs of one grid, and at sigma 1.25 the deconvolution
multiplies FFT rounding by up to ~10 at the edge modes.  Accuracy against the oracle and direct sums is covered by the
2D and 3D cases of the parity tests, which run the pruned path (the
default).  Cases: odd and even mode counts, sigma 1.25, a batch split into a
full and a partial FFT batch (ntransf 3, maxbatchsize 2), execute_grad's
derivative modes, and 3D plan groups (2D groups: the group tests)."""
import numpy as np
import pytest

from tests import gpu_util as G
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = [
    # dtype, nmodes (x, y, z), M, tol, ntransf, extra options
    (np.float64, (24, 20, 17), 20_000, 1e-9, 3, dict(maxbatchsize=2)),
    (np.float32, (32, 30, 28), 40_000, 1e-5, 1, {}),
    (np.float64, (21, 16, 18), 10_000, 1e-6, 1, dict(upsampfac=1.25)),
    (np.float32, (40, 9, 33), 30_000, 1e-4, 2, dict(gpu_method=1, maxbatchsize=2)),
    (np.float32, (64, 48), 20_000, 1e-6, 3, dict(maxbatchsize=2)),
    (np.float64, (70, 33), 10_000, 1e-9, 1, {}),
    (np.float32, (45, 64), 10_000, 1e-5, 1, dict(upsampfac=1.25)),
]


def _bar(dtype, full=None, ref64=None):
    if dtype == np.float64:
        return 1e-12
    return max(1e-6, 2 * rel_l2(full, ref64)) if ref64 is not None else 1e-6


def _as64(a):
    return a.astype(np.complex128 if np.iscomplexobj(a) else np.float64)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{np.dtype(c[0]).name}-{c[1]}-t{c[4]}")
def test_type1_pruned_vs_full(torch_dev, case):
    dtype, nm, M, tol, ntr, opts = case
    print(f"pruned vs full type 1 {np.dtype(dtype).name} {nm} ntr {ntr} {opts}", flush=True)
    pts = G.clustered_pts(M, len(nm), dtype, seed=5)
    c = G.strengths(M, dtype, seed=5, ntr=ntr)
    a, pa = G.run_type1(torch_dev, pts, c, nm, tol, dtype, ntr=ntr, **opts)
    b, pb = G.run_type1(torch_dev, pts, c, nm, tol, dtype, ntr=ntr, full_fft=True, **opts)
    assert pa.info()["nf"] == pb.info()["nf"]
    ref64 = None
    if dtype == np.float32:
        ref64, p64 = G.run_type1(torch_dev, [_as64(p) for p in pts], _as64(c), nm, tol,
                                 np.float64, ntr=ntr, full_fft=True, **opts)
        assert p64.info()["ns"] == pb.info()["ns"]
    err = rel_l2(a, b)
    bar = _bar(dtype, b, ref64)
    print(f"   rel-l2 {err:.3e} (bar {bar:.3e})", flush=True)
    assert err <= bar


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{np.dtype(c[0]).name}-{c[1]}-t{c[4]}")
def test_type2_pruned_vs_full(torch_dev, case):
    dtype, nm, M, tol, ntr, opts = case
    print(f"pruned vs full type 2 {np.dtype(dtype).name} {nm} ntr {ntr} {opts}", flush=True)
    pts = G.uniform_pts(M, len(nm), dtype, seed=6)
    fk = G.modes(tuple(nm[::-1]), dtype, seed=6, ntr=ntr)
    a, _ = G.run_type2(torch_dev, pts, fk, tol, dtype, ntr=ntr, **opts)
    b, pb = G.run_type2(torch_dev, pts, fk, tol, dtype, ntr=ntr, full_fft=True, **opts)
    ref64 = None
    if dtype == np.float32:
        ref64, p64 = G.run_type2(torch_dev, [_as64(p) for p in pts], _as64(fk), tol, np.float64,
                                 ntr=ntr, full_fft=True, **opts)
        assert p64.info()["ns"] == pb.info()["ns"]
    err = rel_l2(a, b)
    bar = _bar(dtype, b, ref64)
    print(f"   rel-l2 {err:.3e} (bar {bar:.3e})", flush=True)
    assert err <= bar


@pytest.mark.parametrize("dtype,tol,nm", [(np.float64, 1e-9, (20, 18, 15)),
                                          (np.float32, 1e-5, (20, 18, 15)),
                                          (np.float32, 1e-6, (40, 36))])
def test_execute_grad_pruned_vs_full(torch_dev, dtype, tol, nm):
    """The derivative modes (i * isign * k_d weights in the first scatter)."""
    import torch
    from cufinufft_amd import cufinufft
    M, ntr, dim = 8000, 2, len(nm)
    print(f"pruned vs full execute_grad {np.dtype(dtype).name} {nm}", flush=True)
    pts = G.uniform_pts(M, dim, dtype, seed=9)
    fk = G.to_dev(G.modes(tuple(nm[::-1]), dtype, seed=9, ntr=ntr), torch_dev)
    out = []
    for full in (False, True):
        plan = cufinufft(2, nm[::-1], n_trans=ntr, eps=tol, dtype=dtype, full_fft=full)
        plan.set_pts(*[G.to_dev(p, torch_dev) for p in pts[::-1]])
        c = torch.zeros(M * ntr, dtype=fk.dtype, device=torch_dev)
        dc = torch.zeros((dim, M * ntr), dtype=fk.dtype, device=torch_dev)
        plan.execute_grad(c, dc, fk)
        torch.cuda.synchronize()
        out.append((c.cpu().numpy(), dc.cpu().numpy()))
    assert rel_l2(out[0][0], out[1][0]) <= _bar(dtype)
    for d in range(dim):
        err = rel_l2(out[0][1][d], out[1][1][d])
        print(f"   d/dx{d} rel-l2 {err:.3e}", flush=True)
        assert err <= _bar(dtype)


@pytest.mark.parametrize("nufft_type", [1, 2])
def test_group_3d_pruned_vs_full(torch_dev, nufft_type):
    """A 3D plan group (each member's FFT stage on its branch stream) against
    the members executed one by one with the full FFT."""
    import torch
    from cufinufft_amd import cufinufft
    from cufinufft_amd.cufinufft import cufinufft_group
    dtype, tol = np.float32, 1e-5
    shapes = [(24, 24, 24), (30, 20, 16), (17, 19, 21)]
    print(f"3d group type {nufft_type}", flush=True)
    plans, cs, fks, refs = [], [], [], []
    for i, nm in enumerate(shapes):
        M = 5000 + 1000 * i
        pts = G.clustered_pts(M, 3, dtype, seed=20 + i)
        dpts = [G.to_dev(p, torch_dev) for p in pts[::-1]]
        # group members run the SM method (type 2 defaults to GM)
        p = cufinufft(nufft_type, nm[::-1], eps=tol, dtype=dtype, gpu_method=2)
        p.set_pts(*dpts)
        q = cufinufft(nufft_type, nm[::-1], eps=tol, dtype=dtype, gpu_method=2, full_fft=True)
        q.set_pts(*dpts)
        if nufft_type == 1:
            c = G.to_dev(G.strengths(M, dtype, seed=20 + i), torch_dev)
            fk = torch.zeros(nm[::-1], dtype=c.dtype, device=torch_dev)
            ref = torch.zeros_like(fk)
            q.execute(c, ref)
        else:
            fk = G.to_dev(G.modes(tuple(nm[::-1]), dtype, seed=20 + i), torch_dev)
            c = torch.zeros(M, dtype=fk.dtype, device=torch_dev)
            ref = torch.zeros_like(c)
            q.execute(ref, fk)
        plans.append(p)
        cs.append(c)
        fks.append(fk)
        refs.append(ref)
    g = cufinufft_group(plans, cs, fks)
    for use_graph in (False, True, True):
        g.execute(use_graph)
        torch.cuda.synchronize()
        for i in range(len(plans)):
            got = (fks[i] if nufft_type == 1 else cs[i]).cpu().numpy()
            err = rel_l2(got, refs[i].cpu().numpy())
            assert err <= _bar(dtype), (i, use_graph, err)
