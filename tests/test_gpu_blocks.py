"""Method 4 (owner-computes blocks) against the oracle and the direct sum.

The reference's method 4 (block gather, spread3d_wrapper.cu:393-889) plays
the same role: blocks own their output cells and points are replicated into
every block their support touches.  Cases cover:
* a support wider than a block (ns = 10 > 8: three blocks per axis), a block
  spanning the whole axis (one block holds two periodic frames of a point);
* blocks with one subproblem (plain stores) and with many (small
  maxsubprobsize: atomic adds of the touched box);
* both fixed-point widths of FP32 plans and the FP64 tile; batched transforms.
Bars: spread alone against the oracle's FP64 sum as for the SM spread (FP32
8e-7 / FP64 1e-12, plus the least-squares scale check); full type 1 against the direct
sum with the bar of the other methods.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import gpu_util as G
from tests.conftest import rel_l2, scale_err

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _ob(ob):
    return dict(gpu_method=4, gpu_obinsizex=ob[0], gpu_obinsizey=ob[1], gpu_obinsizez=ob[2],
                gpu_binsizex=ob[0], gpu_binsizey=ob[1], gpu_binsizez=ob[2])


SI_CASES = [
    # dim, dtype, nmodes = grid (x, y, z), M, tol, blocks, maxsub, distribution
    (2, np.float32, (64, 48), 5000, 1e-5, (16, 16, 1), 1024, "uniform"),
    (2, np.float32, (64, 48), 5000, 1e-5, (32, 8, 1), 16, "clustered"),
    (3, np.float32, (24, 32, 16), 4000, 1e-6, (8, 8, 8), 32, "uniform"),     # 32-bit tile
    (3, np.float32, (24, 32, 16), 4000, 1e-6, (8, 16, 16), 4096, "clustered"),  # 64-bit tile
    (3, np.float64, (24, 24, 32), 3000, 1e-9, (8, 8, 8), 64, "clustered"),   # ns 10 > block
    (3, np.float64, (24, 24, 24), 3000, 1e-9, (24, 8, 8), 1024, "uniform"),  # whole-axis block
    (1, np.float64, (100,), 2000, 1e-8, (20, 1, 1), 1024, "uniform"),
    (1, np.float32, (128,), 3000, 1e-5, (128, 1, 1), 64, "clustered"),       # one block
]


@pytest.mark.parametrize("case", SI_CASES, ids=lambda c: f"{c[0]}d-{np.dtype(c[1]).name}-"
                         f"{'x'.join(map(str, c[5][:c[0]]))}-ms{c[6]}")
def test_spread_only(torch_dev, case):
    dim, dtype, nm, M, tol, ob, maxsub, dist = case
    _flush(f"blocks spread-only {case}")
    pts = (G.uniform_pts if dist == "uniform" else G.clustered_pts)(M, dim, dtype, 51)
    c = G.strengths(M, dtype, seed=51)
    fw, plan = G.run_type1(torch_dev, pts, c, nm, tol, dtype, gpu_spreadinterponly=1,
                           gpu_maxsubprobsize=maxsub, **_ob(ob))
    assert plan.info()["method"] == 4
    ref = O.spread(pts, c, list(nm), O.setup_spreader(tol, dtype == np.float32))
    single = dtype == np.float32
    assert rel_l2(fw, ref) < (8e-7 if single else 1e-12)
    assert scale_err(fw, ref) < (3e-7 if single else 1e-12)


NU_CASES = [
    (2, np.float32, (256, 256), 100_000, 1e-5, (64, 64, 1), 1024, "uniform"),
    # nf = 100 is no multiple of the 8-cell blocks: rounded up to 120 as the
    # reference's method 4 does (contrib/common.cpp:31-32)
    (3, np.float32, (50, 50, 50), 20_000, 1e-5, (8, 8, 8), 1024, "uniform"),
    (3, np.float32, (32, 32, 32), 50_000, 1e-6, (16, 16, 8), 32, "clustered"),
    (3, np.float32, (32, 32, 32), 50_000, 1e-5, (32, 16, 8), 4096, "clustered"),
    (3, np.float64, (32, 24, 28), 30_000, 1e-9, (16, 16, 12), 256, "clustered"),  # nf 64x48x60
]


@pytest.mark.parametrize("case", NU_CASES, ids=lambda c: f"{c[0]}d-{np.dtype(c[1]).name}-"
                         f"{c[4]:g}-ms{c[6]}")
def test_type1(torch_dev, case):
    dim, dtype, nm, M, tol, ob, maxsub, dist = case
    _flush(f"blocks type 1 {case}")
    pts = (G.uniform_pts if dist == "uniform" else G.clustered_pts)(M, dim, dtype, 52)
    c = G.strengths(M, dtype, seed=52)
    fk, plan = G.run_type1(torch_dev, pts, c, nm, tol, dtype, gpu_maxsubprobsize=maxsub,
                           **_ob(ob))
    kidx = np.random.default_rng(8).choice(int(np.prod(nm)), 800, replace=False)
    direct = O.dirft_type1([p.astype(np.float64) for p in pts], c, list(nm), +1, kidx)
    P = O.Plan(1, list(nm), tol, single=dtype == np.float32, blocks=ob[:dim])
    assert tuple(plan.info()["nf"][:dim]) == tuple(P.nf)
    ofk = P.type1(pts, c)
    oerr = rel_l2(ofk.ravel()[kidx], direct)
    err = rel_l2(fk.ravel()[kidx], direct)
    _flush(f"   gpu err {err:.3e} oracle err {oerr:.3e}")
    assert err <= max(2 * tol, 1.25 * oerr)
    # against the FP64 oracle pipeline: tol/5, but not below the ~5.5e-7 that
    # FP32 kernel values alone cost in 3D (CPU emulation of the float ops)
    assert rel_l2(fk, ofk) <= max(tol / 5, 1e-6 if dtype == np.float32 else 0.0)


def test_batched_and_resetpts(torch_dev):
    """ntransf = 3 through maxbatchsize = 2, then a second setpts with more
    points (the copy arrays grow)."""
    import torch
    from cufinufft_amd import cufinufft
    dtype, nm, tol, ntr = np.float32, (40, 48), 1e-5, 3
    _flush("blocks batched + re-setpts")
    plan = cufinufft(1, nm[::-1], n_trans=ntr, eps=tol, dtype=dtype, maxbatchsize=2,
                     **_ob((16, 16, 1)))
    P = O.Plan(1, list(nm), tol, single=True)
    for seed, M in ((1, 3000), (2, 9000)):
        pts = G.clustered_pts(M, 2, dtype, seed)
        c = G.strengths(M, dtype, seed=seed, ntr=ntr)
        plan.set_pts(*[G.to_dev(p, torch_dev) for p in pts[::-1]])
        fk = torch.zeros((ntr,) + nm[::-1], dtype=torch.complex64, device=torch_dev)
        plan.execute(G.to_dev(c, torch_dev), fk)
        out = fk.cpu().numpy()
        for t in range(ntr):
            assert rel_l2(out[t], P.type1(pts, c[t * M:(t + 1) * M])) <= tol / 5
