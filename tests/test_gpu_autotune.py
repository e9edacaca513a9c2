"""The autotuner end to end on small shapes: it times the candidate option
sets (methods 1, 2 and 4, bin and subproblem sizes, then the polynomial
kernel and sigma = 1.25 on the re-timed winner) on the caller's points and
returns options that plan; the tuned plan's output must meet the same bar as
the other full transforms against the direct sum, and a second call with the
same shape is served from the cache."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import gpu_util as G
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nufft_type", [1, 2])
def test_tune_then_transform(torch_dev, nufft_type):
    from cufinufft_amd import autotune
    # nf = 128 x 128: the 2D method-4 block candidates tile it
    dim, dtype, nm, M, tol = 2, np.float32, (64, 64), 20000, 1e-5
    print(f"autotune type {nufft_type}", flush=True)
    pts = G.clustered_pts(M, dim, dtype, seed=31)
    dpts = [G.to_dev(p, torch_dev) for p in pts[::-1]]
    best, table = autotune.tune(nufft_type, tuple(nm[::-1]), dpts, tol, dtype=dtype,
                                use_cache=True)
    assert table and all(t > 0 for _, t in table)
    methods = {o.get("gpu_method") for o, _ in table}
    assert {1, 2}.issubset(methods)
    if nufft_type == 1:
        assert 4 in methods
    # the winner came out of the re-timed stage, so it was planned and timed
    assert any(o == best for o, _ in table)
    again, table2 = autotune.tune(nufft_type, tuple(nm[::-1]), dpts, tol, dtype=dtype,
                                  use_cache=True)
    assert again == best and table2 == []

    P = O.Plan(nufft_type, list(nm), tol, single=True)
    if nufft_type == 1:
        c = G.strengths(M, dtype, seed=31)
        fk, _ = G.run_type1(torch_dev, pts, c, nm, tol, dtype, **best)
        kidx = np.random.default_rng(3).choice(int(np.prod(nm)), 1000, replace=False)
        direct = O.dirft_type1([p.astype(np.float64) for p in pts], c, list(nm), +1, kidx)
        oerr = rel_l2(P.type1(pts, c).ravel()[kidx], direct)
        err = rel_l2(fk.ravel()[kidx], direct)
    else:
        fk = G.modes(tuple(nm[::-1]), dtype, seed=31)
        cc, _ = G.run_type2(torch_dev, pts, fk, tol, dtype, **best)
        pidx = np.random.default_rng(3).choice(M, 1000, replace=False)
        direct = O.dirft_type2([p.astype(np.float64) for p in pts], fk, list(nm), -1, pidx)
        oerr = rel_l2(P.type2(pts, fk)[pidx], direct)
        err = rel_l2(cc[pidx], direct)
    print(f"   tuned {best}: gpu err {err:.3e} oracle err {oerr:.3e}", flush=True)
    assert err <= max(2.0 * tol, 1.25 * oerr)


def test_tune_group(torch_dev):
    """Group tuning: every candidate option set is planned for all members
    and the group timed as one launch; the chosen set is one of those timed,
    and the group run with it matches its members' own executes."""
    import torch
    from cufinufft_amd import autotune, cufinufft
    from cufinufft_amd.cufinufft import cufinufft_group
    print("autotune plan group", flush=True)
    tol = 1e-5
    shapes = [((48, 40), 6000), ((64, 64), 15000), ((40, 56), 3000)]
    data = []
    for i, (nm, M) in enumerate(shapes):
        pts = G.uniform_pts(M, 2, np.float32, seed=40 + i)
        c = G.to_dev(G.strengths(M, np.float32, seed=40 + i), torch_dev)
        data.append((nm, [G.to_dev(p, torch_dev) for p in pts[::-1]], c))

    def make(opts):
        plans, outs = [], []
        for nm, dpts, _ in data:
            p = cufinufft(1, nm[::-1], eps=tol, dtype=np.float32, **opts)
            p.set_pts(*dpts)
            plans.append(p)
            outs.append(torch.zeros(nm[::-1], dtype=torch.complex64, device=torch_dev))
        return plans, outs

    def build(opts):
        plans, outs = make(opts)
        g = cufinufft_group(plans, [d[2] for d in data], outs)
        return lambda: g.execute(True)

    cands = autotune.group_candidates(2)[:4]
    best, table = autotune.tune_group(build, cands, reps=2)
    assert best in [o for o, _ in table] and len(table) == len(cands)
    plans, outs = make(best)
    grp = cufinufft_group(plans, [d[2] for d in data], outs)
    grp.execute(True)
    torch.cuda.synchronize()
    for p, (nm, _, c), o in zip(plans, data, outs):
        ref = torch.zeros_like(o)
        p.execute(c, ref)
        torch.cuda.synchronize()
        assert rel_l2(o.cpu().numpy(), ref.cpu().numpy()) < 1e-6
