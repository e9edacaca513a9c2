"""Plan groups: many independent plans of uneven N and M, spread / interp in
one persistent launch, eager and hipGraph-replayed, against each member's own
execute and the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import gpu_util as G
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def members(dev, typ, dtype, tol, n, seed):
    import torch
    from cufinufft_amd import cufinufft
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        N = (int(np.exp(rng.uniform(np.log(24), np.log(96)))),
             int(np.exp(rng.uniform(np.log(24), np.log(96)))))
        M = int(np.exp(rng.uniform(np.log(200), np.log(20000))))
        ntr = 1 + (i % 3 == 0)
        pts = G.uniform_pts(M, 2, dtype, seed + i) if i % 2 else G.clustered_pts(M, 2, dtype, seed + i)
        plan = cufinufft(typ, N[::-1], n_trans=ntr, eps=tol, dtype=dtype, maxbatchsize=ntr,
                         gpu_method=2, gpu_binsizex=16, gpu_binsizey=16)
        plan.set_pts(*[G.to_dev(p, dev) for p in pts[::-1]])
        if typ == 1:
            c = G.to_dev(G.strengths(M, dtype, seed + i, ntr), dev)
            fk = torch.zeros(((ntr,) if ntr > 1 else ()) + N[::-1], dtype=c.dtype, device=dev)
        else:
            fk = G.to_dev(G.modes(N[::-1], dtype, seed + i, ntr), dev)
            c = torch.zeros(M * ntr, dtype=fk.dtype, device=dev)
        out.append(dict(plan=plan, c=c, fk=fk, N=N, M=M, ntr=ntr, pts=pts))
    return out


def individual(m, typ):
    import torch
    if typ == 1:
        ref = torch.zeros_like(m["fk"])
        m["plan"].execute(m["c"], ref)
    else:
        ref = torch.zeros_like(m["c"])
        m["plan"].execute(ref, m["fk"])
    torch.cuda.synchronize()
    return ref.cpu().numpy()


@pytest.mark.parametrize("typ,dtype,tol", [(1, np.float32, 1e-6), (2, np.float32, 1e-6),
                                           (1, np.float64, 1e-10), (2, np.float64, 1e-9)])
def test_group_matches_members(torch_dev, typ, dtype, tol):
    import torch
    from cufinufft_amd.cufinufft import cufinufft_group
    print(f"group type {typ} {np.dtype(dtype).name}", flush=True)
    ms = members(torch_dev, typ, dtype, tol, 9, 40 + typ)
    refs = [individual(m, typ) for m in ms]
    grp = cufinufft_group([m["plan"] for m in ms], [m["c"] for m in ms], [m["fk"] for m in ms])
    outs = [m["fk"] if typ == 1 else m["c"] for m in ms]
    for o in outs:
        o.zero_()
    grp.execute()
    torch.cuda.synchronize()
    bar = 1e-5 if dtype == np.float32 else 1e-12
    for m, o, r in zip(ms, outs, refs):
        assert rel_l2(o.cpu().numpy(), r) < bar
    # one member against the oracle as well
    m = ms[1]
    P = O.Plan(typ, list(m["N"]), tol, single=dtype == np.float32)
    if typ == 1:
        got = outs[1].cpu().numpy()
        exp = P.type1(m["pts"], m["c"].cpu().numpy()[:m["M"]])
        got = got[0] if m["ntr"] > 1 else got
    else:
        got = outs[1].cpu().numpy()[:m["M"]]
        f0 = m["fk"].cpu().numpy()
        exp = P.type2(m["pts"], f0[0] if m["ntr"] > 1 else f0)
    assert rel_l2(got, exp) < tol / 5 if dtype == np.float64 else rel_l2(got, exp) < 2e-6


@pytest.mark.parametrize("typ", [1, 2])
def test_group_graph_replays(torch_dev, typ):
    """Capture once, replay four times with eager work and new input values
    between replays; every replay must equal the eager result for its inputs."""
    import torch
    from cufinufft_amd.cufinufft import cufinufft_group
    print(f"group graph type {typ}", flush=True)
    dtype, tol = np.float32, 1e-6
    ms = members(torch_dev, typ, dtype, tol, 7, 60 + typ)
    grp = cufinufft_group([m["plan"] for m in ms], [m["c"] for m in ms], [m["fk"] for m in ms])
    ins = [m["c"] if typ == 1 else m["fk"] for m in ms]
    outs = [m["fk"] if typ == 1 else m["c"] for m in ms]
    base = [t.clone() for t in ins]
    for rep, scale in enumerate((1.0, 2.0, -0.5, 3.0)):
        for t, b in zip(ins, base):
            t.copy_(b * scale)             # new inputs, in place (the graph reads these buffers)
        eager = [individual(m, typ) for m in ms]
        for o in outs:
            o.fill_(float("nan"))          # stale values must not survive
        torch.cuda.synchronize()
        grp.execute(use_graph=True)
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert rel_l2(o.cpu().numpy(), e) < 1e-5, (rep, scale)
        _ = torch.randn(1000, device=torch_dev).sum().item()  # unrelated eager work


def test_plan_execute_under_torch_graph_capture(torch_dev):
    """A single plan's execute captured with torch.cuda.graph (the plan
    follows torch's capture stream) and replayed."""
    import torch
    from cufinufft_amd import cufinufft
    print("torch graph capture of execute", flush=True)
    dtype, tol, N, M = np.float32, 1e-5, (64, 48), 5000
    pts = G.uniform_pts(M, 2, dtype, 9)
    plan = cufinufft(1, N[::-1], eps=tol, dtype=dtype)
    plan.set_pts(*[G.to_dev(p, torch_dev) for p in pts[::-1]])
    c = G.to_dev(G.strengths(M, dtype, 9), torch_dev)
    fk = torch.zeros(N[::-1], dtype=c.dtype, device=torch_dev)
    plan.execute(c, fk)              # warm-up outside the capture
    torch.cuda.synchronize()
    ref = fk.clone()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            plan.execute(c, fk)
    torch.cuda.current_stream().wait_stream(s)
    for scale in (1.0, 2.0, 0.5):
        c.mul_(scale)
        ref = ref * scale
        fk.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert rel_l2(fk.cpu().numpy(), ref.cpu().numpy()) < 1e-5
        _ = torch.ones(10, device=torch_dev).sum().item()


def test_group_mixed_tolerances(torch_dev):
    """Members of one kernel width may differ in tol and maxsubprobsize: here
    tol 5e-5 with 1024-point subproblems next to tol 1e-5 with 512-point ones
    (both ns = 6).  The group's fixed-point tile must be sized for the
    strictest member's 32-bit bound, so the 5e-5 member's 537..1024-point
    jobs take 64-bit cells with room for them in the LDS."""
    import torch
    from cufinufft_amd import cufinufft
    from cufinufft_amd.cufinufft import cufinufft_group
    print("group mixed tolerances", flush=True)
    dtype, N = np.float32, (48, 40)
    ms = []
    for i, (tol, maxsub) in enumerate(((5e-5, 1024), (1e-5, 512), (5e-5, 1024))):
        M = 30_000
        pts = G.clustered_pts(M, 2, dtype, 70 + i, ncl=2, sigma=0.05)
        plan = cufinufft(1, N[::-1], eps=tol, dtype=dtype, maxbatchsize=1, gpu_method=2,
                         gpu_binsizex=16, gpu_binsizey=16, gpu_maxsubprobsize=maxsub)
        plan.set_pts(*[G.to_dev(p, torch_dev) for p in pts[::-1]])
        c = G.to_dev(G.strengths(M, dtype, 70 + i), torch_dev)
        fk = torch.zeros(N[::-1], dtype=c.dtype, device=torch_dev)
        ms.append(dict(plan=plan, c=c, fk=fk, pts=pts, tol=tol))
    refs = [individual(m, 1) for m in ms]
    grp = cufinufft_group([m["plan"] for m in ms], [m["c"] for m in ms], [m["fk"] for m in ms])
    for m in ms:
        m["fk"].zero_()
    grp.execute()
    torch.cuda.synchronize()
    for m, r in zip(ms, refs):
        got = m["fk"].cpu().numpy()
        assert rel_l2(got, r) < 1e-5
        P = O.Plan(1, list(N), m["tol"], single=True)
        assert rel_l2(got, P.type1(m["pts"], m["c"].cpu().numpy())) < m["tol"] / 5
