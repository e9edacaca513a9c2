"""cufinufftf_setpts_noalloc: setpts whose sort scratch comes from the
library's fixed device array instead of an allocation (callers that cannot
allocate: graph capture, serving loops).  Called through the C ABI and
checked against cufinufftf_setpts on the same points, on two calls in a row
(the second with fewer points), each followed by the ordinary execute:

* the bin sort's outputs (per-bin counts, bin starts) are bit-exact with
  setpts' and the sorted index holds the same point set per bin;
* the type-1 result of the no-allocation plan matches the ordinary plan's to
  1e-6 relative l2 (same sort, so only the global float atomics of the
  flush may add in another order);
* sizes beyond what is allocated return code 13 and leave the plan usable:
  more points than the plan's point arrays hold (no earlier setpts sized
  them), and a point count whose sort scratch exceeds the 64 MiB array
  (1e7 points: 8 bytes of keys each)."""
import ctypes

import numpy as np
import pytest

from tests import gpu_util as G
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

N = (40, 36, 32)
TOL = 1e-5
OPTS = dict(gpu_method=2, gpu_binsizex=8, gpu_binsizey=8, gpu_binsizez=8,
            gpu_maxsubprobsize=1024)


def _plan(lib):
    from cufinufft_amd import _cufinufft as C
    opts = C.NufftOpts()
    assert lib.cufinufftf_default_opts(1, 3, ctypes.byref(opts)) == 0
    for k, v in OPTS.items():
        setattr(opts, k, v)
    plan = ctypes.c_void_p()
    nm = (ctypes.c_int * 3)(*N)
    assert lib.cufinufftf_makeplan(1, 3, nm, 1, 1, ctypes.c_float(TOL), 1, ctypes.byref(plan),
                                   ctypes.byref(opts)) == 0
    return plan


def _sort(lib, plan, M, nbins):
    bs = np.zeros(nbins, np.int32)
    st = np.zeros(nbins, np.int32)
    idx = np.zeros(M, np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    assert lib.cufinufftf_get_sort(plan, p(bs), p(st), p(idx), None, None) == 0
    return bs, st, idx


def _nbins(lib, plan):
    """Bins of the plan: get_info's entries 4-6 are the bin counts per axis."""
    info = (ctypes.c_int * 16)()
    assert lib.cufinufftf_get_info(plan, info) == 0
    return int(info[4]) * int(info[5]) * int(info[6])


def test_setpts_noalloc(torch_dev):
    import torch
    from cufinufft_amd import _cufinufft as C
    lib = C.lib
    f = lib.cufinufftf_setpts_noalloc
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    for fn in ("cufinufftf_setpts", "cufinufftf_execute", "cufinufftf_get_sort",
               "cufinufftf_get_info", "cufinufftf_destroy"):
        getattr(lib, fn).restype = ctypes.c_int
    lib.cufinufftf_setpts.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 7
    lib.cufinufftf_execute.argtypes = [ctypes.c_void_p] * 3
    lib.cufinufftf_get_sort.argtypes = [ctypes.c_void_p] * 6
    lib.cufinufftf_get_info.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.cufinufftf_destroy.argtypes = [ctypes.c_void_p]

    ref, na, fresh = _plan(lib), _plan(lib), _plan(lib)
    Mmax = 300_000
    # size the no-allocation plan's point arrays once with the ordinary setpts
    big = [G.to_dev(p, torch_dev) for p in G.uniform_pts(Mmax, 3, np.float32, 80)]
    assert lib.cufinufftf_setpts(Mmax, *[b.data_ptr() for b in big], 0, None, None, None,
                                 na) == 0
    torch.cuda.synchronize()
    for call, M in enumerate((250_000, 90_000)):
        print(f"no-alloc setpts call {call}: M={M}", flush=True)
        pts = G.clustered_pts(M, 3, np.float32, 81 + call, ncl=4, sigma=0.3)
        d = [G.to_dev(p, torch_dev) for p in pts]
        c = G.to_dev(G.strengths(M, np.float32, 81 + call), torch_dev)
        assert lib.cufinufftf_setpts(M, *[x.data_ptr() for x in d], 0, None, None, None,
                                     ref) == 0
        assert f(M, *[x.data_ptr() for x in d], na) == 0
        outs = []
        for plan in (ref, na):
            fk = torch.zeros(N[::-1], dtype=torch.complex64, device=torch_dev)
            assert lib.cufinufftf_execute(c.data_ptr(), fk.data_ptr(), plan) == 0
            torch.cuda.synchronize()
            outs.append(fk.cpu().numpy())
        nbins = _nbins(lib, na)
        assert nbins == _nbins(lib, ref) > 0
        s_ref = _sort(lib, ref, M, nbins)
        s_na = _sort(lib, na, M, nbins)
        np.testing.assert_array_equal(s_na[0], s_ref[0])
        np.testing.assert_array_equal(s_na[1], s_ref[1])
        for b in np.flatnonzero(s_ref[0])[:2000]:
            a0, n0 = s_ref[1][b], s_ref[0][b]
            assert set(s_ref[2][a0:a0 + n0]) == set(s_na[2][a0:a0 + n0])
        err = rel_l2(outs[1], outs[0])
        print(f"   no-alloc vs setpts {err:.3e}", flush=True)
        assert err <= 1e-6
    # beyond the plan's point capacity: a plan never sized by setpts
    assert f(1000, *[b.data_ptr() for b in big], fresh) == 13
    # beyond the scratch: more points than 64 MiB of sort buffers hold (the
    # plan itself holds them, so only the scratch bound applies)
    Mhuge = 10_000_000
    gen = torch.Generator(device=torch_dev)
    gen.manual_seed(82)
    huge = [(torch.rand(Mhuge, generator=gen, device=torch_dev) - 0.5) * (2 * np.pi)
            for _ in range(3)]
    assert lib.cufinufftf_setpts(Mhuge, *[h.data_ptr() for h in huge], 0, None, None, None,
                                 fresh) == 0
    assert f(Mhuge, *[h.data_ptr() for h in huge], fresh) == 13
    # the refused call left the plan's earlier points in place
    c = torch.zeros(Mhuge, dtype=torch.complex64, device=torch_dev)
    fk = torch.zeros(N[::-1], dtype=torch.complex64, device=torch_dev)
    assert lib.cufinufftf_execute(c.data_ptr(), fk.data_ptr(), fresh) == 0
    torch.cuda.synchronize()
    for plan in (ref, na, fresh):
        assert lib.cufinufftf_destroy(plan) == 0
