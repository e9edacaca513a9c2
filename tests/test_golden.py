"""Golden fixtures (tests/golden/*.npz): FP64 direct sums and exact bin-sort
arrays of seeded inputs, including period-edge and repeated points.

CPU: the oracle reproduces every fixture (bins bit-exact, transforms within
the algorithm's error).  GPU: the HIP library does the same through the C ABI.
"""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["dtype"] = np.dtype(str(d["dtype"]))
    d["type"] = int(d["type"])
    d["tol"] = float(d["tol"])
    d["maxsub"] = int(d["maxsub"])
    d["pts"] = [np.ascontiguousarray(p) for p in d["pts"]]
    return d


def ids(p):
    return os.path.basename(p)[:-4]


def oracle_result(d):
    nm = list(d["n_modes"])
    P = O.Plan(d["type"], nm, d["tol"], single=d["dtype"] == np.float32)
    if d["type"] == 1:
        return P.type1(d["pts"], d["input"])
    return P.type2(d["pts"], d["input"])


def test_fixtures_present():
    assert len(FIXTURES) >= 7


@pytest.mark.parametrize("path", FIXTURES, ids=ids)
def test_oracle_reproduces_fixture(path):
    d = load(path)
    r = O.bin_sort(d["pts"], list(d["nf"]), list(d["bins"]), d["maxsub"])
    np.testing.assert_array_equal(r["binsize"], d["binsize"])
    np.testing.assert_array_equal(r["binstart"], d["binstart"])
    np.testing.assert_array_equal(r["substart"], d["substart"])
    np.testing.assert_array_equal(r["binof"], d["binof"])
    assert rel_l2(oracle_result(d), d["expected"]) < 3 * d["tol"]


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=ids)
def test_library_reproduces_fixture(torch_dev, path):
    from cufinufft_amd import cufinufft
    from tests import gpu_util as G
    d = load(path)
    print(f"golden {ids(path)}", flush=True)
    nm = tuple(int(v) for v in d["n_modes"])
    dim = len(nm)
    bins = list(d["bins"]) + [1] * (3 - dim)
    opts = dict(gpu_binsizex=int(bins[0]), gpu_binsizey=int(bins[1]), gpu_binsizez=int(bins[2]),
                gpu_maxsubprobsize=d["maxsub"])
    oerr = rel_l2(oracle_result(d), d["expected"])
    for method in (1, 2):
        if d["type"] == 1:
            got, plan = G.run_type1(torch_dev, d["pts"], d["input"], nm, d["tol"], d["dtype"],
                                    gpu_method=method, **opts)
        else:
            got, plan = G.run_type2(torch_dev, d["pts"], d["input"], d["tol"], d["dtype"],
                                    gpu_method=method, **opts)
        err = rel_l2(got, d["expected"])
        assert err <= max(2 * d["tol"], 1.25 * oerr), (method, err, oerr)
        if method == 2:
            s = plan.sort_arrays()
            np.testing.assert_array_equal(s["binsize"], d["binsize"])
            np.testing.assert_array_equal(s["binstart"], d["binstart"])
            np.testing.assert_array_equal(s["substart"], d["substart"])
            np.testing.assert_array_equal(d["binof"][s["idx"]],
                                          np.repeat(np.arange(d["binsize"].size), d["binsize"]))
