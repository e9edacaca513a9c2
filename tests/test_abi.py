"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/cufinufft.h declares, keeps the cufinufft_opts layout, fills the
reference's defaults, and returns the reference's error codes for invalid
arguments (all validated before the first HIP call)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cufinufft.h")


@pytest.fixture(scope="module")
def L():
    from cufinufft_amd import _cufinufft
    return _cufinufft


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(cufinufftf?_\w+)\s*\(", txt)))


def test_exports_every_declared_symbol(L):
    names = declared_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L.lib, n), n
    assert set(names) == set(L.EXPORTED)
    for core in ("cufinufft_default_opts", "cufinufft_makeplan", "cufinufftf_makeplan",
                 "cufinufft_setpts", "cufinufftf_setpts", "cufinufft_execute",
                 "cufinufftf_execute", "cufinufft_destroy", "cufinufftf_destroy"):
        assert core in names


def test_opts_layout(L):
    O = L.NufftOpts
    assert ctypes.sizeof(O) == 64
    assert O.upsampfac.offset == 0
    ints = [f for f, _ in O._fields_[1:]]
    assert len(ints) == 13
    for i, f in enumerate(ints):
        assert getattr(O, f).offset == 8 + 4 * i


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("typ", [1, 2])
def test_default_opts(L, typ, dim):
    o = L.NufftOpts()
    assert L._default_opts(typ, dim, o) == 0
    assert o.upsampfac == 2.0
    assert o.gpu_method == (2 if typ == 1 else 1)
    assert o.gpu_sort == 1 and o.gpu_maxsubprobsize == 1024
    assert (o.gpu_binsizex, o.gpu_binsizey, o.gpu_binsizez) == (-1, -1, -1)
    assert o.gpu_kerevalmeth == 0 and o.gpu_spreadinterponly == 0
    assert o.gpu_device_id == 0 and o.gpu_nstreams == 0
    assert L.lib.cufinufftf_default_opts(typ, dim, ctypes.byref(o)) == 0


def test_default_opts_type3_not_implemented(L):
    assert L._default_opts(3, 2, L.NufftOpts()) == 1


def _mk(L, typ=1, dim=2, nm=(16, 16, 1), ntr=1, tol=1e-6, single=True, **kw):
    o = L.NufftOpts()
    L._default_opts(1 if typ not in (1, 2) else typ, max(1, min(dim, 3)), o)
    for k, v in kw.items():
        setattr(o, k, v)
    plan = ctypes.c_void_p()
    modes = (ctypes.c_int * 3)(*nm)
    f = L._make_planf if single else L._make_plan
    ier = f(typ, dim, modes, 1, ntr, tol, 1, ctypes.byref(plan), o)
    return ier, plan


@pytest.mark.parametrize("kw,code", [
    (dict(typ=3), 1),                              # type 3 not implemented
    (dict(dim=4), 1),                              # bad dimension
    (dict(gpu_method=5), 2),                       # bad method
    (dict(gpu_method=3), 2),                       # Paul method is not built
    (dict(upsampfac=0.9), 7),                      # ERR_UPSAMPFAC_TOO_SMALL
    (dict(upsampfac=1.25, gpu_kerevalmeth=1), 8),  # HORNER_WRONG_BETA
    (dict(ntr=0), 9),                              # no transforms
    (dict(nm=(0, 16, 1)), 9),                      # no modes
    (dict(gpu_spreadinterponly=1, nm=(8, 8, 1)), 3),  # grid smaller than 2 ns
    (dict(gpu_binsizex=0), 1),                     # invalid bin size
    (dict(typ=2, gpu_method=4), 2),                # block method spreads only (type 1)
    # block method: nf is rounded to a multiple of the block, so blocks must
    # be 2,3,5-smooth (14 is not: the reference's search would never end),
    # and they must hold whole bins
    (dict(dim=3, nm=(16, 16, 16), gpu_method=4, gpu_obinsizex=14, gpu_binsizex=7), 1),
    (dict(dim=3, nm=(16, 16, 16), gpu_method=4, gpu_obinsizex=8, gpu_binsizex=3), 1),
    # block of 64^3 cells: its LDS tile cannot fit
    (dict(dim=3, nm=(32, 32, 32), gpu_method=4, gpu_obinsizex=64, gpu_obinsizey=64,
          gpu_obinsizez=64, gpu_binsizex=8, gpu_binsizey=8, gpu_binsizez=8), 12),
])
def test_makeplan_error_codes(L, kw, code):
    ier, plan = _mk(L, **kw)
    assert ier == code
    assert not plan.value


def test_makeplan_ext_flags(L):
    """Planning flags: 1 accuracy-first, 2 tile slots, 4 full FFT; any other
    bit is rejected with 1 before the device is touched."""
    o = L.NufftOpts()
    L._default_opts(1, 2, o)
    plan = ctypes.c_void_p()
    modes = (ctypes.c_int * 3)(16, 16, 1)
    assert L._make_planf_ext(1, 2, modes, 1, 1, 1e-6, 1, ctypes.byref(plan), o, 8) == 1
    assert L._make_planf_ext(1, 2, modes, 1, 1, 1e-6, 1, ctypes.byref(plan), o, 16 | 1) == 1
    assert not plan.value
    assert L.TILE_SLOTS == 2 and L.ACCURACY_FIRST == 1 and L.FULL_FFT == 4


def test_null_plan_calls(L):
    assert L._destroy_plan(None) == 1
    assert L._destroy_planf(None) == 1
    assert L._exec_planf(None, None, None) == 1
    assert L._set_ptsf(10, None, None, None, 0, None, None, None, None) == 1
    assert L._set_stream(None, None) == 1


def test_python_class_option_errors():
    from cufinufft_amd import cufinufft
    with pytest.raises(TypeError) as err:
        cufinufft(1, (8, 8), foo="bar")
    assert "Invalid option 'foo'" in err.value.args[0]
    with pytest.raises(TypeError):
        cufinufft(1, (8, 8), dtype=np.int32)


def test_roctx_ranges_linked(L):
    """makeplan / setpts / execute open roctx ranges (the reference's NVTX
    ranges, include/profile.h:14-15): the library imports the push / pop
    entry points from the roctx library it links."""
    import subprocess
    path = L.lib._name
    dyn = subprocess.run(["nm", "-D", "--undefined-only", path], capture_output=True,
                         text=True, check=True).stdout
    assert "roctxRangePushA" in dyn and "roctxRangePop" in dyn
    needed = subprocess.run(["readelf", "-d", path], capture_output=True, text=True,
                            check=True).stdout
    assert "librocprofiler-sdk-roctx" in needed
