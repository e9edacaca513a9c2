"""Host-side index checks of setpts and the SM spread (tests/cpu_index):
the library's index helpers compiled for the CPU with AddressSanitizer and
UndefinedBehaviorSanitizer, driven with the points of a workload.  Every sort
destination must be hit once and the result sorted and stable; every bin,
subproblem, tap window, round-robin slot, LDS byte and flushed grid offset
must stay inside its allocation.  Small sizes here; the 1e8-point runs of
the 3D type-1 workloads use the same binary with their sizes."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_index")


@pytest.fixture(scope="module")
def index_check():
    if not (shutil.which("hipcc") or os.path.exists("/opt/rocm/bin/hipcc")):
        pytest.skip("hipcc not available")
    subprocess.check_call(["make", "-s", "-C", HERE, "index_check"])
    return os.path.join(HERE, "index_check")


CASES = [
    # prec dim nf1 nf2 nf3 ns M bins maxsub tol dist
    "f 3 64 64 64 6 300000 8 8 8 1024 1e-5 mixed",      # 32-bit cells, multi-chunk drains
    "f 3 64 48 40 7 200000 16 8 8 64 1e-6 mixed",       # 64-bit cells, 256 threads
    "f 3 48 48 48 7 200000 16 8 8 2048 1e-6 uniform",   # 64-bit cells, 1024 threads
    "d 3 64 64 64 10 200000 8 8 8 1024 1e-9 mixed",
    "d 3 40 40 40 11 100000 8 8 8 1024 4e-10 mixed",    # accuracy-first width
    "f 2 512 512 1 7 400000 32 32 1 1024 1e-6 uniform",
    "f 2 1024 1024 1 6 300000 64 64 1 8192 1e-5 mixed",
    "d 1 8192 1 1 9 100000 1024 1 1 1024 1e-8 mixed",
    "f 3 30 30 20 9 100000 8 8 8 1024 1e-5 mixed",      # sigma 1.25 widths
    "d 3 30 30 30 15 50000 4 4 4 1024 1e-9 mixed",
    # tile slots with partial last bins (gather covering rule)
    "f 3 33 40 26 6 50000 8 8 8 256 1e-5 mixed",
    "d 2 70 45 1 10 30000 16 16 1 2048 1e-9 mixed",
    "f 1 300 1 1 7 8000 32 1 1 2048 1e-6 mixed",
    # keys wide enough for two and three counting passes
    "f 3 256 256 256 6 200000 4 4 4 1024 1e-5 mixed",
    "f 3 512 512 512 6 100000 2 2 2 1024 1e-5 uniform",
]


@pytest.mark.parametrize("case", CASES)
def test_index_check(index_check, case):
    r = subprocess.run([index_check] + case.split(), capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failed checks" in r.stdout


@pytest.fixture(scope="module")
def prune_check():
    if not (shutil.which("hipcc") or os.path.exists("/opt/rocm/bin/hipcc")):
        pytest.skip("hipcc not available")
    subprocess.check_call(["make", "-s", "-C", HERE, "prune_check"])
    return os.path.join(HERE, "prune_check")


PRUNE_CASES = [
    # ms0 ms1 ms2 nf0 nf1 nf2 ntr iflag
    "6 5 3 12 10 8 2 1",     # odd and even mode counts, two transforms
    "5 4 4 10 9 8 1 -1",     # odd grid length
    "8 3 1 8 6 4 2 1",       # modes filling an axis (ms = nf), one mode on another
    "7 7 6 16 15 12 1 1",
    "9 8 8 12 10 10 3 -1",   # sigma 1.25-like grids: few non-mode cells
    # 2D (ms2 = nf2 = 1)
    "6 5 1 12 10 1 2 1",
    "7 4 1 16 9 1 3 -1",
    "8 8 1 8 10 1 1 1",      # modes filling the x axis
]


@pytest.mark.parametrize("case", PRUNE_CASES)
def test_pruned_fft_moves(prune_check, case):
    """The pruned 2D / 3D FFT's transposes (the library's own parameter
    builders and offset maps), applied element by element with naive DFTs in
    place of hipFFT: type 1 modes and type 2 grids (plain and with each
    derivative weight) match direct sums to 1e-12; every buffer access in
    bounds and every output element written once."""
    r = subprocess.run([prune_check] + case.split(), capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failed checks" in r.stdout
