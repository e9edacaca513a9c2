"""Build libcufinufft.so in-tree with hipcc for gfx950.

Usage: python cufinufft_amd/build.py [-j N]   (a file script: it never imports
the package, which needs the library it builds)

Each translation unit under csrc/ is compiled to an object in parallel, then
linked with hipFFT into cufinufft_amd/libcufinufft.so (the library the ctypes
wrapper loads).  Every call recompiles.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcufinufft.so")
ARCH = os.environ.get("CUFINUFFT_AMD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -packed-fp32-ops: keep the compiler from forming packed FP32 VALU ops.
FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + os.path.join(HERE, "..", "include"),
]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def compile_one(src):
    obj = os.path.join(BUILD, src.rsplit(".", 1)[0] + ".o")
    cmd = [HIPCC, "-x", "hip", *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
    return obj, time.time() - t0


def build(jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, min(os.cpu_count() or 4, 16)))
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for obj, dt in ex.map(compile_one, srcs):
            objs.append(obj)
            if verbose:
                print(f"  compiled {os.path.basename(obj)} in {dt:.1f}s", flush=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp,
           "-L/opt/rocm/lib", "-lhipfft", "-lrocprofiler-sdk-roctx",
           "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"  linked {LIB}", flush=True)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    build(a.j)


if __name__ == "__main__":
    sys.exit(main())
