"""Write the golden fixtures tests/golden/*.npz.

cuFINUFFT v1.3 ships no golden vectors (its tests recompute a crude direct
sum at run time and check one entry), so these fixtures are built from the
definition: inputs are seeded synthetic points / strengths / modes, expected
outputs are the FP64 direct sums (contrib/dirft2d.cpp:7-74 and its 1D/3D
analogues) over every mode or point, and the expected bin-sort arrays are the
exact counting sort of the reference's RESCALE / bin chain.  Both come from
the oracle (oracle/), which tests/test_oracle.py pins independently.

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

CASES = [
    # name, type, nmodes (x,y,z), M, tol, dtype, distribution, bins, maxsub
    ("t1_2d_f32", 1, (32, 24), 2000, 1e-5, np.float32, "uniform", (8, 8), 64),
    ("t2_2d_f64", 2, (20, 30), 1000, 1e-10, np.float64, "uniform", (16, 8), 32),
    ("t1_3d_f32_clustered", 1, (16, 16, 16), 3000, 1e-6, np.float32, "clustered", (8, 8, 4), 100),
    ("t2_3d_f64_clustered", 2, (12, 10, 8), 1500, 1e-9, np.float64, "clustered", (4, 4, 4), 16),
    ("t1_1d_f64", 1, (50,), 500, 1e-12, np.float64, "uniform", (32,), 20),
    ("t2_1d_f32", 2, (64,), 700, 1e-4, np.float32, "uniform", (16,), 50),
    ("t1_2d_f64_edges", 1, (16, 18), 64, 1e-8, np.float64, "edges", (8, 8), 8),
]


def points(dim, M, dtype, dist, rng):
    if dist == "uniform":
        P = rng.uniform(-np.pi, np.pi, (dim, M))
    elif dist == "clustered":
        cen = rng.uniform(-np.pi, np.pi, (3, dim))
        lab = rng.integers(0, 3, M)
        P = np.mod((cen[lab] + 0.25 * rng.standard_normal((M, dim))).T + np.pi, 2 * np.pi) - np.pi
    else:  # period edges, the origin, repeated points, one period out on either side
        base = np.array([-np.pi, np.nextafter(np.pi, 0), 0.0, 1e-300, -1e-300, 0.5, 0.5,
                         -2 * np.pi + 0.1, 2 * np.pi - 0.1, 3.0, -3.0])
        P = np.stack([np.resize(np.roll(base, d), M) +
                      np.resize([0, 0, 1e-3, -1e-3], M) * d for d in range(dim)])
    return [np.ascontiguousarray(p).astype(dtype) for p in P]


def main():
    for i, (name, typ, nm, M, tol, dtype, dist, bins, maxsub) in enumerate(CASES):
        rng = np.random.default_rng(100 + i)
        dim = len(nm)
        pts = points(dim, M, dtype, dist, rng)
        cd = np.complex64 if dtype == np.float32 else np.complex128
        p64 = [p.astype(np.float64) for p in pts]
        ntot = int(np.prod(nm))
        if typ == 1:
            inp = (rng.standard_normal(M) + 1j * rng.standard_normal(M)).astype(cd)
            exp = O.dirft_type1(p64, inp, list(nm), +1, np.arange(ntot)).reshape(nm[::-1])
        else:
            inp = (rng.standard_normal(nm[::-1]) + 1j * rng.standard_normal(nm[::-1])).astype(cd)
            exp = O.dirft_type2(p64, inp, list(nm), -1, np.arange(M))
        sp = O.setup_spreader(tol, dtype == np.float32)
        nf = [O.set_nf_type12(m, sp.ns) for m in nm]
        srt = O.bin_sort(pts, nf, bins, maxsub)
        out = os.path.join(HERE, name + ".npz")
        np.savez_compressed(
            out, type=typ, n_modes=np.array(nm), tol=tol, dtype=np.dtype(dtype).name,
            iflag=(1 if typ == 1 else -1), bins=np.array(bins), maxsub=maxsub,
            nf=np.array(nf), ns=sp.ns, pts=np.stack(pts), input=inp, expected=exp,
            binsize=srt["binsize"], binstart=srt["binstart"], substart=srt["substart"],
            binof=srt["binof"])
        print(f"wrote {out} ({os.path.getsize(out)} bytes)")


if __name__ == "__main__":
    main()
