"""Writes the committed FP64 fixtures of tests/golden/ from the oracle.

The reference ships no test vectors (it is MATLAB + CUDA and cannot run here),
so these fixtures are FP64 oracle outputs for two small seeded workloads, kept
as regression anchors: tests check that the oracle still reproduces them and
that the GPU path matches them at the stated tolerance.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from xrd_tomography_amd.synthetic import make_problem  # noqa: E402


def main():
    for name, flags in (("reference", 3), ("small", 0)):
        p = make_problem(name, flags=flags)
        N, C, A = O.forward(p, nthreads=4)
        x = np.full(p.n_cols, 10.0)
        y = np.random.default_rng(1).poisson(1e3 * (N + C) / (N + C).max()).astype(np.float64)
        c = C * (1e3 / (N + C).max())
        As = A * (1e3 / (N + C).max())
        xm = O.mlem(As, y, c, x, 10)
        out = os.path.join(HERE, f"{name}_flags{flags}.npz")
        rows0 = np.arange(0, len(N), p.geometry.n_views)  # the rows of view 0
        np.savez_compressed(out, N=N, C=C, A_view0=A[rows0], A_colsum=A.sum(0), y=y, mlem_x10=xm,
                            spectrum=p.flat("spectrum"), segmentation=p.flat("segmentation"))
        print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
