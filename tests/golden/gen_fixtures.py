"""Writes tests/golden/tiny_cauchy.npz: a small exported H-matrix (N=256,
Cauchy kernel 1/(1+(x-y)^2) on LinRange(0,10,256), MaxBlock 64, MinBlock 16,
MaxRank 8, ACA, weak admissibility), one seeded vector and panel, and the
FP64 oracle's products (H x, H' x, H X) plus the dense kernel product.

Usage: python tests/golden/gen_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402
from oracle import oracle  # noqa: E402
from tests.helpers import build, cauchy  # noqa: E402


def main():
    hm = __graft_entry__.package()
    par, c, H = build(hm, np.linspace(0.0, 10.0, 256), cauchy, 64, 16, 8)
    f = hm.flatten(H)
    rng = np.random.default_rng(2024)
    x = rng.standard_normal(256)
    X = rng.standard_normal((256, 3))
    A = hm.FullMat(par.Kernel, c.X, c.X)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiny_cauchy.npz")
    np.savez_compressed(out, blocks=f.blocks, data=f.data, shape=np.array([f.nrows, f.ncols]),
                        x=x, X=X, y=oracle.matvec(f, x), yt=oracle.matvec_t(f, x),
                        Y=oracle.matmat(f, X), y_dense=A @ x, points=np.asarray(c.X))
    print("wrote", out, f.nblocks, "leaves")


if __name__ == "__main__":
    main()
