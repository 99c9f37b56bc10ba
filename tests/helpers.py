"""Workload builders shared by the tests (seeded, deterministic)."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def cauchy(x, y):
    return 1.0 / (1.0 + (x - y) ** 2)


def expk(x, y):
    return np.exp(-np.abs(x - y))


def build(hm, X, kernel, MaxBlock, MinBlock, MaxRank, cons="aca2", adj="weak", eps=1e-10):
    par = hm.Params()
    par.Geom = np.asarray(X, dtype=np.float64)
    par.Kernel = kernel
    par.MaxBlock, par.MinBlock, par.MaxRank = MaxBlock, MinBlock, MaxRank
    par.ConsMethod = cons
    par.εComp = eps
    N = par.Geom.shape[0]
    if adj == "weak":
        par.Adj = sp.identity(N, format="csr")
    elif adj == "strong":
        par.Adj = sp.diags([1, 1, 1], [-1, 0, 1], shape=(N, N), format="csr")
    else:
        par.Adj = adj
    c, H = hm.NewHmat(par)
    return par, c, H


def c1(hm, cons="aca2"):
    """Reference test case: LinRange(0,10,2000), Cauchy, MaxBlock 500, MinBlock 64, MaxRank 5
    (test/harithm.jl:1-12)."""
    return build(hm, np.linspace(0.0, 10.0, 2000), cauchy, 500, 64, 5, cons)


def sorted_uniform(n, lo, hi, seed):
    return np.sort(np.random.default_rng(seed).uniform(lo, hi, n))


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))



def c4_members(hm, nmat=64, nlo=1000, nhi=40000, seed=6):
    """Uneven batch of independent H-matrices (see hmatrix_amd.workloads)."""
    from hmatrix_amd.workloads import uneven_batch
    return uneven_batch(hm, nmat=nmat, nlo=nlo, nhi=nhi, seed=seed)
