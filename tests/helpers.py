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
    """Uneven batch: N log-uniform in [nlo, nhi], kernel 1/(1+((x-y)/l)^2) with
    l ~ U[0.2, 5], MaxRank in {4, 8, 16}, sorted uniform points on [0, 100],
    MinBlock 64, MaxBlock N/4, weak admissibility."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nmat):
        N = int(round(np.exp(rng.uniform(np.log(nlo), np.log(nhi)))))
        l = float(rng.uniform(0.2, 5.0))
        mr = int(rng.choice([4, 8, 16]))
        X = np.sort(rng.uniform(0.0, 100.0, N))
        kern = (lambda L: (lambda x, y: 1.0 / (1.0 + ((x - y) / L) ** 2)))(l)
        par, c, H = build(hm, X, kern, -1, 64, mr)
        out.append(hm.flatten(H))
    return out
