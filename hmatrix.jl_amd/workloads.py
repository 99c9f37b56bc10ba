"""Seeded synthetic workloads (1-D point clouds) used by the benchmark and tests."""
from __future__ import annotations

import numpy as np


def build(hm, X, kernel, MaxBlock: int, MinBlock: int, MaxRank: int, cons: str = "aca2",
          adj: str = "weak", eps: float = 1e-10):
    """``NewHmat`` on points X with weak (identity) or strong (tridiagonal)
    admissibility; returns ``(params, cluster, H)``."""
    import scipy.sparse as sp
    par = hm.Params()
    par.Geom = np.asarray(X, dtype=np.float64)
    par.Kernel = kernel
    par.MaxBlock, par.MinBlock, par.MaxRank = MaxBlock, MinBlock, MaxRank
    par.ConsMethod = cons
    par.εComp = eps
    N = par.Geom.shape[0]
    if isinstance(adj, str) and adj == "weak":
        par.Adj = sp.identity(N, format="csr")
    elif isinstance(adj, str) and adj == "strong":
        par.Adj = sp.diags([1, 1, 1], [-1, 0, 1], shape=(N, N), format="csr")
    else:
        par.Adj = adj
    c, H = hm.NewHmat(par)
    return par, c, H


def uneven_batch(hm, nmat: int = 64, nlo: int = 1000, nhi: int = 40000, seed: int = 6):
    """Independent H-matrices: N log-uniform in [nlo, nhi], kernel
    1/(1+((x-y)/l)^2) with l ~ U[0.2, 5], MaxRank in {4, 8, 16}, sorted
    uniform points on [0, 100], MinBlock 64, MaxBlock N/4, weak
    admissibility.  Returns the flattened members."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nmat):
        N = int(round(np.exp(rng.uniform(np.log(nlo), np.log(nhi)))))
        l = float(rng.uniform(0.2, 5.0))
        mr = int(rng.choice([4, 8, 16]))
        X = np.sort(rng.uniform(0.0, 100.0, N))
        kern = (lambda L: (lambda x, y: 1.0 / (1.0 + ((x - y) / L) ** 2)))(l)
        par, c, H = build(hm, X, kern, -1, 64, mr)
        out.append(hm.flatten(H))
    return out
