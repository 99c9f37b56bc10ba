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


def comb_flat(hm, seed: int, nrows: int = 41, nblk: int = 70):
    """A block table with many narrow leaves per row: column blocks of width
    3..12, each cut at its own row into two leaves (dense or low rank, rank
    up to 6), so a 16-row strip crosses well over a hundred leaves."""
    rng = np.random.default_rng(seed)
    widths = rng.integers(3, 13, size=nblk)
    blocks, chunks, off, c0 = [], [], 0, 0
    for w in widths:
        cut = int(rng.integers(1, nrows))
        for r0, m in ((0, cut), (cut, nrows - cut)):
            if rng.random() < 0.5:
                C = rng.standard_normal((m, w))
                blocks.append((0, r0, c0, m, w, 0, off, 0))
                chunks.append(C.ravel(order="F"))
                off += C.size
            else:
                k = int(rng.integers(1, 7))
                A, B = rng.standard_normal((m, k)), rng.standard_normal((w, k))
                blocks.append((1, r0, c0, m, w, k, off, off + A.size))
                chunks += [A.ravel(order="F"), B.ravel(order="F")]
                off += A.size + B.size
        c0 += int(w)
    return hm.FlatHmat(nrows, c0, np.array(blocks, dtype=np.int64), np.concatenate(chunks))


def wide_rank_flat(hm, seed: int):
    """A 3000 x 2900 block table of large leaves with random factors of rank
    1, 16, 17, 20, 33 and 48: rank columns in several groups of 16, B factors
    longer than one 1024-row projection chunk, dense leaves between them."""
    rng = np.random.default_rng(seed)
    rows, cols = [0, 700, 2100, 3000], [0, 1500, 1600, 2900]
    ranks = iter(rng.permutation([1, 16, 17, 20, 33, 48]))
    blocks, chunks, off = [], [], 0
    for bi in range(3):
        for bj in range(3):
            r0, c0 = rows[bi], cols[bj]
            m, n = rows[bi + 1] - r0, cols[bj + 1] - c0
            if bi == bj or (bi, bj) == (2, 0):
                C = rng.standard_normal((m, n)) / np.sqrt(n)
                blocks.append((0, r0, c0, m, n, 0, off, 0))
                chunks.append(C.ravel(order="F"))
                off += C.size
            else:
                k = int(next(ranks))
                A, B = rng.standard_normal((m, k)), rng.standard_normal((n, k)) / np.sqrt(n * k)
                blocks.append((1, r0, c0, m, n, k, off, off + A.size))
                chunks += [A.ravel(order="F"), B.ravel(order="F")]
                off += A.size + B.size
    return hm.FlatHmat(3000, 2900, np.array(blocks, dtype=np.int64), np.concatenate(chunks))


def random_partition_flat(hm, seed: int, nrows: int, ncols: int, side=(24, 400), p_dense: float = 0.4,
                          kmax: int = 40):
    """A seeded random recursive partition of an nrows x ncols matrix (not
    an H-matrix construction: any exact tiling is a valid block table).  A
    block larger than side[1] on a side always splits; smaller ones split
    with probability 0.65 while both halves keep side[0] rows / columns, and
    one split in twenty cuts anywhere (slivers of a single row or column).
    Leaves are dense or low rank (rank 1..kmax, also above the 16 of one
    MFMA group and the 32 of one K-step) with random factors scaled so the
    product's entries stay O(1)."""
    rng = np.random.default_rng(seed)
    blocks, chunks = [], []
    off = 0

    def cut(dim):
        if dim >= 2 and rng.random() < 0.05:
            return int(rng.integers(1, dim))
        return int(rng.integers(side[0], dim - side[0] + 1))

    stack = [(0, 0, nrows, ncols)]
    while stack:
        r0, c0, m, n = stack.pop()
        can_r, can_c = m >= 2 * side[0], n >= 2 * side[0]
        if max(m, n) > side[1] or ((can_r or can_c) and rng.random() < 0.65):
            if m > side[1] or (can_r and not n > side[1] and (not can_c or rng.random() < 0.5)):
                a = cut(m)
                stack += [(r0, c0, a, n), (r0 + a, c0, m - a, n)]
            else:
                a = cut(n)
                stack += [(r0, c0, m, a), (r0, c0 + a, m, n - a)]
            continue
        if rng.random() < p_dense:
            C = rng.standard_normal((m, n)) / np.sqrt(n)
            blocks.append((0, r0, c0, m, n, 0, off, 0))
            chunks.append(C.ravel(order="F"))
            off += C.size
        else:
            k = int(rng.integers(1, kmax + 1))
            A, B = rng.standard_normal((m, k)), rng.standard_normal((n, k)) / np.sqrt(n * k)
            blocks.append((1, r0, c0, m, n, k, off, off + A.size))
            chunks += [A.ravel(order="F"), B.ravel(order="F")]
            off += A.size + B.size
    return hm.FlatHmat(nrows, ncols, np.array(blocks, dtype=np.int64), np.concatenate(chunks))


def vjp_closed_form(f, x, g):
    """Gradient of <g, H x> over the arena, leaf by leaf in numpy (FP64):
    dense C -> g x', low-rank A -> g (B'x)', B -> x (A'g)'."""
    out = np.zeros_like(f.data)
    for kind, r0, c0, m, n, k, oa, ob in f.blocks:
        gr, xc = g[r0:r0 + m], x[c0:c0 + n]
        if kind == 0:
            out[oa:oa + m * n] = np.outer(gr, xc).ravel(order="F")
        elif k > 0:
            A = f.data[oa:oa + m * k].reshape((m, k), order="F")
            B = f.data[ob:ob + n * k].reshape((n, k), order="F")
            out[oa:oa + m * k] = np.outer(gr, B.T @ xc).ravel(order="F")
            out[ob:ob + n * k] = np.outer(xc, A.T @ gr).ravel(order="F")
    return out

