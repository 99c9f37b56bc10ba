"""Pins the FP64 oracle (oracle/hmat_oracle.c) without trusting its own formulas.

Every check compares the oracle with something computed another way:
brute-force dense reconstruction in numpy, the closed-form kernel matrix,
exact low-rank structure, linearity, the adjoint identity and central finite
differences.  The planted-mistake tests show that each check fails for a
plausible error (dropped leaf kind, transposed factor, wrong sign, wrong
index), so a wrong oracle could not pass.
"""
import numpy as np
import pytest

from tests.helpers import build, c1, cauchy, expk, sorted_uniform, relerr


def random_flat(hm, seed, N=173, M=141):
    """A hand-made block table: dense and low-rank leaves with random factors,
    uneven sizes, k = 0 leaf included, tiling an N x M matrix."""
    rng = np.random.default_rng(seed)
    rows = [0, (N * 37) // 173, (N * 90) // 173, N]
    cols = [0, (M * 51) // 141, (M * 51) // 141 + 1, M]
    blocks, chunks, off = [], [], 0
    kinds = rng.integers(0, 2, size=(3, 3))
    kinds[1, 1] = 1
    for bi in range(3):
        for bj in range(3):
            r0, c0 = rows[bi], cols[bj]
            m, n = rows[bi + 1] - r0, cols[bj + 1] - c0
            if kinds[bi, bj] == 0:
                C = rng.standard_normal((m, n))
                blocks.append((0, r0, c0, m, n, 0, off, 0))
                chunks.append(C.ravel(order="F"))
                off += C.size
            else:
                k = 0 if (bi, bj) == (2, 0) else int(rng.integers(1, 6))
                A = rng.standard_normal((m, k))
                B = rng.standard_normal((n, k))
                blocks.append((1, r0, c0, m, n, k, off, off + A.size))
                chunks += [A.ravel(order="F"), B.ravel(order="F")]
                off += A.size + B.size
    return hm.FlatHmat(N, M, np.array(blocks, dtype=np.int64), np.concatenate(chunks))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_matvec_equals_dense_bruteforce(hm, orc, seed):
    f = random_flat(hm, seed)
    D = hm.to_dense(f)                     # numpy reconstruction, A @ B.T per leaf
    x = np.random.default_rng(seed + 10).standard_normal(f.ncols)
    y0 = np.random.default_rng(seed + 20).standard_normal(f.nrows)
    assert relerr(orc.matvec(f, x), D @ x) < 1e-14
    assert relerr(orc.matvec(f, x, r=y0, s=-0.75), y0 - 0.75 * (D @ x)) < 1e-14
    v = np.random.default_rng(seed + 30).standard_normal(f.nrows)
    assert relerr(orc.matvec_t(f, v), D.T @ v) < 1e-14
    X = np.random.default_rng(seed + 40).standard_normal((f.ncols, 4))
    assert relerr(orc.matmat(f, X), D @ X) < 1e-14
    assert np.allclose(orc.to_dense(f), D, rtol=0, atol=1e-13)


def test_to_dense_three_ways(hm, orc):
    par, c, H = build(hm, np.linspace(0, 5, 300), cauchy, 80, 16, 6)
    f = hm.flatten(H)
    Dt = hm.to_fmat(H)                     # recursive 2x2 concatenation (utils.jl:162-192)
    assert np.allclose(orc.to_dense(f), Dt, rtol=0, atol=1e-15)
    assert np.allclose(hm.to_dense(f), Dt, rtol=0, atol=1e-15)


def test_reference_matmul_case(hm, orc):
    """test/harithm.jl:1-12: A*y ≈ H*y (isapprox rtol sqrt(eps)) with the dense FullMat.
    Built with rank-based admissibility (ConsMethod 'svd'), which resolves the
    Cauchy kernel to 1e-10."""
    par, c, H = c1(hm, "svd")
    f = hm.flatten(H)
    A = hm.FullMat(par.Kernel, c.X, c.X)
    y = np.random.default_rng(0).random(2000)
    assert np.linalg.norm(A @ y - orc.matvec(f, y)) <= np.sqrt(np.finfo(float).eps) * np.linalg.norm(A @ y)


def test_exp_kernel_exact_rank1(hm, orc):
    """exp(-|x-y|) off-diagonal blocks in 1-D are exactly e^{-x} e^{y}: the
    H-matrix product equals the dense kernel product to rounding."""
    X = sorted_uniform(3000, 0.0, 30.0, 3)
    par, c, H = build(hm, X, expk, 750, 64, 24)
    f = hm.flatten(H)
    A = hm.FullMat(par.Kernel, c.X, c.X)
    y = np.random.default_rng(1).standard_normal(3000)
    assert relerr(orc.matvec(f, y), A @ y) < 1e-13


def test_linearity_and_adjoint_identity(hm, orc):
    par, c, H = build(hm, sorted_uniform(2048, 0, 20, 5), cauchy, 512, 64, 12)
    f = hm.flatten(H)
    rng = np.random.default_rng(2)
    x, z, w = rng.standard_normal((3, 2048))
    a, b = 0.3, -1.7
    lhs = orc.matvec(f, a * x + b * z)
    assert relerr(lhs, a * orc.matvec(f, x) + b * orc.matvec(f, z)) < 1e-14
    ip1 = orc.matvec(f, x) @ w
    ip2 = x @ orc.matvec_t(f, w)
    assert abs(ip1 - ip2) <= 1e-13 * np.linalg.norm(orc.matvec(f, x)) * np.linalg.norm(w)


def test_adjoint_gradients_finite_differences(hm, orc):
    """L = 0.5||Hx-b||^2 is quadratic in each single parameter, so central
    differences with h = 1e-6 are exact up to rounding."""
    f = random_flat(hm, 5, N=40, M=33)
    rng = np.random.default_rng(3)
    x, bb = rng.standard_normal(f.ncols), rng.standard_normal(f.nrows)
    L, res, dx, grad = orc.adjoint(f, x, bb)
    D = hm.to_dense(f)
    assert np.isclose(L, 0.5 * np.sum((D @ x - bb) ** 2), rtol=1e-13)
    assert relerr(dx, D.T @ (D @ x - bb)) < 1e-13
    h = 1e-6
    for j in rng.choice(f.ncols, 6, replace=False):
        e = np.zeros(f.ncols)
        e[j] = h
        Lp = orc.adjoint(f, x + e, bb)[0]
        Lm = orc.adjoint(f, x - e, bb)[0]
        assert abs((Lp - Lm) / (2 * h) - dx[j]) < 1e-6 * max(1.0, abs(dx[j]))
    covered = np.zeros(f.data.size, bool)
    for kind, r0, c0, m, n, k, oa, ob in f.blocks:
        covered[oa:oa + (m * n if kind == 0 else m * k)] = True
        if kind == 1:
            covered[ob:ob + n * k] = True
    for p in rng.choice(np.nonzero(covered)[0], 12, replace=False):
        fp = hm.FlatHmat(f.nrows, f.ncols, f.blocks, f.data.copy())
        fm = hm.FlatHmat(f.nrows, f.ncols, f.blocks, f.data.copy())
        fp.data[p] += h
        fm.data[p] -= h
        g = (orc.adjoint(fp, x, bb)[0] - orc.adjoint(fm, x, bb)[0]) / (2 * h)
        assert abs(g - grad[p]) < 1e-6 * max(1.0, abs(grad[p]))


# --- planted mistakes: the checks above must reject a wrong oracle ---------

def _wrong_matvec(f, x, mistake):
    y = np.zeros(f.nrows)
    for kind, r0, c0, m, n, k, oa, ob in f.blocks:
        if kind == 0:
            if mistake == "drop_dense":
                continue
            C = f.data[oa:oa + m * n].reshape((m, n), order="F")
            y[r0:r0 + m] += C @ x[c0:c0 + n]
        else:
            A = f.data[oa:oa + m * k].reshape((m, k), order="F")
            B = f.data[ob:ob + n * k].reshape((n, k), order="F")
            if mistake == "swap_AB" and m == n:
                A, B = B, A
            t = B.T @ x[c0:c0 + n]
            sgn = -1.0 if mistake == "sign" else 1.0
            off = 1 if mistake == "index" else 0
            rr = slice(r0 + off, r0 + off + m) if r0 + off + m <= f.nrows else slice(r0, r0 + m)
            y[rr] += sgn * (A @ t)
    return y


@pytest.mark.parametrize("mistake", ["drop_dense", "swap_AB", "sign", "index"])
def test_planted_mistakes_are_caught(hm, mistake):
    par, c, H = build(hm, np.linspace(0.0, 10.0, 1024), cauchy, 256, 32, 8)   # square leaves
    f = hm.flatten(H)
    x = np.random.default_rng(9).standard_normal(f.ncols)
    D = hm.to_dense(f)
    assert relerr(_wrong_matvec(f, x, "none"), D @ x) < 1e-14
    assert relerr(_wrong_matvec(f, x, mistake), D @ x) > 1e-6


def test_golden_fixture(hm, orc):
    """Committed vectors (tests/golden/gen_fixtures.py) are reproduced."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "tiny_cauchy.npz"))
    f = hm.FlatHmat(int(z["shape"][0]), int(z["shape"][1]), z["blocks"], z["data"])
    assert relerr(orc.matvec(f, z["x"]), z["y"]) < 1e-15
    assert relerr(orc.matvec_t(f, z["x"]), z["yt"]) < 1e-15
    assert relerr(orc.matmat(f, z["X"]), z["Y"]) < 1e-15
    # the H-matrix (MaxRank 8) approximates the kernel matrix
    assert relerr(z["y"], z["y_dense"]) < 1e-6


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_matvec_bitwise_equal(hm, orc, threads):
    """The all-core CPU baseline computes exactly the serial oracle's sums:
    bitwise equal on a hierarchical matrix (rows split unevenly across
    threads, leaves crossing the splits) and on the hand-made table with a
    k = 0 leaf, with s != 1 and a nonzero starting r."""
    par, c, H = build(hm, sorted_uniform(3001, 0, 20, 9), cauchy, 700, 48, 10)
    rng = np.random.default_rng(threads)
    for f in (hm.flatten(H), random_flat(hm, threads)):
        v = rng.standard_normal(f.ncols)
        r0 = rng.standard_normal(f.nrows)
        assert np.array_equal(orc.matvec_mt(f, v, threads, r0, -0.75), orc.matvec(f, v, r0, -0.75))
