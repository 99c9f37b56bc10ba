"""GPU parity of the FP64 multi-right-hand-side product (f64 MFMA, one pass
over the FP64 factors per panel of 32 columns) against the FP64 oracle.

The reference multiplies H by a matrix column by column
(src/harithm.jl:317-323), each column one pass of hmat_matvec!
(src/harithm.jl:340-354); the oracle restates that order.  The MFMA path
sums each 32-column K-step in another order, so the bound is relative
1e-13 (Frobenius norm over all columns), the same as for the
single-column FP64 product; it is also compared with the plan's own
single-column FP64 path (no batched layout).
"""
import numpy as np
import pytest

from tests.helpers import build, c1, cauchy, comb_flat, expk, relerr, sorted_uniform, wide_rank_flat

pytestmark = pytest.mark.gpu

TOL64 = 1e-13


def _plan(hm, f, cfg=None):
    d = hm.DeviceHmat(f, "f64", cfg)
    info = d.info()
    assert info["batched"] == 1 and info["strip_shape"] == 3, info
    return d


def _mm(d, n_in, X, transpose=False, alpha=1.0, beta=0.0, Y0=None):
    import torch
    Xd = d.new_cols(n_in, X.shape[1])
    Xd.copy_(torch.from_numpy(X))
    Yd = None
    if Y0 is not None:
        Yd = d.new_cols(Y0.shape[0], Y0.shape[1])
        Yd.copy_(torch.from_numpy(Y0))
    out = d.matmat(Xd, Yd, alpha=alpha, beta=beta, transpose=transpose)
    return out.cpu().numpy()


@pytest.fixture(scope="module")
def c1f(hm):
    par, c, H = c1(hm, "aca2")
    return hm.flatten(H)


@pytest.mark.parametrize("nrhs", [2, 13, 32, 45, 70])
def test_c1_fp64_batched(hm, orc, c1f, nrhs):
    d = _plan(hm, c1f)
    X = np.random.default_rng(nrhs).standard_normal((c1f.ncols, nrhs))
    assert relerr(_mm(d, c1f.ncols, X), orc.matmat(c1f, X)) <= TOL64
    V = np.random.default_rng(100 + nrhs).standard_normal((c1f.nrows, nrhs))
    ref_t = np.stack([orc.matvec_t(c1f, V[:, i]) for i in range(nrhs)], 1)
    assert relerr(_mm(d, c1f.nrows, V, transpose=True), ref_t) <= TOL64
    d.close()


def test_fp64_batched_alpha_beta_and_single_layout(hm, orc, c1f):
    """Y = beta Y + alpha H X, and equality with the same plan family
    built without the batched layout (panels of 8 columns, VALU)."""
    rng = np.random.default_rng(11)
    X = rng.standard_normal((c1f.ncols, 32))
    Y0 = rng.standard_normal((c1f.nrows, 32))
    d = _plan(hm, c1f)
    got = _mm(d, c1f.ncols, X, alpha=-1.5, beta=0.5, Y0=Y0)
    assert relerr(got, 0.5 * Y0 - 1.5 * orc.matmat(c1f, X)) <= TOL64
    d.close()
    d1 = hm.DeviceHmat(c1f, "f64", {"no_batched": True})
    assert d1.info()["batched"] == 0
    one = _mm(d1, c1f.ncols, X, alpha=-1.5, beta=0.5, Y0=Y0)
    d1.close()
    assert relerr(got, one) <= TOL64


@pytest.mark.parametrize("seed", [0, 1])
def test_fp64_batched_many_narrow_leaves(hm, orc, seed):
    """Strips crossing over a hundred narrow leaves: chained strip records,
    dense pieces of 3..12 columns, panels of many rank columns."""
    f = comb_flat(hm, seed, nblk=70)
    d = _plan(hm, f)
    X = np.random.default_rng(seed).standard_normal((f.ncols, 32))
    assert relerr(_mm(d, f.ncols, X), orc.matmat(f, X)) <= TOL64
    V = np.random.default_rng(seed + 7).standard_normal((f.nrows, 32))
    ref_t = np.stack([orc.matvec_t(f, V[:, i]) for i in range(32)], 1)
    assert relerr(_mm(d, f.nrows, V, transpose=True), ref_t) <= TOL64
    d.close()


def test_fp64_batched_wide_ranks(hm, orc):
    """Leaves of rank 17..48 (several 16-column projection groups per leaf
    in every super-tile it crosses) and B factors over several super-tiles
    (partial sums folded by the reduction)."""
    f = wide_rank_flat(hm, 2)
    d = _plan(hm, f)
    X = np.random.default_rng(50).standard_normal((f.ncols, 32))
    assert relerr(_mm(d, f.ncols, X), orc.matmat(f, X)) <= TOL64
    V = np.random.default_rng(51).standard_normal((f.nrows, 32))
    ref = np.stack([orc.matvec_t(f, V[:, i]) for i in range(32)], 1)
    assert relerr(_mm(d, f.nrows, V, transpose=True), ref) <= TOL64
    d.close()


@pytest.mark.parametrize("minblock", [64, 128])
def test_fp64_batched_strong_admissibility(hm, orc, minblock):
    """N = 20000 Cauchy, strong (tridiagonal) admissibility: with MinBlock
    128 the bands' X windows exceed the LDS window and the row pass reads
    X from memory."""
    X = sorted_uniform(20000, 0.0, 60.0, 21)
    par, c, H = build(hm, X, cauchy, 2048, minblock, 12, adj="strong")
    f = hm.flatten(H)
    d = _plan(hm, f)
    R = np.random.default_rng(22).standard_normal((f.ncols, 32))
    assert relerr(_mm(d, f.ncols, R), orc.matmat(f, R)) <= TOL64
    d.close()


def test_fp64_batched_padded_leading_dimensions(hm, orc, c1f):
    """ldx = n + 37 (odd: 8-byte aligned columns), ldy = n + 5, 33 columns
    (two panels); padding rows of Y untouched."""
    import torch
    d = _plan(hm, c1f)
    n, nrhs = c1f.ncols, 33
    X = np.random.default_rng(33).standard_normal((n, nrhs))
    Xb = torch.zeros((nrhs, n + 37), dtype=torch.float64, device="cuda")
    Yb = torch.full((nrhs, n + 5), 7.0, dtype=torch.float64, device="cuda")
    Xd, Yd = Xb[:, :n].t(), Yb[:, :n].t()
    Xd.copy_(torch.from_numpy(X))
    d.matmat(Xd, Yd)
    assert relerr(Yd.cpu().numpy(), orc.matmat(c1f, X)) <= TOL64
    assert bool((Yb[:, n:] == 7.0).all())
    d.close()


def test_fp64_batched_graph_replay(hm, c1f):
    """A captured FP64 H X of 32 columns replays bitwise equal to the eager
    product, three times, with eager work between the replays."""
    import torch
    d = _plan(hm, c1f)
    X = d.new_cols(c1f.ncols, 32)
    Y = d.new_cols(c1f.nrows, 32)
    g = d.graph(X, Y)
    for rep in range(3):
        X.copy_(torch.randn(c1f.ncols, 32, dtype=torch.float64, generator=torch.Generator().manual_seed(rep)))
        ref = d.matmat(X).clone()
        torch.cuda.synchronize()
        Y.fill_(float("nan"))
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(Y, ref), rep
    g.close()
    d.close()
