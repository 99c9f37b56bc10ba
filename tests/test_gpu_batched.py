"""GPU parity of the batched multi-right-hand-side path (MFMA) against the FP64 oracle.

Tolerances are relative Frobenius norms of the error over all right-hand sides:
split BF16 and FP32 storage (3 BF16 MFMA terms, FP32 accumulation) <= 2e-5,
plain BF16 storage <= 1e-2.
"""
import numpy as np
import pytest

from tests.helpers import build, c1, cauchy, expk, sorted_uniform, relerr

pytestmark = pytest.mark.gpu

TOL = {"bf16x2": 2e-5, "f32": 2e-5, "bf16": 1e-2}


def _run(hm, f, dtype, X, transpose=False):
    import torch
    d = hm.DeviceHmat(f, dtype)
    n_in = f.nrows if transpose else f.ncols
    Xd = d.new_cols(n_in, X.shape[1])
    Xd.copy_(torch.from_numpy(X).float())
    Y = d.matmat(Xd, transpose=transpose)
    out = Y.cpu().numpy().astype(np.float64)
    d.close()
    return out


@pytest.fixture(scope="module")
def c1f(hm):
    par, c, H = c1(hm, "aca2")
    return hm.flatten(H)


@pytest.mark.parametrize("dtype", ["bf16x2", "f32", "bf16"])
@pytest.mark.parametrize("nrhs", [32, 5, 40, 97])
def test_c1_batched(hm, orc, c1f, dtype, nrhs):
    X = np.random.default_rng(nrhs).standard_normal((c1f.ncols, nrhs))
    assert relerr(_run(hm, c1f, dtype, X), orc.matmat(c1f, X)) <= TOL[dtype]


@pytest.mark.parametrize("dtype", ["bf16x2", "bf16"])
def test_c1_batched_transpose(hm, orc, c1f, dtype):
    X = np.random.default_rng(7).standard_normal((c1f.nrows, 32))
    ref = np.stack([orc.matvec_t(c1f, X[:, i]) for i in range(32)], 1)
    assert relerr(_run(hm, c1f, dtype, X, transpose=True), ref) <= TOL[dtype]


def test_batched_alpha_beta(hm, orc, c1f):
    import torch
    d = hm.DeviceHmat(c1f, "bf16x2")
    rng = np.random.default_rng(11)
    X = rng.standard_normal((c1f.ncols, 32))
    Y0 = rng.standard_normal((c1f.nrows, 32))
    Xd = d.new_cols(c1f.ncols, 32)
    Xd.copy_(torch.from_numpy(X).float())
    Yd = d.new_cols(c1f.nrows, 32)
    Yd.copy_(torch.from_numpy(Y0).float())
    d.matmat(Xd, Yd, alpha=-1.5, beta=0.5)
    ref = 0.5 * Y0 - 1.5 * orc.matmat(c1f, X)
    assert relerr(Yd.cpu().numpy(), ref) <= 2e-5


def test_exp_kernel_large_factors(hm, orc):
    """exp(-|x-y|) on a wide interval: ACA factors span e^{+-100}; the library
    balances them so that FP32/BF16 copies stay finite."""
    X = sorted_uniform(16384, 0.0, 400.0, 3)
    par, c, H = build(hm, X, expk, 4096, 64, 24)
    f = hm.flatten(H)
    R = np.random.default_rng(4).standard_normal((f.ncols, 32))
    got = _run(hm, f, "bf16x2", R)
    assert np.isfinite(got).all()
    assert relerr(got, orc.matmat(f, R)) <= 2e-5


def test_ragged_small(hm, orc):
    """N smaller than one projection super-tile, uneven leaves, strong admissibility."""
    X = sorted_uniform(300, 0.0, 7.0, 12)
    par, c, H = build(hm, X, cauchy, 100, 16, 6, adj="strong")
    f = hm.flatten(H)
    R = np.random.default_rng(5).standard_normal((f.ncols, 3))
    assert relerr(_run(hm, f, "bf16x2", R), orc.matmat(f, R)) <= 2e-5


def test_kernel_timing_counts_launches(hm, orc, c1f):
    """hm_plan_timing records one event set per 32-column panel; the phase
    sums are positive and the product is unchanged."""
    import torch
    d = hm.DeviceHmat(c1f, "bf16x2")
    X = np.random.default_rng(12).standard_normal((c1f.ncols, 40))
    Xd = d.new_cols(c1f.ncols, 40)
    Xd.copy_(torch.from_numpy(X).float())
    d.timing(True)
    Y = d.matmat(Xd)
    Y = d.matmat(Xd)
    r = d.timing_read()
    d.timing(False)
    assert r["launches"] == 4                      # 2 products x 2 panels (32 + 8)
    assert r["rows_ms"] > 0 and r["proj_ms"] > 0 and r["reduce_ms"] >= 0
    assert d.timing_read()["launches"] == 0
    assert relerr(Y.cpu().numpy().astype(np.float64), orc.matmat(c1f, X)) <= TOL["bf16x2"]
    d.close()


@pytest.mark.parametrize("dtype", ["bf16x2", "f32", "bf16"])
@pytest.mark.parametrize("seed,nblk", [(0, 70), (1, 70), (2, 200)])
def test_many_leaves_per_strip(hm, orc, dtype, seed, nblk):
    """More than a hundred narrow leaves cross each 16-row strip: the row
    pass walks several chained descriptor records per strip.  With 200
    column blocks (about 1500 columns of dense pieces per workgroup) no X
    window fits in LDS and the row pass reads X straight from memory."""
    from tests.helpers import comb_flat
    f = comb_flat(hm, seed, nblk=nblk)
    X = np.random.default_rng(seed).standard_normal((f.ncols, 32))
    assert relerr(_run(hm, f, dtype, X), orc.matmat(f, X)) <= TOL[dtype]
    V = np.random.default_rng(seed + 5).standard_normal((f.nrows, 32))
    ref = np.stack([orc.matvec_t(f, V[:, i]) for i in range(32)], 1)
    assert relerr(_run(hm, f, dtype, V, transpose=True), ref) <= TOL[dtype]


def test_two_plans_two_streams(hm, orc, c1f):
    """Two plans of one matrix, each on its own default-priority stream, with
    products interleaved and in flight together: each plan's scratch belongs to
    its stream, so both results match the oracle."""
    import torch
    S = 2
    plans = [hm.DeviceHmat(c1f, "bf16x2") for _ in range(S)]
    streams = [torch.cuda.Stream() for _ in range(S)]
    rng = np.random.default_rng(21)
    Xs = [rng.standard_normal((c1f.ncols, 32)) for _ in range(S)]
    Xd = [plans[i].new_cols(c1f.ncols, 32) for i in range(S)]
    Yd = [plans[i].new_cols(c1f.nrows, 32) for i in range(S)]
    for i in range(S):
        Xd[i].copy_(torch.from_numpy(Xs[i]).float())
    torch.cuda.synchronize()
    for _ in range(5):
        for i in range(S):
            plans[i].matmat(Xd[i], Yd[i], stream=streams[i])
    torch.cuda.synchronize()
    for i in range(S):
        assert relerr(Yd[i].cpu().numpy().astype(np.float64), orc.matmat(c1f, Xs[i])) <= TOL["bf16x2"]
        plans[i].close()


@pytest.mark.parametrize("dtype", ["bf16x2", "f64"])
def test_padded_leading_dimensions(hm, orc, c1f, dtype):
    """X and Y as column-major views with leading dimensions larger than the
    row count (ldx = n + 37, ldy = n + 5) and 33 right-hand sides (two
    panels); the padding rows of Y are not written."""
    import torch
    d = hm.DeviceHmat(c1f, dtype)
    td = torch.float64 if dtype == "f64" else torch.float32
    n, nrhs = c1f.ncols, 33
    X = np.random.default_rng(33).standard_normal((n, nrhs))
    Xb = torch.zeros((nrhs, n + 37), dtype=td, device="cuda")
    Yb = torch.full((nrhs, n + 5), 7.0, dtype=td, device="cuda")
    Xd, Yd = Xb[:, :n].t(), Yb[:, :n].t()
    assert Xd.stride() == (1, n + 37) and Yd.stride() == (1, n + 5)
    Xd.copy_(torch.from_numpy(X).to(td))
    d.matmat(Xd, Yd)
    tol = 1e-13 if dtype == "f64" else TOL[dtype]
    assert relerr(Yd.cpu().numpy().astype(np.float64), orc.matmat(c1f, X)) <= tol
    assert bool((Yb[:, n:] == 7.0).all())
    d.close()


@pytest.mark.parametrize("dtype", ["bf16x2", "f32"])
def test_wide_ranks_batched(hm, orc, dtype):
    """Batched H X and H' X on leaves of rank 17..48 (several 16-column
    groups per leaf in every projection super-tile they cross)."""
    from tests.helpers import wide_rank_flat
    f = wide_rank_flat(hm, 2)
    X = np.random.default_rng(50).standard_normal((f.ncols, 32))
    assert relerr(_run(hm, f, dtype, X), orc.matmat(f, X)) <= TOL[dtype]
    V = np.random.default_rng(51).standard_normal((f.nrows, 32))
    ref = np.stack([orc.matvec_t(f, V[:, i]) for i in range(32)], 1)
    assert relerr(_run(hm, f, dtype, V, transpose=True), ref) <= TOL[dtype]


def test_two_dimensional_points(hm, orc):
    """3000 points in the unit square (Lloyd 2-means cluster tree),
    exp(-|x - y|) with the Euclidean distance, MaxRank 20: the leaves are
    less regular than in 1-D (ranks up to 20, uneven sizes).  FP64 H x and
    H' x against the oracle to 1e-13, batched split-BF16 H X to 2e-5."""
    import scipy.sparse as sp
    import torch
    rng = np.random.default_rng(3)
    par = hm.Params()
    par.Geom = rng.random((3000, 2))
    par.Kernel = lambda x, y: np.exp(-np.sqrt(((x - y) ** 2).sum(-1)))
    par.MaxBlock, par.MinBlock, par.MaxRank = 500, 64, 20
    par.ConsMethod = "aca2"
    par.Adj = sp.identity(3000, format="csr")
    c, H = hm.NewHmat(par)
    f = hm.flatten(H)
    x = rng.standard_normal(3000)
    d = hm.DeviceHmat(f, "f64")
    xd = torch.from_numpy(x).cuda()
    assert relerr(d.matvec(xd).cpu().numpy(), orc.matvec(f, x)) <= 1e-13
    assert relerr(d.matvec(xd, transpose=True).cpu().numpy(), orc.matvec_t(f, x)) <= 1e-13
    d.close()
    X = rng.standard_normal((3000, 32))
    assert relerr(_run(hm, f, "bf16x2", X), orc.matmat(f, X)) <= TOL["bf16x2"]


@pytest.mark.parametrize("minblock", [64, 128])
def test_strong_admissibility_mid_size(hm, orc, minblock):
    """N = 20000 Cauchy with strong (tridiagonal) admissibility: three dense
    leaves per block row; with MinBlock 128 the workgroups' X windows exceed
    the LDS window and the row pass reads X straight from memory."""
    X = sorted_uniform(20000, 0.0, 60.0, 21)
    par, c, H = build(hm, X, cauchy, 2048, minblock, 12, adj="strong")
    f = hm.flatten(H)
    R = np.random.default_rng(22).standard_normal((f.ncols, 32))
    assert relerr(_run(hm, f, "bf16x2", R), orc.matmat(f, R)) <= TOL["bf16x2"]
