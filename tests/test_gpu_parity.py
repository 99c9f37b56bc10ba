"""GPU parity of the FP64 plain product against the FP64 oracle (<= 1e-13 relative)."""
import numpy as np
import pytest

from tests.helpers import c1, build, cauchy, expk, sorted_uniform, relerr

pytestmark = pytest.mark.gpu

TOL64 = 1e-13


@pytest.fixture(scope="module")
def c1_aca(hm):
    par, c, H = c1(hm, "aca2")
    return H, hm.flatten(H)


def test_c1_matvec_f64(hm, orc, c1_aca):
    H, f = c1_aca
    y = np.random.default_rng(0).random(f.ncols)
    got = H * y
    assert relerr(got, orc.matvec(f, y)) <= TOL64


def test_c1_alpha_beta_accumulate(hm, orc, c1_aca):
    import torch
    H, f = c1_aca
    d = H.device("f64")
    rng = np.random.default_rng(1)
    x, r0 = rng.standard_normal(f.ncols), rng.standard_normal(f.nrows)
    xd = torch.from_numpy(x).cuda()
    yd = torch.from_numpy(r0).cuda()
    d.matvec(xd, yd, alpha=-1.0, beta=1.0)          # hmat_matvec!(r, H, v, -1.0)
    ref = orc.matvec(f, x, r=r0, s=-1.0)
    assert relerr(yd.cpu().numpy(), ref) <= TOL64
    yd2 = torch.full((f.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    d.matvec(xd, yd2, alpha=2.5, beta=0.0)           # beta = 0 never reads y
    assert relerr(yd2.cpu().numpy(), orc.matvec(f, x, s=2.5)) <= TOL64


def test_c1_transpose_f64(hm, orc, c1_aca):
    H, f = c1_aca
    v = np.random.default_rng(2).standard_normal((3, f.nrows))
    got = v * H
    for i in range(3):
        assert relerr(got[i], orc.matvec_t(f, v[i])) <= TOL64


def test_c1_matmat_f64(hm, orc, c1_aca):
    H, f = c1_aca
    X = np.random.default_rng(3).standard_normal((f.ncols, 5))
    assert relerr(H * X, orc.matmat(f, X)) <= TOL64


def test_c2_scale_f64(hm, orc):
    """N=65536 sorted uniform points on [0,100], Cauchy, MaxBlock 1024, MaxRank 16."""
    X = sorted_uniform(65536, 0.0, 100.0, 1)
    par, c, H = build(hm, X, cauchy, 1024, 64, 16)
    f = hm.flatten(H)
    y = np.random.default_rng(2).standard_normal(f.ncols)
    got = H * y
    assert relerr(got, orc.matvec(f, y)) <= TOL64


@pytest.mark.parametrize("dtype,tol", [("f32", 1e-6), ("bf16x2", 2e-5), ("bf16", 1e-2)])
def test_c1_storage_dtypes(hm, orc, c1_aca, dtype, tol):
    import torch
    H, f = c1_aca
    d = hm.DeviceHmat(f, dtype)
    x = np.random.default_rng(4).standard_normal(f.ncols)
    y = d.matvec(torch.from_numpy(x).float().cuda()).cpu().numpy()
    assert relerr(y, orc.matvec(f, x)) <= tol


@pytest.mark.parametrize("seed", [0, 1])
def test_wide_ranks_f64_and_f32(hm, orc, seed):
    """Leaves of rank 17..48 (several 16-column projection groups, each with
    its own compiled column count) and B factors longer than one projection
    chunk: H x, H' x and the adjoint in FP64 (<= 1e-13), H x in FP32."""
    import torch
    from tests.helpers import wide_rank_flat
    f = wide_rank_flat(hm, seed)
    rng = np.random.default_rng(seed + 40)
    x, v = rng.standard_normal(f.ncols), rng.standard_normal(f.nrows)
    d = hm.DeviceHmat(f, "f64")
    xd, vd = torch.from_numpy(x).cuda(), torch.from_numpy(v).cuda()
    assert relerr(d.matvec(xd).cpu().numpy(), orc.matvec(f, x)) <= TOL64
    assert relerr(d.matvec(vd, transpose=True).cpu().numpy(), orc.matvec_t(f, v)) <= TOL64
    L, res, dx, g = d.adjoint(xd, vd)
    Lr, resr, dxr, gr = orc.adjoint(f, x, v)
    assert abs(L - Lr) <= TOL64 * Lr
    assert relerr(dx.cpu().numpy(), dxr) <= TOL64 and relerr(g.cpu().numpy(), gr) <= TOL64
    d.close()
    d32 = hm.DeviceHmat(f, "f32")
    y32 = d32.matvec(torch.from_numpy(x).float().cuda()).cpu().numpy()
    assert relerr(y32, orc.matvec(f, x)) <= 1e-6
    d32.close()


def test_product_argument_errors(hm, c1_aca):
    """Entry points reject bad arguments with a negative status and a message,
    before any kernel runs: overlapping x / y, short leading dimensions,
    negative nrhs, panels whose extent overflows the address range,
    non-finite alpha, and adjoint outputs overlapping its
    inputs or each other."""
    import ctypes
    import torch
    from hmatrix_amd import _native
    H, f = c1_aca
    L = _native.lib()
    d = hm.DeviceHmat(f, "f64")
    n = f.nrows
    buf = torch.zeros(4 * n, dtype=torch.float64, device="cuda")
    P = lambda off: ctypes.c_void_p(buf.data_ptr() + 8 * off)

    def err(rc, text):
        assert rc < 0 and text in L.hm_last_error().decode()

    err(L.hm_matvec(d._h, P(0), P(n // 2), 1.0, 0.0, None), "overlap")
    err(L.hm_matmat(d._h, 2, P(0), n - 1, P(2 * n), n, 1.0, 0.0, None), "leading dimension")
    err(L.hm_matmat(d._h, -1, P(0), n, P(2 * n), n, 1.0, 0.0, None), "nrhs")
    err(L.hm_matmat(d._h, 1 << 40, P(0), 1 << 40, P(2 * n), n, 1.0, 0.0, None), "address range")
    err(L.hm_matvec(d._h, P(0), P(2 * n), float("nan"), 0.0, None), "finite")
    err(L.hm_adjoint(d._h, P(0), P(n), P(n // 2), P(3 * n), None, None, None), "overlap")    # res over x
    err(L.hm_adjoint(d._h, P(0), P(n), P(2 * n), P(2 * n + 7), None, None, None), "overlap")  # dx over res
    err(L.hm_adjoint(d._h, P(0), P(n), P(2 * n), P(3 * n), ctypes.c_void_p(buf.data_ptr()), None, None),
        "grad")
    # the plan still works after the rejected calls
    x = torch.from_numpy(np.random.default_rng(6).standard_normal(n)).cuda()
    assert np.isfinite(d.matvec(x).cpu().numpy()).all()
    d.close()


@pytest.mark.parametrize("dtype,tol", [("f64", TOL64), ("f32", 1e-6), ("bf16x2", 2e-5)])
@pytest.mark.parametrize("nrhs", [2, 8, 13])
def test_multi_rhs_single_layout(hm, orc, dtype, tol, nrhs):
    """Several right-hand sides on the single-right-hand-side layout (panels of
    eight columns per pass over the factors; the FP64 H*X path, and FP32 /
    split-BF16 plans built without the batched layout): H X, H' X and
    Y = beta Y + alpha H X with padded leading dimensions, against the oracle
    column by column.  Leaves of rank up to 48 (a projection item takes
    its rank columns 4 at a time, re-reading its 8-column X slice for each
    group) and multi-chunk B factors."""
    import torch
    from tests.helpers import wide_rank_flat
    f = wide_rank_flat(hm, nrhs)
    d = hm.DeviceHmat(f, dtype, {"no_batched": True})
    td = torch.float64 if dtype == "f64" else torch.float32
    rng = np.random.default_rng(nrhs)
    X = rng.standard_normal((f.ncols, nrhs))
    Y0 = rng.standard_normal((f.nrows, nrhs))
    Xb = torch.zeros((nrhs, f.ncols + 3), dtype=td, device="cuda")
    Yb = torch.zeros((nrhs, f.nrows + 5), dtype=td, device="cuda")
    Xd, Yd = Xb[:, :f.ncols].t(), Yb[:, :f.nrows].t()
    Xd.copy_(torch.from_numpy(X).to(td))
    Yd.copy_(torch.from_numpy(Y0).to(td))
    d.matmat(Xd, Yd, alpha=0.75, beta=-2.0)
    ref = -2.0 * Y0 + 0.75 * orc.matmat(f, X)
    assert relerr(Yd.cpu().numpy().astype(np.float64), ref) <= tol
    V = rng.standard_normal((f.nrows, nrhs))
    Vd = d.new_cols(f.nrows, nrhs)
    Vd.copy_(torch.from_numpy(V).to(td))
    got = d.matmat(Vd, transpose=True).cpu().numpy().astype(np.float64)
    ref_t = np.stack([orc.matvec_t(f, V[:, i]) for i in range(nrhs)], 1)
    assert relerr(got, ref_t) <= tol
    d.close()
