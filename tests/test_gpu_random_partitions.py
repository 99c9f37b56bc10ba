"""GPU parity on seeded random exact tilings of rectangular matrices
(tests/helpers.py random_partition_flat): leaves of every size from single
rows or columns to 400, dense or low rank with ranks 1-40, nrows != ncols,
neither a multiple of the 16-row strip or the 8-column octet.  Every storage
(FP64 <= 1e-13, FP32 and split BF16 <= 2e-5, BF16 <= 1e-2 relative) runs
H*X and H'X against the FP64 oracle for 1, 7 and 33 right-hand sides (one
column, a partial panel, a full panel plus one), y = 2 H X - 0.5 Y and
w = -1.5 H'Z + 0.25 W once.
The reference's product accepts any such tiling (src/harithm.jl:340-354
walks whatever leaves the tree holds)."""
import numpy as np
import pytest

from tests.helpers import random_partition_flat, vjp_closed_form

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-13, "f32": 2e-5, "bf16x2": 2e-5, "bf16": 1e-2}


def _relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# (seed, leaf side bounds, dense share, largest rank): the default mix, many
# small leaves of low rank, and large leaves that span many 256-column
# projection super-tiles and 16-row strips
SHAPES = [(101, (24, 400), 0.4, 40), (102, (24, 400), 0.4, 40), (103, (24, 400), 0.4, 40),
          (104, (24, 400), 0.4, 40), (105, (8, 128), 0.7, 8), (106, (64, 1500), 0.2, 40)]


@pytest.mark.parametrize("seed,side,p_dense,kmax", SHAPES)
def test_random_partition_products(hm, orc, seed, side, p_dense, kmax):
    import torch
    rng = np.random.default_rng(seed)
    nr, nc = int(rng.integers(300, 3000)), int(rng.integers(300, 3000))
    f = random_partition_flat(hm, seed, nr, nc, side, p_dense, kmax)
    print(f"seed {seed}: {nr} x {nc}, {f.nblocks} leaves", flush=True)
    data = np.random.default_rng(seed + 1000)
    for dtype, tol in TOL.items():
        d = hm.DeviceHmat(f, dtype)
        assert d.info()["batched"] == 1
        for nrhs in (1, 7, 33):
            X = data.standard_normal((nc, nrhs))
            Xd = d.new_cols(nc, nrhs)
            Xd.copy_(torch.from_numpy(X).to(Xd.dtype))
            Y = d.matmat(Xd).cpu().numpy().astype(np.float64)
            assert _relerr(Y, orc.matmat(f, X)) <= tol, (dtype, nrhs, "H")
            Z = data.standard_normal((nr, nrhs))
            Zd = d.new_cols(nr, nrhs)
            Zd.copy_(torch.from_numpy(Z).to(Zd.dtype))
            W = d.matmat(Zd, transpose=True).cpu().numpy().astype(np.float64)
            ref = np.stack([orc.matvec_t(f, Z[:, i]) for i in range(nrhs)], 1)
            assert _relerr(W, ref) <= tol, (dtype, nrhs, "H'")
        X = data.standard_normal((nc, 32))
        Y0 = data.standard_normal((nr, 32))
        Xd = d.new_cols(nc, 32)
        Xd.copy_(torch.from_numpy(X).to(Xd.dtype))
        Yd = d.new_cols(nr, 32)
        Yd.copy_(torch.from_numpy(Y0).to(Yd.dtype))
        d.matmat(Xd, Yd, alpha=2.0, beta=-0.5)
        ref = 2.0 * orc.matmat(f, X) - 0.5 * Y0
        assert _relerr(Yd.cpu().numpy().astype(np.float64), ref) <= 2 * tol, (dtype, "alpha/beta")
        Z = data.standard_normal((nr, 32))
        W0 = data.standard_normal((nc, 32))
        Zd = d.new_cols(nr, 32)
        Zd.copy_(torch.from_numpy(Z).to(Zd.dtype))
        Wd = d.new_cols(nc, 32)
        Wd.copy_(torch.from_numpy(W0).to(Wd.dtype))
        d.matmat(Zd, Wd, alpha=-1.5, beta=0.25, transpose=True)
        ref = -1.5 * np.stack([orc.matvec_t(f, Z[:, i]) for i in range(32)], 1) + 0.25 * W0
        assert _relerr(Wd.cpu().numpy().astype(np.float64), ref) <= 2 * tol, (dtype, "alpha/beta H'")
        d.close()


@pytest.mark.parametrize("seed,side,p_dense,kmax", SHAPES)
def test_random_partition_vector_forms(hm, orc, seed, side, p_dense, kmax):
    """The single-right-hand-side entry points on the same tables: hm_matvec
    and hm_matvec_t with alpha/beta (FP64 <= 1e-13, FP32 <= 2e-6), and in
    FP64 the adjoint (loss, residual, dx = H'(Hx - b) and the factor
    gradients against the oracle's closed form) and hm_vjp (H'g and the
    gradient of <g, Hx> against the per-leaf closed form), all <= 1e-13."""
    import torch
    rng = np.random.default_rng(seed)
    nr, nc = int(rng.integers(300, 3000)), int(rng.integers(300, 3000))
    f = random_partition_flat(hm, seed, nr, nc, side, p_dense, kmax)
    data = np.random.default_rng(seed + 2000)
    x, z = data.standard_normal(nc), data.standard_normal(nr)
    y0, w0 = data.standard_normal(nr), data.standard_normal(nc)
    for dtype, tol in (("f64", 1e-13), ("f32", 2e-6)):
        d = hm.DeviceHmat(f, dtype)
        dt = torch.float64 if dtype == "f64" else torch.float32
        y = d.matvec(torch.from_numpy(x).to(dt).cuda()).cpu().numpy().astype(np.float64)
        assert _relerr(y, orc.matvec(f, x)) <= tol, (dtype, "H")
        yd = torch.from_numpy(y0).to(dt).cuda()
        d.matvec(torch.from_numpy(x).to(dt).cuda(), yd, alpha=-1.5, beta=0.5)
        assert _relerr(yd.cpu().numpy().astype(np.float64), 0.5 * y0 - 1.5 * orc.matvec(f, x)) <= 2 * tol
        wd = torch.from_numpy(w0).to(dt).cuda()
        d.matvec(torch.from_numpy(z).to(dt).cuda(), wd, alpha=2.0, beta=-1.0, transpose=True)
        assert _relerr(wd.cpu().numpy().astype(np.float64), 2.0 * orc.matvec_t(f, z) - w0) <= 2 * tol
        d.close()
    d = hm.DeviceHmat(f, "f64")
    xd, zd = torch.from_numpy(x).cuda(), torch.from_numpy(z).cuda()
    L, res, dx, g = d.adjoint(xd, zd)
    Lr, resr, dxr, gr = orc.adjoint(f, x, z)
    assert abs(L - Lr) <= 1e-13 * Lr
    assert _relerr(res.cpu().numpy(), resr) <= 1e-13
    assert _relerr(dx.cpu().numpy(), dxr) <= 1e-13
    assert _relerr(g.cpu().numpy(), gr) <= 1e-13
    dxv, dg = d.vjp(xd, zd)
    assert _relerr(dxv.cpu().numpy(), orc.matvec_t(f, z)) <= 1e-13
    assert _relerr(dg.cpu().numpy(), vjp_closed_form(f, x, z)) <= 1e-13
    d.close()


@pytest.mark.parametrize("one", [False, True])
def test_random_partition_batch(hm, orc, one):
    """A many-matrix plan over six random tilings of different shapes (FP32,
    all members per launch; three launches or the one-launch form with its
    per-member counters): every member's slice of H x and of H' z against
    the oracle (<= 2e-6), over products in a row."""
    import torch
    members = []
    for i, (seed, side, p_dense, kmax) in enumerate(SHAPES):
        rng = np.random.default_rng(seed + 500)
        nr, nc = int(rng.integers(50, 1500)), int(rng.integers(50, 1500))
        members.append(random_partition_flat(hm, seed + 500, nr, nc, side, p_dense, kmax))
    d = hm.DeviceHmat.batch(members, "f32", {"one_launch": one})
    assert d.info()["fused"] == int(one)
    data = np.random.default_rng(77)
    for rep in range(2):
        xs = [data.standard_normal(m.ncols) for m in members]
        zs = [data.standard_normal(m.nrows) for m in members]
        y = d.matvec(torch.from_numpy(np.concatenate(xs)).float().cuda()).cpu().numpy()
        w = d.matvec(torch.from_numpy(np.concatenate(zs)).float().cuda(), transpose=True).cpu().numpy()
        for i, m in enumerate(members):
            yi = y[d.row_offsets[i]:d.row_offsets[i + 1]].astype(np.float64)
            wi = w[d.col_offsets[i]:d.col_offsets[i + 1]].astype(np.float64)
            assert _relerr(yi, orc.matvec(m, xs[i])) <= 2e-6, (rep, i, "H")
            assert _relerr(wi, orc.matvec_t(m, zs[i])) <= 2e-6, (rep, i, "H'")
    assert d.info()["wait_faults"] == 0
    d.close()

