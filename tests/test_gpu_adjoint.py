"""Adjoint of L = 0.5||Hx - b||^2 on the GPU: against the FP64 oracle's closed
form (<= 1e-13), the gradient of the same loss with the dense FP64 kernel
matrix A in place of H (kernel_full, src/cluster.jl:258-275), and central
finite differences with h = 1e-6 (FP64) evaluated through the GPU product
itself.

The central difference (L(p+h) - L(p-h)) / 2h is evaluated in the identical
form (r+ - r-).(r+ + r-) / 4h, with r+- the GPU residuals at p +- h: at
N = 65536 L is ~1e7, and subtracting the two losses directly would lose
~1e-3 of the difference to rounding."""
import numpy as np
import pytest

from tests.helpers import build, c1, cauchy, sorted_uniform, relerr, vjp_closed_form

pytestmark = pytest.mark.gpu


def _dev(hm, f):
    return hm.DeviceHmat(f, "f64")


def _cd(rp, rm, h):
    """(L(p+h) - L(p-h)) / 2h from the two residual vectors."""
    rp, rm = rp.cpu().numpy(), rm.cpu().numpy()
    return float((rp - rm) @ (rp + rm)) / (4 * h)


@pytest.fixture(scope="module")
def c5(hm):
    """N=65536 sorted uniform points on [0,100] (seed 7), Cauchy, MaxBlock 1024, MaxRank 16."""
    X = sorted_uniform(65536, 0.0, 100.0, 7)
    par, c, H = build(hm, X, cauchy, 1024, 64, 16)
    return np.asarray(c.X), hm.flatten(H)


# H approximates the Cauchy kernel matrix A of these points to ~1e-12
# (ACA to 1e-10 per block; measured on the oracle: ||H'r - A'r|| / ||A'r||
# = 1.0e-12 and ||(Hx - b) - (Ax - b)|| / ||Ax - b|| = 8.9e-12), so gradients
# of the H loss and of the dense loss agree to this bound.
DENSE_TOL = 1e-10


def _dense_apply(Xc, v):
    """A v with A_ij = 1/(1 + (x_i - x_j)^2), FP64, matrix-free in row
    blocks on the device (torch as an FP64 calculator, not the product)."""
    import torch
    xg = torch.from_numpy(Xc).cuda()
    vg = torch.from_numpy(v).cuda()
    out = torch.empty_like(xg)
    for s in range(0, xg.numel(), 2048):
        K = 1.0 / (1.0 + (xg[s:s + 2048, None] - xg[None, :]) ** 2)
        out[s:s + 2048] = K @ vg
    return out.cpu().numpy()


def test_adjoint_matches_oracle_c1(hm, orc):
    import torch
    par, c, H = c1(hm, "aca2")
    f = hm.flatten(H)
    rng = np.random.default_rng(8)
    x, b = rng.standard_normal(f.ncols), rng.standard_normal(f.nrows)
    d = _dev(hm, f)
    L, res, dx, g = d.adjoint(torch.from_numpy(x).cuda(), torch.from_numpy(b).cuda())
    Lr, resr, dxr, gr = orc.adjoint(f, x, b)
    assert abs(L - Lr) <= 1e-13 * Lr
    assert relerr(res.cpu().numpy(), resr) <= 1e-13
    assert relerr(dx.cpu().numpy(), dxr) <= 1e-13
    assert relerr(g.cpu().numpy(), gr) <= 1e-13


def test_adjoint_c5_oracle_dense_and_fd(hm, orc, c5):
    import torch
    Xc, f = c5
    rng = np.random.default_rng(8)
    x, b = rng.standard_normal(f.ncols), rng.standard_normal(f.nrows)
    d = _dev(hm, f)
    xd, bd = torch.from_numpy(x).cuda(), torch.from_numpy(b).cuda()
    L, res, dx, g = d.adjoint(xd, bd)
    Lr, resr, dxr, gr = orc.adjoint(f, x, b)
    dx, g = dx.cpu().numpy(), g.cpu().numpy()
    assert abs(L - Lr) <= 1e-13 * Lr                # 32 partial sums of 2048 residuals
    assert relerr(res.cpu().numpy(), resr) <= 1e-13
    assert relerr(dx, dxr) <= 1e-13 and relerr(g, gr) <= 1e-13
    # dense A' gradient on 64 sampled columns: dL/dx_j = sum_i A_ij (A x - b)_i
    # with A the FP64 kernel matrix (kernel_full columns), against the GPU's
    # H'(H x - b); also A' applied to the GPU's own residual
    cols = rng.choice(f.ncols, 64, replace=False)
    Acols = cauchy(Xc[:, None], Xc[cols][None, :])             # kernel_full columns
    res_dense = _dense_apply(Xc, x) - b
    e_full = relerr(dx[cols], Acols.T @ res_dense)
    e_res = relerr(dx[cols], Acols.T @ res.cpu().numpy())
    print(f"C5 dx vs dense A'(Ax - b): {e_full:.2e}; vs A' res: {e_res:.2e}")
    assert e_full <= DENSE_TOL and e_res <= DENSE_TOL
    # central differences in x through the GPU product (L is quadratic: exact up to rounding)
    h = 1e-6
    for j in cols[:8]:
        e = torch.zeros_like(xd)
        e[j] = h
        rp = d.adjoint(xd + e, bd, grad=False, loss=False)[1]
        rm = d.adjoint(xd - e, bd, grad=False, loss=False)[1]
        assert abs(_cd(rp, rm, h) - dx[j]) <= 1e-6 * max(1.0, abs(dx[j]))
    # central differences in factor entries (dense C, A, B) on sampled positions
    blocks = f.blocks
    picks = []
    for kind in (0, 1):
        idx = np.nonzero(blocks[:, 0] == kind)[0]
        for q in rng.choice(idx, 3, replace=False):
            kd, r0, c0, m, n, k, oa, ob = blocks[q]
            picks.append(oa + rng.integers(0, m * (n if kd == 0 else k)))
            if kd == 1:
                picks.append(ob + rng.integers(0, n * k))
    for p in picks:
        vals = []
        for sgn in (1, -1):
            data = f.data.copy()
            data[p] += sgn * h
            dp = hm.DeviceHmat(hm.FlatHmat(f.nrows, f.ncols, f.blocks, data), "f64")
            vals.append(dp.adjoint(xd, bd, grad=False, loss=False)[1])
            dp.close()
        fd = _cd(vals[0], vals[1], h)
        assert abs(fd - g[p]) <= 1e-6 * max(1.0, abs(g[p]))


def test_torch_autograd_x(hm, orc):
    import torch
    par, c, H = c1(hm, "aca2")
    f = hm.flatten(H)
    d = _dev(hm, f)
    x = torch.from_numpy(np.random.default_rng(3).standard_normal(f.ncols)).cuda().requires_grad_(True)
    y = hm.hmatvec(d, x)
    w = torch.from_numpy(np.random.default_rng(4).standard_normal(f.nrows)).cuda()
    (y * w).sum().backward()
    assert relerr(x.grad.cpu().numpy(), orc.matvec_t(f, w.cpu().numpy())) <= 1e-13


@pytest.mark.parametrize("dtype,tol", [("f64", 1e-13), ("f32", 1e-5)])
def test_adjoint_reused_buffers_ragged(hm, orc, dtype, tol):
    """N = 5003 (three uneven loss partitions, leaves of uneven size): a
    second call into the first call's (now dirty) output buffers, with other
    inputs, matches the oracle as well."""
    import torch
    X = sorted_uniform(5003, 0.0, 30.0, 14)
    par, c, H = build(hm, X, cauchy, 700, 48, 8)
    f = hm.flatten(H)
    d = hm.DeviceHmat(f, dtype)
    td = torch.float64 if dtype == "f64" else torch.float32
    rng = np.random.default_rng(15)
    out = None
    for _ in range(2):
        x, b = rng.standard_normal(f.ncols), rng.standard_normal(f.nrows)
        L, res, dx, g = d.adjoint(torch.from_numpy(x).to("cuda", td), torch.from_numpy(b).to("cuda", td),
                                  out=out)
        Lr, resr, dxr, gr = orc.adjoint(f, x, b)
        assert abs(L - Lr) <= tol * Lr
        assert relerr(res.cpu().numpy().astype(np.float64), resr) <= tol
        assert relerr(dx.cpu().numpy().astype(np.float64), dxr) <= tol
        assert relerr(g.cpu().numpy().astype(np.float64), gr) <= tol
        out = (res, dx, g)
    d.close()


def test_vjp_closed_form_and_central_difference(hm, orc):
    """hm_vjp on C1 (FP64): dx = H'g against the oracle, the factor gradient
    against the per-leaf closed form, and <grad, d> against the central
    difference of <g, H(p +- h d) x> through the oracle (exact for a
    bilinear A B': the h^2 terms cancel)."""
    import torch
    par, c, H = c1(hm, "aca2")
    f = hm.flatten(H)
    d = hm.DeviceHmat(f, "f64")
    rng = np.random.default_rng(21)
    x, g = rng.standard_normal(f.ncols), rng.standard_normal(f.nrows)
    dx, dg = d.vjp(torch.from_numpy(x).cuda(), torch.from_numpy(g).cuda())
    dx, dg = dx.cpu().numpy(), dg.cpu().numpy()
    assert relerr(dx, orc.matvec_t(f, g)) <= 1e-13
    ref = vjp_closed_form(f, x, g)
    assert relerr(dg, ref) <= 1e-13
    h = 1e-3
    for seed in range(3):
        dirn = np.random.default_rng(100 + seed).standard_normal(f.data.size)
        fp = hm.FlatHmat(f.nrows, f.ncols, f.blocks, f.data + h * dirn)
        fm = hm.FlatHmat(f.nrows, f.ncols, f.blocks, f.data - h * dirn)
        cd = float(g @ (orc.matvec(fp, x) - orc.matvec(fm, x))) / (2 * h)
        assert abs(float(dg @ dirn) - cd) <= 1e-10 * abs(cd), seed
    d.close()


@pytest.mark.parametrize("dtype,tol", [("f64", 1e-13), ("f32", 2e-5)])
def test_torch_autograd_factors(hm, orc, dtype, tol):
    """hmatvec(dev, x, data): backward gives x.grad = H'w and data.grad = the
    factor tangent (closed form), for L = <w, H x>; without data the arena
    gets no gradient and x.grad is unchanged."""
    import torch
    par, c, H = c1(hm, "aca2")
    f = hm.flatten(H)
    d = hm.DeviceHmat(f, dtype)
    rng = np.random.default_rng(22)
    xn, wn = rng.standard_normal(f.ncols), rng.standard_normal(f.nrows)
    td = d.vec_dtype
    x = torch.from_numpy(xn).to("cuda", td).requires_grad_(True)
    data = torch.from_numpy(f.data.copy()).cuda().requires_grad_(True)
    w = torch.from_numpy(wn).to("cuda", td)
    (hm.hmatvec(d, x, data) * w).sum().backward()
    xs = x.detach().double().cpu().numpy()
    assert relerr(x.grad.double().cpu().numpy(), orc.matvec_t(f, wn)) <= tol
    assert data.grad.dtype == torch.float64
    assert relerr(data.grad.cpu().numpy(), vjp_closed_form(f, xs, wn.astype(np.float64))) <= tol
    x2 = torch.from_numpy(xn).to("cuda", td).requires_grad_(True)
    (hm.hmatvec(d, x2) * w).sum().backward()
    assert torch.equal(x2.grad, x.grad)
    d.close()


def test_vjp_argument_errors(hm):
    import torch
    from hmatrix_amd._native import HmError
    par, c, H = c1(hm, "aca2")
    f = hm.flatten(H)
    d = hm.DeviceHmat(f, "f64", {"no_transpose": True})
    x = torch.zeros(f.ncols, dtype=torch.float64, device="cuda")
    with pytest.raises(HmError):
        d.vjp(x, torch.zeros(f.nrows, dtype=torch.float64, device="cuda"))
    d.close()
    d = hm.DeviceHmat(f, "f64")
    with pytest.raises(ValueError):
        hm.hmatvec(d, x, torch.zeros(3, dtype=torch.float64, device="cuda"))
    d.close()
