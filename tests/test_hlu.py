"""Host H-LU (hmatrix_amd.hlu, FP64) against the reference's own LU tests
(test/lu.jl) and the dense arithmetic it approximates.

* "lu" (test/lu.jl:2-15): C1 points, MaxRank 25; max |L U - H[P, :]| < 1e-8
* "solve" (test/lu.jl:17-32): H + 10 I; H \\ y equals the dense solve of the
  same matrix to Julia's isapprox (rtol sqrt(eps))
* "lu2" (test/lu.jl:34-49): 50 x 50 grid, kernel 1/(100 + |x - y|^2),
  MaxBlock 256, MinBlock 64, MaxRank 3, H + 10 I; max |L U - H[P, :]| < 1e-8
* "solve2-dense" (test/lu.jl:51-73): that matrix's H \\ y against the dense
  kernel matrix's solve (the reference prints the error; the bound here is
  the H-matrix approximation error)
* hmul / hadd against the products and sums of the dense blocks (closed
  form up to the recompression cut 1e-10), so a wrong block, transposed
  factor or sign fails
"""
import numpy as np
import pytest

from tests.helpers import build, c1, cauchy, relerr

SQRT_EPS = np.sqrt(np.finfo(np.float64).eps)


def _grid_case(hm):
    x = np.linspace(0.0, 10.0, 50)
    Xg, Yg = np.meshgrid(x, x)
    X = np.stack([Xg.ravel(), Yg.ravel()], 1)
    kern = lambda a, b: 1.0 / (100.0 + np.sum((a - b) ** 2, axis=-1))
    return build(hm, X, kern, 256, 64, 3)


def test_lu_reference_case(hm):
    from hmatrix_amd import hlu
    par, c, H = build(hm, np.linspace(0.0, 10.0, 2000), cauchy, 500, 64, 25)
    LU = hm.lu(H)
    D = hm.to_fmat(H)
    F = hm.to_fmat(LU)
    assert np.abs(hlu.getl(F) @ hlu.getu(F) - D[LU.P, :]).max() < 1e-8
    assert len(H.P) == 0                      # lu works on a copy


def test_solve_reference_case(hm):
    par, c, H = build(hm, np.linspace(0.0, 10.0, 2000), cauchy, 500, 64, 25)
    G = H + 10 * hm.I
    y = np.random.default_rng(0).random(2000)
    x1 = np.linalg.solve(hm.to_fmat(G), y)
    hm.lu_inplace(G)
    x2 = hm.hsolve(G, y)
    assert relerr(x2, x1) <= SQRT_EPS


def test_lu2_grid_case(hm):
    from hmatrix_amd import hlu
    par, c, H = _grid_case(hm)
    G = H + 10 * hm.I
    LU = hm.lu(G)
    D = hm.to_fmat(G)
    F = hm.to_fmat(LU)
    assert np.abs(hlu.getl(F) @ hlu.getu(F) - D[LU.P, :]).max() < 1e-8


def test_solve2_dense_kernel(hm):
    par, c, H = _grid_case(hm)
    G = H + 10 * hm.I
    Xc = np.asarray(c.X)
    dH = par.Kernel(Xc[:, None, :], Xc[None, :, :]) + 10 * np.eye(Xc.shape[0])
    y = np.random.default_rng(1).random(Xc.shape[0])
    x1 = np.linalg.solve(dH, y)
    x2 = hm.hsolve(hm.lu(G), y)
    x3 = np.linalg.solve(hm.to_fmat(G), y)
    # the LU solves the H-matrix to rounding; both differ from the dense
    # kernel's solve by the H-matrix approximation error
    assert relerr(x2, x3) <= 1e-10
    assert np.linalg.norm(x1 - x2) / np.linalg.norm(y) <= 10 * np.linalg.norm(x1 - x3) / np.linalg.norm(y) + 1e-12


def test_hmul_hadd_dense_closed_form(hm):
    """hmul of every block pair of the C1 tree (dense x dense, dense x
    low-rank, low-rank x H, H x H, ...) and hadd against the dense blocks."""
    par, c, H = build(hm, np.linspace(0.0, 10.0, 600), cauchy, 150, 32, 8)
    blocks = [H, H.children[0][0], H.children[0][1], H.children[1][0], H.children[1][1],
              H.children[0][0].children[0][0], H.children[0][0].children[0][1]]
    npairs = 0
    for a in blocks:
        for b in blocks:
            if not a.t.same_range(b.s):        # hmul pairs blocks over one cluster
                continue
            npairs += 1
            P = hm.hmul(a, b)
            ref = hm.to_fmat(a) @ hm.to_fmat(b)
            assert (P.m, P.n) == ref.shape
            assert relerr(hm.to_fmat(P), ref) <= 1e-9, (a.shape, b.shape)
    assert npairs >= 10
    S = hm.hadd(H, H)
    assert relerr(hm.to_fmat(S), 2 * hm.to_fmat(H)) <= 1e-9     # recompressed sums, cut 1e-10


def test_factorized_solve_rejects_and_reuses(hm):
    """hsolve needs factors; solve_ir refuses a matrix that holds its own
    factors (its residual needs the matrix)."""
    par, c, H = c1(hm, "aca2")
    G = H + 10 * hm.I
    with pytest.raises(ValueError):
        hm.hsolve(G, np.ones(2000))
    F = hm.factorized(G)
    assert hm.factorized(G) is F and len(G.P) == 0
    hm.lu_inplace(G)
    with pytest.raises(ValueError):
        hm.solve_ir(G, np.ones(2000), inner="hlu")
