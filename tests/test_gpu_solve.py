"""Solvers on the GPU product: the reference's solve test (test/lu.jl:17-32:
H += 10I, H\\y ≈ dense\\y) through iterative refinement, and the Krylov
wrappers of src/tools.jl."""
import numpy as np
import pytest

from tests.helpers import c1, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shifted(hm):
    par, c, H = c1(hm, "svd")
    G = H + 10 * hm.I
    A = hm.FullMat(par.Kernel, c.X, c.X) + 10 * np.eye(2000)
    return G, A


def test_solve_ir_matches_dense(hm, shifted):
    G, A = shifted
    y = np.random.default_rng(0).random(2000)
    x, hist = hm.solve_ir(G, y, tol=1e-12, inner="f32")
    assert hist[-1] <= 1e-12
    assert relerr(x, np.linalg.solve(A, y)) <= 1e-10      # H approximates A to ~1e-13


@pytest.mark.parametrize("inner", ["bf16x2", "bf16"])
def test_solve_ir_low_precision_inner(hm, shifted, inner):
    G, A = shifted
    y = np.random.default_rng(1).standard_normal(2000)
    x, hist = hm.solve_ir(G, y, tol=1e-12, inner=inner)
    assert hist[-1] <= 1e-12 and relerr(x, np.linalg.solve(A, y)) <= 1e-10


def test_krylov_wrappers(hm, shifted):
    G, A = shifted
    y = np.random.default_rng(2).random(2000)
    x0 = np.linalg.solve(A, y)
    seen = []
    x = hm.gmres(G, y, tol=1e-10, callback=lambda r: seen.append(r))
    assert relerr(x, x0) <= 1e-8 and len(seen) > 0
    assert relerr(hm.bicgstab(G, y, tol=1e-10), x0) <= 1e-8


def test_solve_ir_hlu_reference_case(hm):
    """test/lu.jl:17-32 (C1, MaxRank 25, H + 10 I, H\\y ~ dense\\y) with the
    reference's own correction: FP64 residuals on the GPU, corrections from
    the host H-LU factors (lu!, src/hlu.jl:465-515, and \\, 553-561).  The
    H-LU is exact up to its 1e-10 recompressions, so one or two sweeps reach
    the bound."""
    from tests.helpers import build, cauchy
    par, c, H = build(hm, np.linspace(0.0, 10.0, 2000), cauchy, 500, 64, 25)
    G = H + 10 * hm.I
    y = np.random.default_rng(3).random(2000)
    x, hist = hm.solve_ir(G, y, tol=1e-12, inner="hlu")
    dense = np.linalg.solve(hm.to_fmat(G), y)
    print(f"H-LU refinement: residual history {hist}")
    assert hist[-1] <= 1e-12 and len(hist) - 1 <= 3
    assert relerr(x, dense) <= 1e-10


def test_hlu_preconditioner_against_plain_gmres(hm):
    """C1 + 0.01 I (condition number ~5e4): GMRES on the GPU product needs
    at least 5x the iterations without the H-LU preconditioner (measured on
    the host: 45 against 2), and H-LU refinement still reaches 1e-12."""
    par, c, H = c1(hm, "aca2")
    G = H + 0.01 * hm.I
    y = np.random.default_rng(4).random(2000)
    plain, prec = [], []
    x0 = hm.gmres(G, y, tol=1e-10, restart=50, maxiter=20, callback=lambda r: plain.append(r))
    x1 = hm.gmres(G, y, tol=1e-10, restart=50, maxiter=20, M=G, callback=lambda r: prec.append(r))
    print(f"GMRES iterations: {len(plain)} plain, {len(prec)} with the H-LU preconditioner")
    assert len(plain) >= 5 * len(prec)
    ref = np.linalg.solve(hm.to_fmat(G), y)
    assert relerr(x1, ref) <= 1e-7 and relerr(x0, ref) <= 1e-5
    x, hist = hm.solve_ir(G, y, tol=1e-12, inner="hlu")
    assert hist[-1] <= 1e-12 and relerr(x, ref) <= 1e-8
