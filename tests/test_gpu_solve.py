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
