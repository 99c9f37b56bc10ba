"""Solvers driving the GPU product.

* ``gmres`` / ``bicgstab``: src/tools.jl:7-65 -- SciPy's Krylov solvers on a
  ``LinearOperator`` whose matvec is the H-matrix product (here the HIP
  product instead of the reference's recursive CPU ``A*x``); same arguments,
  same warnings on breakdown / non-convergence.
* ``solve_ir``: ``H\\b`` by mixed-precision iterative refinement.  The
  residual ``r = b - H x`` is accumulated in FP64 on the GPU (the plain FP64
  product); corrections come either from the host H-LU factorisation in FP64
  (``inner="hlu"``: the reference's ``lu!`` and ``\\``, src/hlu.jl:465-561,
  module ``hlu``) or from restarted GMRES on the device (basis vectors in
  HBM) over a low-precision copy of H (FP32 or split-BF16 storage).
* ``M`` (preconditioner) of ``gmres`` / ``bicgstab`` may be an ``Hmat``: its
  host H-LU factors (computed once if it is not factorised yet) apply
  ``M\\x``.
"""
from __future__ import annotations

import warnings
from typing import Callable, Optional

import numpy as np

from . import hlu
from .cluster import Hmat


def _operator(A, N: int):
    import scipy.sparse.linalg as ssl
    if isinstance(A, np.ndarray):
        return ssl.aslinearoperator(A)
    if isinstance(A, Hmat):
        dev = A.device("f64")
        return ssl.LinearOperator((N, N), matvec=lambda x: dev.mul(np.asarray(x, dtype=np.float64).ravel()),
                                  dtype=np.float64)
    if hasattr(A, "mul") and hasattr(A, "nrows"):        # DeviceHmat
        return ssl.LinearOperator((N, N), matvec=lambda x: A.mul(np.asarray(x, dtype=np.float64).ravel()),
                                  dtype=np.float64)
    if callable(A):
        return ssl.LinearOperator((N, N), matvec=A, dtype=np.float64)
    raise TypeError(f"unsupported operator type {type(A)}")


def _precond(M, N: int):
    import scipy.sparse.linalg as ssl
    if M is None:
        return None
    if isinstance(M, np.ndarray):
        return ssl.LinearOperator((N, N), matvec=lambda x: np.linalg.solve(M, x), dtype=np.float64)
    if isinstance(M, Hmat):
        LU = factorized(M)
        return ssl.LinearOperator((N, N), matvec=lambda x: hlu.solve(LU, np.asarray(x, dtype=np.float64).ravel()),
                                  dtype=np.float64)
    return ssl.LinearOperator((N, N), matvec=M, dtype=np.float64)


def factorized(H: Hmat) -> Hmat:
    """H itself when it holds its H-LU factors (``lu!`` was applied), else
    the factors of a copy, computed once and kept with H."""
    if len(H.P) > 0:
        return H
    LU = getattr(H, "_hlu", None)
    if LU is None:
        LU = hlu.lu(H)
        H._hlu = LU
    return LU


def _report(info: int) -> None:
    if info < 0:
        warnings.warn("illegal input or breakdown")
    elif info > 0:
        warnings.warn("convergence to tolerance not achieved, number of iterations")


def gmres(A, b, x0=None, tol: float = 1e-5, restart: Optional[int] = None, maxiter: Optional[int] = None,
          M=None, callback: Optional[Callable] = None):
    """GMRES on the H-matrix product (src/tools.jl:7-35)."""
    import scipy.sparse.linalg as ssl
    b = np.asarray(b, dtype=np.float64)
    N = b.shape[0]
    y, info = ssl.gmres(_operator(A, N), b, x0=x0, rtol=tol, restart=restart, maxiter=maxiter,
                        M=_precond(M, N), callback=callback, callback_type="pr_norm" if callback else None)
    _report(info)
    return y


def bicgstab(A, b, x0=None, tol: float = 1e-5, maxiter: Optional[int] = None, M=None,
             callback: Optional[Callable] = None):
    """BiCGSTAB on the H-matrix product (src/tools.jl:36-65)."""
    import scipy.sparse.linalg as ssl
    b = np.asarray(b, dtype=np.float64)
    N = b.shape[0]
    y, info = ssl.bicgstab(_operator(A, N), b, x0=x0, rtol=tol, maxiter=maxiter, M=_precond(M, N),
                           callback=callback)
    _report(info)
    return y


def _gmres_device(mv, b, tol: float, restart: int, maxiter: int):
    """Restarted GMRES on the device: ``mv`` maps a device vector to a new
    device vector of the same dtype.  Arnoldi with classical Gram-Schmidt
    applied twice (two matrix-vector products with the basis per step); the
    (restart+1) x restart Hessenberg matrix and its Givens rotations live on
    the host, one column (restart+1 numbers) read back per step.  Stops when
    the residual estimate reaches ``tol * ||b||``.  Returns ``(x, converged)``.
    """
    import torch
    from scipy.linalg import solve_triangular
    n = b.shape[0]
    x = torch.zeros_like(b)
    bnorm = float(torch.linalg.norm(b))
    if bnorm == 0.0:
        return x, True
    target = tol * bnorm
    V = torch.empty((restart + 1, n), dtype=b.dtype, device=b.device)
    for _ in range(maxiter):
        r = b - mv(x) if _ else b.clone()
        beta = float(torch.linalg.norm(r))
        if beta <= target:
            return x, True
        V[0] = r / beta
        H = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        k = restart
        for j in range(restart):
            w = mv(V[j])
            Vj = V[: j + 1]
            h = Vj @ w
            w = w - Vj.T @ h
            h2 = Vj @ w
            w = w - Vj.T @ h2
            hn = torch.linalg.norm(w)
            col = torch.cat([h + h2, hn.reshape(1)]).double().cpu().numpy()
            H[: j + 2, j] = col
            if col[-1] > 0:
                V[j + 1] = w / hn
            for i in range(j):
                a, c = H[i, j], H[i + 1, j]
                H[i, j], H[i + 1, j] = cs[i] * a + sn[i] * c, -sn[i] * a + cs[i] * c
            d = float(np.hypot(H[j, j], H[j + 1, j]))
            cs[j], sn[j] = (H[j, j] / d, H[j + 1, j] / d) if d > 0 else (1.0, 0.0)
            H[j, j], H[j + 1, j] = d, 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            if abs(g[j + 1]) <= target or col[-1] == 0:
                k = j + 1
                break
        y = solve_triangular(H[:k, :k], g[:k])
        x += V[:k].T @ torch.from_numpy(y).to(b.device, b.dtype)
        if abs(g[k]) <= target:
            return x, True
    return x, float(torch.linalg.norm(b - mv(x))) <= target


def solve_ir(H, b, tol: float = 1e-12, inner: str = "f32", inner_tol: float = 1e-4, maxiter: int = 30,
             restart: int = 50):
    """Solve ``H x = b`` by iterative refinement.

    Each sweep: ``r = b - H x`` with the FP64 GPU product, then a correction
    ``d ~ H \\ r`` and ``x += d``.  ``inner="hlu"``: d from the host FP64
    H-LU factors of H (``lu!`` and ``\\``, src/hlu.jl:465-561; computed
    once and kept with H, which must be an ``Hmat``).  Otherwise ``inner`` is
    a device storage ("f32", "bf16x2", ...) and d comes from restarted GMRES
    (relative tolerance ``inner_tol``, at most 20 restarts) on that copy of
    H, with the Krylov basis on the device.  Stops when
    ``||r|| <= tol * ||b||``.  Returns ``(x, residual_history)``.
    """
    import torch
    from .device import DeviceHmat
    if isinstance(H, Hmat) and len(H.P) > 0:
        raise ValueError("H holds its own H-LU factors (lu! was applied): solve_ir needs the matrix for "
                         "the residual; keep an unfactorised copy, or use hsolve")
    dev64 = H.device("f64") if isinstance(H, Hmat) else H
    LU = None
    if inner == "hlu":
        if not isinstance(H, Hmat):
            raise TypeError("inner='hlu' needs the host Hmat (its H-LU factors are computed on the host)")
        LU = factorized(H)
        devlo = None
    elif isinstance(H, Hmat):
        devlo = H.device(inner)
    else:
        devlo = DeviceHmat(dev64.flat, inner)
    b = np.asarray(b, dtype=np.float64)
    N = b.shape[0]
    bd = torch.from_numpy(b).to(dev64.device)
    xd = torch.zeros(N, dtype=torch.float64, device=dev64.device)
    rd = torch.empty_like(bd)
    nb = float(np.linalg.norm(b))
    hist = []
    for _ in range(maxiter + 1):
        rd.copy_(bd)
        dev64.matvec(xd, rd, alpha=-1.0, beta=1.0)          # r = b - H x, FP64
        hist.append(float(torch.linalg.norm(rd)) / nb if nb > 0 else 0.0)
        if hist[-1] <= tol:
            break
        if LU is not None:
            d = torch.from_numpy(hlu.solve(LU, rd.cpu().numpy())).to(xd.device)
        else:
            d, _ok = _gmres_device(devlo.matvec, rd.to(devlo.vec_dtype), inner_tol, restart, 20)
        xd += d.to(torch.float64)
    else:
        warnings.warn("iterative refinement did not reach the tolerance")
    return xd.cpu().numpy(), hist
