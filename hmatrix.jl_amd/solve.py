"""Solvers driving the GPU product.

* ``gmres`` / ``bicgstab``: src/tools.jl:7-65 -- SciPy's Krylov solvers on a
  ``LinearOperator`` whose matvec is the H-matrix product (here the HIP
  product instead of the reference's recursive CPU ``A*x``); same arguments,
  same warnings on breakdown / non-convergence.
* ``solve_ir``: ``H\\b`` by mixed-precision iterative refinement.  The
  residual ``r = b - H x`` is accumulated in FP64 on the GPU (the plain FP64
  product); corrections come from GMRES on a low-precision copy of H (FP32 or
  split-BF16 storage).  The reference's ``\\`` (src/hlu.jl:553-561) needs an
  H-LU factorisation (``lu!``, src/hlu.jl:465-515), which is not part of this
  library; refinement reaches the same solution to the requested tolerance.
"""
from __future__ import annotations

import warnings
from typing import Callable, Optional

import numpy as np

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
        raise NotImplementedError("an Hmat preconditioner needs M\\x (H-LU), which this library does not provide")
    return ssl.LinearOperator((N, N), matvec=M, dtype=np.float64)


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


def solve_ir(H, b, tol: float = 1e-12, inner: str = "f32", inner_tol: float = 1e-4, maxiter: int = 30,
             restart: int = 50):
    """Solve ``H x = b`` by iterative refinement.

    Each sweep: ``r = b - H x`` with the FP64 GPU product, then ``H d = r``
    by GMRES (relative tolerance ``inner_tol``) on the ``inner`` storage copy
    of H, then ``x += d``.  Stops when ``||r|| <= tol * ||b||``.
    Returns ``(x, residual_history)``.
    """
    import scipy.sparse.linalg as ssl
    import torch
    from .device import DeviceHmat
    dev64 = H.device("f64") if isinstance(H, Hmat) else H
    if isinstance(H, Hmat):
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
    lo = ssl.LinearOperator((N, N), dtype=np.float64,
                            matvec=lambda v: devlo.matvec(torch.from_numpy(np.asarray(v, np.float64).ravel())
                                                          .to(devlo.device, devlo.vec_dtype)).cpu().numpy()
                            .astype(np.float64))
    for _ in range(maxiter + 1):
        rd.copy_(bd)
        dev64.matvec(xd, rd, alpha=-1.0, beta=1.0)          # r = b - H x, FP64
        r = rd.cpu().numpy()
        hist.append(float(np.linalg.norm(r)) / nb if nb > 0 else 0.0)
        if hist[-1] <= tol:
            break
        d, info = ssl.gmres(lo, r, rtol=inner_tol, restart=restart, maxiter=20)
        xd += torch.from_numpy(d).to(xd.device)
    else:
        warnings.warn("iterative refinement did not reach the tolerance")
    return xd.cpu().numpy(), hist
