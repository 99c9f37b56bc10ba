"""Low-rank compression of admissible blocks (host, FP64).

Mirrors, by behaviour:

* ``aca2``  partial-pivoting ACA ............ src/hlu.jl:31-57 (1-D), 60-111 (d-D)
* ``compress`` / ``rank_truncate`` .......... src/hlu.jl:144-190
* ``_rk_matrix`` (separable alpha/beta) ..... src/cluster.jl:171-192
* ``bbfmm`` Chebyshev interpolation ......... src/hlu.jl:20-29, 113-132 and
  deps/src/lowrank.cpp:24-100, 186-222 (Lagrange basis on Chebyshev nodes,
  ``U = S_x K(nodes_x, nodes_y)``, ``V = S_y``)

Every factorisation returns ``(U, V)`` with the block approximated by
``U @ V.T`` (the reference's ``A * B'``).
"""
from __future__ import annotations

import numpy as np


def _row(f, x, Y):
    """``[f(x, y) for y in Y]`` vectorised; x is one point."""
    if Y.ndim == 1:
        return np.asarray(f(np.asarray(x)[None], Y), dtype=np.float64).reshape(Y.shape[0])
    return np.asarray(f(np.asarray(x)[None, :], Y), dtype=np.float64).reshape(Y.shape[0])


def _col(f, X, y):
    """``[f(x, y) for x in X]`` vectorised; y is one point."""
    if X.ndim == 1:
        return np.asarray(f(X, np.asarray(y)[None]), dtype=np.float64).reshape(X.shape[0])
    return np.asarray(f(X, np.asarray(y)[None, :]), dtype=np.float64).reshape(X.shape[0])


def aca2(f, X, Y, Rrank: int, eps: float = 1e-10):
    """Adaptive cross approximation with partial pivoting.

    Builds at most ``Rrank + 1`` crosses (hlu.jl:37, 71).  Stops when a pivot
    is exactly zero (returns the crosses so far, hlu.jl:42-43) or when
    ``|a_r| |b_r| / (|a_1| |b_1|) <= eps`` (hlu.jl:51-53, 104-106).

    Pivoting follows the textbook rule (the one hlu.jl:102-103 uses in the
    d-D variant): the column pivot is the largest entry of the *residual*
    row, and the next row pivot is the largest entry of the new ``a_r`` among
    rows not used yet.  The reference's 1-D variant picks the column pivot on
    the raw kernel row and the next row from ``B`` masked by a column mask
    (hlu.jl:38-39, 49-50), which repeats pivots and can index past ``m``;
    that is not reproduced.
    """
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    m, n = X.shape[0], Y.shape[0]
    R = Rrank + 1
    A = np.zeros((m, R))
    B = np.zeros((n, R))
    used = np.zeros(m, dtype=bool)
    tau = 0
    nrm0 = None
    for r in range(R):
        row = _row(f, X[tau], Y)
        if r:
            row = row - B[:, :r] @ A[tau, :r]
        used[tau] = True
        sigma = int(np.argmax(np.abs(row)))
        delta = row[sigma]
        if delta == 0.0:
            return np.asfortranarray(A[:, :r]), np.asfortranarray(B[:, :r])
        col = _col(f, X, Y[sigma])
        if r:
            col = col - A[:, :r] @ B[sigma, :r]
        A[:, r] = col / delta
        B[:, r] = row
        na, nb = np.linalg.norm(A[:, r]), np.linalg.norm(B[:, r])
        if nrm0 is None:
            nrm0 = na * nb
        if na * nb / nrm0 <= eps:
            return np.asfortranarray(A[:, :r + 1]), np.asfortranarray(B[:, :r + 1])
        free = np.nonzero(~used)[0]
        if free.size == 0:
            return np.asfortranarray(A[:, :r + 1]), np.asfortranarray(B[:, :r + 1])
        tau = int(free[np.argmax(np.abs(A[free, r]))])
    return np.asfortranarray(A), np.asfortranarray(B)


def rank_truncate(S, eps_trunc: float = 1e-10) -> int:
    """Number of singular values with ``S / S[0] > eps_trunc`` (hlu.jl:144-154)."""
    S = np.asarray(S)
    if S.size == 0 or S[0] == 0:
        return 0
    idx = np.nonzero(S / S[0] > eps_trunc)[0]
    return int(idx[-1]) + 1 if idx.size else 0


def compress(C, par):
    """Truncated factorisation of a dense block (hlu.jl:159-190).

    A block with ``sum(|C|) < 1e-5`` becomes rank-1 zero factors
    (hlu.jl:163-167).  ``CompMethod == "svd"`` truncates an SVD at
    ``εTrunc``; the reference uses a randomised SVD for square blocks
    (LowRankApprox ``psvd``), here a full LAPACK SVD is used for both shapes.
    """
    C = np.asarray(C, dtype=np.float64)
    if np.abs(C).sum() < 1e-5:
        return np.zeros((C.shape[0], 1), order="F"), np.zeros((C.shape[1], 1), order="F")
    method = par.CompMethod
    if method == "svd":
        U, S, Vt = np.linalg.svd(C, full_matrices=False)
        k = rank_truncate(S, par.εTrunc)
        assert k > 0
        A = U[:, :k]
        B = (S[:k, None] * Vt[:k, :]).T
        return np.asfortranarray(A), np.asfortranarray(B)
    if method == "aca":
        raise NotImplementedError(
            "CompMethod 'aca' calls the fully pivoted C++ ACA, which the reference "
            "cannot reach either (compress passes 3 arguments to a 2-argument "
            "method, hlu.jl:181)")
    if method == "rrqr":
        import scipy.linalg as sl
        Q, R, p = sl.qr(C, mode="economic", pivoting=True)
        d = np.abs(np.diag(R))
        k = int(np.sum(d > par.εComp * d[0])) if d.size and d[0] > 0 else 1
        ip = np.empty_like(p)
        ip[p] = np.arange(p.size)
        return np.asfortranarray(Q[:, :k]), np.asfortranarray(R[:k, ip].T)
    raise ValueError(f"Method {method} not implemented yet")


def _rk_matrix(alpha, beta, x, y):
    """Separable factors ``U[:, i] = α_i(x - x̄)``, ``V[:, i] = β_i(y - x̄)``
    with ``x̄`` the mean of ``x`` (src/cluster.jl:171-192)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xbar = x.mean(axis=0)
    t0 = x - xbar
    t = y - xbar
    r = len(alpha)
    U = np.zeros((x.shape[0], r), order="F")
    V = np.zeros((y.shape[0], r), order="F")
    for i in range(r):
        if x.ndim == 1:
            U[:, i] = alpha[i](t0)
            V[:, i] = beta[i](t)
        elif x.shape[1] == 2:
            U[:, i] = alpha[i](t0[:, 0], t0[:, 1])
            V[:, i] = beta[i](t[:, 0], t[:, 1])
        else:
            raise NotImplementedError("Not implemented yet")
    return U, V


def _cheb_nodes(r: int, a: float, b: float) -> np.ndarray:
    i = np.arange(r)
    return np.cos(np.pi * (2.0 * i + 1.0) / (2.0 * r)) * (b - a) / 2.0 + (a + b) / 2.0


def _lagrange(x: np.ndarray, nodes: np.ndarray) -> np.ndarray:
    """``P[i, j] = prod_{k != j} (x_i - c_k) / (c_j - c_k)`` (lowrank.cpp:47-60)."""
    r = nodes.shape[0]
    P = np.ones((x.shape[0], r))
    for j in range(r):
        for k in range(r):
            if k != j:
                P[:, j] *= (x - nodes[k]) / (nodes[j] - nodes[k])
    return P


def bbfmm(f, X, Y, Rrank: int):
    """Chebyshev-interpolation low rank (hlu.jl:20-29, 113-132).

    1-D: ``r = Rrank`` nodes on ``[min X, max X]`` and ``[min Y, max Y]``;
    returns ``U = S_x K(c_x, c_y)`` (m x r) and ``V = S_y`` (n x r).
    2-D: tensor nodes, ``r = min(Rrank, floor(sqrt(max(m, n))))`` per axis
    (hlu.jl:24-26), rank ``r^2``.
    """
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    if X.ndim == 1:
        r = Rrank
        cx = _cheb_nodes(r, X.min(), X.max())
        cy = _cheb_nodes(r, Y.min(), Y.max())
        K = np.asarray(f(cx[:, None], cy[None, :]), dtype=np.float64)
        return np.asfortranarray(_lagrange(X, cx) @ K), np.asfortranarray(_lagrange(Y, cy))
    if X.shape[1] != 2:
        raise NotImplementedError("bbfmm supports 1-D and 2-D geometry")
    r = Rrank
    if r * r > max(X.shape[0], Y.shape[0]):
        r = int(np.floor(np.sqrt(max(X.shape[0], Y.shape[0]))))
    def basis(P):
        c0 = _cheb_nodes(r, P[:, 0].min(), P[:, 0].max())
        c1 = _cheb_nodes(r, P[:, 1].min(), P[:, 1].max())
        L0, L1 = _lagrange(P[:, 0], c0), _lagrange(P[:, 1], c1)
        S = (L0[:, :, None] * L1[:, None, :]).reshape(P.shape[0], r * r)
        nodes = np.stack(np.meshgrid(c0, c1, indexing="ij"), -1).reshape(r * r, 2)
        return S, nodes
    Sx, nx = basis(X)
    Sy, ny = basis(Y)
    K = np.asarray(f(nx[:, None, :], ny[None, :, :]), dtype=np.float64)
    return np.asfortranarray(Sx @ K), np.asfortranarray(Sy)
