"""Host H-LU factorisation and solve (FP64): the reference's ``lu!`` / ``\\``.

This is the host side of ``H\\y``: the factorisation stays in FP64 on the
host, as in the reference, and serves as the correction step of the GPU
iterative refinement (``solve.solve_ir(..., inner="hlu")``) or as a
preconditioner (``solve.gmres(..., M=hlu_solver(LU))``).

Block arithmetic on the tree of ``cluster.Hmat`` (leaves: dense ``C``, or
``A @ B.T``), following the reference:

* ``lu`` / ``lu_inplace``      -- src/hlu.jl:465-521 (``lu!``, ``lu``): LU of
  [1,1], rows of [1,2] permuted by its pivots, [1,2] <- L11 \\ [1,2],
  [2,1] <- [2,1] / U11, [2,2] -= [2,1] [1,2], LU of [2,2], rows of [2,1]
  permuted by its pivots; dense leaves by LAPACK getrf with partial pivoting
* ``hmat_trisolve``            -- src/hlu.jl:299-440
* ``mat_full_solve``           -- src/hlu.jl:236-296
* ``hmat_solve`` / ``solve``   -- src/hlu.jl:524-561 (``hmat_solve!``, ``\\``)
* ``hmul``, ``hmat_add`` and helpers -- src/harithm.jl:1-314, 374-420
  (recompression of sums of low-rank blocks by QR + truncated SVD,
  ``compress`` of src/hlu.jl:159-190)

Two branches of the reference cannot run as written and are restated with
the arithmetic they intend: ``hmat_add!`` with a dense ``a`` and a
hierarchical ``b`` (src/harithm.jl:124-131) adds ``(1 + s) C + s H``; here
``C + s H``.  ``full_mat_mul`` / ``mat_full_mul`` recurse with plain arrays
where they need ``Hmat`` arguments when neither cluster is a leaf
(src/harithm.jl:168-192, 208-231); here the block is formed densely as in
their leaf branch.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sl

from .cluster import Hmat, copy_hmat
from .lowrank import rank_truncate

# relative singular-value cut of every recompression (Hparams.εTrunc default,
# src/utils.jl:8-26)
EPS_TRUNC = 1e-10


# ----------------------------------------------------------------------------
# dense helpers
# ----------------------------------------------------------------------------

def _compress(C: np.ndarray, eps: float = EPS_TRUNC):
    """Truncated SVD of a dense block (src/hlu.jl:159-190, "svd" method):
    blocks with sum |C| < 1e-5 become rank-1 zeros."""
    C = np.asarray(C, dtype=np.float64)
    if C.size == 0 or np.abs(C).sum() < 1e-5:
        return np.zeros((C.shape[0], 1)), np.zeros((C.shape[1], 1))
    U, S, Vt = np.linalg.svd(C, full_matrices=False)
    k = max(rank_truncate(S, eps), 1)
    return U[:, :k], (S[:k, None] * Vt[:k, :]).T


def _rk_add(A1, B1, A2, B2, eps: float = EPS_TRUNC):
    """A1 B1' + A2 B2' recompressed (src/harithm.jl:374-396): thin QR of
    [A1 A2] and [B1 B2], truncated SVD of R_A R_B'."""
    if A2.shape[1] == 0:
        return A1, B1
    if A1.shape[1] == 0:
        return A2, B2
    QA, RA = np.linalg.qr(np.hstack([A1, A2]))
    QB, RB = np.linalg.qr(np.hstack([B1, B2]))
    U, V = _compress(RA @ RB.T, eps)
    return QA @ U, QB @ V


def _dense(H: Hmat) -> np.ndarray:
    """to_fmat (src/utils.jl:162-192)."""
    if H.is_fullmatrix:
        return np.array(H.C, dtype=np.float64)
    if H.is_rkmatrix:
        return np.asarray(H.A) @ np.asarray(H.B).T
    m = H.children[0][0].m
    n = H.children[0][0].n
    out = np.empty((H.m, H.n))
    out[:m, :n] = _dense(H.children[0][0])
    out[:m, n:] = _dense(H.children[0][1])
    out[m:, :n] = _dense(H.children[1][0])
    out[m:, n:] = _dense(H.children[1][1])
    return out


def _mul(H: Hmat, V: np.ndarray) -> np.ndarray:
    """H @ V on the host (the recursion of src/harithm.jl:340-354)."""
    if H.is_fullmatrix:
        return np.asarray(H.C) @ V
    if H.is_rkmatrix:
        return np.asarray(H.A) @ (np.asarray(H.B).T @ V)
    m = H.children[0][0].m
    n = H.children[0][0].n
    top = _mul(H.children[0][0], V[:n]) + _mul(H.children[0][1], V[n:])
    bot = _mul(H.children[1][0], V[:n]) + _mul(H.children[1][1], V[n:])
    return np.concatenate([top, bot], axis=0)


def _tmul(H: Hmat, V: np.ndarray) -> np.ndarray:
    """H' @ V (transpose_hmat_mat_mul, src/harithm.jl:243-260)."""
    if H.is_fullmatrix:
        return np.asarray(H.C).T @ V
    if H.is_rkmatrix:
        return np.asarray(H.B) @ (np.asarray(H.A).T @ V)
    m = H.children[0][0].m
    V1 = _tmul(H.children[0][0], V[:m]) + _tmul(H.children[1][0], V[m:])
    V2 = _tmul(H.children[0][1], V[:m]) + _tmul(H.children[1][1], V[m:])
    return np.concatenate([V1, V2], axis=0)


def _rk(A, B, s=None, t=None) -> Hmat:
    return Hmat(is_rkmatrix=True, A=np.asarray(A), B=np.asarray(B), m=A.shape[0], n=B.shape[0], s=s, t=t)


def _full(C, s=None, t=None) -> Hmat:
    C = np.asarray(C)
    return Hmat(is_fullmatrix=True, C=C, m=C.shape[0], n=C.shape[1], s=s, t=t)


# ----------------------------------------------------------------------------
# H-arithmetic (src/harithm.jl)
# ----------------------------------------------------------------------------

def hmat_full_add(a: Hmat, b: np.ndarray, scalar: float) -> None:
    """a += scalar * b, b dense, format of a kept (src/harithm.jl:4-32)."""
    if a.is_fullmatrix:
        a.C = a.C + scalar * b
    elif a.is_rkmatrix:
        a.A, a.B = _compress(a.A @ a.B.T + scalar * b)
    else:
        m = a.children[0][0].m
        n = a.children[0][0].n
        hmat_full_add(a.children[0][0], b[:m, :n], scalar)
        hmat_full_add(a.children[0][1], b[:m, n:], scalar)
        hmat_full_add(a.children[1][0], b[m:, :n], scalar)
        hmat_full_add(a.children[1][1], b[m:, n:], scalar)


def _compress_low_rank(A, B):
    """Recompression of A B' (src/harithm.jl:49-56)."""
    QA, RA = np.linalg.qr(A)
    QB, RB = np.linalg.qr(B)
    A1, A2 = _compress(RA @ RB.T)
    return QA @ A1, QB @ A2


def rkmat_hmat_add(A, B, scalar: float, H: Hmat):
    """Low-rank factors of A B' + scalar * H (src/harithm.jl:58-91)."""
    if H.is_fullmatrix:
        return _compress(A @ B.T + scalar * H.C)
    if H.is_rkmatrix:
        return _rk_add(A, B, scalar * H.A, H.B)
    m = H.children[0][0].m
    n = H.children[0][0].n
    A11, B11 = rkmat_hmat_add(A[:m], B[:n], scalar, H.children[0][0])
    A12, B12 = rkmat_hmat_add(A[:m], B[n:], scalar, H.children[0][1])
    A21, B21 = rkmat_hmat_add(A[m:], B[:n], scalar, H.children[1][0])
    A22, B22 = rkmat_hmat_add(A[m:], B[n:], scalar, H.children[1][1])
    z = np.zeros
    A0 = np.block([[A11, A12, z((A11.shape[0], A21.shape[1])), z((A11.shape[0], A22.shape[1]))],
                   [z((A21.shape[0], A11.shape[1])), z((A21.shape[0], A12.shape[1])), A21, A22]])
    B0 = np.block([[B11, z((B11.shape[0], B12.shape[1])), B21, z((B11.shape[0], B22.shape[1]))],
                   [z((B12.shape[0], B11.shape[1])), B12, z((B12.shape[0], B21.shape[1])), B22]])
    return _compress_low_rank(A0, B0)


def hmat_add(a: Hmat, b: Hmat, scalar: float = 1.0) -> None:
    """a += scalar * b, format of a kept (src/harithm.jl:95-151)."""
    if b.is_fullmatrix:
        hmat_full_add(a, b.C, scalar)
    elif a.is_fullmatrix and b.is_rkmatrix:
        if b.A.size:
            a.C = a.C + scalar * (b.A @ b.B.T)
    elif a.is_fullmatrix and b.is_hmat:
        a.C = a.C + scalar * _dense(b)
    elif a.is_rkmatrix and b.is_rkmatrix:
        a.A, a.B = _rk_add(a.A, a.B, scalar * b.A, b.B)
    elif a.is_rkmatrix and b.is_hmat:
        a.A, a.B = rkmat_hmat_add(a.A, a.B, scalar, b)
    elif a.is_hmat and b.is_rkmatrix:
        m = a.children[0][0].m
        n = a.children[0][0].n
        hmat_add(a.children[0][0], _rk(b.A[:m], b.B[:n]), scalar)
        hmat_add(a.children[1][0], _rk(b.A[m:], b.B[:n]), scalar)
        hmat_add(a.children[0][1], _rk(b.A[:m], b.B[n:]), scalar)
        hmat_add(a.children[1][1], _rk(b.A[m:], b.B[n:]), scalar)
    elif a.is_hmat and b.is_hmat:
        for i in range(2):
            for j in range(2):
                hmat_add(a.children[i][j], b.children[i][j], scalar)
    else:
        raise ValueError("hmat_add: invalid operands")


def hadd(a: Hmat, b: Hmat) -> Hmat:
    """a + b in the format of a (src/harithm.jl:153-158)."""
    c = copy_hmat(a)
    hmat_add(c, b, 1.0)
    return c


def hmul(a: Hmat, b: Hmat) -> Hmat:
    """a * b (src/harithm.jl:262-314)."""
    if a.is_fullmatrix:
        # full_mat_mul (src/harithm.jl:164-202)
        if b.is_fullmatrix:
            H = _full(a.C @ b.C)
        elif b.is_rkmatrix:
            H = _rk(a.C @ b.A, b.B)
        else:
            H = _full(a.C @ _dense(b))
    elif b.is_fullmatrix:
        # mat_full_mul (src/harithm.jl:204-241)
        if a.is_rkmatrix:
            H = _rk(a.A, b.C.T @ a.B)
        else:
            H = _full(_mul(a, b.C))
    elif a.is_rkmatrix and b.is_rkmatrix:
        H = _rk(a.A, b.B @ (b.A.T @ a.B))
    elif a.is_rkmatrix and b.is_hmat:
        H = _rk(a.A, _tmul(b, a.B))
    elif a.is_hmat and b.is_rkmatrix:
        H = _rk(_mul(a, b.A), b.B)
    elif a.is_hmat and b.is_hmat:
        H = Hmat(is_hmat=True)
        H.children = [[None, None], [None, None]]
        for i in range(2):
            for j in range(2):
                H1 = hmul(a.children[i][0], b.children[0][j])
                H2 = hmul(a.children[i][1], b.children[1][j])
                H.children[i][j] = hadd(H1, H2)
    else:
        raise ValueError("hmul: undefined operands")
    H.s, H.t, H.m, H.n = a.s, b.t, a.m, b.n
    return H


def hmat_sub_mul(c: Hmat, a: Hmat, b: Hmat) -> None:
    """c -= a * b (src/hlu.jl:230-234)."""
    hmat_add(c, hmul(a, b), -1.0)


# ----------------------------------------------------------------------------
# triangular solves (src/hlu.jl:236-440)
# ----------------------------------------------------------------------------

def mat_full_solve(a: Hmat, b: np.ndarray, unitdiag: bool, ul: str = "L") -> None:
    """In place: b <- L(a) \\ b (ul "L") or b <- b / U(a) (ul "U"), with a an
    H-matrix or dense leaf holding its LU factors (src/hlu.jl:236-296)."""
    if a.is_rkmatrix:
        raise ValueError("A should not be a low-rank matrix")
    if ul == "L":
        if a.is_fullmatrix:
            b[...] = sl.solve_triangular(a.C, b, lower=True, unit_diagonal=unitdiag)
        else:
            n = a.children[0][0].m
            b1 = np.array(b[:n])
            b2 = np.array(b[n:])
            mat_full_solve(a.children[0][0], b1, unitdiag, ul)
            b2 -= _mul(a.children[1][0], b1)
            mat_full_solve(a.children[1][1], b2, unitdiag, ul)
            b[:n] = b1
            b[n:] = b2
    else:
        if a.is_fullmatrix:
            b[...] = sl.solve_triangular(a.C, b.T, lower=False, trans="T", unit_diagonal=unitdiag).T
        else:
            n = a.children[0][0].m
            b1 = np.array(b[:, :n])
            b2 = np.array(b[:, n:])
            mat_full_solve(a.children[0][0], b1, unitdiag, ul)
            b2 -= _tmul(a.children[0][1], b1.T).T          # b1 * a12
            mat_full_solve(a.children[1][1], b2, unitdiag, ul)
            b[:, :n] = b1
            b[:, n:] = b2


def hmat_trisolve(a: Hmat, b: Hmat, islower: bool, unitdiag: bool) -> None:
    """b <- L(a) \\ b (islower) or b <- b / U(a), in place
    (src/hlu.jl:299-440)."""
    if a.is_rkmatrix:
        raise ValueError("A should not be a low-rank matrix")
    if islower:
        if a.is_fullmatrix and b.is_fullmatrix:
            b.C = sl.solve_triangular(a.C, b.C, lower=True, unit_diagonal=unitdiag)
        elif a.is_fullmatrix and b.is_rkmatrix:
            if b.A.shape[0]:
                b.A = sl.solve_triangular(a.C, b.A, lower=True, unit_diagonal=unitdiag)
        elif a.is_hmat and b.is_hmat:
            a11, a21, a22 = a.children[0][0], a.children[1][0], a.children[1][1]
            b11, b12, b21, b22 = b.children[0][0], b.children[0][1], b.children[1][0], b.children[1][1]
            hmat_trisolve(a11, b11, islower, unitdiag)
            hmat_trisolve(a11, b12, islower, unitdiag)
            hmat_sub_mul(b21, a21, b11)
            hmat_sub_mul(b22, a21, b12)
            hmat_trisolve(a22, b21, islower, unitdiag)
            hmat_trisolve(a22, b22, islower, unitdiag)
        elif a.is_hmat and b.is_fullmatrix:
            b.C = np.array(b.C)
            mat_full_solve(a, b.C, unitdiag)
        elif a.is_hmat and b.is_rkmatrix:
            b.A = np.array(b.A)
            mat_full_solve(a, b.A, unitdiag)
        else:
            raise ValueError("This is never used")
    else:
        if a.is_fullmatrix and b.is_fullmatrix:
            b.C = sl.solve_triangular(a.C, b.C.T, lower=False, trans="T", unit_diagonal=unitdiag).T
        elif a.is_fullmatrix and b.is_rkmatrix:
            if b.A.shape[0]:
                b.B = sl.solve_triangular(a.C, b.B, lower=False, trans="T", unit_diagonal=unitdiag)
        elif a.is_hmat and b.is_fullmatrix:
            b.C = np.array(b.C)
            mat_full_solve(a, b.C, unitdiag, "U")
        elif a.is_hmat and b.is_rkmatrix:
            Bt = np.array(b.B.T)
            mat_full_solve(a, Bt, unitdiag, "U")
            b.B = Bt.T
        elif a.is_hmat and b.is_hmat:
            a11, a12, a22 = a.children[0][0], a.children[0][1], a.children[1][1]
            b11, b12, b21, b22 = b.children[0][0], b.children[0][1], b.children[1][0], b.children[1][1]
            hmat_trisolve(a11, b11, islower, unitdiag)
            hmat_trisolve(a11, b21, islower, unitdiag)
            hmat_sub_mul(b12, b11, a12)
            hmat_sub_mul(b22, b21, a12)
            hmat_trisolve(a22, b12, islower, unitdiag)
            hmat_trisolve(a22, b22, islower, unitdiag)
        else:
            raise ValueError("Invalid")


def permute_hmat(H: Hmat, P: np.ndarray) -> None:
    """Rows of H in the order P (0-based) (src/hlu.jl:446-463)."""
    if H.is_fullmatrix:
        H.C = np.asarray(H.C)[P, :]
    elif H.is_rkmatrix:
        H.A = np.asarray(H.A)[P, :]
    else:
        m = H.children[0][0].m
        P1 = P[:m]
        P2 = P[m:] - m
        permute_hmat(H.children[0][0], P1)
        permute_hmat(H.children[0][1], P1)
        permute_hmat(H.children[1][0], P2)
        permute_hmat(H.children[1][1], P2)


def _getrf(C: np.ndarray):
    """LU with partial pivoting in place of a dense block; returns the row
    order p (0-based) with C_in[p, :] = L U."""
    lu, piv = sl.lu_factor(np.asarray(C, dtype=np.float64), check_finite=False)
    p = np.arange(lu.shape[0])
    for i, j in enumerate(piv):
        p[i], p[j] = p[j], p[i]
    return lu, p


def lu_inplace(H: Hmat) -> Hmat:
    """``lu!`` (src/hlu.jl:465-515): H holds L\\U afterwards and H.P the
    row order, H[P, :] = L U."""
    if len(H.P) > 0:
        return H                  # already factorised: the reference reuses it
    if H.is_rkmatrix:
        raise ValueError("H should not be a low-rank matrix")
    H.invalidate_device()
    if H.is_fullmatrix:
        H.C, H.P = _getrf(H.C)
    else:
        h11, h12, h21, h22 = H.children[0][0], H.children[0][1], H.children[1][0], H.children[1][1]
        lu_inplace(h11)
        permute_hmat(h12, h11.P)
        hmat_trisolve(h11, h12, True, True)
        hmat_trisolve(h11, h21, False, False)
        hmat_sub_mul(h22, h21, h12)
        lu_inplace(h22)
        permute_hmat(h21, h22.P)
        H.P = np.concatenate([h11.P, h22.P + h11.m])
    return H


def lu(H: Hmat) -> Hmat:
    """A factorised copy of H (src/hlu.jl:517-521)."""
    return lu_inplace(copy_hmat(H))


def hmat_solve(a: Hmat, y: np.ndarray, lower: bool = True) -> None:
    """In place: y <- L \\ y (lower) or U \\ y with the factors held in a
    (src/hlu.jl:524-551); the off-diagonal blocks act through host products
    (the reference's hmat_matvec! with s = -1)."""
    if a.is_rkmatrix:
        raise ValueError("a cannot be a low-rank matrix")
    if a.is_fullmatrix:
        y[...] = sl.solve_triangular(a.C, y, lower=lower, unit_diagonal=lower)
        return
    m = a.children[0][0].m
    y1, y2 = y[:m], y[m:]
    if lower:
        hmat_solve(a.children[0][0], y1, True)
        y2 -= _mul(a.children[1][0], y1)
        hmat_solve(a.children[1][1], y2, True)
    else:
        hmat_solve(a.children[1][1], y2, False)
        y1 -= _mul(a.children[0][1], y2)
        hmat_solve(a.children[0][0], y1, False)


def solve(a: Hmat, y: np.ndarray) -> np.ndarray:
    """``a \\ y`` for a factorised a (src/hlu.jl:553-561)."""
    if len(a.P) == 0:
        raise ValueError("a has not been factorized")
    w = np.array(y, dtype=np.float64)[a.P]
    hmat_solve(a, w, True)
    hmat_solve(a, w, False)
    return w


def getl(A: np.ndarray, unitdiag: bool = True) -> np.ndarray:
    """Lower factor of a dense L\\U array (src/hlu.jl:193-199)."""
    L = np.tril(A)
    if unitdiag:
        np.fill_diagonal(L, 1.0)
    return L


def getu(A: np.ndarray, unitdiag: bool = False) -> np.ndarray:
    """Upper factor of a dense L\\U array (src/hlu.jl:201-207)."""
    U = np.triu(A)
    if unitdiag:
        np.fill_diagonal(U, 1.0)
    return U
