"""Flattened ("exported") H-matrix: a block table plus one FP64 factor arena.

This is the format both the device library and the CPU oracle consume, so
GPU parity is measured "on identical exported factors".

Block table: int64 array of shape (nblocks, 8), one row per leaf, columns

    kind, row0, col0, m, n, k, offA, offB

* ``kind`` 0 = dense leaf ``C`` (m x n, column-major at ``data[offA:]``);
  ``kind`` 1 = low-rank leaf, block = ``A @ B.T`` with ``A`` m x k at
  ``offA`` and ``B`` n x k at ``offB``, both column-major (Julia's layout of
  ``Hmat.A`` / ``Hmat.B``, src/cluster.jl:22-24).
* ``row0``/``col0`` are 0-based offsets of the leaf in cluster order.
* Rows are in the depth-first order of the reference's recursion
  ``hmat_matvec!`` (children [1,1], [1,2], [2,1], [2,2]; harithm.jl:346-352),
  so a sequential walk over the table accumulates in the reference's order.

On disk (``save_factors`` / ``load_factors``): one ``.npz`` with arrays
``blocks``, ``data``, ``shape`` = [nrows, ncols], ``perm`` (cluster order ->
input order, may be empty) and ``version``.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from .cluster import Hmat

KIND_DENSE = 0
KIND_RK = 1
FORMAT_VERSION = 1

COL_KIND, COL_ROW0, COL_COL0, COL_M, COL_N, COL_K, COL_OFFA, COL_OFFB = range(8)


@dataclass
class FlatHmat:
    nrows: int
    ncols: int
    blocks: np.ndarray          # (nblocks, 8) int64
    data: np.ndarray            # float64 arena
    perm: Optional[np.ndarray] = None

    @property
    def nblocks(self) -> int:
        return int(self.blocks.shape[0])

    def stored_elements(self) -> int:
        """Σ_dense m·n + Σ_rk (m+n)·k: the factor elements one product reads."""
        b = self.blocks
        dense = b[:, COL_KIND] == KIND_DENSE
        return int((b[dense, COL_M] * b[dense, COL_N]).sum()
                   + ((b[~dense, COL_M] + b[~dense, COL_N]) * b[~dense, COL_K]).sum())

    def counts(self):
        b = self.blocks
        nd = int((b[:, COL_KIND] == KIND_DENSE).sum())
        return nd, self.nblocks - nd

    def validate(self) -> None:
        """Raise ValueError unless every leaf lies inside the matrix and the
        arena, and the leaves tile the matrix exactly once."""
        b = self.blocks
        if b.ndim != 2 or b.shape[1] != 8:
            raise ValueError("blocks must be (nblocks, 8)")
        if np.any((b[:, COL_KIND] != KIND_DENSE) & (b[:, COL_KIND] != KIND_RK)):
            raise ValueError("unknown block kind")
        if np.any(b[:, 1:6] < 0):
            raise ValueError("negative offset or size")
        if np.any(b[:, COL_ROW0] + b[:, COL_M] > self.nrows) or np.any(b[:, COL_COL0] + b[:, COL_N] > self.ncols):
            raise ValueError("block outside the matrix")
        dense = b[:, COL_KIND] == KIND_DENSE
        endA = np.where(dense, b[:, COL_OFFA] + b[:, COL_M] * b[:, COL_N], b[:, COL_OFFA] + b[:, COL_M] * b[:, COL_K])
        endB = np.where(dense, 0, b[:, COL_OFFB] + b[:, COL_N] * b[:, COL_K])
        if np.any(endA > self.data.size) or np.any(endB > self.data.size):
            raise ValueError("factor offset outside the arena")
        if int((b[:, COL_M] * b[:, COL_N]).sum()) != self.nrows * self.ncols:
            raise ValueError("leaves do not tile the matrix")


def flatten(H: Hmat, perm: Optional[np.ndarray] = None) -> FlatHmat:
    """Depth-first flattening of an ``Hmat`` tree (order of harithm.jl:346-352)."""
    rows = []
    chunks = []
    off = [0]

    def put(a: np.ndarray) -> int:
        a = np.asarray(a, dtype=np.float64)
        o = off[0]
        chunks.append(a.ravel(order="F"))
        off[0] += a.size
        return o

    stack = [(H, 0, 0)]
    while stack:
        h, r0, c0 = stack.pop()
        if h.is_fullmatrix:
            C = np.asarray(h.C)
            if C.shape != (h.m, h.n):
                raise ValueError(f"dense leaf has shape {C.shape}, expected {(h.m, h.n)}")
            rows.append((KIND_DENSE, r0, c0, h.m, h.n, 0, put(C), 0))
        elif h.is_rkmatrix:
            A, B = np.asarray(h.A), np.asarray(h.B)
            k = A.shape[1] if A.ndim == 2 else 0
            if A.shape != (h.m, k) or B.shape != (h.n, k):
                raise ValueError(f"low-rank leaf has A {A.shape}, B {B.shape} for a {h.m}x{h.n} block")
            oa = put(A)
            ob = put(B)
            rows.append((KIND_RK, r0, c0, h.m, h.n, k, oa, ob))
        elif h.is_hmat:
            c11 = h.children[0][0]
            m, n = c11.m, c11.n
            # push in reverse so the pop order is [1,1], [1,2], [2,1], [2,2]
            stack.append((h.children[1][1], r0 + m, c0 + n))
            stack.append((h.children[1][0], r0 + m, c0))
            stack.append((h.children[0][1], r0, c0 + n))
            stack.append((h.children[0][0], r0, c0))
        else:
            raise ValueError("Invalid H")
    blocks = np.asarray(rows, dtype=np.int64).reshape(-1, 8)
    data = np.concatenate(chunks) if chunks else np.zeros(0)
    if perm is None and H.s is not None and getattr(H.s, "P", None) is not None:
        perm = np.asarray(H.s.P, dtype=np.int64)
    f = FlatHmat(H.m, H.n, blocks, np.ascontiguousarray(data), perm)
    f.validate()
    return f


def to_dense(f: FlatHmat) -> np.ndarray:
    """Dense reconstruction of the flattened matrix (utils.jl:162-192)."""
    out = np.zeros((f.nrows, f.ncols))
    for kind, r0, c0, m, n, k, oa, ob in f.blocks:
        if kind == KIND_DENSE:
            out[r0:r0 + m, c0:c0 + n] = f.data[oa:oa + m * n].reshape((m, n), order="F")
        else:
            A = f.data[oa:oa + m * k].reshape((m, k), order="F")
            B = f.data[ob:ob + n * k].reshape((n, k), order="F")
            out[r0:r0 + m, c0:c0 + n] = A @ B.T
    return out


def to_fmat(H: Hmat) -> np.ndarray:
    """``Array(H)`` / ``to_fmat(H)`` (utils.jl:162-192), on the host."""
    if H.is_fullmatrix:
        return np.array(H.C, dtype=np.float64)
    if H.is_rkmatrix:
        return np.asarray(H.A) @ np.asarray(H.B).T
    top = np.hstack([to_fmat(H.children[0][0]), to_fmat(H.children[0][1])])
    bot = np.hstack([to_fmat(H.children[1][0]), to_fmat(H.children[1][1])])
    return np.vstack([top, bot])


Array = to_fmat


def info(H: Hmat):
    """``(#dense, #rk, depth, stored/(m n))`` (utils.jl:83-110)."""
    dmat = rkmat = level = 0
    ratio = 0.0
    stack = [(H, 1)]
    while stack:
        h, l = stack.pop()
        if h.is_fullmatrix:
            dmat += 1
            level = max(level, l)
            ratio += h.m * h.n
        elif h.is_rkmatrix:
            rkmat += 1
            level = max(level, l)
            ratio += h.A.size + h.B.size
        else:
            for row in h.children:
                for c in row:
                    stack.append((c, l + 1))
    return dmat, rkmat, level, ratio / H.m / H.n


def save_factors(path, f: FlatHmat) -> None:
    np.savez(path, blocks=f.blocks, data=f.data,
             shape=np.asarray([f.nrows, f.ncols], dtype=np.int64),
             perm=np.zeros(0, np.int64) if f.perm is None else np.asarray(f.perm, np.int64),
             version=np.asarray([FORMAT_VERSION], dtype=np.int64))


def load_factors(path) -> FlatHmat:
    with np.load(path, allow_pickle=False) as z:
        if int(z["version"][0]) != FORMAT_VERSION:
            raise ValueError(f"factor file version {int(z['version'][0])} != {FORMAT_VERSION}")
        perm = z["perm"]
        f = FlatHmat(int(z["shape"][0]), int(z["shape"][1]),
                     np.ascontiguousarray(z["blocks"], dtype=np.int64),
                     np.ascontiguousarray(z["data"], dtype=np.float64),
                     perm if perm.size else None)
    f.validate()
    return f
