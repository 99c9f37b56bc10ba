"""Cluster tree, block partition and the ``Hmat`` host tree.

Host side, FP64, as in the reference.  This module is the producer of every
array the device path reads; it mirrors:

* ``Cluster`` / ``Hmat`` structs ............ src/cluster.jl:7-34
* ``NewHmat`` ............................... src/cluster.jl:36-43
* ``__c1`` block classification ............. src/cluster.jl:48-133
* ``compute_geom_info`` (admissibility) ..... src/cluster.jl:140-169
* ``construct_cluster`` / ``bisect_cluster``  src/cluster.jl:196-256
* ``kernel_full`` / ``FullMat`` ............. src/cluster.jl:45, 258-275

Index conventions (Python-natural, documented once here): ``Cluster.P`` holds
0-based point indices; ``Cluster.s`` is the 0-based first position of the
cluster in cluster order and ``Cluster.e`` the 0-based last one (inclusive),
so ``e - s + 1 == N`` exactly as ``s:e`` covers ``N`` points in the reference.

Kernel convention: ``Hparams.Kernel(x, y)`` is called on broadcast numpy
arrays.  For 1-D geometry ``x`` has shape (m, 1) and ``y`` shape (1, n); for
d-D geometry ``x`` is (m, 1, d) and ``y`` is (1, n, d), and the kernel reduces
the last axis itself, e.g. ``lambda x, y: 1/(100 + ((x - y)**2).sum(-1))``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, List, Optional

import numpy as np

from .params import Hparams, Params
from . import lowrank


# ----------------------------------------------------------------------------
# geometry helpers
# ----------------------------------------------------------------------------

def _as_points(X) -> np.ndarray:
    """(N,) for 1-D input, (N, d) for d >= 2; a (N, 1) array is treated as 1-D."""
    X = np.asarray(X, dtype=np.float64)
    if X.ndim == 2 and X.shape[1] == 1:
        X = X[:, 0]
    if X.ndim not in (1, 2):
        raise ValueError(f"Geom must be (N,) or (N, d); got shape {X.shape}")
    return np.ascontiguousarray(X)


def kernel_full(f, X, Y) -> np.ndarray:
    """Dense block ``A[i, j] = f(X[i], Y[j])`` (src/cluster.jl:258-275)."""
    X = _as_points(X)
    Y = _as_points(Y)
    if X.ndim == 1:
        A = f(X[:, None], Y[None, :])
    else:
        A = f(X[:, None, :], Y[None, :, :])
    A = np.asarray(A, dtype=np.float64)
    if A.shape != (X.shape[0], Y.shape[0]):
        A = np.broadcast_to(A, (X.shape[0], Y.shape[0])).copy()
    return A


def FullMat(F, X, Y) -> np.ndarray:
    """``FullMat(F, X, Y) = kernel_full(F, X, Y)`` (src/cluster.jl:45)."""
    return kernel_full(F, X, Y)


# ----------------------------------------------------------------------------
# cluster tree
# ----------------------------------------------------------------------------

@dataclass(eq=False)
class Cluster:
    """Node of the cluster tree (src/cluster.jl:7-19)."""

    X: Any = None                    # points of this cluster, in cluster order
    P: Any = None                    # 0-based indices into the user's Geom
    left: Optional["Cluster"] = None
    right: Optional["Cluster"] = None
    m: int = 0                       # |left child|
    n: int = 0                       # |right child|
    N: int = 0
    isleaf: bool = False
    s: int = 0                       # 0-based first position (inclusive)
    e: int = -1                      # 0-based last position (inclusive)
    id: int = 0

    def same_range(self, other: "Cluster") -> bool:
        """``==`` of the reference compares index ranges (cluster.jl:135-137)."""
        return self.s == other.s and self.e == other.e


def _bisect_1d(x: np.ndarray) -> np.ndarray:
    """Optimal 2-means split of 1-D points; returns a bool mask, True = right.

    The reference calls sklearn ``KMeans(n_clusters=2, random_state=0)``
    (cluster.jl:196-205).  In one dimension a converged 2-means partition is
    a threshold split; this function takes the threshold that minimises the
    within-cluster sum of squares over all split positions (exact, so the tree
    does not depend on a random initialisation).  The left child is the
    cluster with the smaller centroid.
    """
    n = x.shape[0]
    order = np.argsort(x, kind="stable")
    v = x[order] - x.mean()
    cs = np.cumsum(v)
    S = cs[-1]
    i = np.arange(1, n)                      # left size
    score = cs[:-1] ** 2 / i + (S - cs[:-1]) ** 2 / (n - i)
    split = int(np.argmax(score)) + 1
    mask = np.zeros(n, dtype=bool)
    mask[order[split:]] = True
    return mask


def _bisect_nd(X: np.ndarray) -> np.ndarray:
    """Lloyd 2-means for d-D points started from a median split along the
    widest coordinate; True = right (the cluster whose centroid is larger
    along that coordinate)."""
    n = X.shape[0]
    ax = int(np.argmax(X.max(0) - X.min(0)))
    order = np.argsort(X[:, ax], kind="stable")
    lab = np.zeros(n, dtype=bool)
    lab[order[n // 2:]] = True
    for _ in range(100):
        c0 = X[~lab].mean(0)
        c1 = X[lab].mean(0)
        d0 = ((X - c0) ** 2).sum(1)
        d1 = ((X - c1) ** 2).sum(1)
        new = d1 < d0
        if new.all() or (~new).all() or np.array_equal(new, lab):
            break
        lab = new
    if X[lab, ax].mean() < X[~lab, ax].mean():
        lab = ~lab
    return lab


def bisect_cluster(X: np.ndarray):
    """Split a cluster's points in two (src/cluster.jl:196-205).

    Returns ``(X1, P1, X2, P2)`` with ``P1``/``P2`` 0-based positions into
    ``X`` in ascending order, as ``findall(L .== 0)`` gives in the reference.
    """
    X = np.asarray(X, dtype=np.float64)
    mask = _bisect_1d(X) if X.ndim == 1 else _bisect_nd(X)
    if mask.all() or (~mask).all():          # identical points: split by position
        mask = np.zeros(X.shape[0], dtype=bool)
        mask[X.shape[0] // 2:] = True
    P1 = np.nonzero(~mask)[0]
    P2 = np.nonzero(mask)[0]
    return X[P1], P1, X[P2], P2


def construct_cluster(X, Nleaf: int):
    """Build the cluster tree (src/cluster.jl:216-256).

    Ids are assigned in pre-order starting at 1; leaves have ``N <= Nleaf``.
    Returns ``(ncluster, root)``; ``root.P`` is the permutation from cluster
    order to input order and ``root.X == X[root.P]``.
    """
    X = _as_points(X)
    N = X.shape[0]
    if N == 0:
        raise ValueError("Geom is empty")
    counter = [0]

    def down(idx: np.ndarray, s: int) -> Cluster:
        counter[0] += 1
        c = Cluster(N=int(idx.shape[0]), s=s, e=s + int(idx.shape[0]) - 1, id=counter[0])
        if c.N <= Nleaf:
            c.isleaf = True
            c.P = idx
            return c
        _, P1, _, P2 = bisect_cluster(X[idx])
        c.left = down(idx[P1], s)
        c.right = down(idx[P2], s + P1.shape[0])
        c.m, c.n = c.left.N, c.right.N
        c.P = np.concatenate([c.left.P, c.right.P])
        return c

    import sys
    old = sys.getrecursionlimit()
    sys.setrecursionlimit(max(old, 10000))
    try:
        root = down(np.arange(N, dtype=np.int64), 0)
    finally:
        sys.setrecursionlimit(old)

    Xp = X[root.P]
    Pp = root.P

    def views(c: Cluster) -> None:
        c.X = Xp[c.s:c.e + 1]
        c.P = Pp[c.s:c.e + 1]
        if not c.isleaf:
            views(c.left)
            views(c.right)

    views(root)
    return counter[0], root


def leaves(c: Cluster) -> List[Cluster]:
    out: List[Cluster] = []
    stack = [c]
    while stack:
        x = stack.pop()
        if x.isleaf:
            out.append(x)
        else:
            stack.append(x.right)
            stack.append(x.left)
    return out


# ----------------------------------------------------------------------------
# admissibility (compute_geom_info)
# ----------------------------------------------------------------------------

class _Adjacency:
    """``Neighbors[s.id, t.id] != 0`` of the reference (cluster.jl:140-169).

    The reference marks a cluster pair as neighbours when some nonzero
    ``Adj[i, j]`` has ``i`` in ``s`` and ``j`` in ``t`` (all ancestors are
    marked through ``p2c``).  Clusters are contiguous ranges in cluster
    order, so the same predicate is a range query on ``Adj[P, P]``: any stored
    column index of rows ``s.s..s.e`` inside ``[t.s, t.e]``.
    """

    def __init__(self, Adj, P: np.ndarray):
        import scipy.sparse as sp
        if Adj is None:
            raise ValueError(
                "Hparams.Adj is nothing: NewHmat needs a sparse point adjacency "
                "(e.g. scipy.sparse.identity(N) for weak admissibility); the "
                "reference fails on Adj[c.P, c.P] in the same case (cluster.jl:141)")
        A = sp.csr_matrix(Adj)
        if A.shape != (P.shape[0], P.shape[0]):
            raise ValueError(f"Hparams.Adj has shape {A.shape}, expected {(P.shape[0],) * 2}")
        A = A[P][:, P].tocsr()
        A.eliminate_zeros()
        A.sort_indices()
        self.indptr = A.indptr
        self.indices = A.indices

    def neighbours(self, s: Cluster, t: Cluster) -> bool:
        cols = self.indices[self.indptr[s.s]:self.indptr[s.e + 1]]
        if cols.size == 0:
            return False
        return bool(np.any((cols >= t.s) & (cols <= t.e)))


# ----------------------------------------------------------------------------
# Hmat
# ----------------------------------------------------------------------------

class Hmat:
    """H-matrix node (src/cluster.jl:21-34).

    A leaf is full (``C``, m x n) or low rank (block = ``A @ B.T`` with ``A``
    m x k and ``B`` n x k); an inner node has a 2 x 2 ``children`` list.
    Products (``H * y``, ``H * X``, ``v * H``) run on the GPU through the HIP
    library; see ``hmatrix_amd.device``.
    """

    __array_ufunc__ = None          # make numpy defer ``ndarray * Hmat`` to us

    def __init__(self, **kw):
        self.A = np.zeros((0, 0))
        self.B = np.zeros((0, 0))
        self.C = np.zeros((0, 0))
        self.P = np.zeros(0, dtype=np.int64)
        self.is_rkmatrix = False
        self.is_fullmatrix = False
        self.is_hmat = False
        self.m = 0
        self.n = 0
        self.children: List[List["Hmat"]] = []
        self.s: Optional[Cluster] = None
        self.t: Optional[Cluster] = None
        self._device_cache = {}
        for k, v in kw.items():
            setattr(self, k, v)

    # --- shape -------------------------------------------------------------
    @property
    def shape(self):
        return (self.m, self.n)

    def size(self, i: Optional[int] = None):
        """``size(H)`` / ``size(H, i)`` with Julia's 1-based ``i`` (utils.jl:69-81)."""
        if i is None:
            return (self.m, self.n)
        if i == 1:
            return self.m
        if i == 2:
            return self.n
        raise ValueError(f"size(Hmat, Int64): invalid dimension {i}")

    # --- products (device) -------------------------------------------------
    def device(self, dtype: str = "f64"):
        """Device-resident copy of this H-matrix in the given storage dtype."""
        from .device import DeviceHmat
        d = self._device_cache.get(dtype)
        if d is None:
            d = DeviceHmat.from_hmat(self, dtype=dtype)
            self._device_cache[dtype] = d
        return d

    def invalidate_device(self) -> None:
        """Drop cached device copies after the host factors were changed."""
        self._device_cache = {}

    def __mul__(self, v):
        if isinstance(v, Hmat):
            raise NotImplementedError(
                "Hmat * Hmat is not provided (the reference's method recurses "
                "on itself, harithm.jl:315)")
        if np.isscalar(v):
            raise NotImplementedError("scalar scaling of an Hmat is outside the matvec path")
        return self.device("f64").mul(np.asarray(v, dtype=np.float64))

    __matmul__ = __mul__

    def __rmul__(self, v):
        """``v * H`` for a p x m array ``v`` (harithm.jl:331-337) -> p x n."""
        if np.isscalar(v):
            raise NotImplementedError("scalar scaling of an Hmat is outside the matvec path")
        v = np.asarray(v, dtype=np.float64)
        if v.ndim != 2:
            raise ValueError("v * H needs a 2-D array (p x m); use H.T * y for a vector")
        return self.device("f64").rmul(v)

    __rmatmul__ = __rmul__

    def __add__(self, other):
        """``H + u*I``: u added to the diagonal of the diagonal dense leaves
        (src/harithm.jl:457-474); returns a new Hmat."""
        if isinstance(other, UniformScaling):
            if self.m != self.n:
                raise ValueError("H + u*I needs a square H")
            G = copy_hmat(self)
            _add_diag(G, float(other.value))
            return G
        return NotImplemented

    __radd__ = __add__

    @property
    def T(self):
        from .device import TransposedHmat
        return TransposedHmat(self)

    def __repr__(self) -> str:
        kind = "full" if self.is_fullmatrix else "rk" if self.is_rkmatrix else "hmat"
        return f"Hmat({self.m}x{self.n}, {kind})"


# ----------------------------------------------------------------------------
# NewHmat / __c1
# ----------------------------------------------------------------------------

def _lowrank_block(par: Params, s: Cluster, t: Cluster):
    """Factors of an admissible block by ``ConsMethod`` (cluster.jl:72-98)."""
    if par.ConsMethod == "aca2":
        return lowrank.aca2(par.Kernel, s.X, t.X, par.MaxRank - 1, par.εComp)
    if par.ConsMethod == "separate":
        return lowrank._rk_matrix(par.α, par.β, s.X, t.X)
    if par.ConsMethod == "bbfmm":
        return lowrank.bbfmm(par.Kernel, s.X, t.X, par.MaxRank)
    raise AssertionError(par.ConsMethod)


def __c1(c: Cluster, par: Params, MaxBlock: int, adj: Optional[_Adjacency]) -> Hmat:
    """Recursive 2 x 2 block partition (src/cluster.jl:48-133).

    Classification order is the reference's: full if either side is small or
    a leaf; hmat if larger than ``MaxBlock``; hmat on the diagonal; otherwise
    low rank when the clusters are not neighbours (ConsMethod aca2 / separate
    / bbfmm) or when the SVD rank is at most ``MaxRank`` (any other
    ConsMethod: dense + ``compress``).
    """
    named = par.ConsMethod in ("aca2", "separate", "bbfmm")

    def helper(H: Hmat, s: Cluster, t: Cluster) -> None:
        H.m, H.n, H.s, H.t = s.N, t.N, s, t
        U = V = None
        if H.m <= par.MinBlock or H.n <= par.MinBlock or s.isleaf or t.isleaf:
            H.is_fullmatrix = True
        elif H.m > MaxBlock or H.n > MaxBlock:
            H.is_hmat = True
        elif s.same_range(t):
            H.is_hmat = True           # s, t are not leaves here (cluster.jl:65-70)
        elif named:
            if not adj.neighbours(s, t):
                H.is_rkmatrix = True
                U, V = _lowrank_block(par, s, t)
                if par.ConsMethod == "aca2" and par.aca_force and U.shape[1] > par.MaxRank:
                    H.is_rkmatrix, H.is_hmat = False, True
            else:
                H.is_hmat = True
        else:
            A = kernel_full(par.Kernel, s.X, t.X)
            U, V = lowrank.compress(A, par)
            k = U.shape[1]
            if k <= par.MaxRank:
                H.is_rkmatrix = True
            else:
                H.is_hmat = True
            if par.verbose:
                print(f"{A.shape}, {k}({par.MaxRank}), {H.is_rkmatrix} ")

        if H.is_fullmatrix:
            H.C = kernel_full(par.Kernel, s.X, t.X)
        elif H.is_rkmatrix:
            H.A = np.asfortranarray(U, dtype=np.float64)
            H.B = np.asfortranarray(V, dtype=np.float64)
        else:
            H.children = [[Hmat(), Hmat()], [Hmat(), Hmat()]]
            helper(H.children[0][0], s.left, t.left)
            helper(H.children[0][1], s.left, t.right)
            helper(H.children[1][0], s.right, t.left)
            helper(H.children[1][1], s.right, t.right)

    h = Hmat()
    helper(h, c, c)
    return h


def NewHmat(params: Optional[Params] = None):
    """Build ``(c, H)`` from ``Hparams`` (src/cluster.jl:36-43).

    ``params`` defaults to the global ``Hparams``; it is read, never written.
    """
    par = Hparams if params is None else params
    if par.Geom is None:
        raise ValueError("Hparams.Geom is nothing")
    if par.Kernel is None and par.ConsMethod != "separate":
        raise ValueError("Hparams.Kernel is nothing")
    X = _as_points(par.Geom)
    N = X.shape[0]
    MaxBlock = N // 4 if par.MaxBlock == -1 else par.MaxBlock
    ncluster, c = construct_cluster(X, par.MinBlock)
    named = par.ConsMethod in ("aca2", "separate", "bbfmm")
    adj = _Adjacency(par.Adj, c.P) if named else None
    H = __c1(c, par, MaxBlock, adj)
    return c, H


class UniformScaling:
    """``u*I`` for ``H + u*I`` (Julia's ``LinearAlgebra.I``)."""

    __array_ufunc__ = None

    def __init__(self, value: float = 1.0):
        self.value = value

    def __mul__(self, u):
        return UniformScaling(self.value * u)

    __rmul__ = __mul__


I = UniformScaling(1.0)


def copy_hmat(H: Hmat) -> Hmat:
    """Deep copy of the leaves (utils.jl:129-159); cluster references shared."""
    G = Hmat(m=H.m, n=H.n, s=H.s, t=H.t, is_rkmatrix=H.is_rkmatrix, is_fullmatrix=H.is_fullmatrix,
             is_hmat=H.is_hmat, P=np.array(H.P, copy=True))
    if H.is_fullmatrix:
        G.C = np.array(H.C, copy=True)
    elif H.is_rkmatrix:
        G.A, G.B = np.array(H.A, copy=True, order="F"), np.array(H.B, copy=True, order="F")
    else:
        G.children = [[copy_hmat(c) for c in row] for row in H.children]
    return G


def _add_diag(H: Hmat, u: float) -> None:
    if H.is_fullmatrix:
        if H.m != H.n:
            raise ValueError("a diagonal leaf is not square")
        H.C[np.diag_indices(H.m)] += u
    elif H.is_rkmatrix:
        raise ValueError("H should not be a low-rank matrix")   # harithm.jl:462
    else:
        _add_diag(H.children[0][0], u)
        _add_diag(H.children[1][1], u)
