"""ctypes front-end of the FP64 CPU oracle (hmat_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product path never imports
this module.  Each function follows the reference file:line cited in
hmat_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int64
_D = ctypes.c_double


def build() -> str:
    """Compile liboracle.so from hmat_oracle.c (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.or_matvec.argtypes = [_I, _P, _P, _P, _P, _D]
        L.or_matvec.restype = None
        L.or_matvec_mt.argtypes = [_I, _P, _P, _P, _P, _D, _I, ctypes.c_int]
        L.or_matvec_mt.restype = None
        L.or_matmat.argtypes = [_I, _P, _P, _I, _P, _I, _P, _I, _D]
        L.or_matmat.restype = None
        L.or_matvec_t.argtypes = [_I, _P, _P, _P, _P, _D]
        L.or_matvec_t.restype = None
        L.or_to_dense.argtypes = [_I, _I, _I, _P, _P, _P]
        L.or_to_dense.restype = None
        L.or_adjoint.argtypes = [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]
        L.or_adjoint.restype = _D
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _arrs(f):
    b = np.ascontiguousarray(f.blocks, dtype=np.int64)
    d = np.ascontiguousarray(f.data, dtype=np.float64)
    return b, d


def matvec(f, v, r=None, s: float = 1.0) -> np.ndarray:
    """r += s*H*v (r = 0 when not given)."""
    b, d = _arrs(f)
    v = np.ascontiguousarray(v, dtype=np.float64)
    assert v.shape == (f.ncols,)
    r = np.zeros(f.nrows) if r is None else np.array(r, dtype=np.float64, copy=True)
    lib().or_matvec(b.shape[0], _p(b), _p(d), _p(v), _p(r), s)
    return r


def matvec_mt(f, v, threads: int, r=None, s: float = 1.0) -> np.ndarray:
    """r += s*H*v on `threads` OpenMP threads; bitwise equal to matvec."""
    b, d = _arrs(f)
    v = np.ascontiguousarray(v, dtype=np.float64)
    assert v.shape == (f.ncols,) and threads >= 1
    r = np.zeros(f.nrows) if r is None else np.array(r, dtype=np.float64, copy=True)
    lib().or_matvec_mt(b.shape[0], _p(b), _p(d), _p(v), _p(r), s, f.nrows, int(threads))
    return r


def matmat(f, X, s: float = 1.0) -> np.ndarray:
    """H*X column by column (harithm.jl:317-323)."""
    b, d = _arrs(f)
    X = np.asfortranarray(X, dtype=np.float64)
    R = np.zeros((f.nrows, X.shape[1]), order="F")
    lib().or_matmat(b.shape[0], _p(b), _p(d), X.shape[1], _p(X), max(1, f.ncols), _p(R),
                    max(1, f.nrows), s)
    return R


def matvec_t(f, v, r=None, s: float = 1.0) -> np.ndarray:
    """r += s*H'*v."""
    b, d = _arrs(f)
    v = np.ascontiguousarray(v, dtype=np.float64)
    assert v.shape == (f.nrows,)
    r = np.zeros(f.ncols) if r is None else np.array(r, dtype=np.float64, copy=True)
    lib().or_matvec_t(b.shape[0], _p(b), _p(d), _p(v), _p(r), s)
    return r


def to_dense(f) -> np.ndarray:
    b, d = _arrs(f)
    out = np.zeros((f.nrows, f.ncols), order="F")
    lib().or_to_dense(f.nrows, f.ncols, b.shape[0], _p(b), _p(d), _p(out))
    return out


def adjoint(f, x, bvec):
    """Returns (L, res, dx, grad) for L = 0.5*||H x - b||^2."""
    b, d = _arrs(f)
    x = np.ascontiguousarray(x, dtype=np.float64)
    bvec = np.ascontiguousarray(bvec, dtype=np.float64)
    res = np.zeros(f.nrows)
    dx = np.zeros(f.ncols)
    grad = np.zeros_like(d)
    L = lib().or_adjoint(f.nrows, f.ncols, b.shape[0], _p(b), _p(d), _p(x), _p(bvec), _p(res),
                         _p(dx), _p(grad))
    return L, res, dx, grad
