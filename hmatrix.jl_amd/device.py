"""Device-resident H-matrix and its products (the drop-in for ``H*y``).

``DeviceHmat`` owns one plan of libhmat_amd.so: the exported factors packed
in a storage precision and uploaded once.  Products take and return torch
tensors on the plan's device (torch is used only for device memory and the
stream handle) or, through ``mul``/``rmul``, numpy arrays like the
reference's ``H*y`` / ``v*H``:

* ``Base.:*(H, v)``   src/harithm.jl:325-329  -> ``DeviceHmat.mul`` (1-D)
* ``Base.:*(H, X)``   src/harithm.jl:317-323  -> ``DeviceHmat.mul`` (2-D)
* ``hmat_matvec!``    src/harithm.jl:340-354  -> ``DeviceHmat.matvec(alpha=s, beta=1)``
* ``Base.:*(v, H)``   src/harithm.jl:331-337  -> ``DeviceHmat.rmul``
* ``transpose_hmat_mat_mul`` harithm.jl:243-260 -> ``DeviceHmat.matmat_t``
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _native
from .flat import FlatHmat, flatten


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: H-matrix products run only on the HIP device")
    return torch


def _stream_handle(stream) -> ctypes.c_void_p:
    torch = _torch()
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


class DeviceHmat:
    """One H-matrix resident on the GPU in a given storage precision."""

    def __init__(self, flat: FlatHmat, dtype: str = "f64", config: Optional[dict] = None,
                 _batch=None):
        """Upload ``flat`` in storage ``dtype``.

        ``config`` (optional) sets ``hm_config`` fields: ``proj_chunk``,
        ``row_waves``, ``no_transpose``, ``no_batched``, ``strip_shape``
        (0 auto, 1 narrow, 2 wide), ``one_launch`` (many-matrix plans: one
        launch per single-right-hand-side product instead of three).
        """
        torch = _torch()
        if dtype not in _native.STORAGE:
            raise ValueError(f"dtype must be one of {sorted(_native.STORAGE)}")
        L = _native.lib()
        torch.cuda.init()
        self.flat = flat
        self.dtype = dtype
        self.config = dict(config or {})
        self.vec_dtype = torch.float64 if dtype == "f64" else torch.float32
        self.device = torch.device("cuda", torch.cuda.current_device())
        cfg = _native.HmConfig(int(self.config.get("proj_chunk", 0)), int(self.config.get("row_waves", 0)),
                               (_native.HM_NO_TRANSPOSE if self.config.get("no_transpose") else 0)
                               | (_native.HM_NO_BATCHED if self.config.get("no_batched") else 0)
                               | (_native.HM_ONE_LAUNCH if self.config.get("one_launch") else 0),
                               int(self.config.get("strip_shape", 0)))
        h = ctypes.c_void_p()
        if _batch is None:
            b = np.ascontiguousarray(flat.blocks, dtype=np.int64)
            d = np.ascontiguousarray(flat.data, dtype=np.float64)
            rc = L.hm_plan_create_ex(flat.nrows, flat.ncols, b.shape[0],
                                     b.ctypes.data_as(ctypes.c_void_p), d.ctypes.data_as(ctypes.c_void_p),
                                     d.size, _native.STORAGE[dtype], ctypes.byref(cfg), ctypes.byref(h))
            _native.check(rc, "hm_plan_create_ex")
            self.nrows, self.ncols = flat.nrows, flat.ncols
            self.ndata = int(flat.data.size)
        else:
            members = _batch
            n = len(members)
            arr = lambda vals: (ctypes.c_int64 * n)(*vals)
            keep_b = [np.ascontiguousarray(m.blocks, dtype=np.int64) for m in members]
            keep_d = [np.ascontiguousarray(m.data, dtype=np.float64) for m in members]
            pb = (ctypes.c_void_p * n)(*[k.ctypes.data for k in keep_b])
            pd = (ctypes.c_void_p * n)(*[k.ctypes.data for k in keep_d])
            rc = L.hm_plan_create_batch(n, arr([m.nrows for m in members]), arr([m.ncols for m in members]),
                                        arr([k.shape[0] for k in keep_b]), pb, pd, arr([k.size for k in keep_d]),
                                        _native.STORAGE[dtype], ctypes.byref(cfg), ctypes.byref(h))
            _native.check(rc, "hm_plan_create_batch")
            self.nrows = sum(m.nrows for m in members)
            self.ncols = sum(m.ncols for m in members)
            self.ndata = int(sum(k.size for k in keep_d))
            self.row_offsets = np.concatenate([[0], np.cumsum([m.nrows for m in members])])
            self.col_offsets = np.concatenate([[0], np.cumsum([m.ncols for m in members])])
        self._h = h

    @classmethod
    def batch(cls, flats, dtype: str = "f32", config: Optional[dict] = None) -> "DeviceHmat":
        """One plan for many independent H-matrices (hm_plan_create_batch):
        one product launch covers all of them; member i acts on the slice
        ``col_offsets[i]:col_offsets[i+1]`` of the concatenated vector."""
        return cls(None, dtype, config, _batch=list(flats))

    @classmethod
    def from_hmat(cls, H, dtype: str = "f64") -> "DeviceHmat":
        return cls(flatten(H), dtype)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _native.lib().hm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        out = (ctypes.c_int64 * len(_native.INFO_FIELDS))()
        _native.check(_native.lib().hm_plan_info(self._h, out, len(out)), "hm_plan_info")
        return dict(zip(_native.INFO_FIELDS, list(out)))

    def timing(self, enable: bool = True) -> None:
        """Record HIP events around the projection, reduction and row-pass
        kernels of every later product on this plan (not graph replays)."""
        _native.check(_native.lib().hm_plan_timing(self._h, 1 if enable else 0), "hm_plan_timing")

    def timing_read(self) -> dict:
        """Summed kernel milliseconds since the last read, per phase, and the
        number of launch sequences they cover; waits for the recorded events."""
        ms = (ctypes.c_double * 3)()
        n = ctypes.c_int64(0)
        _native.check(_native.lib().hm_plan_timing_read(self._h, ms, 3, ctypes.byref(n)),
                      "hm_plan_timing_read")
        return {"proj_ms": ms[0], "reduce_ms": ms[1], "rows_ms": ms[2], "launches": n.value}

    # ------------------------------------------------------------------
    def _vec(self, t, n: int, name: str):
        torch = _torch()
        if not isinstance(t, torch.Tensor) or t.device != self.device or t.dtype != self.vec_dtype:
            raise TypeError(f"{name} must be a {self.vec_dtype} tensor on {self.device}")
        if t.dim() != 1 or t.shape[0] != n or (n > 1 and t.stride(0) != 1):
            raise ValueError(f"{name} must be a contiguous vector of length {n}")
        return ctypes.c_void_p(t.data_ptr())

    def _mat(self, t, n: int, name: str):
        """Column-major (n x nrhs) tensor -> (pointer, nrhs, ld)."""
        torch = _torch()
        if not isinstance(t, torch.Tensor) or t.device != self.device or t.dtype != self.vec_dtype:
            raise TypeError(f"{name} must be a {self.vec_dtype} tensor on {self.device}")
        if t.dim() != 2 or t.shape[0] != n:
            raise ValueError(f"{name} must have {n} rows")
        if t.shape[1] > 1 and t.stride(0) != 1:
            raise ValueError(f"{name} must be column-major (stride(0) == 1)")
        ld = t.stride(1) if t.shape[1] > 1 else max(1, n)
        return ctypes.c_void_p(t.data_ptr()), t.shape[1], ld

    def new_cols(self, n: int, nrhs: int):
        """Zero column-major (n x nrhs) tensor on the device."""
        torch = _torch()
        return torch.zeros((nrhs, n), dtype=self.vec_dtype, device=self.device).t()

    def matvec(self, x, y=None, alpha: float = 1.0, beta: float = 0.0, stream=None, transpose=False):
        """y = beta*y + alpha*H*x (or H' with transpose=True), device tensors."""
        torch = _torch()
        m, n = (self.ncols, self.nrows) if transpose else (self.nrows, self.ncols)
        px = self._vec(x, n, "x")
        if y is None:
            y = torch.zeros(m, dtype=self.vec_dtype, device=self.device)
            beta = 0.0
        py = self._vec(y, m, "y")
        fn = "hm_matvec_t" if transpose else "hm_matvec"
        rc = getattr(_native.lib(), fn)(self._h, px, py, float(alpha), float(beta), _stream_handle(stream))
        _native.check(rc, fn)
        return y

    def matmat(self, X, Y=None, alpha: float = 1.0, beta: float = 0.0, stream=None, transpose=False):
        """Y = beta*Y + alpha*H*X for column-major device matrices."""
        m, n = (self.ncols, self.nrows) if transpose else (self.nrows, self.ncols)
        px, nrhs, ldx = self._mat(X, n, "X")
        if Y is None:
            Y = self.new_cols(m, nrhs)
            beta = 0.0
        py, nrhs2, ldy = self._mat(Y, m, "Y")
        if nrhs2 != nrhs:
            raise ValueError("X and Y have different column counts")
        fn = "hm_matmat_t" if transpose else "hm_matmat"
        rc = getattr(_native.lib(), fn)(self._h, nrhs, px, ldx, py, ldy, float(alpha), float(beta),
                                        _stream_handle(stream))
        _native.check(rc, fn)
        return Y

    def matmat_t(self, X, Y=None, alpha: float = 1.0, beta: float = 0.0, stream=None):
        return self.matmat(X, Y, alpha, beta, stream, transpose=True)

    def graph(self, X, Y, alpha: float = 1.0, beta: float = 0.0, transpose: bool = False):
        """Capture ``Y = beta*Y + alpha*op(H)*X`` on these buffers into a HIP
        graph (hm_graph_create); returns a ``CapturedProduct``."""
        m, n = (self.ncols, self.nrows) if transpose else (self.nrows, self.ncols)
        if X.dim() == 1:
            px, nrhs, ldx = self._vec(X, n, "X"), 1, max(1, n)
            py, ldy = self._vec(Y, m, "Y"), max(1, m)
        else:
            px, nrhs, ldx = self._mat(X, n, "X")
            py, nrhs2, ldy = self._mat(Y, m, "Y")
            if nrhs2 != nrhs:
                raise ValueError("X and Y have different column counts")
        g = ctypes.c_void_p()
        rc = _native.lib().hm_graph_create(self._h, _native.HM_OP_T if transpose else _native.HM_OP_N, nrhs,
                                           px, ldx, py, ldy, float(alpha), float(beta), ctypes.byref(g))
        _native.check(rc, "hm_graph_create")
        return CapturedProduct(g, self, X, Y)

    def adjoint(self, x, b, grad: bool = True, loss: bool = True, stream=None, out=None):
        """Gradients of L = 0.5*||H x - b||^2 on the device.

        Returns ``(L, res, dx, g)``: ``res = H x - b``, ``dx = H' res`` and
        ``g`` the factor gradients in the exported arena layout (``None``
        when ``grad`` is False; ``L`` is ``None`` when ``loss`` is False).
        ``out = (res, dx, g)`` reuses device buffers from an earlier call
        (``g`` may be ``None`` when ``grad`` is False): the product writes
        every element of res and dx and every arena element that belongs to
        a leaf; a new ``g`` starts as zeros so arena gaps read as zero.
        """
        torch = _torch()
        px = self._vec(x, self.ncols, "x")
        pb = self._vec(b, self.nrows, "b")
        if out is not None:
            res, dx, g = out
            self._vec(res, self.nrows, "res")
            self._vec(dx, self.ncols, "dx")
            if grad:
                if g is None:
                    raise ValueError("out: g is None but grad=True")
                self._vec(g, self.ndata, "g")
        else:
            res = torch.empty(self.nrows, dtype=self.vec_dtype, device=self.device)
            dx = torch.empty(self.ncols, dtype=self.vec_dtype, device=self.device)
            g = torch.zeros(self.ndata, dtype=self.vec_dtype, device=self.device) if grad else None
        L = ctypes.c_double(0.0)
        rc = _native.lib().hm_adjoint(self._h, px, pb, ctypes.c_void_p(res.data_ptr()),
                                      ctypes.c_void_p(dx.data_ptr()),
                                      ctypes.c_void_p(g.data_ptr()) if grad else None,
                                      ctypes.byref(L) if loss else None, _stream_handle(stream))
        _native.check(rc, "hm_adjoint")
        return (L.value if loss else None), res, dx, g

    def vjp(self, x, g, grad: bool = True, stream=None):
        """Pullback of ``y = H x`` for the cotangent ``g`` (hm_vjp).

        Returns ``(dx, dg)``: ``dx = H' g`` and ``dg`` the gradient of
        ``<g, H x>`` with respect to every arena entry in the exported
        layout (dense C: ``g x'``, low-rank A: ``g (B'x)'``, B: ``x (A'g)'``;
        zeros outside the leaves), or ``None`` when ``grad`` is False.
        """
        torch = _torch()
        px = self._vec(x, self.ncols, "x")
        pg = self._vec(g, self.nrows, "g")
        dx = torch.empty(self.ncols, dtype=self.vec_dtype, device=self.device)
        dg = torch.zeros(self.ndata, dtype=self.vec_dtype, device=self.device) if grad else None
        rc = _native.lib().hm_vjp(self._h, px, pg, ctypes.c_void_p(dx.data_ptr()),
                                  ctypes.c_void_p(dg.data_ptr()) if grad else None, _stream_handle(stream))
        _native.check(rc, "hm_vjp")
        return dx, dg

    # ------------------------------------------------------------------
    # numpy in, numpy out (the reference's H*y / H*X / v*H)
    def mul(self, v: np.ndarray) -> np.ndarray:
        torch = _torch()
        v = np.asarray(v)
        if v.ndim == 1:
            if v.shape[0] != self.ncols:
                raise ValueError(f"H is {self.nrows}x{self.ncols}; vector has {v.shape[0]} entries")
            x = torch.from_numpy(np.ascontiguousarray(v)).to(self.device, self.vec_dtype)
            return self.matvec(x).cpu().numpy().astype(np.float64)
        if v.ndim == 2:
            if v.shape[0] != self.ncols:
                raise ValueError(f"H is {self.nrows}x{self.ncols}; matrix has {v.shape[0]} rows")
            X = self.new_cols(self.ncols, v.shape[1])
            X.copy_(torch.from_numpy(np.asarray(v)).to(self.device, self.vec_dtype))
            return self.matmat(X).cpu().numpy().astype(np.float64)
        raise ValueError("H * v needs a vector or a matrix")

    def rmul(self, v: np.ndarray) -> np.ndarray:
        """v (p x m) * H -> p x n."""
        torch = _torch()
        v = np.asarray(v)
        if v.ndim != 2 or v.shape[1] != self.nrows:
            raise ValueError(f"v * H needs a p x {self.nrows} array")
        X = self.new_cols(self.nrows, v.shape[0])
        X.copy_(torch.from_numpy(np.ascontiguousarray(v.T)).to(self.device, self.vec_dtype))
        return self.matmat_t(X).cpu().numpy().T.astype(np.float64)


class CapturedProduct:
    """A product captured into a HIP graph; ``launch()`` replays it."""

    def __init__(self, handle, plan: DeviceHmat, X, Y):
        self._g = handle
        self._keep = (plan, X, Y)       # the graph refers to these buffers

    def launch(self, stream=None) -> None:
        _native.check(_native.lib().hm_graph_launch(self._g, _stream_handle(stream)), "hm_graph_launch")

    def close(self) -> None:
        if getattr(self, "_g", None) is not None and self._g.value:
            _native.lib().hm_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TransposedHmat:
    """``H.T``: products with the transpose, ``H.T * y``."""

    __array_ufunc__ = None

    def __init__(self, H):
        self.H = H

    @property
    def shape(self):
        return (self.H.n, self.H.m)

    def __mul__(self, v):
        v = np.asarray(v, dtype=np.float64)
        d = self.H.device("f64")
        if v.ndim == 1:
            return d.rmul(v[None, :])[0]
        return d.rmul(v.T).T

    __matmul__ = __mul__


def hmatvec(dev: DeviceHmat, x, data=None):
    """Differentiable ``H x`` for torch autograd.

    The gradient with respect to x is ``H' g``.  ``data`` (optional) is a
    tensor standing for the plan's factor arena in the exported layout (the
    values the plan was built from, e.g. ``torch.from_numpy(flat.data)``
    with ``requires_grad``): its gradient is the factor tangent of the
    product, ``dC = g x'``, ``dA = g (B'x)'``, ``dB = x (A'g)'`` per leaf
    (hm_vjp).  The product itself always uses the plan's own copy.
    """
    if data is not None and data.numel() != dev.ndata:
        raise ValueError(f"data has {data.numel()} entries; the plan's arena has {dev.ndata}")
    return _HMatVec.apply(x, data, dev)


def _make_fn():
    import torch

    class HMatVec(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, data, dev):
            ctx.dev = dev
            ctx.data_meta = None if data is None else (data.dtype, data.device, data.shape)
            x = x.contiguous()
            ctx.save_for_backward(x)
            return dev.matvec(x)

        @staticmethod
        def backward(ctx, g):
            (x,) = ctx.saved_tensors
            g = g.contiguous().to(ctx.dev.vec_dtype)
            want = ctx.data_meta is not None and ctx.needs_input_grad[1]
            if not want:
                return ctx.dev.matvec(g, transpose=True), None, None
            dx, dg = ctx.dev.vjp(x, g, grad=True)
            dtype, device, shape = ctx.data_meta
            return dx, dg.to(device=device, dtype=dtype).reshape(shape), None

    return HMatVec


class _Lazy:
    fn = None

    def apply(self, *a):
        if _Lazy.fn is None:
            _Lazy.fn = _make_fn()
        return _Lazy.fn.apply(*a)


_HMatVec = _Lazy()
