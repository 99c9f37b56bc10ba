"""ctypes binding of libhmat_amd.so (include/hmat_amd.h).

The library is built in-tree by ``__graft_entry__.build()`` (hipcc,
gfx950).  Loading fails loudly when it is missing: there is no CPU
fallback for any product.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhmat_amd.so")

HM_F64, HM_F32, HM_BF16, HM_BF16X2 = 0, 1, 2, 3
STORAGE = {"f64": HM_F64, "f32": HM_F32, "bf16": HM_BF16, "bf16x2": HM_BF16X2}
STATUS = {0: "HM_OK", -1: "HM_EINVAL", -2: "HM_EHIP", -3: "HM_ENOMEM", -4: "HM_EUNSUPPORTED"}
INFO_FIELDS = ("nrows", "ncols", "ndense", "nrk", "dense_elems", "rk_elems", "storage",
               "device_bytes", "factor_bytes", "maxrank", "proj_chunk", "row_waves", "flags", "strip_shape",
               "batched", "fused", "wait_faults")
HM_NO_TRANSPOSE, HM_NO_BATCHED, HM_ONE_LAUNCH = 1, 2, 4
HM_STRIPS_AUTO, HM_STRIPS_NARROW, HM_STRIPS_WIDE = 0, 1, 2
HM_OP_N, HM_OP_T = 0, 1


class HmConfig(ctypes.Structure):
    """struct hm_config (include/hmat_amd.h)."""
    _fields_ = [("proj_chunk", ctypes.c_int32), ("row_waves", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("strip_shape", ctypes.c_int32)]

# Every symbol include/hmat_amd.h declares, with its ctypes signature.
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
SIGNATURES = {
    "hm_plan_create": (ctypes.c_int, [_I64, _I64, _I64, _P, _P, _I64, ctypes.c_int, ctypes.POINTER(_P)]),
    "hm_plan_create_ex": (ctypes.c_int, [_I64, _I64, _I64, _P, _P, _I64, ctypes.c_int, _P,
                                         ctypes.POINTER(_P)]),
    "hm_plan_create_batch": (ctypes.c_int, [_I64, _P, _P, _P, _P, _P, _P, ctypes.c_int, _P,
                                            ctypes.POINTER(_P)]),
    "hm_plan_destroy": (ctypes.c_int, [_P]),
    "hm_plan_info": (ctypes.c_int, [_P, ctypes.POINTER(_I64), ctypes.c_int]),
    "hm_matvec": (ctypes.c_int, [_P, _P, _P, _D, _D, _P]),
    "hm_matvec_static": (ctypes.c_int, [_P, _P, _P, _D, _D, _P]),
    "hm_matmat": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64, _D, _D, _P]),
    "hm_matvec_t": (ctypes.c_int, [_P, _P, _P, _D, _D, _P]),
    "hm_matmat_t": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64, _D, _D, _P]),
    "hm_adjoint": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, ctypes.POINTER(_D), _P]),
    "hm_vjp": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "hm_graph_create": (ctypes.c_int, [_P, ctypes.c_int, _I64, _P, _I64, _P, _I64, _D, _D,
                                       ctypes.POINTER(_P)]),
    "hm_autotune": (ctypes.c_int, [_I64, _I64, _I64, _P, _P, _I64, ctypes.c_int, _I64, ctypes.c_int, _P,
                                   ctypes.POINTER(_D)]),
    "hm_graph_launch": (ctypes.c_int, [_P, _P]),
    "hm_graph_destroy": (ctypes.c_int, [_P]),
    "hm_plan_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "hm_plan_timing_read": (ctypes.c_int, [_P, ctypes.POINTER(_D), ctypes.c_int, ctypes.POINTER(_I64)]),
    "hm_last_error": (ctypes.c_char_p, []),
    "hm_version": (ctypes.c_char_p, []),
}


class HmError(RuntimeError):
    """A negative status from the library, with its message."""

    def __init__(self, code: int, fn: str, msg: str):
        super().__init__(f"{fn} -> {STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libhmat_amd.so once; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, fn: str) -> None:
    if rc != 0:
        raise HmError(rc, fn, lib().hm_last_error().decode())
