r"""hmatrix_amd: MI355X H-matrix products behind Hmatrix.jl's interface.

Host side (FP64, numpy): ``Hparams``, ``NewHmat``, ``FullMat``, ``Hmat``,
``Cluster``, ``aca2``, ``info``, ``Array``/``to_fmat`` -- the names of
src/cluster.jl:1-5, src/utils.jl:2-6 and src/hlu.jl:7 -- and the H-LU
``lu`` / ``lu_inplace`` (``lu!``) / ``hsolve`` (``\``) of src/hlu.jl:465-561.
Device side (HIP, gfx950): ``H * y``, ``H * X``, ``v * H``, ``H.T * y`` and
``DeviceHmat`` for device-resident products in FP64, FP32, BF16 or split
BF16 storage.

The directory is named ``hmatrix.jl_amd``; it is imported as
``hmatrix_amd`` (see ``__graft_entry__.package()``).
"""
from .params import Hparams, Params
from .cluster import (Cluster, Hmat, NewHmat, FullMat, kernel_full, construct_cluster, bisect_cluster,
                      UniformScaling, I, copy_hmat)
from .solve import gmres, bicgstab, solve_ir, factorized
from .hlu import lu, lu_inplace, solve as hsolve, hmul, hadd, getl, getu
from .lowrank import aca2, compress, rank_truncate, bbfmm
from .flat import (FlatHmat, flatten, to_dense, to_fmat, Array, info, save_factors, load_factors,
                   KIND_DENSE, KIND_RK)

__all__ = [
    "Hparams", "Params", "Cluster", "Hmat", "NewHmat", "FullMat", "kernel_full",
    "construct_cluster", "bisect_cluster", "aca2", "compress", "rank_truncate", "bbfmm",
    "FlatHmat", "flatten", "to_dense", "to_fmat", "Array", "info", "save_factors",
    "load_factors", "KIND_DENSE", "KIND_RK", "DeviceHmat", "hmatvec", "UniformScaling", "I",
    "copy_hmat", "gmres", "bicgstab", "solve_ir", "factorized", "lu", "lu_inplace", "hsolve", "hmul", "hadd",
    "getl", "getu",
]


def __getattr__(name):
    if name in ("DeviceHmat", "TransposedHmat", "hmatvec"):
        from . import device
        return getattr(device, name)
    raise AttributeError(name)
