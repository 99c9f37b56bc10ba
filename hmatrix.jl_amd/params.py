"""Global construction parameters, mirroring Hmatrix.jl's ``Hparams``.

Reference: src/utils.jl:8-26 (``@with_kw mutable struct Params`` and the
global ``Hparams = Params()``).  Field names and defaults are kept, including
the Greek ones (``εComp``, ``α``, ``β``, ``εTrunc``), so code written against
the reference reads the same:

    from hmatrix_amd import Hparams, NewHmat
    Hparams.Geom = np.linspace(0, 10, 2000)
    Hparams.Kernel = lambda x, y: 1 / (1 + (x - y) ** 2)

Differences from the reference, all deliberate:

* ``NewHmat`` does not write back into ``Hparams`` (the reference rewrites
  ``MaxBlock`` when it is -1, permutes ``Adj`` in place and stores
  ``Neighbors``; cluster.jl:37-39, 141, 168).  A second ``NewHmat`` call on a
  different geometry therefore sees the user's values, not the previous run's.
* ``Kernel`` is called on whole numpy arrays (broadcasting) instead of one
  scalar pair at a time; see ``cluster.kernel_full`` for the convention.
* ``Adj=None`` raises a ``ValueError`` with a message, where the reference
  fails with a MethodError on ``nothing[c.P, c.P]`` (cluster.jl:141).
"""
from __future__ import annotations

import copy as _copy
from dataclasses import dataclass, field, fields
from typing import Any, Callable, List, Optional


@dataclass
class Params:
    """Construction parameters (src/utils.jl:8-26)."""

    Geom: Any = None            # points, (N,) or (N, d) float64
    Adj: Any = None             # scipy.sparse (N, N) point adjacency
    Neighbors: Any = None       # kept for API parity; NewHmat does not set it
    εComp: float = 1e-10        # ACA stopping tolerance (hlu.jl:51-53)
    MaxRank: int = 16
    MaxBlock: int = -1          # -1 -> N // 4 (cluster.jl:37-39)
    MinBlock: int = 64
    Kernel: Optional[Callable] = None
    α: Optional[List[Callable]] = None
    β: Optional[List[Callable]] = None
    CompMethod: str = "svd"
    ConsMethod: str = "bbfmm"
    εTrunc: float = 1e-10       # SVD truncation (hlu.jl:175)
    verbose: bool = False
    aca_force: bool = False

    def copy(self) -> "Params":
        return _copy.copy(self)

    def reset(self) -> None:
        """Restore every field to its default (tests use this between cases)."""
        for f in fields(self):
            setattr(self, f.name, f.default)


Hparams = Params()
