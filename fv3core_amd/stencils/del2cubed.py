"""HyperdiffusionDamping (fv3core/stencils/del2cubed.py:78-181) on the GPU."""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..utils.grid import DampingCoefficients
from ..utils.stencil import StencilFactory
from ._workspace import Workspace


class HyperdiffusionDamping:
    """Fortran del2_cubed: min(3, nmax) del-2 passes with corner averaging."""

    def __init__(self, stencil_factory: StencilFactory, damping_coefficients: DampingCoefficients,
                 rarea, nmax: int):
        self._gi = stencil_factory.grid_indexing
        self._del6_u = damping_coefficients.del6_u
        self._del6_v = damping_coefficients.del6_v
        self._rarea = rarea
        self._ntimes = min(3, nmax)
        self._work = Workspace(2)

    def __call__(self, qdel, cd: float):
        _lib.require_device(qdel, self._rarea)
        nk = self._gi.domain[2]
        d = _lib.make_domain(self._gi, qdel, nk, self._rarea,
                             others2=(self._del6_u, self._del6_v))
        work, slab = self._work.get(qdel, nk)
        scalar = ctypes.c_float if qdel.dtype == torch.float32 else ctypes.c_double
        _lib.check(_lib.fn("fv3_del2cubed", qdel.dtype)(
            ctypes.byref(d), _lib.ptr(self._del6_u), _lib.ptr(self._del6_v), _lib.ptr(self._rarea),
            ctypes.c_int32(self._ntimes), scalar(cd), _lib.ptr(qdel), _lib.ptr(work),
            ctypes.c_int64(slab), ctypes.c_void_p(_lib.stream_handle())))
