"""AdjustNegativeTracerMixingRatio (fv3core/stencils/neg_adj3.py:279-372) on the GPU.

``fv3_neg_adj3`` runs fix_neg_water, fillq on graupel and rain,
fix_water_vapor_down and fix_neg_cloud in one kernel, one lane per column,
in three walks over the levels; the result equals the four stencils run one
after the other.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..utils.stencil import StencilFactory


class AdjustNegativeTracerMixingRatio:
    """Fortran neg_adj3: moves mass between the water species of a column so
    that none stays negative, heating or cooling pt by the latent heat
    released or taken."""

    def __init__(self, stencil_factory: StencilFactory, check_negative: bool,
                 hydrostatic: bool):
        self._gi = stencil_factory.grid_indexing
        if check_negative:
            raise NotImplementedError("Unimplemented namelist value check_negative=True")
        if hydrostatic:
            raise NotImplementedError("Unimplemented namelist hydrostatic=True")

    def __call__(self, qvapor, qliquid, qrain, qsnow, qice, qgraupel, qcld, pt, delp, delz,
                 peln):
        """Updates the seven species and pt in place (delz and peln are
        accepted for the reference's signature and not read)."""
        fields = (qvapor, qliquid, qrain, qsnow, qice, qgraupel, qcld, pt, delp)
        _lib.require_device(*fields)
        for t in fields:
            if t.dtype != torch.float64:
                raise TypeError("neg_adj3 runs in FP64: pass float64 storages")
        d = _lib.make_domain(self._gi, pt, self._gi.domain[2], others3=fields)
        _lib.check(_lib.lib().fv3_neg_adj3_f64(
            ctypes.byref(d), *(_lib.ptr(t) for t in fields),
            ctypes.c_void_p(_lib.stream_handle())))
