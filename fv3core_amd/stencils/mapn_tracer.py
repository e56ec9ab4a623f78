"""MapNTracer (fv3core/stencils/mapn_tracer.py:10-83) on the GPU.

Each tracer goes through MapSingle (iv = 0) from pe1 to pe2, the sixth one
with kord 9 whatever the configured kord (mapn_tracer.py:41-42, index 5 is
qgraupel), then FillNegativeTracerValues when ``fill`` is set.
"""
from __future__ import annotations

from typing import Dict

from ..utils import gt4py_utils as utils
from ..utils.stencil import StencilFactory
from .fillz import FillNegativeTracerValues
from .map_single import MapSingle


class MapNTracer:
    """Fortran mapn_tracer; ``__call__(pe1, pe2, dp2, tracers, q_min)``."""

    def __init__(self, stencil_factory: StencilFactory, kord: int, nq: int, i1: int, i2: int,
                 j1: int, j2: int, fill: bool):
        grid_indexing = stencil_factory.grid_indexing
        self._nk = grid_indexing.domain[2]
        self._nq = int(nq)
        kord_tracer = [kord] * self._nq
        if self._nq > 5:
            kord_tracer[5] = 9
        self._list_of_remap_objects = [
            MapSingle(stencil_factory, kord_tracer[i], 0, i1, i2, j1, j2)
            for i in range(len(kord_tracer))
        ]
        self._fill_negative_tracers = bool(fill)
        if fill:
            self._fillz = FillNegativeTracerValues(
                stencil_factory, i2 - i1 + 1, j2 - j1 + 1, self._nk, self._nq)

    def __call__(self, pe1, pe2, dp2, tracers: Dict[str, object], q_min: float):
        for i, q in enumerate(utils.tracer_variables[0:self._nq]):
            self._list_of_remap_objects[i](tracers[q], pe1, pe2)
        if self._fill_negative_tracers:
            self._fillz(dp2, tracers)
