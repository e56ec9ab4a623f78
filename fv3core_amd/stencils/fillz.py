"""FillNegativeTracerValues (fv3core/stencils/fillz.py:101-150) on the GPU.

``fv3_fillz`` runs fix_tracer (fillz.py:12-98) with one lane per column and
one grid plane per tracer: negative layers borrow from the layer above, then
from the layer below, and a column that still needed a fix is rescaled below
its top layer so that its mass is kept.
"""
from __future__ import annotations

import ctypes
from typing import Any, Dict

from .. import _lib
from ..utils import gt4py_utils as utils
from ..utils.stencil import StencilFactory

MAX_TRACERS = 9


def tracer_pointers(tracers, names):
    """ctypes array of device pointers of the named tracers, checked for one layout."""
    arr = (ctypes.c_void_p * MAX_TRACERS)()
    for n, name in enumerate(names):
        t = tracers[name]
        _lib.require_device(t)
        arr[n] = t.data_ptr()
    return arr


class FillNegativeTracerValues:
    """Fortran fillz; ``__call__(dp2, tracers)`` fixes the first nq tracers in place."""

    def __init__(self, stencil_factory: StencilFactory, im: int, jm: int, km: int, nq: int):
        if nq > MAX_TRACERS:
            raise ValueError(f"at most {MAX_TRACERS} tracers (got {nq})")
        self._gi = stencil_factory.grid_indexing
        self._im, self._jm, self._km = int(im), int(jm), int(km)
        self._nq = int(nq)

    def __call__(self, dp2, tracers: Dict[str, Any]):
        names = utils.tracer_variables[0:self._nq]
        tracer_list = [tracers[name] for name in names]
        _lib.require_device(dp2, *tracer_list)
        d = _lib.make_domain(self._gi, dp2, self._km, others3=tracer_list)
        d.iec = d.isc + self._im - 1
        d.jec = d.jsc + self._jm - 1
        ptrs = tracer_pointers(tracers, names)
        _lib.check(_lib.lib().fv3_fillz_f64(
            ctypes.byref(d), ctypes.c_int32(self._nq), ptrs, _lib.ptr(dp2),
            ctypes.c_void_p(_lib.stream_handle())))
        return tracer_list
