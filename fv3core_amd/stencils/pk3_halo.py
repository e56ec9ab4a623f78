"""PK3Halo (fv3core/stencils/pk3_halo.py:35-68) and edge_pe (pe_halo.py:6-33) on the GPU."""
from __future__ import annotations

import ctypes

from .. import _lib
from ..utils.stencil import StencilFactory


class PK3Halo:
    """Fortran pk3_halo: pk3 on the 2-wide halo ring from summed delp."""

    def __init__(self, stencil_factory: StencilFactory):
        self._gi = stencil_factory.grid_indexing

    def __call__(self, pk3, delp, ptop: float, akap: float):
        _lib.require_device(pk3, delp)
        d = _lib.make_domain(self._gi, pk3, self._gi.domain[2] + 1)
        _lib.make_domain(self._gi, delp, self._gi.domain[2], others3=(pk3,))
        _lib.check(_lib.lib().fv3_pk3_halo_f64(ctypes.byref(d), _lib.ptr(pk3), _lib.ptr(delp),
                                               ctypes.c_double(ptop), ctypes.c_double(akap),
                                               ctypes.c_void_p(_lib.stream_handle())))


def edge_pe(grid_indexing, pe, delp, ptop: float):
    """pe_halo: pe on the 1-wide halo ring (interface levels)."""
    _lib.require_device(pe, delp)
    d = _lib.make_domain(grid_indexing, pe, grid_indexing.domain[2] + 1)
    _lib.make_domain(grid_indexing, delp, grid_indexing.domain[2], others3=(pe,))
    _lib.check(_lib.lib().fv3_pe_halo_f64(ctypes.byref(d), _lib.ptr(pe), _lib.ptr(delp),
                                          ctypes.c_double(ptop),
                                          ctypes.c_void_p(_lib.stream_handle())))
