"""CubedToLatLon (fv3core/stencils/c2l_ord.py:74-141) on the GPU.

The D-grid winds are interpolated to cell centres and rotated to
latitude-longitude components with the cell-centre transform a11..a22
(fv3core/grid/geometry.py:587-601).  4th order (``ord4_transform``) runs on
the compute box after a vector halo update of u and v, with 2nd-order
interpolation on the tile-edge rows and columns; 2nd order (``c2l_ord2``)
runs on the compute box grown by one cell.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..utils.stencil import StencilFactory


class CubedToLatLon:
    """Fortran c2l_ord2 / c2l_ord4."""

    def __init__(self, stencil_factory: StencilFactory, grid_data, order: int):
        if order not in (2, 4):
            raise ValueError(f"order must be 2 or 4 (got {order})")
        self._gi = stencil_factory.grid_indexing
        self._n_halo = self._gi.n_halo
        self._order = int(order)
        missing = [n for n in ("dx", "dy", "a11", "a12", "a21", "a22") if not grid_data.has(n)]
        if missing:
            raise ValueError(f"grid_data lacks the metric terms {missing}")
        self._dx, self._dy = grid_data.dx, grid_data.dy
        self._a = (grid_data.a11, grid_data.a12, grid_data.a21, grid_data.a22)

    def __call__(self, u, v, ua, va, comm):
        """ua, va <- latitude-longitude winds of the D-grid winds u, v (whose
        halos the 4th-order form updates first through ``comm``)."""
        us = u if isinstance(u, torch.Tensor) else u.storage
        vs = v if isinstance(v, torch.Tensor) else v.storage
        _lib.require_device(us, vs, ua, va, self._dx, *self._a)
        for t in (us, vs, ua, va):
            if t.dtype != torch.float64:
                raise TypeError("CubedToLatLon runs in FP64: pass float64 storages")
        if self._order == 4:
            comm.vector_halo_update(u, v, n_points=self._n_halo)
        d = _lib.make_domain(self._gi, ua, self._gi.domain[2], ref2=self._dx,
                             others3=(us, vs, va), others2=(self._dy,) + self._a)
        _lib.check(_lib.lib().fv3_cubed_to_latlon_f64(
            ctypes.byref(d), ctypes.c_int32(self._order), _lib.ptr(us), _lib.ptr(vs),
            _lib.ptr(self._dx), _lib.ptr(self._dy), *(_lib.ptr(a) for a in self._a),
            _lib.ptr(ua), _lib.ptr(va), ctypes.c_void_p(_lib.stream_handle())))
