"""AGrid2BGridFourthOrder (fv3core/stencils/a2b_ord4.py:454-695) on the GPU.

``fv3_a2b_ord4`` runs the corner extrapolations and tile-edge lines, the
volume means and the fourth-order interpolation as three launches.  The
longitudes/latitudes (agrid1/2, bgrid1/2) and edge weights (edge_w/e/s/n)
come from ``grid_data`` (the reference reads them from its global grid).
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..utils.grid import GridData
from ..utils.stencil import StencilFactory, Z_DIM, Z_INTERFACE_DIM
from ._workspace import Workspace

A2B_METRICS = ("dxa", "dya", "agrid1", "agrid2", "bgrid1", "bgrid2", "edge_w", "edge_e",
               "edge_s", "edge_n")


class AGrid2BGridFourthOrder:
    """Fortran a2b_ord4: cell-centre values to cell corners."""

    def __init__(self, stencil_factory: StencilFactory, grid_data: GridData, grid_type,
                 z_dim=Z_DIM, replace: bool = False):
        if grid_type >= 3:
            raise NotImplementedError("grid_type >= 3 is not supported")
        self._gi = stencil_factory.grid_indexing
        self._grid = grid_data
        self.replace = bool(replace)
        self._nk = self._gi.domain[2] + (1 if z_dim == Z_INTERFACE_DIM else 0)
        self._work = Workspace(3)

    def __call__(self, qin, qout):
        """qout <- qin on the B grid (and qin too when ``replace``)."""
        _lib.require_device(qin, qout)
        if qin.dtype != torch.float64:
            raise TypeError("a2b_ord4 runs in FP64: pass float64 storages")
        met, ref2 = _lib.metrics_struct(self._grid, qin.dtype, A2B_METRICS)
        d = _lib.make_domain(self._gi, qin, self._nk, ref2, others3=(qout,))
        work, slab = self._work.get(qin, self._nk)
        _lib.check(
            _lib.lib().fv3_a2b_ord4_f64(
                ctypes.byref(d), ctypes.byref(met), ctypes.c_int32(int(self.replace)),
                _lib.ptr(qin), _lib.ptr(qout), _lib.ptr(work), ctypes.c_int64(slab),
                ctypes.c_void_p(_lib.stream_handle()),
            )
        )
