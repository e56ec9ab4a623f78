"""DGrid2AGrid2CGridVectors (fv3core/stencils/d2a2c_vect.py:374-651) on the GPU.

The reference's eleven stencils run as five launches of ``fv3_d2a2c_vect``
(temporaries, x corner fills, uc/utc, y corner fills, vc/vtc).
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..utils.grid import GridData
from ..utils.stencil import StencilFactory
from ._workspace import Workspace

D2A2C_METRICS = ("cosa_s", "rsin2", "cosa_u", "cosa_v", "rsin_u", "rsin_v", "dxa", "dya",
                 "sin_sg1", "sin_sg2", "sin_sg3", "sin_sg4")


class DGrid2AGrid2CGridVectors:
    """Fortran d2a2c_vect: D-grid winds to A-grid and C-grid winds."""

    def __init__(self, stencil_factory: StencilFactory, grid_data: GridData, nested: bool,
                 grid_type: int, dord4: bool):
        if grid_type >= 3:
            raise NotImplementedError("unimplemented grid_type >= 3")
        if nested:
            raise NotImplementedError("nested grids are not supported")
        if not dord4:
            raise NotImplementedError("only the fourth-order (dord4) interpolation is built")
        self._grid_indexing = stencil_factory.grid_indexing
        self._grid_data = grid_data
        self._nk = self._grid_indexing.domain[2]
        self._work = Workspace(2)

    def __call__(self, uc, vc, u, v, ua, va, utc, vtc):
        """uc, vc, ua, va, utc, vtc <- interpolations of the D-grid winds u, v."""
        _lib.require_device(uc, vc, u, v, ua, va, utc, vtc)
        if u.dtype != torch.float64:
            raise TypeError("d2a2c_vect runs in FP64: pass float64 storages")
        met, ref2 = _lib.metrics_struct(self._grid_data, u.dtype, D2A2C_METRICS)
        d = _lib.make_domain(self._grid_indexing, u, self._nk, ref2,
                             others3=(uc, vc, v, ua, va, utc, vtc))
        work, slab = self._work.get(u, self._nk)
        _lib.check(
            _lib.lib().fv3_d2a2c_vect_f64(
                ctypes.byref(d), ctypes.byref(met), _lib.ptr(uc), _lib.ptr(vc), _lib.ptr(u),
                _lib.ptr(v), _lib.ptr(ua), _lib.ptr(va), _lib.ptr(utc), _lib.ptr(vtc),
                _lib.ptr(work), ctypes.c_int64(slab), ctypes.c_void_p(_lib.stream_handle()),
            )
        )
