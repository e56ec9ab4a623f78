"""CGridShallowWaterDynamics (fv3core/stencils/c_sw.py:379-659) on the GPU.

One call of ``fv3_c_sw`` runs the reference's sequence (d2a2c_vect,
divergence_corner, geoadjust, corner fills, C-grid fluxes of delp/pt/w,
kinetic energy, vorticity, C-grid momentum update) as ten launches.  The
component owns ``delpc`` and ``ptc`` and returns them, as the reference
does (c_sw.py:405-410, 659).  The Coriolis parameter ``fC`` comes from
``grid_data`` (the reference reads it from its global grid object).
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..utils import gt4py_utils as utils
from ..utils.grid import GridData
from ..utils.stencil import StencilFactory
from ._workspace import Workspace
from .d2a2c_vect import D2A2C_METRICS

C_SW_METRICS = D2A2C_METRICS + (
    "dx", "dy", "dxc", "dyc", "rdxc", "rdyc", "rarea", "rarea_c", "cos_sg1", "cos_sg2",
    "cos_sg3", "cos_sg4", "sina_u", "sina_v", "fC",
)


def batch_shape(grid_indexing, grid_data: GridData):
    """max_shape with the leading tile axis of the grid data, if it has one."""
    lead = tuple(grid_data.dx.shape[:-2])
    return lead + tuple(grid_indexing.max_shape)


class CGridShallowWaterDynamics:
    """Fortran c_sw: advances the C-grid winds by half a time step."""

    def __init__(self, stencil_factory: StencilFactory, grid_data: GridData, nested: bool,
                 grid_type: int, nord: int):
        if grid_type >= 3:
            raise NotImplementedError("unimplemented grid_type >= 3")
        if nested:
            raise NotImplementedError("nested grids are not supported")
        self._grid_indexing = stencil_factory.grid_indexing
        self.grid_data = grid_data
        self._nord = int(nord)
        self._nk = self._grid_indexing.domain[2]
        shape = batch_shape(self._grid_indexing, grid_data)
        self.delpc = utils.zeros(shape, dtype=grid_data.dx.dtype, device=grid_data.dx.device)
        self.ptc = utils.zeros(shape, dtype=grid_data.dx.dtype, device=grid_data.dx.device)
        self._work = Workspace(7)

    def __call__(self, delp, pt, u, v, w, uc, vc, ua, va, ut, vt, divgd, omga, dt2: float):
        """Returns (delpc, ptc); uc, vc, ua, va, ut, vt, divgd, omga in place."""
        _lib.require_device(delp, pt, u, v, w, uc, vc, ua, va, ut, vt, divgd, omga)
        if delp.dtype not in (torch.float64, torch.float32):
            raise TypeError("c_sw runs on float64 (or, in the mixed form, float32) storages")
        if self.grid_data.dx.dtype != delp.dtype:
            raise TypeError("c_sw: the grid data and the storages must share one dtype")
        scalar = ctypes.c_double if delp.dtype == torch.float64 else ctypes.c_float
        met, ref2 = _lib.metrics_struct(self.grid_data, delp.dtype, C_SW_METRICS)
        d = _lib.make_domain(self._grid_indexing, delp, self._nk, ref2,
                             others3=(pt, u, v, w, uc, vc, ua, va, ut, vt, divgd, omga,
                                      self.delpc, self.ptc))
        work, slab = self._work.get(delp, self._nk)
        _lib.check(
            _lib.fn("fv3_c_sw", delp.dtype)(
                ctypes.byref(d), ctypes.byref(met), ctypes.c_int32(self._nord),
                scalar(dt2), _lib.ptr(delp), _lib.ptr(pt), _lib.ptr(u), _lib.ptr(v),
                _lib.ptr(w), _lib.ptr(uc), _lib.ptr(vc), _lib.ptr(ua), _lib.ptr(va),
                _lib.ptr(ut), _lib.ptr(vt), _lib.ptr(divgd), _lib.ptr(omga),
                _lib.ptr(self.delpc), _lib.ptr(self.ptc), _lib.ptr(work), ctypes.c_int64(slab),
                ctypes.c_void_p(_lib.stream_handle()),
            )
        )
        return self.delpc, self.ptc
