"""NonHydrostaticPressureGradient (fv3core/stencils/nh_p_grad.py:87-206) on the GPU.

Four a2b_ord4 interpolations (pp, pk3 and gz replaced by their B-grid
values, delp into a scratch), the top-level reset with wk = pk3 differences,
then the u and v updates, all from one ``fv3_nh_p_grad`` call.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..utils.grid import GridData
from ..utils.stencil import StencilFactory
from ._workspace import Workspace
from .a2b_ord4 import A2B_METRICS


class NonHydrostaticPressureGradient:
    """Fortran nh_p_grad."""

    def __init__(self, stencil_factory: StencilFactory, grid_data: GridData, grid_type):
        if grid_type >= 3:
            raise NotImplementedError("grid_type >= 3 is not supported")
        self._gi = stencil_factory.grid_indexing
        self._grid = grid_data
        self.nk = self._gi.domain[2]
        self._work = Workspace(6)

    def __call__(self, u, v, pp, gz, pk3, delp, dt: float, ptop: float, akap: float):
        """u, v updated; pp, gz, pk3 replaced on the B grid (as the reference does)."""
        _lib.require_device(u, v, pp, gz, pk3, delp)
        # u and v are FP64, or FP32 in the mixed form; the rest is FP64
        for t in (pp, gz, pk3, delp):
            if t.dtype != torch.float64:
                raise TypeError("nh_p_grad takes pp, gz, pk3 and delp in FP64")
        if u.dtype != v.dtype or u.dtype not in (torch.float64, torch.float32):
            raise TypeError("u and v must share one dtype (float64 or float32)")
        name = "fv3_nh_p_grad_f64" if u.dtype == torch.float64 else "fv3_nh_p_grad_mixed"
        met, ref2 = _lib.metrics_struct(self._grid, torch.float64, A2B_METRICS + ("rdx", "rdy"))
        d = _lib.make_domain(self._gi, pp, self.nk + 1, ref2, others3=(gz, pk3))
        d.nk = self.nk
        _lib.make_domain(self._gi, u, self.nk, ref2, others3=(v, delp, pp))
        ptk = ptop ** akap
        work, slab = self._work.get(pp, self.nk + 1)
        _lib.check(
            getattr(_lib.lib(), name)(
                ctypes.byref(d), ctypes.byref(met), _lib.ptr(u), _lib.ptr(v), _lib.ptr(pp),
                _lib.ptr(gz), _lib.ptr(pk3), _lib.ptr(delp), ctypes.c_double(dt),
                ctypes.c_double(ptk), _lib.ptr(work), ctypes.c_int64(slab),
                ctypes.c_void_p(_lib.stream_handle()),
            )
        )
