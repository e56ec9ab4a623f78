"""RiemannSolverC (fv3core/stencils/riem_solver_c.py:67-165) on the GPU.

precompute, sim1_solver and finalize run as one FP64 kernel
(``fv3_riem_solver_c``), one column per lane.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..utils.stencil import StencilFactory
from ._workspace import Workspace


class RiemannSolverC:
    """Fortran Riem_Solver_C."""

    def __init__(self, stencil_factory: StencilFactory, p_fac):
        self._grid_indexing = stencil_factory.grid_indexing
        self._p_fac = float(p_fac)
        self._nk = self._grid_indexing.domain[2] + 1
        self._work = Workspace(3)

    def __call__(self, dt2: float, cappa, ptop: float, hs, ws, ptc, q_con, delpc, gz, pef, w3):
        """gz, pef <- nonhydrostatic C-grid solve (w3 is read only)."""
        _lib.require_device(cappa, hs, ws, ptc, q_con, delpc, gz, pef, w3)
        for t in (gz, pef):
            if t.dtype != torch.float64:
                raise TypeError("the Riemann solvers integrate gz and pef in FP64")
        # cappa, ptc, q_con, delpc and w3 are FP64, or FP32 in the mixed form
        bulk = cappa.dtype
        if bulk not in (torch.float64, torch.float32) or any(
                t.dtype != bulk for t in (ptc, q_con, delpc, w3)):
            raise TypeError("cappa, ptc, q_con, delpc and w3 must share one dtype")
        name = "fv3_riem_solver_c_f64" if bulk == torch.float64 else "fv3_riem_solver_c_mixed"
        d = _lib.make_domain(self._grid_indexing, gz, self._nk, hs,
                             others3=(cappa, ptc, q_con, delpc, pef, w3), others2=(ws,))
        work, slab = self._work.get(gz, self._nk)
        _lib.check(
            getattr(_lib.lib(), name)(
                ctypes.byref(d), ctypes.c_double(dt2), ctypes.c_double(ptop),
                ctypes.c_double(self._p_fac), _lib.ptr(cappa), _lib.ptr(hs), _lib.ptr(ws),
                _lib.ptr(ptc), _lib.ptr(q_con), _lib.ptr(delpc), _lib.ptr(gz), _lib.ptr(pef),
                _lib.ptr(w3), _lib.ptr(work), ctypes.c_int64(slab),
                ctypes.c_void_p(_lib.stream_handle()),
            )
        )
