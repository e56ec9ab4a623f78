"""RiemannSolver3 (fv3core/stencils/riem_solver3.py:106-242) on the GPU.

precompute, sim1_solver and finalize run as one FP64 kernel
(``fv3_riem_solver3``), one column per lane.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from .._config import RiemannConfig
from ..utils.stencil import StencilFactory
from ._workspace import Workspace


class RiemannSolver3:
    """Fortran Riem_Solver3 (semi-implicit, a_imp > 0.999)."""

    def __init__(self, stencil_factory: StencilFactory, config: RiemannConfig):
        if config.a_imp <= 0.999:
            raise NotImplementedError("a_imp <= 0.999 is not implemented")
        self._grid_indexing = stencil_factory.grid_indexing
        self._config = config
        self._nk = self._grid_indexing.domain[2] + 1
        self._work = Workspace(3)

    def __call__(self, last_call: bool, dt: float, cappa, ptop: float, zs, wsd, delz, q_con,
                 delp, pt, zh, pe, ppe, pk3, pk, peln, w):
        fields = (cappa, delz, q_con, delp, pt, zh, pe, ppe, pk3, pk, peln, w)
        _lib.require_device(zs, wsd, *fields)
        for t in (delz, zh, pe, ppe, pk3, pk, peln):
            if t.dtype != torch.float64:
                raise TypeError("the Riemann solvers integrate the heights and pressures in FP64")
        # cappa, q_con, delp, pt and w are FP64, or FP32 in the mixed form
        bulk = cappa.dtype
        if bulk not in (torch.float64, torch.float32) or any(
                t.dtype != bulk for t in (q_con, delp, pt, w)):
            raise TypeError("cappa, q_con, delp, pt and w must share one dtype")
        name = "fv3_riem_solver3_f64" if bulk == torch.float64 else "fv3_riem_solver3_mixed"
        d = _lib.make_domain(self._grid_indexing, zh, self._nk, zs,
                             others3=fields, others2=(wsd,))
        work, slab = self._work.get(zh, self._nk)
        c = self._config
        _lib.check(
            getattr(_lib.lib(), name)(
                ctypes.byref(d), int(bool(last_call)), ctypes.c_double(dt),
                ctypes.c_double(ptop), ctypes.c_double(c.p_fac), int(bool(c.use_logp)),
                ctypes.c_double(c.beta), _lib.ptr(cappa), _lib.ptr(zs), _lib.ptr(wsd),
                _lib.ptr(delz), _lib.ptr(q_con), _lib.ptr(delp), _lib.ptr(pt), _lib.ptr(zh),
                _lib.ptr(pe), _lib.ptr(ppe), _lib.ptr(pk3), _lib.ptr(pk), _lib.ptr(peln),
                _lib.ptr(w), _lib.ptr(work), ctypes.c_int64(slab),
                ctypes.c_void_p(_lib.stream_handle()),
            )
        )
