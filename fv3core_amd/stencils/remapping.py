"""LagrangianToEulerian (fv3core/stencils/remapping.py:205-628) on the GPU.

``fv3_lagrangian_to_eulerian`` runs the whole step on one stream: a column
kernel for the pressures and the moist pt conversion, the MapSingle remaps
of pt, the tracers, w and delz, fillz, a column kernel for delz/peln/pkz,
ua/pe and the final pt, then the remaps of u and v on their staggered
pressures.  With do_sat_adj the saturation adjustment (SatAdjust3d) runs
before the final pt step, which ``fv3_remap_final_pt`` then applies.  The component owns its scratch (pe1, pe2, dp2, pn2, the u/v
pressures and the remap workspace), as the reference's does.
"""
from __future__ import annotations

import ctypes
from typing import Dict

import numpy as np
import torch

from .. import _lib
from ..constants import CONSV_MIN
from ..utils import gt4py_utils as utils
from ..utils.stencil import StencilFactory
from .._config import SatAdjustConfig
from ._workspace import Workspace
from .fillz import MAX_TRACERS
from .saturation_adjustment import SatAdjust3d


class RemapParams(ctypes.Structure):
    """Mirror of ``fv3_remap_params``."""

    _fields_ = [
        ("kord_tm", ctypes.c_int32), ("kord_tr", ctypes.c_int32), ("kord_wz", ctypes.c_int32),
        ("kord_mt", ctypes.c_int32), ("nq", ctypes.c_int32), ("fill", ctypes.c_int32),
        ("last_step", ctypes.c_int32), ("defer_pt", ctypes.c_int32), ("ptop", ctypes.c_double),
        ("akap", ctypes.c_double), ("zvir", ctypes.c_double),
    ]


_FIELDS = ("pt", "delp", "delz", "peln", "u", "v", "w", "ua", "cappa", "q_con", "pkz", "pk",
           "pe", "ps", "wsd", "ak", "bk")


class RemapState(ctypes.Structure):
    """Mirror of ``fv3_remap_state``."""

    _fields_ = [("tracers", ctypes.c_void_p * MAX_TRACERS)] + [
        (n, ctypes.c_void_p) for n in _FIELDS]


def _column(a, device, n):
    """A K column (FloatFieldK) as a contiguous float64 device vector of n values."""
    if isinstance(a, torch.Tensor):
        t = a.detach()
        if t.device != device or t.dtype != torch.float64 or not t.is_contiguous():
            t = t.to(device=device, dtype=torch.float64).contiguous()
    else:
        t = torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float64)),
                            device=device)
    if t.dim() != 1 or t.shape[0] < n:
        raise ValueError(f"a K column needs {n} values, got shape {tuple(t.shape)}")
    return t


class LagrangianToEulerian:
    """Fortran Lagrangian_to_Eulerian."""

    def __init__(self, stencil_factory: StencilFactory, config, area_64, nq, pfull):
        grid_indexing = stencil_factory.grid_indexing
        if config.kord_tm >= 0:
            raise NotImplementedError("map ppm, untested mode where kord_tm >= 0")
        if getattr(config, "hydrostatic", False):
            raise NotImplementedError("Hydrostatic is not implemented")
        if nq > MAX_TRACERS:
            raise ValueError(f"at most {MAX_TRACERS} tracers (got {nq})")
        for kord in (config.kord_tm, config.kord_tr, config.kord_wz, config.kord_mt):
            if abs(kord) > 10:
                raise AssertionError(f"kord {kord} not implemented.")
        self._gi = grid_indexing
        self._config = config
        self._nq = int(nq)
        self._t_min = 184.0
        self._do_sat_adjust = bool(config.do_sat_adj)
        self.kmp = grid_indexing.domain[2] - 1
        pf = pfull.detach().cpu().numpy() if isinstance(pfull, torch.Tensor) else np.asarray(pfull)
        for k in range(pf.shape[0]):
            if pf[k] > 10.0e2:
                self.kmp = k
                break
        self._saturation_adjustment = SatAdjust3d(
            stencil_factory, getattr(config, "sat_adjust", SatAdjustConfig()), area_64, self.kmp)
        self._work = Workspace(15)

    def __call__(self, tracers: Dict[str, "torch.Tensor"], pt, delp, delz, peln, u, v, w, ua, va,
                 cappa, q_con, q_cld, pkz, pk, pe, hs, te0_2d, ps, wsd, omga, ak, bk, pfull, dp1,
                 ptop: float, akap: float, zvir: float, last_step: bool, consv_te: float,
                 mdt: float, bdt: float, do_adiabatic_init: bool, nq: int):
        """Remap the deformed Lagrangian layers onto the hybrid levels (all in place)."""
        if last_step and not do_adiabatic_init:
            if consv_te > CONSV_MIN:
                raise NotImplementedError(
                    "We do not support consv_te > 0.001 because that would trigger an allReduce")
            elif consv_te < -CONSV_MIN:
                raise NotImplementedError(
                    f"Unimplemented/untested case consv({consv_te})  < -CONSV_MIN({-CONSV_MIN})")
        fast_mp_consv = not do_adiabatic_init and consv_te > CONSV_MIN
        if self._do_sat_adjust and fast_mp_consv:
            # the column energy sum after the adjustment (sum_te, remapping.py:199-202,
            # 607-611) is not built
            raise NotImplementedError("consv_te > CONSV_MIN with do_sat_adj is not implemented")
        nz = self._gi.domain[2]
        names = utils.tracer_variables
        fields3 = [pt, delp, delz, peln, u, v, w, ua, cappa, q_con, pkz, pk, pe]
        tr = [tracers[n] for n in names[:max(self._nq, 6)]]
        _lib.require_device(*fields3, *tr, ps, wsd)
        d = _lib.make_domain(self._gi, pt, nz, ps, others3=fields3 + tr, others2=(wsd,))
        if pt.shape[-1] < nz + 1:
            raise ValueError("the remapped fields need nz + 1 levels")
        dev = pt.device
        p = RemapParams(kord_tm=int(self._config.kord_tm), kord_tr=int(self._config.kord_tr),
                        kord_wz=int(self._config.kord_wz), kord_mt=int(self._config.kord_mt),
                        nq=self._nq, fill=int(bool(self._config.fill)),
                        last_step=int(bool(last_step)), defer_pt=int(self._do_sat_adjust),
                        ptop=float(ptop),
                        akap=float(akap), zvir=float(zvir))
        akd = _column(ak, dev, nz + 1)
        bkd = _column(bk, dev, nz + 1)
        s = RemapState()
        for n, t in enumerate(tr):
            s.tracers[n] = t.data_ptr()
        for name, t in zip(_FIELDS, fields3 + [ps, wsd, akd, bkd]):
            setattr(s, name, t.data_ptr())
        work, slab = self._work.get(pt, nz + 1)
        _lib.check(_lib.lib().fv3_lagrangian_to_eulerian_f64(
            ctypes.byref(d), ctypes.byref(p), ctypes.byref(s), _lib.ptr(work),
            ctypes.c_int64(slab), ctypes.c_void_p(_lib.stream_handle())))
        if self._do_sat_adjust:
            # remapping.py:581-606: the adjustment between the remap and the
            # final pt step, with dp1 as its energy scratch
            self._saturation_adjustment(
                dp1, tracers["qvapor"], tracers["qliquid"], tracers["qice"], tracers["qrain"],
                tracers["qsnow"], tracers["qgraupel"], q_cld, hs, peln, delp, delz, q_con, pt,
                pkz, cappa, zvir, mdt, fast_mp_consv, last_step, akap, self.kmp)
            _lib.check(_lib.lib().fv3_remap_final_pt_f64(
                ctypes.byref(d), ctypes.byref(p), ctypes.byref(s),
                ctypes.c_void_p(_lib.stream_handle())))
