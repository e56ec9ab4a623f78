"""SatAdjust3d (fv3core/stencils/saturation_adjustment.py:899-1024) on the GPU.

``fv3_sat_adjust`` runs the satadjust stencil (:554-896) pointwise on the
compute domain from level kmp down, in FP64.  The saturation vapour-pressure
tables the stencil evaluates from closed forms at each lookup are evaluated
once per table index into device tables (``fv3_sat_adjust_tables``) when the
component is first called on a device.
"""
from __future__ import annotations

import ctypes
import math

import torch

from .. import _lib
from .._config import SatAdjustConfig
from ..constants import C_LIQ, CP_AIR, CP_VAP, CV_AIR, CV_VAP, HLV, TICE
from ..utils.stencil import StencilFactory


class SatAdjustParams(ctypes.Structure):
    """Mirror of ``fv3_satadj_params``."""

    _fields_ = [(n, ctypes.c_int32) for n in (
        "hydrostatic", "rad_snow", "rad_rain", "rad_graupel", "tintqs", "icloud_f", "do_qa",
        "consv_te", "last_step")] + [(n, ctypes.c_double) for n in (
            "sat_adj0", "ql_gen", "qs_mlt", "ql0_max", "t_sub", "qi_gen", "qi_lim", "qi0_max",
            "dw_ocean", "dw_land", "cld_min", "sdt", "zvir", "fac_i2s", "c_air", "c_vap", "mdt",
            "fac_r2g", "fac_smlt", "fac_l2r", "fac_imlt", "d0_vap", "lv00", "fac_v2l",
            "fac_l2v")]


class SatAdjust3d:
    """Fortran fv_sat_adj (fast saturation adjustment inside the remap)."""

    def __init__(self, stencil_factory: StencilFactory, config: SatAdjustConfig, area_64, kmp):
        self._gi = stencil_factory.grid_indexing
        self._config = config
        self._area_64 = area_64
        self._kmp = int(kmp)
        if config.icloud_f not in (0, 1, 2):
            raise ValueError(f"icloud_f must be 0, 1 or 2 (got {config.icloud_f})")
        self._tables = {}

    def _tables_on(self, device):
        key = str(device)
        if key not in self._tables:
            n = int(_lib.lib().fv3_sat_adjust_table_entries())
            t = torch.empty(4 * n, dtype=torch.float64, device=device)
            _lib.check(_lib.lib().fv3_sat_adjust_tables_f64(
                _lib.ptr(t), ctypes.c_void_p(_lib.stream_handle())))
            self._tables[key] = t
        return self._tables[key]

    def __call__(self, te, qvapor, qliquid, qice, qrain, qsnow, qgraupel, qcld, hs, peln, delp,
                 delz, q_con, pt, pkz, cappa, r_vir: float, mdt: float, fast_mp_consv: bool,
                 last_step: bool, akap: float, kmp: int):
        """Every argument storage is updated in place (te on fast_mp_consv)."""
        cfg = self._config
        fields3 = (te, qvapor, qliquid, qice, qrain, qsnow, qgraupel, qcld, peln, delp, delz,
                   q_con, pt, pkz, cappa)
        _lib.require_device(*fields3, hs, self._area_64)
        for t in fields3 + (hs, self._area_64):
            if t.dtype != torch.float64:
                raise TypeError("SatAdjust3d runs in FP64: pass float64 storages")
        nz = self._gi.domain[2]
        d = _lib.make_domain(self._gi, pt, nz, hs, others3=fields3, others2=(self._area_64,))
        sdt = 0.5 * mdt  # half remapping time step (:960)
        if cfg.hydrostatic:
            c_air, c_vap = CP_AIR, CP_VAP
        else:
            c_air, c_vap = CV_AIR, CV_VAP
        d0_vap = c_vap - C_LIQ
        p = SatAdjustParams(
            hydrostatic=int(bool(cfg.hydrostatic)), rad_snow=int(bool(cfg.rad_snow)),
            rad_rain=int(bool(cfg.rad_rain)), rad_graupel=int(bool(cfg.rad_graupel)),
            tintqs=int(bool(cfg.tintqs)), icloud_f=int(cfg.icloud_f), do_qa=1,
            consv_te=int(bool(fast_mp_consv)), last_step=int(bool(last_step)),
            sat_adj0=cfg.sat_adj0, ql_gen=cfg.ql_gen, qs_mlt=cfg.qs_mlt, ql0_max=cfg.ql0_max,
            t_sub=cfg.t_sub, qi_gen=cfg.qi_gen, qi_lim=cfg.qi_lim, qi0_max=cfg.qi0_max,
            dw_ocean=cfg.dw_ocean, dw_land=cfg.dw_land, cld_min=cfg.cld_min, sdt=sdt,
            zvir=float(r_vir), fac_i2s=1.0 - math.exp(-mdt / cfg.tau_i2s), c_air=c_air,
            c_vap=c_vap, mdt=float(mdt), fac_r2g=1.0 - math.exp(-mdt / cfg.tau_r2g),
            fac_smlt=1.0 - math.exp(-mdt / cfg.tau_smlt),
            fac_l2r=1.0 - math.exp(-mdt / cfg.tau_l2r),
            fac_imlt=1.0 - math.exp(-sdt / cfg.tau_imlt), d0_vap=d0_vap,
            lv00=HLV - d0_vap * TICE, fac_v2l=1.0 - math.exp(-sdt / cfg.tau_v2l),
            fac_l2v=min(cfg.sat_adj0, 1.0 - math.exp(-sdt / cfg.tau_l2v)))
        tables = self._tables_on(pt.device)
        _lib.check(_lib.lib().fv3_sat_adjust_f64(
            ctypes.byref(d), ctypes.c_int32(int(kmp)), ctypes.byref(p), _lib.ptr(tables),
            _lib.ptr(te), _lib.ptr(qvapor), _lib.ptr(qliquid), _lib.ptr(qice), _lib.ptr(qrain),
            _lib.ptr(qsnow), _lib.ptr(qgraupel), _lib.ptr(qcld), _lib.ptr(self._area_64),
            _lib.ptr(hs), _lib.ptr(peln), _lib.ptr(delp), _lib.ptr(delz), _lib.ptr(q_con),
            _lib.ptr(pt), _lib.ptr(pkz), _lib.ptr(cappa), ctypes.c_void_p(_lib.stream_handle())))
