"""DGridShallowWaterLagrangianDynamics (fv3core/stencils/d_sw.py:606-1131) on the GPU.

One ``fv3_d_sw`` call runs the reference's whole D-grid step: flux
preparation (fxadv), the four fv_tp_2d transports with their del-n
damping, the flux capacitor, heat dissipation, the kinetic energy from
PPM-advected winds, divergence damping, vorticity transport and the wind
update.  The column namelist is computed on the host by
:func:`get_column_namelist`, as in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from .._config import DGridShallowWaterLagrangianDynamicsConfig
from ..utils.grid import DampingCoefficients, GridData
from ..utils.stencil import StencilFactory
from ._workspace import Workspace
from .delnflux import calc_damp

DCON_THRESHOLD = 1e-5

D_SW_METRICS = (
    "dx", "dy", "dxa", "dya", "dxc", "dyc", "rdx", "rdy", "rdxa", "rdya", "area", "rarea",
    "rarea_c", "sin_sg1", "sin_sg2", "sin_sg3", "sin_sg4", "cosa", "cosa_u", "cosa_v",
    "cosa_s", "sina_u", "sina_v", "rsin_u", "rsin_v", "rsina", "rsin2", "f0", "del6_u",
    "del6_v", "divg_u", "divg_v", "agrid1", "agrid2", "bgrid1", "bgrid2", "edge_w", "edge_e",
    "edge_s", "edge_n",
)

_STATE = ("delpc", "delp", "ptc", "pt", "u", "v", "w", "uc", "vc", "ua", "va", "divgd", "mfx",
          "mfy", "cx", "cy", "crx", "cry", "xfx", "yfx", "q_con", "zh", "heat_source",
          "diss_est")


class _Params(ctypes.Structure):
    """Mirror of ``fv3_dsw_params``."""

    _fields_ = [
        ("hord_dp", ctypes.c_int32), ("hord_tm", ctypes.c_int32), ("hord_mt", ctypes.c_int32),
        ("hord_vt", ctypes.c_int32), ("hydrostatic", ctypes.c_int32),
        ("nord_v4", ctypes.c_int32 * 4), ("nord_w4", ctypes.c_int32 * 4),
        ("nord_t4", ctypes.c_int32 * 4), ("div_k0", ctypes.c_int32),
        ("div_nord", ctypes.c_int32), ("transport_fp32", ctypes.c_int32),
        ("dddmp", ctypes.c_double), ("da_min_c", ctypes.c_double),
        ("dd8", ctypes.c_double), ("d_con", ctypes.c_double),
        ("damp_vt_deln", ctypes.c_void_p), ("damp_t_deln", ctypes.c_void_p),
        ("damp_w_nosg", ctypes.c_void_p), ("damp_v_nosg", ctypes.c_void_p),
        ("damp_w", ctypes.c_void_p), ("damp_vt", ctypes.c_void_p), ("ke_bg", ctypes.c_void_p),
        ("d_con_col", ctypes.c_void_p), ("d2_divg", ctypes.c_void_p),
    ]


class _State(ctypes.Structure):
    """Mirror of ``fv3_dsw_state``."""

    _fields_ = [(n, ctypes.c_void_p) for n in _STATE]


def get_column_namelist(config: DGridShallowWaterLagrangianDynamicsConfig, npz: int):
    """Per-level damping columns (d_sw.py:512-563), numpy arrays of npz + 1."""
    names = ("ke_bg", "d_con", "nord", "nord_v", "nord_w", "nord_t", "damp_vt", "damp_w",
             "damp_t", "d2_divg")
    col = {n: np.zeros(npz + 1) for n in names}
    for n in ("ke_bg", "d_con", "nord"):
        col[n][:] = getattr(config, n)
    col["d2_divg"][:] = min(0.2, config.d2_bg)
    col["nord_v"][:] = min(2, col["nord"][0])
    col["nord_w"][:] = col["nord_v"][0]
    col["nord_t"][:] = col["nord_v"][0]
    col["damp_vt"][:] = config.vtdm4 if config.do_vort_damp else 0
    col["damp_w"][:] = col["damp_vt"][0]
    col["damp_t"][:] = col["damp_vt"][0]

    def set_low_kvals(k):
        for n in ("nord", "nord_w", "d_con"):
            col[n][k] = 0
        col["damp_w"][k] = col["d2_divg"][k]

    def lowest_kvals(k):
        set_low_kvals(k)
        if config.do_vort_damp:
            col["nord_v"][k] = 0
            col["damp_vt"][k] = 0.5 * col["d2_divg"][k]

    if npz == 1 or config.n_sponge < 0:
        col["d2_divg"][0] = config.d2_bg
    else:
        col["d2_divg"][0] = max(0.01, config.d2_bg, config.d2_bg_k1)
        lowest_kvals(0)
        if config.d2_bg_k2 > 0.01:
            col["d2_divg"][1] = max(config.d2_bg, config.d2_bg_k2)
            lowest_kvals(1)
        if config.d2_bg_k2 > 0.05:
            col["d2_divg"][2] = max(config.d2_bg, 0.2 * config.d2_bg_k2)
            set_low_kvals(2)
    return col


class DGridShallowWaterLagrangianDynamics:
    """Fortran d_sw: full-timestep D-grid dynamics."""

    def __init__(self, stencil_factory: StencilFactory, grid_data: GridData,
                 damping_coefficients: DampingCoefficients, column_namelist, nested: bool,
                 stretched_grid: bool, config: DGridShallowWaterLagrangianDynamicsConfig):
        if config.grid_type >= 3:
            raise NotImplementedError("ubke and vbke only implemented for grid_type < 3")
        if config.inline_q:
            raise NotImplementedError("inline_q not yet implemented")
        if config.d_ext > 0:
            raise NotImplementedError("untested d_ext > 0. need to call a2b_ord2")
        if nested:
            raise NotImplementedError("nested grids are not supported")
        if stretched_grid:
            raise NotImplementedError("stretched grids are not supported")
        if config.do_skeb:
            raise NotImplementedError("do_skeb (stochastic backscatter) is not built")
        if config.do_f3d and not config.hydrostatic:
            raise NotImplementedError("do_f3d is not built")
        col = {k: np.asarray(v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v,
                             dtype=np.float64) for k, v in column_namelist.items()}
        if not (col["damp_vt"] > DCON_THRESHOLD).all():
            raise ValueError("damp_vt must exceed the d_con threshold on every level")
        if not (col["damp_w"] > DCON_THRESHOLD).all():
            raise ValueError("damp_w must exceed the d_con threshold on every level")
        for h in (config.hord_dp, config.hord_tm, config.hord_mt, config.hord_vt):
            if h not in (5, 6, 7, 8):
                raise NotImplementedError(f"hord={h} is not supported (5, 6, 7, 8)")
        self._gi = stencil_factory.grid_indexing
        self.grid = grid_data
        self._config = config
        npz = self._gi.domain[2]
        self._nk = npz
        da_min, da_min_c = damping_coefficients.da_min, damping_coefficients.da_min_c
        # DelnFlux damping (calc_damp with da_min) and DelnFluxNoSG columns
        # (calc_damp with da_min_c, d_sw.py:789-810)
        self._cols = {
            "damp_vt_deln": calc_damp(col["nord_v"], col["damp_vt"], da_min, npz),
            "damp_t_deln": calc_damp(col["nord_t"], col["damp_t"], da_min, npz),
            "damp_w_nosg": calc_damp(col["nord_w"], col["damp_w"], da_min_c, npz),
            "damp_v_nosg": calc_damp(col["nord_v"], col["damp_vt"], da_min_c, npz),
            "damp_w": col["damp_w"][:npz], "damp_vt": col["damp_vt"][:npz],
            "ke_bg": col["ke_bg"][:npz], "d_con_col": col["d_con"][:npz],
            "d2_divg": col["d2_divg"][:npz],
        }
        for name in ("nord_v", "nord_w", "nord_t"):
            if max(col[name][:4]) > 2:
                raise NotImplementedError(f"{name} > 2 needs a 4-cell halo")
        # divergence damping order (divergence_damping.py:204-219)
        nonzero_k, nonzero_nord = 0, int(config.nord)
        for k in range(len(col["nord"])):
            if col["nord"][k] > 0:
                nonzero_k, nonzero_nord = k, int(col["nord"][k])
                break
        if nonzero_nord > 3:
            raise NotImplementedError("divergence damping order above 3")
        p = _Params()
        p.hord_dp, p.hord_tm = config.hord_dp, config.hord_tm
        p.hord_mt, p.hord_vt = config.hord_mt, config.hord_vt
        p.hydrostatic = int(bool(config.hydrostatic))
        for name in ("nord_v", "nord_w", "nord_t"):
            getattr(p, name + "4")[:] = [int(v) for v in col[name][:4]]
        p.div_k0 = nonzero_k
        p.div_nord = nonzero_nord
        p.dddmp = float(config.dddmp)
        p.da_min_c = float(da_min_c)
        p.dd8 = float((da_min_c * config.d4_bg) ** (nonzero_nord + 1))
        p.d_con = float(config.d_con)
        self._params = p
        self._dev_cols = {}
        self._work = Workspace(22)
        # mixed precision: FP32 arithmetic in the fv_tp_2d calls
        self.transport_fp32 = False

    def _columns(self, device, dtype):
        key = (str(device), dtype)
        if key not in self._dev_cols:
            self._dev_cols[key] = {k: torch.as_tensor(np.ascontiguousarray(v), dtype=dtype,
                                                      device=device)
                                   for k, v in self._cols.items()}
        return self._dev_cols[key]

    def __call__(self, delpc, delp, ptc, pt, u, v, w, uc, vc, ua, va, divgd, mfx, mfy, cx, cy,
                 crx, cry, xfx, yfx, q_con, zh, heat_source, diss_est, dt: float):
        """D-grid step of dt; every argument storage is updated in place."""
        fields = dict(delpc=delpc, delp=delp, ptc=ptc, pt=pt, u=u, v=v, w=w, uc=uc, vc=vc,
                      ua=ua, va=va, divgd=divgd, mfx=mfx, mfy=mfy, cx=cx, cy=cy, crx=crx,
                      cry=cry, xfx=xfx, yfx=yfx, q_con=q_con, zh=zh, heat_source=heat_source,
                      diss_est=diss_est)
        if delp.dtype == torch.float32 and zh is not None and zh.dtype != torch.float32:
            # the FP32 form does not read zh (the mixed step keeps it FP64)
            fields["zh"] = None
        present = [f for f in fields.values() if f is not None]
        _lib.require_device(*present)
        if delp.dtype not in (torch.float64, torch.float32):
            raise TypeError("d_sw runs on float64 (or, in the mixed form, float32) storages")
        if self.grid.dx.dtype != delp.dtype:
            raise TypeError("d_sw: the grid data and the storages must share one dtype")
        met, ref2 = _lib.metrics_struct(self.grid, delp.dtype, D_SW_METRICS)
        d = _lib.make_domain(self._gi, delp, self._nk + 1, ref2, others3=present)
        d.nk = self._nk
        cols = self._columns(delp.device, delp.dtype)
        p = self._params
        p.transport_fp32 = int(bool(self.transport_fp32))
        for k, t in cols.items():
            setattr(p, k, t.data_ptr())
        st = _State()
        for k, t in fields.items():
            setattr(st, k, t.data_ptr() if t is not None else None)
        work, slab = self._work.get(delp, self._nk + 1)
        scalar = ctypes.c_double if delp.dtype == torch.float64 else ctypes.c_float
        _lib.check(
            _lib.fn("fv3_d_sw", delp.dtype)(
                ctypes.byref(d), ctypes.byref(met), ctypes.byref(p), ctypes.byref(st),
                scalar(dt), _lib.ptr(work), ctypes.c_int64(slab),
                ctypes.c_void_p(_lib.stream_handle()),
            )
        )
