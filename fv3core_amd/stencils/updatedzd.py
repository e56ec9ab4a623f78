"""UpdateHeightOnDGrid (fv3core/stencils/updatedzd.py:190-393) on the GPU.

The spline constants come from the reference's FORWARD recurrence on the
host; ``fv3_update_dz_d`` interpolates the Courant numbers and area fluxes
to the interfaces, transports the height (fv_tp_2d + DelnFluxNoSG) and
applies the fluxes with the DZ_MIN floor.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from .. import _lib
from ..utils.grid import DampingCoefficients, GridData
from ..utils.stencil import StencilFactory
from ._workspace import Workspace


def cubic_spline_constants(dp0, nk: int):
    """cubic_spline_interpolation_constants (updatedzd.py:121-145): gk, beta, gamma."""
    dp0 = np.asarray(dp0, dtype=np.float64)
    gk, beta, gamma = np.zeros(nk), np.zeros(nk), np.zeros(nk)
    gk[0] = dp0[1] / dp0[0]
    beta[0] = gk[0] * (gk[0] + 0.5)
    gamma[0] = (1.0 + gk[0] * (gk[0] + 1.5)) / beta[0]
    for k in range(1, nk - 1):
        gk[k] = dp0[k - 1] / dp0[k]
        beta[k] = 2.0 + 2.0 * gk[k] - gamma[k - 1]
        gamma[k] = gk[k] / beta[k]
    return gk, beta, gamma


class UpdateHeightOnDGrid:
    """Fortran updatedzd."""

    def __init__(self, stencil_factory: StencilFactory, damping_coefficients: DampingCoefficients,
                 grid_data: GridData, grid_type: int, hord_tm: int, dp0, column_namelist,
                 k_bounds):
        self._gi = stencil_factory.grid_indexing
        col = {k: np.asarray(v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v,
                             dtype=np.float64) for k, v in column_namelist.items()}
        if any(col["damp_vt"][k] <= 1e-5 for k in range(len(k_bounds))):
            raise NotImplementedError("damp <= 1e-5 in column_cols is untested")
        if hord_tm not in (5, 6, 7, 8):
            raise NotImplementedError(f"hord_tm={hord_tm} is not supported (5, 6, 7, 8)")
        if max(col["nord_v"][:4]) > 2:
            raise NotImplementedError("nord_v > 2 needs a 4-cell halo")
        if grid_type >= 3:
            raise NotImplementedError("grid_type >= 3 is not supported")
        self._grid = grid_data
        self._hord = int(hord_tm)
        self._nk = self._gi.domain[2] + 1
        dp0 = dp0.detach().cpu().numpy() if isinstance(dp0, torch.Tensor) else dp0
        gk, beta, gamma = cubic_spline_constants(dp0, self._nk)
        self._host = {"gk": gk, "beta": beta, "gamma": gamma,
                      "damp": np.ascontiguousarray(col["damp_vt"][:self._nk])}
        self._nord4 = (ctypes.c_int32 * 4)(*[int(v) for v in col["nord_v"][:4]])
        self._dev = {}
        self._work = Workspace(9)

    def _cols(self, device):
        key = str(device)
        if key not in self._dev:
            self._dev[key] = {k: torch.as_tensor(v, dtype=torch.float64, device=device)
                              for k, v in self._host.items()}
        return self._dev[key]

    def __call__(self, surface_height, height, courant_number_x, courant_number_y, x_area_flux,
                 y_area_flux, ws, dt: float):
        """height (interface levels) and ws in place."""
        _lib.require_device(surface_height, height, courant_number_x, courant_number_y,
                            x_area_flux, y_area_flux, ws)
        met, ref2 = _lib.metrics_struct(self._grid, height.dtype,
                                        ("area", "dxa", "dya", "rarea", "del6_u", "del6_v"))
        d = _lib.make_domain(self._gi, height, self._nk, ref2,
                             others3=(courant_number_x, courant_number_y, x_area_flux,
                                      y_area_flux), others2=(surface_height, ws))
        c = self._cols(height.device)
        work, slab = self._work.get(height, self._nk)
        # the Courant numbers and area fluxes are FP64, or the mixed-precision
        # step's FP32 storages
        bulk = courant_number_x.dtype
        if bulk not in (torch.float64, torch.float32) or any(
                t.dtype != bulk for t in (courant_number_y, x_area_flux, y_area_flux)):
            raise TypeError("crx, cry, xfx and yfx must share one dtype, FP64 or FP32")
        name = "fv3_update_dz_d_f64" if bulk == torch.float64 else "fv3_update_dz_d_mixed"
        _lib.check(
            getattr(_lib.lib(), name)(
                ctypes.byref(d), ctypes.byref(met), ctypes.c_int32(self._hord), self._nord4,
                _lib.ptr(c["damp"]), _lib.ptr(c["gk"]), _lib.ptr(c["beta"]),
                _lib.ptr(c["gamma"]), _lib.ptr(surface_height), _lib.ptr(height),
                _lib.ptr(courant_number_x), _lib.ptr(courant_number_y), _lib.ptr(x_area_flux),
                _lib.ptr(y_area_flux), _lib.ptr(ws), ctypes.c_double(dt), _lib.ptr(work),
                ctypes.c_int64(slab), ctypes.c_void_p(_lib.stream_handle()),
            )
        )
