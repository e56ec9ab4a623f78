"""DelnFlux and DelnFluxNoSG (fv3core/stencils/delnflux.py:901-1207) on the GPU.

Both call the fused kernel ``fv3_delnflux``: d2 preamble, corner copies,
the del-2 flux and the ``nmax`` higher-order passes run in LDS per block.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from .. import _lib
from ..utils.grid import DampingCoefficients
from ..utils.stencil import StencilFactory


def _column(values, n: int) -> np.ndarray:
    a = values.detach().cpu().numpy() if isinstance(values, torch.Tensor) else np.asarray(values)
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    if a.size < n:
        raise ValueError(f"column has {a.size} levels, {n} are needed")
    return a[:n]


def calc_damp(nord, damp_c, da_min: float, nk: int) -> np.ndarray:
    """calc_damp (delnflux.py:24-26): (damp_c * da_min) ** (nord + 1) per level."""
    return (_column(damp_c, nk) * da_min) ** (_column(nord, nk) + 1)


def _nord_bands(nord) -> list:
    col = _column(nord, 4)
    bands = [int(v) for v in col]
    if max(bands) > 3:
        raise ValueError("nord must be less than 3")
    return bands


class _DampColumn:
    """A per-level coefficient column cached on the device per dtype."""

    def __init__(self, values: np.ndarray):
        self._values = np.asarray(values, dtype=np.float64)
        self._cache = {}

    def get(self, dtype: torch.dtype, device) -> torch.Tensor:
        key = (dtype, str(device))
        if key not in self._cache:
            self._cache[key] = torch.as_tensor(self._values, dtype=dtype, device=device)
        return self._cache[key]

    def __getitem__(self, k):
        return self._values[k]

    def __len__(self):
        return len(self._values)


class DelnFluxNoSG:
    """del6_vt / deln_flux mechanics (delnflux.py:1001-1207)."""

    def __init__(self, stencil_factory: StencilFactory, damping_coefficients: DampingCoefficients,
                 rarea, nord, nk: Optional[int] = None):
        gi = stencil_factory.grid_indexing
        self._grid_indexing = gi
        self._del6_u = damping_coefficients.del6_u
        self._del6_v = damping_coefficients.del6_v
        self._rarea = rarea
        self._nord4 = _nord_bands(nord)
        self._nmax = max(self._nord4)
        if nk is None:
            nk = gi.domain[2]
        if nk <= 3:
            raise NotImplementedError("nk must be more than 3 for DelnFluxNoSG")
        if self._nmax > 2:
            raise NotImplementedError(
                "nord = 3 places the d2 origin outside the 3-cell halo storage"
            )
        self._nk = nk

    def __call__(self, q, fx2, fy2, damp_c, d2=None, mass=None):
        """fx2, fy2 <- del-n fluxes of q (damp_c: per-level coefficient column)."""
        _lib.require_device(q, fx2, fy2, mass, self._rarea)
        if isinstance(damp_c, _DampColumn):
            damp = damp_c.get(q.dtype, q.device)
        elif isinstance(damp_c, torch.Tensor) and damp_c.is_cuda:
            damp = damp_c.to(q.dtype).reshape(-1)
        else:
            damp = torch.as_tensor(_column(damp_c, self._nk), dtype=q.dtype, device=q.device)
        mode = _lib.DELN_STORE if mass is None else _lib.DELN_STORE_MASS
        self._launch(mode, q, mass, damp, fx2, fy2)

    def _launch(self, mode, q, mass, damp, fx, fy):
        d = _lib.make_domain(self._grid_indexing, q, self._nk, self._rarea,
                             others3=(fx, fy, mass), others2=(self._del6_u, self._del6_v))
        nord4 = (ctypes.c_int32 * 4)(*self._nord4)
        _lib.check(
            _lib.fn("fv3_delnflux", q.dtype)(
                ctypes.byref(d), mode, nord4, _lib.ptr(q), _lib.ptr(mass), _lib.ptr(damp),
                _lib.ptr(self._del6_u), _lib.ptr(self._del6_v), _lib.ptr(self._rarea),
                _lib.ptr(fx), _lib.ptr(fy), ctypes.c_void_p(_lib.stream_handle()),
            )
        )


class DelnFlux:
    """Fortran deln_flux (delnflux.py:901-998): adds del-n damping to fluxes."""

    def __init__(self, stencil_factory: StencilFactory, damping_coefficients: DampingCoefficients,
                 rarea, nord, damp_c):
        damp_c_arr = _column(damp_c, len(damp_c) if hasattr(damp_c, "__len__") else 1)
        self._no_compute = False
        if (damp_c_arr <= 1e-4).all():
            self._no_compute = True
        elif (damp_c_arr[:-1] <= 1e-4).any():
            raise NotImplementedError(
                "damp_c currently must be always greater than 10^-4 for delnflux"
            )
        gi = stencil_factory.grid_indexing
        nk = gi.domain[2]
        self._nk = nk
        self._damp = _DampColumn(calc_damp(nord, damp_c, damping_coefficients.da_min, nk))
        self.delnflux_nosg = DelnFluxNoSG(stencil_factory, damping_coefficients, rarea, nord,
                                          nk=nk)

    def __call__(self, q, fx, fy, d2=None, mass=None):
        if self._no_compute:
            return fx, fy
        damp = self._damp.get(q.dtype, q.device)
        mode = _lib.DELN_ADD if mass is None else _lib.DELN_MASS
        _lib.require_device(q, fx, fy, mass)
        self.delnflux_nosg._launch(mode, q, mass, damp, fx, fy)
        return fx, fy
