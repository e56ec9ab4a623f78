"""RayleighDamping (fv3core/stencils/ray_fast.py:130-192) on the GPU.

The per-level quantities (rf, the nudged column mass dm and the two level
masks) depend only on pfull, dp, dt and ptop; they are computed on the host
as in the reference's FORWARD/BACKWARD sweeps, and ``fv3_ray_fast`` applies
them one column per lane.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from .. import _lib
from ..utils.stencil import StencilFactory

SDAY = 86400.0


def rayleigh_columns(pfull, dp, dt, ptop, rf_cutoff, tau):
    """(rf, use_rf, use_nudge, dm) per level (ray_fast.py:25-75)."""
    pfull = np.asarray(pfull, dtype=np.float64)
    dp = np.asarray(dp, dtype=np.float64)
    nk = len(pfull)
    nudge = rf_cutoff + min(100.0, 10.0 * ptop)
    use_rf = pfull < rf_cutoff
    use_n = pfull < nudge
    tau0 = tau * SDAY
    with np.errstate(all="ignore"):
        rfv = dt / tau0 * np.sin(0.5 * math.pi * np.log(rf_cutoff / pfull)
                                 / np.log(rf_cutoff / ptop)) ** 2
        rfv = 1.0 / (1.0 + rfv)
    rf = np.where(use_rf, rfv, 1.0)
    dm = np.zeros(nk)
    acc = 0.0
    for k in range(nk):
        if k == 0:
            acc = dp[0] if use_n[0] else 0.0
        elif use_n[k]:
            acc = acc + dp[k]
        dm[k] = acc
    for k in range(nk - 2, -1, -1):
        if use_n[k]:
            dm[k] = dm[k + 1]
    if not use_rf[0]:
        dm[0] = 0.0  # the reference zeroes dm at the top in that case
    return rf, use_rf, use_n, dm


class RayleighDamping:
    """Fortran ray_fast."""

    def __init__(self, stencil_factory: StencilFactory, rf_cutoff, tau, hydrostatic):
        self._gi = stencil_factory.grid_indexing
        self._rf_cutoff = float(rf_cutoff)
        self._tau = float(tau)
        self._hydrostatic = bool(hydrostatic)
        self._cache = {}

    def __call__(self, u, v, w, dp, pfull, dt: float, ptop: float, ks: int):
        _lib.require_device(u, v, w)
        nz = self._gi.domain[2]
        key = (float(dt), float(ptop), str(u.device), u.dtype)
        if key not in self._cache:
            dp_h = dp.detach().cpu().numpy() if isinstance(dp, torch.Tensor) else np.asarray(dp)
            pf_h = (pfull.detach().cpu().numpy() if isinstance(pfull, torch.Tensor)
                    else np.asarray(pfull))
            rf, use_rf, use_n, dm = rayleigh_columns(pf_h[:nz], dp_h[:nz], dt, ptop,
                                                     self._rf_cutoff, self._tau)
            dev = u.device
            self._cache[key] = (
                torch.as_tensor(rf, dtype=u.dtype, device=dev),
                torch.as_tensor(np.asarray(dp_h[:nz], dtype=np.float64), dtype=u.dtype,
                                device=dev),
                torch.as_tensor(dm, dtype=u.dtype, device=dev),
                torch.as_tensor(use_rf.astype(np.int32), device=dev),
                torch.as_tensor(use_n.astype(np.int32), device=dev),
            )
        rf, dpd, dm, urf, un = self._cache[key]
        d = _lib.make_domain(self._gi, u, nz, others3=(v, w))
        _lib.check(_lib.fn("fv3_ray_fast", u.dtype)(
            ctypes.byref(d), ctypes.c_int32(int(self._hydrostatic)), _lib.ptr(rf), _lib.ptr(dpd),
            _lib.ptr(dm), _lib.ptr(urf), _lib.ptr(un), _lib.ptr(u), _lib.ptr(v), _lib.ptr(w),
            ctypes.c_void_p(_lib.stream_handle())))
