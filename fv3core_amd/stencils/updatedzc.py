"""UpdateGeopotentialHeightOnCGrid (fv3core/stencils/updatedzc.py:106-163) on the GPU.

``fv3_update_dz_c`` evaluates the corner-filled copies of gz on the fly,
advances gz one column per lane and applies the DZ_MIN floor; a second
launch copies the result back into gz (neighbouring columns read the old gz).
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib
from ..utils.stencil import StencilFactory
from ._workspace import Workspace


class UpdateGeopotentialHeightOnCGrid:
    """Fortran updatedzc."""

    def __init__(self, stencil_factory: StencilFactory, area):
        self._gi = stencil_factory.grid_indexing
        self._area = area
        self._work = Workspace(1)
        self._dp = {}

    def _column(self, dp_ref, device):
        if isinstance(dp_ref, torch.Tensor) and dp_ref.is_cuda:
            return dp_ref.to(torch.float64).contiguous()
        key = str(device)
        if key not in self._dp:
            import numpy as np

            self._dp[key] = torch.as_tensor(np.asarray(dp_ref, dtype=np.float64),
                                            device=device)
        return self._dp[key]

    def __call__(self, dp_ref, zs, ut, vt, gz, ws, dt: float):
        """gz (interface levels) and ws in place."""
        _lib.require_device(zs, ut, vt, gz, ws, self._area)
        nk = self._gi.domain[2] + 1
        d = _lib.make_domain(self._gi, gz, nk, self._area, others3=(ut, vt), others2=(zs, ws))
        dp = self._column(dp_ref, gz.device)
        work, slab = self._work.get(gz, nk)
        # ut, vt are FP64, or the mixed-precision step's FP32 storages
        if ut.dtype != vt.dtype or ut.dtype not in (torch.float64, torch.float32):
            raise TypeError("ut and vt must share one dtype, FP64 or FP32")
        name = "fv3_update_dz_c_f64" if ut.dtype == torch.float64 else "fv3_update_dz_c_mixed"
        _lib.check(
            getattr(_lib.lib(), name)(
                ctypes.byref(d), _lib.ptr(self._area), _lib.ptr(zs), _lib.ptr(dp), _lib.ptr(ut),
                _lib.ptr(vt), _lib.ptr(gz), _lib.ptr(ws), ctypes.c_double(dt), _lib.ptr(work),
                ctypes.c_int64(slab), ctypes.c_void_p(_lib.stream_handle()),
            )
        )
