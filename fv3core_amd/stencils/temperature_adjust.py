"""compute_pkz_tempadjust (fv3core/stencils/temperature_adjust.py:8-44) on the GPU."""
from __future__ import annotations

import ctypes

import torch

from .. import _lib


def compute_pkz_tempadjust(grid_indexing, nk: int, delp, delz, cappa, heat_source, pt, pkz,
                           delt_time_factor: float):
    """pt += limited heating / pkz and pkz recomputed, levels [0, nk)."""
    _lib.require_device(delp, delz, cappa, heat_source, pt, pkz)
    if nk <= 0:
        return
    d = _lib.make_domain(grid_indexing, delp, nk, others3=(delz, cappa, heat_source, pt, pkz))
    # delz and pkz FP64; delp, cappa, heat_source and pt FP64 or (mixed form) FP32
    if delz.dtype != torch.float64 or pkz.dtype != torch.float64:
        raise TypeError("delz and pkz must be float64")
    bulk = delp.dtype
    if bulk not in (torch.float64, torch.float32) or any(
            t.dtype != bulk for t in (cappa, heat_source, pt)):
        raise TypeError("delp, cappa, heat_source and pt must share one dtype")
    name = "fv3_pkz_tempadjust_f64" if bulk == torch.float64 else "fv3_pkz_tempadjust_mixed"
    _lib.check(getattr(_lib.lib(), name)(
        ctypes.byref(d), _lib.ptr(delp), _lib.ptr(delz), _lib.ptr(cappa), _lib.ptr(heat_source),
        _lib.ptr(pt), _lib.ptr(pkz), ctypes.c_double(delt_time_factor),
        ctypes.c_void_p(_lib.stream_handle())))
