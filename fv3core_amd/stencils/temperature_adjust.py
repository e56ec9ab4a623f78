"""compute_pkz_tempadjust (fv3core/stencils/temperature_adjust.py:8-44) on the GPU."""
from __future__ import annotations

import ctypes

from .. import _lib


def compute_pkz_tempadjust(grid_indexing, nk: int, delp, delz, cappa, heat_source, pt, pkz,
                           delt_time_factor: float):
    """pt += limited heating / pkz and pkz recomputed, levels [0, nk)."""
    _lib.require_device(delp, delz, cappa, heat_source, pt, pkz)
    if nk <= 0:
        return
    d = _lib.make_domain(grid_indexing, delp, nk, others3=(delz, cappa, heat_source, pt, pkz))
    _lib.check(_lib.lib().fv3_pkz_tempadjust_f64(
        ctypes.byref(d), _lib.ptr(delp), _lib.ptr(delz), _lib.ptr(cappa), _lib.ptr(heat_source),
        _lib.ptr(pt), _lib.ptr(pkz), ctypes.c_double(delt_time_factor),
        ctypes.c_void_p(_lib.stream_handle())))
