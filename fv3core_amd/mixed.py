"""Storage plan of the mixed-precision acoustic step.

The reference computes in FP64 only (fv3core/utils/typing.py:24-25).  In the
mixed form the bulk horizontal flux and advection stencils (c_sw, d_sw with
its fv_tp_2d transports, del-n damping, divergence damping and a2b_ord4,
Rayleigh damping, del2cubed, the halo exchange of their fields) hold and
compute their fields in FP32, and FP64 is kept for the Riemann solvers'
tridiagonal pressure / geopotential
solve, the pressure-edge and pk / pe integration (pe, peln, pk, pk3, pkc,
pkz), the heights the solvers integrate (zh, gz, delz) and the pressure
gradient on them.

* ``BULK_FIELDS``: state fields stored in FP32.
* Every other field of the acoustic state stays FP64.
* The FP64 solvers take their FP32 inputs directly (``fv3_riem_solver_c_mixed``,
  ``fv3_riem_solver3_mixed``, ``fv3_p_grad_c_mixed``, ``fv3_nh_p_grad_mixed``,
  ``fv3_pkz_tempadjust_mixed``: FP32 loads widened to FP64, FP64 results
  rounded on store), and so do the C- and D-grid height updates
  (``fv3_update_dz_c_mixed``, ``fv3_update_dz_d_mixed``: FP32 winds, Courant
  numbers and area fluxes widened exactly on load); the pressure halos and
  nh_p_grad read an FP64 copy of delp made by ``fv3_cast_f32_to_f64``.
"""
from __future__ import annotations

import ctypes
import types

import torch

from . import _lib
from .utils.grid import GridData

BULK_FIELDS = (
    "u", "v", "w", "delp", "pt", "q_con", "cappa", "uc", "vc", "ua", "va", "omga",
    "mfxd", "mfyd", "cxd", "cyd", "diss_estd",
    # AcousticDynamics' temporaries, which it adds to the state
    "ut", "vt", "heat_source", "divgd", "crx", "xfx", "cry", "yfx", "delpc", "ptc",
)


def _like(t: torch.Tensor, dtype: torch.dtype) -> torch.Tensor:
    """A copy of t in dtype with t's strides (the i-fastest storage layout)."""
    out = torch.empty_like(t, dtype=dtype)
    out.copy_(t)
    return out


def mixed_state(state) -> types.SimpleNamespace:
    """A copy of an FP64 acoustic state with the bulk fields in FP32."""
    out = types.SimpleNamespace()
    for k, v in vars(state).items():
        if isinstance(v, torch.Tensor):
            dtype = torch.float32 if k in BULK_FIELDS else v.dtype
            setattr(out, k, _like(v, dtype))
        else:
            setattr(out, k, v)
    return out


def grid_fp32(grid_data: GridData) -> GridData:
    """The metric terms of grid_data rounded to FP32 (same layouts)."""
    fields = {k: (_like(v, torch.float32) if isinstance(v, torch.Tensor) and v.is_floating_point()
                  else v) for k, v in grid_data._fields.items()}
    return GridData(fields, grid_data._vertical)


def check_mixed_state(state) -> None:
    """Raises TypeError unless the bulk fields are FP32 and the rest FP64."""
    for k, v in vars(state).items():
        if not isinstance(v, torch.Tensor) or not v.is_floating_point():
            continue
        want = torch.float32 if k in BULK_FIELDS else torch.float64
        if v.dtype != want:
            raise TypeError(f"mixed-precision state: {k} must be {want}, got {v.dtype}")


def cast_into(src: torch.Tensor, dst: torch.Tensor) -> torch.Tensor:
    """dst <- src converted (FP32 <-> FP64) over the whole storage, on the
    library's conversion kernel; both must be dense with the same layout."""
    if src.shape != dst.shape or src.stride() != dst.stride():
        raise ValueError("cast_into needs two storages of one layout")
    if not _dense(src):
        raise ValueError("cast_into needs a dense storage")
    n = src.numel()
    stream = ctypes.c_void_p(_lib.stream_handle())
    if src.dtype == torch.float32 and dst.dtype == torch.float64:
        rc = _lib.lib().fv3_cast_f32_to_f64(ctypes.c_int64(n), _lib.ptr(src), _lib.ptr(dst),
                                            stream)
    elif src.dtype == torch.float64 and dst.dtype == torch.float32:
        rc = _lib.lib().fv3_cast_f64_to_f32(ctypes.c_int64(n), _lib.ptr(src), _lib.ptr(dst),
                                            stream)
    else:
        raise TypeError(f"cast_into: {src.dtype} -> {dst.dtype}")
    _lib.check(rc)
    return dst


def _dense(t: torch.Tensor) -> bool:
    """True when t's elements fill one block of t.numel() elements."""
    span = 1 + sum((n - 1) * s for n, s in zip(t.shape, t.stride()))
    return span == t.numel() and min(t.stride()) >= 1
