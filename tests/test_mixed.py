"""Mixed precision: FP32 arithmetic in the fv_tp_2d
transports on FP64 storages, against the plain FP64 form.

Error metric: the reference's 2|a-b|/(|a|+|b|) (test_translate.py:26-33),
with denominators floored at 1e-6 of the field's largest magnitude.  The
gate is 1e-5 in ps, pt and delp after one AcousticDynamics call; every other
output field's error is printed.
"""
import types

import numpy as np
import pytest
import torch

from fv3core_amd._config import DynamicalCoreConfig
from fv3core_amd.workloads import cube_state, transport_inputs

pytestmark = pytest.mark.gpu


def _rel(a, b, floor=1e-6):
    a = a.double()
    b = b.double()
    scale = torch.maximum(a.abs().max(), b.abs().max())
    den = torch.clamp(a.abs() + b.abs(), min=floor * float(scale) + 1e-300)
    return float((2.0 * (a - b).abs() / den).max())


@pytest.mark.parametrize("hord", [5, 6, 8])
@pytest.mark.parametrize("edges", [(False,) * 4, (True,) * 4])
def test_gpu_mixed_transport_close_to_fp64(hord, edges):
    from fv3core_amd import _lib
    from fv3core_amd.adjoint import TransportLinearization
    from fv3core_amd.utils.grid import to_storage
    from fv3core_amd.utils.stencil import GridIndexing
    import ctypes

    n, nz = 48, 8
    d = transport_inputs(n, nz, workload=3)
    gi = GridIndexing((n, n, nz), 3, *edges)
    dev = {k: to_storage(v) for k, v in d.items()}
    lin = TransportLinearization(gi, dev["area"], dev["dxa"], dev["dya"], hord, nz + 1)
    q = dev["q"]
    aux = (dev["crx"], dev["cry"], dev["x_area_flux"], dev["y_area_flux"])
    fx, fy = lin.forward(q, q, *aux)
    mx, my = torch.zeros_like(fx), torch.zeros_like(fy)
    dd = lin._domain(q, q, *aux, mx, my)
    _lib.check(_lib.lib().fv3_fvtp2d_mixed_f64(
        ctypes.byref(dd), hord, _lib.ptr(q), _lib.ptr(q), *[_lib.ptr(a) for a in aux],
        None, None, _lib.ptr(dev["area"]), _lib.ptr(dev["dxa"]), _lib.ptr(dev["dya"]),
        _lib.ptr(mx), _lib.ptr(my), ctypes.c_void_p(_lib.stream_handle())))
    torch.cuda.synchronize()
    h = 3
    ex = _rel(mx[h:h + n + 1, h:h + n], fx[h:h + n + 1, h:h + n])
    ey = _rel(my[h:h + n, h:h + n + 1], fy[h:h + n, h:h + n + 1])
    print(f"hord={hord} edges={edges[0]}: mixed vs FP64 fx {ex:.2e} fy {ey:.2e}")
    assert ex < 1e-5 and ey < 1e-5
    assert ex > 0.0  # the mixed form really computes in FP32


def test_gpu_mixed_acoustic_step_error_budget():
    from fv3core_amd.ensemble import ensemble_acoustic, member_slice

    n, nz = 24, 8
    init = cube_state(n, nz, seed=7)
    cfg = DynamicalCoreConfig(npz=nz, n_split=2, k_split=1, nord=3, dt_atmos=225)
    plain, state = ensemble_acoustic(init, 1, cfg, seed=1)
    mixed_state = member_slice(state, 0)
    mixed, _ = ensemble_acoustic(init, 1, cfg, seed=1)
    mixed.dgrid_shallow_water_lagrangian_dynamics.transport_fp32 = True
    plain(state)
    mixed(mixed_state)
    torch.cuda.synchronize()
    c = (slice(None), slice(3, n + 3), slice(3, n + 3))
    errs = {}
    for k in ("delp", "pt", "delz", "w", "u", "v", "zh", "pe", "pk", "peln", "pkz", "ua", "va",
              "uc", "vc", "omga"):
        a = getattr(mixed_state, k)[c + (slice(0, nz),)]
        b = getattr(state, k)[c + (slice(0, nz),)]
        errs[k] = _rel(a, b)
    ps_m = mixed_state.pe[c + (nz,)]
    ps_p = state.pe[c + (nz,)]
    errs["ps"] = _rel(ps_m, ps_p)
    print("mixed vs FP64 after one acoustic call: " +
          ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    for k in ("ps", "pt", "delp"):
        assert errs[k] <= 1e-5, f"{k}: {errs[k]:.2e} above the 1e-5 gate"
    assert errs["delp"] > 0.0
