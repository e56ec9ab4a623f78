"""The C-ABI library loads and exports every symbol the header declares
(no kernel is launched: this runs without a GPU)."""
import ctypes

import pytest
import torch

from fv3core_amd import _lib


def test_header_symbols_are_exported():
    lib = _lib.lib()
    names = _lib.exported_symbols()
    assert "fv3_fvtp2d_f64" in names and "fv3_delnflux_f64" in names
    for name in names:
        assert hasattr(lib, name), name


def test_version_and_error_channel():
    lib = _lib.lib()
    assert lib.fv3_version().decode().startswith("fv3core_amd")
    assert isinstance(lib.fv3_last_error().decode(), str)


def test_domain_struct_matches_header_layout():
    # 12 int32 then 5 int64: offsets fixed by the C struct layout
    assert ctypes.sizeof(_lib.Domain) == 12 * 4 + 5 * 8
    assert _lib.Domain.sj.offset == 48


def test_argument_errors_are_reported_without_a_gpu():
    lib = _lib.lib()
    d = _lib.Domain()
    d.isc, d.iec, d.jsc, d.jec = 3, 14, 3, 14
    d.i_start, d.i_end, d.j_start, d.j_end = 3, 14, 3, 14
    d.ni = d.nj = 19
    d.nk, d.ntiles = 4, 1
    d.sj, d.sk, d.st, d.sj2, d.st2 = 19, 19 * 19, 0, 19, 0
    rc = lib.fv3_fvtp2d_f64(ctypes.byref(d), 8, *([None] * 13), None)
    assert rc == -1
    assert "NULL" in lib.fv3_last_error().decode()
    rc = lib.fv3_fvtp2d_f64(ctypes.byref(d), 10, *([ctypes.c_void_p(16)] * 13), None)
    assert rc == -2
    d.ni = 10
    rc = lib.fv3_fvtp2d_f64(ctypes.byref(d), 8, *([ctypes.c_void_p(16)] * 13), None)
    assert rc == -1
    assert "storage" in lib.fv3_last_error().decode()
    nord = (ctypes.c_int32 * 4)(3, 3, 3, 3)
    d.ni = 19
    rc = lib.fv3_delnflux_f64(ctypes.byref(d), 0, nord, *([ctypes.c_void_p(16)] * 8), None)
    assert rc == -2


def test_cpu_tensor_is_rejected():
    with pytest.raises(RuntimeError, match="GPU only"):
        _lib.require_device(torch.zeros(3))


def test_riemann_solvers_reject_more_levels_than_their_lds_column_holds():
    """The Riemann solvers keep one column array per wave in LDS and accept
    4 to 128 interface levels; the check runs before any launch."""
    lib = _lib.lib()
    d = _lib.Domain()
    d.isc, d.iec, d.jsc, d.jec = 3, 14, 3, 14
    d.i_start, d.i_end, d.j_start, d.j_end = 3, 14, 3, 14
    d.ni = d.nj = 19
    d.nk, d.ntiles = 129, 1
    d.sj, d.sk, d.st, d.sj2, d.st2 = 19, 19 * 19, 19 * 19 * 130, 19, 19 * 19
    p = ctypes.c_void_p(16)
    slab = ctypes.c_int64(19 * 19 * 130)
    rc = lib.fv3_riem_solver_c_f64(ctypes.byref(d), ctypes.c_double(1.0), ctypes.c_double(300.0),
                                   ctypes.c_double(0.05), *([p] * 10), slab, None)
    assert rc == -2
    assert "128" in lib.fv3_last_error().decode()
    rc = lib.fv3_riem_solver3_f64(ctypes.byref(d), 1, ctypes.c_double(1.0),
                                  ctypes.c_double(300.0), ctypes.c_double(0.05), 0,
                                  ctypes.c_double(0.0), *([p] * 15), slab, None)
    assert rc == -2
    d.nk = 3
    rc = lib.fv3_riem_solver3_f64(ctypes.byref(d), 1, ctypes.c_double(1.0),
                                  ctypes.c_double(300.0), ctypes.c_double(0.05), 0,
                                  ctypes.c_double(0.0), *([p] * 15), slab, None)
    assert rc == -1
