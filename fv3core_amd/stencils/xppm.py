"""XPiecewiseParabolic (fv3core/stencils/xppm.py:283-348) on the GPU."""
from __future__ import annotations

import ctypes

from .. import _lib
from ..utils.stencil import StencilFactory


class XPiecewiseParabolic:
    """Fortran name xppm.  ``__call__(q_in, c, q_mean_advected_through_x_interface)``."""

    _direction = 0

    def __init__(self, stencil_factory: StencilFactory, dxa, grid_type: int, iord, origin,
                 domain):
        assert grid_type < 3
        if abs(iord) not in (5, 6, 7, 8) or iord < 0:
            raise NotImplementedError(f"iord={iord} is not supported (5, 6, 7, 8)")
        self._grid_indexing = stencil_factory.grid_indexing
        self._dxa = dxa
        self._ord = int(iord)
        self._origin = tuple(origin)
        self._domain = tuple(domain)

    def __call__(self, q_in, c, q_mean_advected_through_x_interface):
        out = q_mean_advected_through_x_interface
        _lib.require_device(q_in, c, out, self._dxa)
        k0, nk = self._origin[2], self._domain[2]
        qs, cs, os_ = q_in[..., k0:], c[..., k0:], out[..., k0:]
        d = _lib.make_domain(self._grid_indexing, qs, nk, self._dxa, others3=(cs, os_))
        _lib.check(
            _lib.fn("fv3_ppm_flux", q_in.dtype)(
                ctypes.byref(d), self._direction, self._ord, self._origin[0], self._origin[1],
                self._domain[0], self._domain[1], _lib.ptr(qs), _lib.ptr(cs),
                _lib.ptr(self._dxa), _lib.ptr(os_), ctypes.c_void_p(_lib.stream_handle()),
            )
        )
