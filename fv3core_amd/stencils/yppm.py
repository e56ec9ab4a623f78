"""YPiecewiseParabolic (fv3core/stencils/yppm.py, the y twin of xppm.py)."""
from __future__ import annotations

from .xppm import XPiecewiseParabolic


class YPiecewiseParabolic(XPiecewiseParabolic):
    """Fortran name yppm.  ``__call__(q_in, c, q_mean_advected_through_y_interface)``."""

    _direction = 1

    def __init__(self, stencil_factory, dya, grid_type: int, jord, origin, domain):
        super().__init__(stencil_factory, dya, grid_type, jord, origin, domain)
