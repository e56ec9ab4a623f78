"""FiniteVolumeTransport (fv_tp_2d) on the GPU (fv3core/stencils/fvtp2d.py:96-362).

The class keeps the reference's constructor and ``__call__`` signatures.  The
eight stencils of the reference (inner/outer PPM in x and y, q_i, q_j,
final_fluxes) run as one fused kernel ``fv3_fvtp2d``; optional del-n damping
runs as the fused kernel ``fv3_delnflux`` afterwards, as the reference's
``DelnFlux`` call does.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Optional, Sequence

from .. import _lib
from ..utils import corners
from ..utils import gt4py_utils as utils
from ..utils.grid import DampingCoefficients, GridData
from ..utils.stencil import StencilFactory
from .delnflux import DelnFlux


@dataclasses.dataclass
class CopiedCorners:
    """Storage plus its x- and y-differentiable versions (fvtp2d.py:96-111)."""

    base: object
    x_differentiable: object
    y_differentiable: object


class PreAllocatedCopiedCornersFactory:
    """Creates CopiedCorners with a pre-allocated y storage (fvtp2d.py:114-148)."""

    def __init__(self, stencil_factory: StencilFactory, *, dims: Sequence[str], y_temporary):
        if y_temporary is None:
            y_temporary = utils.make_storage_from_shape(
                stencil_factory.grid_indexing.max_shape,
                origin=stencil_factory.grid_indexing.origin_compute(),
            )
        self._copy_corners_xy = corners.CopyCornersXY(stencil_factory, dims, y_field=y_temporary)

    def __call__(self, field) -> CopiedCorners:
        x_field, y_field = self._copy_corners_xy(field)
        return CopiedCorners(base=field, x_differentiable=x_field, y_differentiable=y_field)


class FiniteVolumeTransport:
    """Fortran fv_tp_2d in three dimensions (fvtp2d.py:151-362)."""

    def __init__(
        self,
        stencil_factory: StencilFactory,
        grid_data: GridData,
        damping_coefficients: DampingCoefficients,
        grid_type: int,
        hord,
        nord=None,
        damp_c=None,
    ):
        idx = stencil_factory.grid_indexing
        self._grid_indexing = idx
        self.grid_data = grid_data
        self._area = grid_data.area
        self._dxa = grid_data.dxa
        self._dya = grid_data.dya
        self._nord = nord
        self._damp_c = damp_c
        ord_outer = hord
        ord_inner = 8 if hord == 10 else hord
        if ord_outer != ord_inner or ord_outer not in (5, 6, 7, 8):
            # hord = 10 makes the outer sweep use iord 10, which the
            # reference's ord8+ path rejects (xppm.py:256)
            raise NotImplementedError(f"hord={hord} is not supported (5, 6, 7, 8)")
        if grid_type >= 3:
            raise NotImplementedError("grid_type >= 3 is not supported")
        self._hord = int(hord)
        # fluxes are computed on nz + 1 levels (domain_compute(add=(1, 1, 1)))
        self._nk = idx.domain[2] + 1
        if (self._nord is not None) and (self._damp_c is not None):
            self.delnflux: Optional[DelnFlux] = DelnFlux(
                stencil_factory=stencil_factory,
                damping_coefficients=damping_coefficients,
                rarea=grid_data.rarea,
                nord=self._nord,
                damp_c=self._damp_c,
            )
        else:
            self.delnflux = None

    def __call__(
        self,
        q: CopiedCorners,
        crx,
        cry,
        x_area_flux,
        y_area_flux,
        q_x_flux,
        q_y_flux,
        x_mass_flux=None,
        y_mass_flux=None,
        mass=None,
    ):
        """Fluxes for horizontal finite-volume transport (PL07, FV3 doc eq. 4.17).

        Args mirror the reference: q (CopiedCorners, in), crx/cry Courant
        numbers, x/y_area_flux (in), q_x_flux/q_y_flux (out), optional
        x/y_mass_flux and mass (for damping of per-mass scalars).
        """
        if (
            self.delnflux is not None
            and mass is None
            and (x_mass_flux is not None or y_mass_flux is not None)
        ):
            raise ValueError(
                "when damping is enabled, mass must be given if mass flux is given"
            )
        if not isinstance(q, CopiedCorners):
            q = CopiedCorners(base=q, x_differentiable=q, y_differentiable=q)
        qx, qy = q.x_differentiable, q.y_differentiable
        _lib.require_device(qx, qy, crx, cry, x_area_flux, y_area_flux, q_x_flux, q_y_flux,
                            x_mass_flux, y_mass_flux, self._area)
        d = _lib.make_domain(
            self._grid_indexing, qx, self._nk, self._area,
            others3=(qy, crx, cry, x_area_flux, y_area_flux, q_x_flux, q_y_flux, x_mass_flux,
                     y_mass_flux),
            others2=(self._dxa, self._dya),
        )
        _lib.check(
            _lib.fn("fv3_fvtp2d", qx.dtype)(
                ctypes.byref(d), self._hord, _lib.ptr(qx), _lib.ptr(qy), _lib.ptr(crx),
                _lib.ptr(cry), _lib.ptr(x_area_flux), _lib.ptr(y_area_flux),
                _lib.ptr(x_mass_flux), _lib.ptr(y_mass_flux), _lib.ptr(self._area),
                _lib.ptr(self._dxa), _lib.ptr(self._dya), _lib.ptr(q_x_flux), _lib.ptr(q_y_flux),
                ctypes.c_void_p(_lib.stream_handle()),
            )
        )
        if self.delnflux is not None:
            self.delnflux(q.base, q_x_flux, q_y_flux, mass=mass)
