"""Corner copies on the GPU (fv3core/utils/corners.py:19-117).

``CopyCornersXY`` keeps the reference's contract: calling it with a field
fills that field's x-corner halo cells in place (the x-differentiable field
shares memory with the base field) and writes a y-differentiable copy into
the y storage given at construction.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

from .. import _lib
from .stencil import StencilFactory


class CopyCornersXY:
    """corners.py:64-117; kernel ``fv3_copy_corners_xy``."""

    def __init__(self, stencil_factory: StencilFactory, dims: Sequence[str], y_field) -> None:
        self._grid_indexing = stencil_factory.grid_indexing
        self._dims = tuple(dims)
        self._y_field = y_field
        # the levels the dims name (corners.py:84-86): nz for Z_DIM, nz + 1
        # for Z_INTERFACE_DIM; levels beyond them are left alone
        _, domain = self._grid_indexing.get_origin_domain(dims=self._dims)
        self._nk = int(domain[2]) if len(domain) > 2 else 1

    def __call__(self, field):
        _lib.require_device(field, self._y_field)
        nk = self._nk
        if field.shape[-1] < nk:
            raise ValueError(f"the dims name {nk} levels but the field has {field.shape[-1]}")
        d = _lib.make_domain(self._grid_indexing, field, nk, others3=(self._y_field,))
        _lib.check(
            _lib.fn("fv3_copy_corners_xy", field.dtype)(
                ctypes.byref(d), _lib.ptr(field), _lib.ptr(field), _lib.ptr(self._y_field),
                ctypes.c_void_p(_lib.stream_handle()),
            )
        )
        return field, self._y_field
