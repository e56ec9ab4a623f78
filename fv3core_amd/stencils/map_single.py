"""MapSingle (fv3core/stencils/map_single.py:69-173) on the GPU.

One kernel (``fv3_map_single``) builds the cs_profile of every column
(RemapProfile, remap_profile.py:498-638) and integrates it over the target
layers (lagrangian_contributions).
"""
from __future__ import annotations

import ctypes
from typing import Optional

from .. import _lib
from ..utils.stencil import StencilFactory
from ._workspace import Workspace


class MapSingle:
    """Fortran map_single; ``__call__(q1, pe1, pe2, qs=None, qmin=0.0) -> q1``."""

    def __init__(self, stencil_factory: StencilFactory, kord: int, mode: int, i1: int, i2: int,
                 j1: int, j2: int):
        if kord > 10:
            raise AssertionError(f"kord {kord} not implemented.")
        if abs(kord) > 10:
            # remap_profile.py:611 runs other limiter forms for kord <= -11;
            # those are not built here
            raise NotImplementedError(f"kord {kord}: only -10 <= kord <= 10 is implemented")
        self._grid_indexing = stencil_factory.grid_indexing
        self._kord = int(kord)
        self._mode = int(mode)
        self._box = (int(i1), int(i2), int(j1), int(j2))
        self._extents = (i2 - i1 + 1, j2 - j1 + 1)
        self._km = self._grid_indexing.domain[2]
        self._work = Workspace(7)

    @property
    def i_extent(self):
        return self._extents[0]

    @property
    def j_extent(self):
        return self._extents[1]

    def __call__(self, q1, pe1, pe2, qs: Optional[object] = None, qmin: float = 0.0):
        _lib.require_device(q1, pe1, pe2, qs)
        d = _lib.make_domain(self._grid_indexing, q1, self._km, qs, others3=(pe1, pe2))
        if pe1.shape[-1] < self._km + 1 or pe2.shape[-1] < self._km + 1:
            raise ValueError("pe1 and pe2 need km + 1 interface levels")
        work, slab = self._work.get(q1, self._km + 1)
        i1, i2, j1, j2 = self._box
        _lib.check(
            _lib.fn("fv3_map_single", q1.dtype)(
                ctypes.byref(d), self._kord, self._mode, i1, i2, j1, j2, _lib.ptr(q1),
                _lib.ptr(pe1), _lib.ptr(pe2), _lib.ptr(qs), ctypes.c_double(qmin),
                _lib.ptr(work), ctypes.c_int64(slab), ctypes.c_void_p(_lib.stream_handle()),
            )
        )
        return q1
