"""Component-owned scratch slabs laid out like the fields they serve."""
from __future__ import annotations

import torch

from .. import _lib


class Workspace:
    """``nslabs`` contiguous slabs, each addressable with a field's strides.

    A slab must hold ``(ntiles - 1) * st + nk * sk`` elements for the
    field layout it is used with; the buffer grows when a call needs more
    and is reused otherwise (the reference's components likewise own their
    temporaries, so a component is not re-entrant).
    """

    def __init__(self, nslabs: int):
        self._nslabs = nslabs
        self._buf = None
        self._slab = 0

    def get(self, ref: torch.Tensor, nk: int):
        ntiles, ni, nj, _, sj, sk, st = _lib.layout3(ref)
        need = (ntiles - 1) * st + nk * sk
        need = (need + 63) // 64 * 64
        if (self._buf is None or self._slab < need or self._buf.dtype != ref.dtype
                or self._buf.device != ref.device):
            self._buf = torch.empty(self._nslabs * need, dtype=ref.dtype, device=ref.device)
            self._slab = need
        return self._buf, self._slab
