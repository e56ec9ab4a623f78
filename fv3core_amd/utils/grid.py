"""Grid metric containers (fv3core/utils/grid.py:474-799), holding torch storages.

``GridData`` exposes the same attribute names as the reference's (``area``,
``rarea``, ``dxa``, ``sin_sg1`` ... ``rsin2``, ``ak``/``bk``/``ptop``), each an
IJ storage ``(N + 2*halo + 1, N + 2*halo + 1)``, optionally with a leading
tile axis ``(T, I, J)`` for batched calls.
"""
from __future__ import annotations

import dataclasses
from typing import Any, Dict, Optional

import numpy as np
import torch

from . import gt4py_utils as utils

HORIZONTAL = (
    "area", "area_64", "rarea", "rarea_c", "dx", "dy", "dxc", "dyc", "dxa", "dya",
    "rdx", "rdy", "rdxc", "rdyc", "rdxa", "rdya",
)
CONTRAVARIANT = (
    "cosa", "cosa_u", "cosa_v", "cosa_s", "sina_u", "sina_v", "rsina", "rsin_u",
    "rsin_v", "rsin2",
)
ANGLE = (
    "sin_sg1", "sin_sg2", "sin_sg3", "sin_sg4", "cos_sg1", "cos_sg2", "cos_sg3", "cos_sg4",
)


@dataclasses.dataclass(frozen=True)
class DampingCoefficients:
    """del6_u, del6_v, da_min, da_min_c (grid.py:571-580)."""

    del6_u: Any
    del6_v: Any
    da_min: float
    da_min_c: float


@dataclasses.dataclass
class VerticalGridData:
    ak: Optional[Any] = None
    bk: Optional[Any] = None
    p_ref: Optional[Any] = None

    @property
    def ptop(self) -> float:
        if self.ak is None:
            raise ValueError("ak is not set")
        return float(self.ak[0])


class GridData:
    """Metric terms by name (grid.py:583-798)."""

    def __init__(self, fields: Dict[str, Any], vertical: Optional[VerticalGridData] = None):
        self._fields = dict(fields)
        self._vertical = vertical if vertical is not None else VerticalGridData()

    def __getattr__(self, name):
        fields = self.__dict__.get("_fields", {})
        if name in fields:
            return fields[name]
        raise AttributeError(f"GridData has no field {name!r}")

    def has(self, name: str) -> bool:
        return name in self._fields

    @property
    def ptop(self):
        return self._vertical.ptop

    @property
    def ak(self):
        return self._vertical.ak

    @property
    def bk(self):
        return self._vertical.bk

    @property
    def p_ref(self):
        return self._vertical.p_ref

    def fields(self) -> Dict[str, Any]:
        return dict(self._fields)


def to_storage(a, dtype=torch.float64, device=None) -> torch.Tensor:
    """Copy a numpy IJ/IJK array into an i-fastest storage."""
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    a = np.asarray(a, dtype=np.float64)
    t = utils.empty(a.shape, dtype=dtype, device=device)
    t.copy_(torch.as_tensor(a, dtype=dtype))
    return t


def grid_from_numpy(metric: Dict[str, Any], dtype=torch.float64, device=None):
    """(GridData, DampingCoefficients) from a dict of numpy metric arrays.

    Missing reciprocal fields are derived (rarea = 1/area, rdx = 1/dx ...).
    """
    m = dict(metric)
    for name, base in (("rarea", "area"), ("rdx", "dx"), ("rdy", "dy"), ("rdxc", "dxc"),
                       ("rdyc", "dyc"), ("rdxa", "dxa"), ("rdya", "dya")):
        if name not in m and base in m:
            m[name] = 1.0 / np.asarray(m[base])
    if "area_64" not in m and "area" in m:
        m["area_64"] = m["area"]
    fields = {}
    for k, v in m.items():
        if isinstance(v, (float, int)):
            continue
        fields[k] = to_storage(v, dtype=dtype, device=device)
    damp = None
    if "del6_u" in fields and "del6_v" in fields:
        damp = DampingCoefficients(
            fields["del6_u"], fields["del6_v"], float(m.get("da_min", 0.0)),
            float(m.get("da_min_c", m.get("da_min", 0.0))),
        )
    return GridData(fields), damp
