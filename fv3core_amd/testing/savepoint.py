"""Savepoint data on disk.

The reference reads Fortran savepoints with serialbox
(``ser.Serializer(ser.OpenModeKind.Read, path, "Generator_rank<N>")``,
tests/savepoint/test_translate.py:228-231): a savepoint has a name (for
example "FvTp2d-In" / "FvTp2d-Out") and holds named fields.  serialbox is a
compiled third-party library that is not part of this build; savepoints are
read here from ``.npz`` files, one per savepoint and rank:

    <directory>/Generator_rank<N>/<savepoint name>.npz   (one array per field)

``serialbox_to_npz`` writes that layout from a serialbox archive where
serialbox is installed; ``write_savepoint`` writes it from arrays (tests,
or data made by another model).
"""
from __future__ import annotations

import os
from typing import Dict, List

import numpy as np


class Savepoint:
    def __init__(self, name: str):
        self.name = name

    def __repr__(self):
        return f"Savepoint({self.name!r})"


class NpzSerializer:
    """Read-only savepoints of one rank (the reader API the translate
    harness uses: ``savepoint_list``, ``fields_at_savepoint``, ``read``)."""

    def __init__(self, directory: str, rank: int = 0, prefix: str = "Generator_rank"):
        self.path = os.path.join(directory, f"{prefix}{rank}")
        if not os.path.isdir(self.path):
            raise FileNotFoundError(f"no savepoints for rank {rank} under {directory}")
        self._cache: Dict[str, Dict[str, np.ndarray]] = {}

    def savepoint_list(self) -> List[Savepoint]:
        names = sorted(f[:-4] for f in os.listdir(self.path) if f.endswith(".npz"))
        return [Savepoint(n) for n in names]

    def _load(self, savepoint) -> Dict[str, np.ndarray]:
        name = savepoint.name if isinstance(savepoint, Savepoint) else str(savepoint)
        if name not in self._cache:
            with np.load(os.path.join(self.path, name + ".npz"), allow_pickle=False) as z:
                self._cache[name] = {k: z[k] for k in z.files}
        return self._cache[name]

    def fields_at_savepoint(self, savepoint) -> List[str]:
        return sorted(self._load(savepoint))

    def read(self, varname: str, savepoint) -> np.ndarray:
        fields = self._load(savepoint)
        if varname not in fields:
            raise KeyError(f"{varname} is not saved at {savepoint}")
        return fields[varname]


def write_savepoint(directory: str, name: str, fields: Dict[str, np.ndarray], rank: int = 0,
                    prefix: str = "Generator_rank") -> str:
    path = os.path.join(directory, f"{prefix}{rank}")
    os.makedirs(path, exist_ok=True)
    out = os.path.join(path, name + ".npz")
    np.savez(out, **{k: np.asarray(v) for k, v in fields.items()})
    return out


def serialbox_to_npz(data_path: str, out_dir: str, rank: int = 0) -> List[str]:
    """Copy every savepoint of one rank of a serialbox archive into the
    ``.npz`` layout (needs the serialbox Python module)."""
    import serialbox as ser  # noqa: F401 (absent in this build's environment)

    serializer = ser.Serializer(ser.OpenModeKind.Read, data_path, f"Generator_rank{rank}")
    written = []
    for sp in serializer.savepoint_list():
        fields = {f: serializer.read(f, sp) for f in serializer.fields_at_savepoint(sp)}
        written.append(write_savepoint(out_dir, sp.name, fields, rank))
    return written
