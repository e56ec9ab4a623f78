"""Host restatement of the six-tile halo exchange for the oracle drivers.

Test infrastructure only.  Applies the index maps of
``fv3core_amd.halo.build_map`` / ``build_sync_map`` with numpy gathers
(``apply_map_numpy``); the maps themselves are pinned against the cube
geometry by the halo tests.  ``HostHalo(state, n)`` refreshes fields of a
dict of [6, I, J, K] arrays in place, by kind: "scalar", "vector_d" and
"vector_c" (pairs of names), "corner"; ``sync()`` joins the D-grid winds on
the shared tile edges; ``tracers(list)`` returns refreshed scalar copies.
"""
from __future__ import annotations

import numpy as np

from fv3core_amd.halo import (CENTER, CORNER, X_FACE, Y_FACE, apply_map_numpy, build_map,
                              build_sync_map)

_STAGGERS = {"scalar": (CENTER,), "vector_d": (Y_FACE, X_FACE), "vector_c": (X_FACE, Y_FACE),
             "corner": (CORNER,)}


class HostHalo:
    def __init__(self, state, n: int, n_halo: int = 3):
        self.state = state
        self.n = n
        self.h = n_halo
        self.maps = {}

    def _map(self, kind, n_halo):
        key = (kind, n_halo)
        if key not in self.maps:
            self.maps[key] = build_map(self.n, _STAGGERS[kind], n_halo, self.h)
        return self.maps[key]

    def __call__(self, kind, names, n_halo=3):
        mp = self._map(kind, n_halo)
        if kind.startswith("vector"):
            for a in range(0, len(names), 2):
                u, v = apply_map_numpy([self.state[names[a]], self.state[names[a + 1]]], mp)
                self.state[names[a]][...] = u
                self.state[names[a + 1]][...] = v
        else:
            for name in names:
                self.state[name][...] = apply_map_numpy([self.state[name]], mp)[0]

    def sync(self):
        mp = build_sync_map(self.n, self.h)
        u, v = apply_map_numpy([self.state["u"], self.state["v"]], mp)
        self.state["u"][...] = u
        self.state["v"][...] = v

    def tracers(self, qs):
        mp = self._map("scalar", self.h)
        return [apply_map_numpy([q], mp)[0] for q in qs]


def acoustic_temporaries(shape):
    """The acoustic step's component temporaries as the reference allocates
    them (dyn_core.py:166-207): zeros, with gz and pk3 at the 1e40 fill."""
    tmp = {k: np.zeros(shape) for k in ("ut", "vt", "gz", "zh", "pem", "pkc", "pk3",
                                         "heat_source", "divgd", "crx", "xfx", "cry", "yfx",
                                         "delpc", "ptc")}
    tmp["gz"][:] = 1.0e40
    tmp["pk3"][:] = 1.0e40
    return tmp


def acoustic_state(st0, init):
    """A copy of the ``cube_state`` state plus phis, wsd, ws3 and the temporaries."""
    S = {k: v.copy() for k, v in st0.items()}
    S["phis"] = init["phis"].copy()
    S["wsd"] = init["wsd"].copy()
    S["ws3"] = np.zeros_like(init["wsd"])
    S.update(acoustic_temporaries(st0["delp"].shape))
    return S
