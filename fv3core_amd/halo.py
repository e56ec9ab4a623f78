"""Six-tile halo update on one GPU: a gather through a precomputed index map.

The reference exchanges halos through fv3gfs-util's CubedSphereCommunicator
over MPI (call sites in fv3core/stencils/dyn_core.py:216-286, 492-747 and
tracer_2d_1l.py:189-192).  Here all six tiles live in one batched storage
``(6, I, J, K)``.  A halo update is one kernel launch (``fv3_halo_update``)
that, for every halo point of every tile and every level, copies the value
of the neighbouring tile's interior point, swapping and negating vector
components where the tile axes rotate.

The map is built on the host from the cube connectivity in
:mod:`fv3core_amd.cube`.  Halo points are addressed in continuous tile
coordinates X = i + ox, Y = j + oy, where (ox, oy) is the point's stagger
(0.5 for a cell centre, 0 for an interface).  Across an edge, the point at
normal distance d and along-edge position s lands in the neighbour at the
same d, with s kept (aligned edges) or reversed (rotated edges).  The
Jacobian of that map tells which neighbour component, and which sign, a
vector component takes.  Points on the shared edge itself are not
exchanged, and neither are the three-tile corner regions of the halo; both
are the tiles' own business, as in the reference.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .cube import neighbour

# staggers
CENTER = (0.5, 0.5)
X_FACE = (0.0, 0.5)   # X_INTERFACE_DIM, Y_DIM (e.g. v on the D-grid, uc on the C-grid)
Y_FACE = (0.5, 0.0)   # X_DIM, Y_INTERFACE_DIM (e.g. u on the D-grid, vc on the C-grid)
CORNER = (0.0, 0.0)   # X_INTERFACE_DIM, Y_INTERFACE_DIM


def _edge_points(n: int, h: int, edge: str, stagger: Tuple[float, float], n_halo: int):
    """(X, Y) tile coordinates of the halo points beyond ``edge``.

    Compute cells are [h, h + n) in both directions, so the tile spans
    X, Y in [h, h + n].
    """
    ox, oy = stagger
    lo, hi = float(h), float(h + n)
    normal = oy if edge in ("north", "south") else ox
    along = ox if edge in ("north", "south") else oy
    if normal == 0.5:
        dists = [0.5 + m for m in range(n_halo)]
    else:
        dists = [1.0 + m for m in range(n_halo)]
    if along == 0.5:
        svals = [0.5 + s for s in range(n)]
    else:
        svals = [float(s) for s in range(n + 1)]
    pts = []
    for dist in dists:
        for s in svals:
            if edge == "east":
                pts.append((hi + dist, lo + s, dist, s))
            elif edge == "west":
                pts.append((lo - dist, lo + s, dist, s))
            elif edge == "north":
                pts.append((lo + s, hi + dist, dist, s))
            else:
                pts.append((lo + s, lo - dist, dist, s))
    return pts


def _into_neighbour(n: int, h: int, edge_b: str, rotated: bool, dist: float, s: float):
    """Neighbour coordinates of the point at normal distance ``dist`` and
    along-edge position ``s`` beyond the neighbour's edge ``edge_b``."""
    lo, hi = float(h), float(h + n)
    sb = (n - s) if rotated else s
    if edge_b == "west":
        return lo + dist, lo + sb
    if edge_b == "east":
        return hi - dist, lo + sb
    if edge_b == "south":
        return lo + sb, lo + dist
    return lo + sb, hi - dist


def _jacobian(n: int, h: int, edge: str, edge_b: str, rotated: bool):
    """d(Xb, Yb)/d(X, Y) of the map across ``edge`` (constant for each edge)."""
    def tile_ds(X, Y):
        lo, hi = float(h), float(h + n)
        if edge == "east":
            return X - hi, Y - lo
        if edge == "west":
            return lo - X, Y - lo
        if edge == "north":
            return Y - hi, X - lo
        return lo - Y, X - lo

    def f(X, Y):
        dist, s = tile_ds(X, Y)
        return np.array(_into_neighbour(n, h, edge_b, rotated, dist, s))

    X0, Y0 = (h + n + 1.0, h + 1.0) if edge == "east" else (h + 1.0, h + 1.0)
    if edge == "west":
        X0, Y0 = h - 1.0, h + 1.0
    elif edge == "north":
        X0, Y0 = h + 1.0, h + n + 1.0
    elif edge == "south":
        X0, Y0 = h + 1.0, h - 1.0
    jx = f(X0 + 1.0, Y0) - f(X0, Y0)
    jy = f(X0, Y0 + 1.0) - f(X0, Y0)
    return np.stack([jx, jy], axis=1)  # columns: images of A's x and y axes


def build_map(n: int, staggers: Sequence[Tuple[float, float]], n_halo: int = 3, h: int = 3,
              ntiles: int = 6) -> np.ndarray:
    """Index map for one update: int32 array of shape (m, 8).

    Columns: dst_tile, dst_i, dst_j, dst_comp, src_tile, src_i, src_j,
    src_comp * 2 + (sign < 0).  ``staggers`` has one entry per component:
    one for a scalar field, two (x then y component) for a vector pair.
    """
    if ntiles != 6:
        raise ValueError("the cube has six tiles")
    rows: List[Tuple[int, ...]] = []
    vector = len(staggers) == 2
    for t in range(6):
        for edge in ("east", "west", "north", "south"):
            tb, edge_b, rot = neighbour(t, edge)
            jac = _jacobian(n, h, edge, edge_b, rot)
            for comp, stg in enumerate(staggers):
                if vector:
                    # A's axis `comp` maps onto B's axis src_comp with a sign
                    col = jac[:, comp]
                    src_comp = int(np.argmax(np.abs(col)))
                    sign = int(np.sign(col[src_comp]))
                    stg_b = staggers[src_comp]
                else:
                    src_comp, sign = 0, 1
                    stg_b = stg if not rot else (stg[1], stg[0])
                for X, Y, dist, s in _edge_points(n, h, edge, stg, n_halo):
                    Xb, Yb = _into_neighbour(n, h, edge_b, rot, dist, s)
                    ib, jb = Xb - stg_b[0], Yb - stg_b[1]
                    i, j = X - stg[0], Y - stg[1]
                    assert abs(ib - round(ib)) < 1e-9 and abs(jb - round(jb)) < 1e-9, (ib, jb)
                    assert abs(i - round(i)) < 1e-9 and abs(j - round(j)) < 1e-9
                    rows.append((t, int(round(i)), int(round(j)), comp, tb, int(round(ib)),
                                 int(round(jb)), src_comp * 2 + (1 if sign < 0 else 0)))
    out = np.array(rows, dtype=np.int32)
    # write order: tile, then j, then i, so that neighbouring lanes store to
    # neighbouring addresses where the halo runs along i
    order = np.lexsort((out[:, 1], out[:, 2], out[:, 3], out[:, 0]))
    return out[order]


def build_sync_map(n: int, h: int = 3) -> np.ndarray:
    """Index map of synchronize_vector_interfaces for a D-grid pair (u, v).

    On a cube edge the two tiles both compute the wind component lying on
    that edge line: u on a north/south edge, v on an east/west edge.  The
    tile for which the edge is its south or west edge keeps its value and
    the neighbour, for which the same line is its north or east edge, takes
    it over, rotated where the tiles' axes rotate (the Fortran dyn_core
    fills u(:, je+1) and v(ie+1, :) from the neighbours' buffers the same
    way).  Rows as in :func:`build_map`.
    """
    staggers = (Y_FACE, X_FACE)
    rows: List[Tuple[int, ...]] = []
    lo, hi = float(h), float(h + n)
    for t in range(6):
        for edge, comp in (("north", 0), ("east", 1)):
            tb, edge_b, rot = neighbour(t, edge)
            jac = _jacobian(n, h, edge, edge_b, rot)
            col = jac[:, comp]
            src_comp = int(np.argmax(np.abs(col)))
            sign = int(np.sign(col[src_comp]))
            stg, stg_b = staggers[comp], staggers[src_comp]
            for m in range(n):
                sv = 0.5 + m
                X, Y = (lo + sv, hi) if edge == "north" else (hi, lo + sv)
                Xb, Yb = _into_neighbour(n, h, edge_b, rot, 0.0, sv)
                i, j = X - stg[0], Y - stg[1]
                ib, jb = Xb - stg_b[0], Yb - stg_b[1]
                assert abs(ib - round(ib)) < 1e-9 and abs(jb - round(jb)) < 1e-9
                rows.append((t, int(round(i)), int(round(j)), comp, tb, int(round(ib)),
                             int(round(jb)), src_comp * 2 + (1 if sign < 0 else 0)))
    out = np.array(rows, dtype=np.int32)
    order = np.lexsort((out[:, 1], out[:, 2], out[:, 3], out[:, 0]))
    return out[order]


def apply_map_numpy(fields: Sequence[np.ndarray], mp: np.ndarray) -> List[np.ndarray]:
    """Host restatement of the gather, used by the tests as the checker."""
    out = [f.copy() for f in fields]
    for row in mp:
        t, i, j, c, tb, ib, jb, sc = (int(v) for v in row)
        src_comp, neg = sc // 2, sc % 2
        v = fields[src_comp][tb, ib, jb]
        out[c][t, i, j] = -v if neg else v
    return out


_MAX_FIELDS = 16


class _Ptrs(ctypes.Structure):
    _fields_ = [("p", ctypes.c_void_p * _MAX_FIELDS)]


class HaloUpdater:
    """Updates the halos of batched 6-tile storages in one launch per call.

    ``kind`` is "scalar" (cell centres), "vector_d" (u on y-interfaces, v on
    x-interfaces), "vector_c" (uc on x-interfaces, vc on y-interfaces) or
    "corner" (x- and y-interface corners, e.g. divgd).  ``start``/``wait``
    keep the reference's split-phase calls; the work is enqueued on the
    current stream at ``start`` and ``wait`` does nothing further.
    """

    _KINDS: Dict[str, Tuple[Tuple[float, float], ...]] = {
        "scalar": (CENTER,),
        "vector_d": (Y_FACE, X_FACE),
        "vector_c": (X_FACE, Y_FACE),
        "corner": (CORNER,),
        "sync_d": (Y_FACE, X_FACE),
    }
    _cache: Dict[Tuple, torch.Tensor] = {}

    def __init__(self, grid_indexing, kind: str = "scalar", n_halo: int = 3):
        if kind not in self._KINDS:
            raise ValueError(f"unknown halo kind {kind}")
        self._gi = grid_indexing
        self._kind = kind
        self._n_halo = n_halo
        n = grid_indexing.domain[0]
        if grid_indexing.domain[1] != n:
            raise NotImplementedError("cube tiles are square")
        self._n = n

    def _device_map(self, device) -> torch.Tensor:
        key = (self._n, self._kind, self._n_halo, str(device))
        if key not in self._cache:
            if self._kind == "sync_d":
                mp = build_sync_map(self._n, self._gi.n_halo)
            else:
                mp = build_map(self._n, self._KINDS[self._kind], self._n_halo, self._gi.n_halo)
            self._cache[key] = torch.as_tensor(mp, device=device)
        return self._cache[key]

    def update(self, *fields):
        """Fields: scalars, or (x, y) component pairs for vector kinds."""
        ncomp = len(self._KINDS[self._kind])
        if len(fields) == 0 or len(fields) % ncomp:
            raise ValueError(f"{self._kind} update needs fields in groups of {ncomp}")
        if len(fields) > _MAX_FIELDS:
            raise ValueError(f"at most {_MAX_FIELDS} fields per update")
        _lib.require_device(*fields)
        ref = fields[0]
        if ref.dim() != 4 or ref.shape[0] % 6 != 0 or ref.shape[0] == 0:
            raise ValueError("halo updates act on batched (6 * members, I, J, K) storages")
        nk = min(f.shape[-1] for f in fields)
        d = _lib.make_domain(self._gi, ref, nk, others3=fields[1:])
        mp = self._device_map(ref.device)
        ptrs = _Ptrs()
        for a, f in enumerate(fields):
            ptrs.p[a] = f.data_ptr()
        _lib.check(
            _lib.fn("fv3_halo_update", ref.dtype)(
                ctypes.byref(d), ctypes.c_int32(mp.shape[0]), _lib.ptr(mp),
                ctypes.c_int32(len(fields)), ctypes.c_int32(ncomp), ctypes.byref(ptrs),
                ctypes.c_void_p(_lib.stream_handle()),
            )
        )

    def start(self, fields):
        self.update(*fields)

    def wait(self):
        return None


class CubedSphereCommunicator:
    """The part of fv3gfs-util's communicator the dycore uses, for six tiles
    held in one batched storage on one GPU (no MPI: updates are gathers)."""

    def __init__(self, grid_indexing):
        self._gi = grid_indexing

    def get_scalar_halo_updater(self, specs=None, n_halo: int = 3):
        return _GroupUpdater(HaloUpdater(self._gi, "scalar", n_halo))

    def get_vector_halo_updater(self, specs_x=None, specs_y=None, kind: str = "vector_d",
                                n_halo: int = 3):
        return _GroupUpdater(HaloUpdater(self._gi, kind, n_halo))

    def vector_halo_update(self, x_quantity, y_quantity, n_points: int):
        """Blocking halo update of a D-grid wind pair (u on y-faces, v on
        x-faces), ``n_points`` deep."""
        key = ("vector_d", int(n_points))
        if not hasattr(self, "_vector_updaters"):
            self._vector_updaters = {}
        upd = self._vector_updaters.get(key)
        if upd is None:
            upd = self._vector_updaters[key] = HaloUpdater(self._gi, "vector_d", int(n_points))
        upd.update(*_GroupUpdater._storages([x_quantity, y_quantity]))

    def synchronize_vector_interfaces(self, x_quantity, y_quantity):
        """Make the cube-edge values of a D-grid wind pair agree between tiles
        (see :func:`build_sync_map`)."""
        if not hasattr(self, "_sync"):
            self._sync = HaloUpdater(self._gi, "sync_d")
        fields = _GroupUpdater._storages([x_quantity, y_quantity])
        self._sync.update(*fields)


class _GroupUpdater:
    """``update(quantities)`` / ``start(quantities)`` + ``wait()`` on a list."""

    def __init__(self, upd: HaloUpdater):
        self._upd = upd

    @staticmethod
    def _storages(quantities):
        return [q if isinstance(q, torch.Tensor) else q.storage for q in quantities]

    def update(self, quantities, quantities_y=None):
        fields = self._storages(quantities)
        if quantities_y is not None:
            ys = self._storages(quantities_y)
            fields = [f for pair in zip(fields, ys) for f in pair]
        for c in range(0, len(fields), 16):
            self._upd.update(*fields[c:c + 16])

    def start(self, quantities, quantities_y=None):
        self.update(quantities, quantities_y)

    def wait(self):
        return None
