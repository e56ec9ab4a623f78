"""Cubed-sphere geometry: tile orientation, connectivity and metric terms.

Tile orientation follows the reference's global grid construction
(fv3core/grid/global_setup.py:29-223 with mirror.py:150-205): tile 0 is the
equal-distance gnomonic face on +x, and the other five faces are rotations of
it.  In Cartesian coordinates (unit cube before normalisation), with a, b the
face coordinates along i and j:

    tile 0: ( 1,  a,  b)    tile 3: (-1, -b, -a)
    tile 1: (-a,  1,  b)    tile 4: ( b, -1, -a)
    tile 2: (-a, -b,  1)    tile 5: ( b,  a, -1)

(the reference's rotations act in a left-handed frame, RIGHT_HAND_GRID =
False, which flips the sense of its x- and y-axis rotations).  From this
geometry each tile's four neighbours follow: even tiles have east -> t+1 and
south -> t-1 aligned, north -> t+2 and west -> t-2 rotated; odd tiles have
north -> t+1 and west -> t-1 aligned, east -> t+2 and south -> t-2 rotated.
``tests/test_cube_halo.py`` checks the halo map built from these rules
against the geometry.

Metric terms are computed from the corner points of each tile's own face
parametrisation extended into the halo (the reference instead fills halo
metrics from the neighbouring tiles; the interior values agree).
"""
from __future__ import annotations

import math
from typing import Dict

import numpy as np

RADIUS = 6.3712e6
OMEGA = 7.2921e-5


def face_coordinates(n: int, n_halo: int = 3) -> np.ndarray:
    """Gnomonic face coordinate of the n+1 grid lines plus halo lines.

    The tile edges are split into equal great-circle arcs (the equal-distance
    grid, global_setup.py:45-56): c_k = sqrt(2) * tan(-alpha + 2*alpha*k/n)
    with alpha = asin(1/sqrt(3)), so that the face spans [-1, 1].
    """
    alpha = math.asin(3 ** -0.5)
    k = np.arange(-n_halo, n + n_halo + 1, dtype=np.float64)
    return math.sqrt(2.0) * np.tan(-alpha + 2.0 * alpha * k / n)


def _face_xyz(t: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    one = np.ones_like(a)
    if t == 0:
        p = (one, a, b)
    elif t == 1:
        p = (-a, one, b)
    elif t == 2:
        p = (-a, -b, one)
    elif t == 3:
        p = (-one, -b, -a)
    elif t == 4:
        p = (b, -one, -a)
    else:
        p = (b, a, -one)
    v = np.stack(p, axis=-1)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def tile_corners(n: int, n_halo: int = 3) -> np.ndarray:
    """Unit vectors of the grid corners, shape (6, n+1+2h, n+1+2h, 3).

    Index [t, i, j] is the corner at storage index (i, j) of tile t, i.e. the
    south-west corner of cell (i, j).
    """
    c = face_coordinates(n, n_halo)
    a, b = np.meshgrid(c, c, indexing="ij")
    return np.stack([_face_xyz(t, a, b) for t in range(6)])


def neighbour(t: int, edge: str):
    """(neighbour tile, its edge, rotated?) across ``edge`` of tile ``t``."""
    if t % 2 == 0:
        table = {"east": ((t + 1) % 6, "west", False), "south": ((t - 1) % 6, "north", False),
                 "north": ((t + 2) % 6, "west", True), "west": ((t - 2) % 6, "north", True)}
    else:
        table = {"north": ((t + 1) % 6, "south", False), "west": ((t - 1) % 6, "east", False),
                 "east": ((t + 2) % 6, "south", True), "south": ((t - 2) % 6, "east", True)}
    return table[edge]


# ---------------------------------------------------------------------------
# metric terms


def _gc_dist(p, q):
    """Great-circle distance on the unit sphere between unit vectors."""
    d = np.linalg.norm(p - q, axis=-1)
    return 2.0 * np.arcsin(np.clip(d / 2.0, 0.0, 1.0))


def _normalize(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _sph_area(p1, p2, p3, p4):
    """Area of the spherical quadrilateral p1-p2-p3-p4 (unit sphere), by the
    excess of its two triangles (L'Huilier-free formula of Van Oosterom)."""
    def tri(a, b, c):
        num = np.abs(np.einsum("...i,...i->...", a, np.cross(b, c)))
        den = 1.0 + np.einsum("...i,...i->...", a, b) + np.einsum("...i,...i->...", b, c) \
            + np.einsum("...i,...i->...", c, a)
        return 2.0 * np.arctan2(num, den)
    return tri(p1, p2, p3) + tri(p1, p3, p4)


def _angle_cos(corner, a, b):
    """cos of the angle at ``corner`` between the great circles to a and b."""
    ta = _normalize(np.cross(np.cross(corner, a), corner))
    tb = _normalize(np.cross(np.cross(corner, b), corner))
    return np.einsum("...i,...i->...", ta, tb)


def tile_metrics(n: int, n_halo: int = 3, radius: float = RADIUS) -> Dict[str, np.ndarray]:
    """FV3 metric terms of one tile's storage index space for each of the 6 tiles.

    Returns arrays of shape (6, n+7, n+7) (for n_halo = 3) with the reference's
    names: dx, dy (D-grid edge lengths), dxc, dyc (C-grid), dxa, dya (A-grid),
    area, area_c, rarea, rarea_c, sin_sg1..4, cos_sg1..4 (cell-corner angles:
    1 west, 2 south, 3 east, 4 north edge midpoint of each cell), cosa_u,
    cosa_v, sina_u, sina_v, rsin_u, rsin_v, cosa_s, rsin2, cosa, sina, rsina,
    del6_u, del6_v, divg_u, divg_v, fC (Coriolis at corners), f0 (centres),
    plus the corner and centre unit vectors (grid3, agrid3).
    """
    h = n_halo
    ng = n + 1 + 2 * h  # corner points
    nc = n + 2 * h  # cells
    P = tile_corners(n, h)  # (6, ng, ng, 3)
    C = _normalize(P[:, :-1, :-1] + P[:, 1:, :-1] + P[:, 1:, 1:] + P[:, :-1, 1:])  # centres
    shp = (6, ng, ng)

    def pad(a, fill=1.0):
        out = np.full(shp, fill)
        out[:, : a.shape[1], : a.shape[2]] = a
        return out

    m = {}
    # D-grid edge lengths: dx along i on y-interfaces (i, j) -> corners (i,j)-(i+1,j)
    m["dx"] = pad(radius * _gc_dist(P[:, :-1, :], P[:, 1:, :]))
    m["dy"] = pad(radius * _gc_dist(P[:, :, :-1], P[:, :, 1:]))
    # A-grid: distance between the midpoints of opposite edges
    mw = _normalize(P[:, :-1, :-1] + P[:, :-1, 1:])
    me = _normalize(P[:, 1:, :-1] + P[:, 1:, 1:])
    ms = _normalize(P[:, :-1, :-1] + P[:, 1:, :-1])
    mn = _normalize(P[:, :-1, 1:] + P[:, 1:, 1:])
    m["dxa"] = pad(radius * _gc_dist(mw, me))
    m["dya"] = pad(radius * _gc_dist(ms, mn))
    # C-grid: dxc(i, j) between centres (i-1, j) and (i, j)
    dxc = np.ones((6, ng, ng))
    dxc[:, 1:nc, :nc] = radius * _gc_dist(C[:, :-1, :], C[:, 1:, :])
    dxc[:, 0, :] = dxc[:, 1, :]
    dxc[:, nc, :] = dxc[:, nc - 1, :]
    dyc = np.ones((6, ng, ng))
    dyc[:, :nc, 1:nc] = radius * _gc_dist(C[:, :, :-1], C[:, :, 1:])
    dyc[:, :, 0] = dyc[:, :, 1]
    dyc[:, :, nc] = dyc[:, :, nc - 1]
    m["dxc"] = dxc
    m["dyc"] = dyc
    area = radius ** 2 * _sph_area(P[:, :-1, :-1], P[:, 1:, :-1], P[:, 1:, 1:], P[:, :-1, 1:])
    m["area"] = pad(area)
    # dual cells around corners (C-grid area) from neighbouring centres
    area_c = np.full(shp, np.nan)
    area_c[:, 1:nc, 1:nc] = radius ** 2 * _sph_area(C[:, :-1, :-1], C[:, 1:, :-1], C[:, 1:, 1:],
                                                     C[:, :-1, 1:])
    for t in range(6):
        a = area_c[t]
        a[0, :] = a[1, :]
        a[nc, :] = a[nc - 1, :]
        a[:, 0] = a[:, 1]
        a[:, nc] = a[:, nc - 1]
    m["area_c"] = area_c
    # corner angles of each cell: sin_sg1..4 / cos_sg1..4 at the midpoints of
    # the west (1), south (2), east (3) and north (4) edges as FV3 uses them:
    # the cosine of the angle between the two coordinate lines through that
    # edge midpoint.
    cos1 = _angle_cos(mw, _normalize(P[:, :-1, 1:]), C)
    cos2 = _angle_cos(ms, _normalize(P[:, 1:, :-1]), C)
    cos3 = _angle_cos(me, _normalize(P[:, 1:, 1:]), C)
    cos4 = _angle_cos(mn, _normalize(P[:, 1:, 1:]), C)
    for k, cs in zip((1, 2, 3, 4), (cos1, cos2, cos3, cos4)):
        cs = np.clip(cs, -0.999, 0.999)
        m[f"cos_sg{k}"] = pad(cs, 0.0)
        m[f"sin_sg{k}"] = pad(np.sqrt(1.0 - cs ** 2), 1.0)
    # angle at corners (cosa) and on C-grid faces
    cosa = np.zeros(shp)
    cosa[:, 1:-1, 1:-1] = _angle_cos(P[:, 1:-1, 1:-1], P[:, 2:, 1:-1], P[:, 1:-1, 2:])
    cosa = np.clip(cosa, -0.999, 0.999)
    m["cosa"] = cosa
    m["sina"] = np.sqrt(1.0 - cosa ** 2)
    m["cosa_u"] = 0.5 * (m["cos_sg3"] + np.roll(m["cos_sg1"], -1, axis=1))
    m["cosa_v"] = 0.5 * (m["cos_sg4"] + np.roll(m["cos_sg2"], -1, axis=2))
    m["cosa_s"] = 0.5 * (m["cos_sg1"] + m["cos_sg3"])  # centre value
    m["sina_u"] = np.sqrt(1.0 - m["cosa_u"] ** 2)
    m["sina_v"] = np.sqrt(1.0 - m["cosa_v"] ** 2)
    m["rsin_u"] = 1.0 / m["sina_u"] ** 2
    m["rsin_v"] = 1.0 / m["sina_v"] ** 2
    m["rsina"] = 1.0 / m["sina"] ** 2
    m["rsin2"] = 1.0 / (1.0 - m["cosa_s"] ** 2)
    m["rarea"] = 1.0 / m["area"]
    m["rarea_c"] = 1.0 / m["area_c"]
    for r, b in (("rdx", "dx"), ("rdy", "dy"), ("rdxc", "dxc"), ("rdyc", "dyc"),
                 ("rdxa", "dxa"), ("rdya", "dya")):
        m[r] = 1.0 / m[b]
    # damping coefficients (fv3core/grid/geometry.py:514-530)
    m["del6_u"] = m["sina_v"] * m["dx"] / m["dyc"]
    m["del6_v"] = m["sina_u"] * m["dy"] / m["dxc"]
    m["divg_u"] = m["sina_v"] * m["dyc"] / m["dx"]
    m["divg_v"] = m["sina_u"] * m["dxc"] / m["dy"]
    # Coriolis on corners and centres (latitude from z)
    m["fC"] = 2.0 * OMEGA * P[..., 2]
    fz = np.zeros(shp)
    fz[:, :nc, :nc] = 2.0 * OMEGA * C[..., 2]
    m["f0"] = fz
    grid3 = P
    agrid3 = np.zeros(shp + (3,))
    agrid3[:, :nc, :nc] = C
    m["grid3"] = grid3
    m["agrid3"] = agrid3
    # longitude / latitude of corners (bgrid) and centres (agrid)
    m["bgrid1"] = np.arctan2(P[..., 1], P[..., 0])
    m["bgrid2"] = np.arcsin(np.clip(P[..., 2], -1.0, 1.0))
    ag1 = np.zeros(shp)
    ag2 = np.zeros(shp)
    ag1[:, :nc, :nc] = np.arctan2(C[..., 1], C[..., 0])
    ag2[:, :nc, :nc] = np.arcsin(np.clip(C[..., 2], -1.0, 1.0))
    m["agrid1"], m["agrid2"] = ag1, ag2
    # cell-centre transform from the covariant cube winds to east/north
    # winds (fv3core/grid/geometry.py:14-61 get_center_vector, :888-905
    # unit_vector_lonlat, :587-601 calculate_grid_z / calculate_grid_a):
    # ec1 / ec2 are the unit tangents at the centre along the great circles
    # through the west/east and south/north edge midpoints, and
    # a = 0.5 * inverse of [[ec1.vlon, ec1.vlat], [ec2.vlon, ec2.vlat]]
    # with determinant sin_sg5 = |ec1 x ec2|.
    ec1 = _normalize(np.cross(C, np.cross(me, mw)))
    ec2 = _normalize(np.cross(C, np.cross(mn, ms)))
    lon = np.arctan2(C[..., 1], C[..., 0])
    lat = np.arcsin(np.clip(C[..., 2], -1.0, 1.0))
    zero = np.zeros_like(lon)
    vlon = np.stack([-np.sin(lon), np.cos(lon), zero], axis=-1)
    vlat = np.stack([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)], axis=-1)
    z11 = np.sum(ec1 * vlon, axis=-1)
    z12 = np.sum(ec1 * vlat, axis=-1)
    z21 = np.sum(ec2 * vlon, axis=-1)
    z22 = np.sum(ec2 * vlat, axis=-1)
    cos5 = np.sum(ec1 * ec2, axis=-1)
    sin5 = np.sqrt(np.maximum(0.0, 1.0 - cos5 ** 2))
    m["a11"] = pad(0.5 * z22 / sin5, 0.0)
    m["a12"] = pad(-0.5 * z12 / sin5, 0.0)
    m["a21"] = pad(-0.5 * z21 / sin5, 0.0)
    m["a22"] = pad(0.5 * z11 / sin5, 0.0)
    # a2b_ord4 edge weights (fv3core/grid/geometry.py:604-713): at a corner
    # point on a tile edge, the weight of the along-edge neighbour behind it,
    # d2 / (d1 + d2) with d1, d2 the distances to the midpoints of the A-grid
    # pairs straddling the edge on either side.  Stored per IJ point.
    for name in ("edge_w", "edge_e", "edge_s", "edge_n"):
        m[name] = np.full(shp, 0.5)
    for t in range(6):
        for name, ia, ib, along_i in (("edge_w", h - 1, h, False), ("edge_e", h + n, h + n - 1, False),
                                      ("edge_s", h - 1, h, True), ("edge_n", h + n, h + n - 1, True)):
            ip = h if ia < ib else h + n  # corner-point index across the edge
            for k in range(h + 1, h + n):
                if along_i:
                    mid0 = _normalize(C[t, k - 1, ia] + C[t, k - 1, ib])
                    mid1 = _normalize(C[t, k, ia] + C[t, k, ib])
                    pt = P[t, k, ip]
                else:
                    mid0 = _normalize(C[t, ia, k - 1] + C[t, ib, k - 1])
                    mid1 = _normalize(C[t, ia, k] + C[t, ib, k])
                    pt = P[t, ip, k]
                d1 = _gc_dist(mid0, pt)
                d2 = _gc_dist(mid1, pt)
                if along_i:
                    m[name][t, k, :] = d2 / (d1 + d2)
                else:
                    m[name][t, :, k] = d2 / (d1 + d2)
    comp = (slice(None), slice(h, h + n), slice(h, h + n))
    m["da_min"] = float(np.min(m["area"][comp]))
    m["da_max"] = float(np.max(m["area"][comp]))
    m["da_min_c"] = float(np.min(m["area_c"][:, h:h + n + 1, h:h + n + 1]))
    m["da_max_c"] = float(np.max(m["area_c"][:, h:h + n + 1, h:h + n + 1]))
    return m
