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

The grid with its halo (:func:`grid_points`) and the metric terms
(:func:`tile_metrics`) follow the reference's MetricTerms
(fv3core/grid/generation.py, geometry.py, gnomonic.py): compute-domain terms
from the tile's own points, edge halos from the neighbouring tiles, the
three-tile corner blocks by the reference's corner fills, and the tile-edge
rules of the C-grid lengths and areas, the supergrid angles and the damping
coefficients.  All of it is numpy set-up work done once per resolution.
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
# the grid with its halo

# generation.py:1346-1350: the whole cube is turned by -180/18 degrees of
# longitude (shift_fac = 18), which moves a cube corner off Japan
LON_SHIFT = -math.pi / 18.0
BIG = 1.0e8  # the reference's big_number / tiny_number placeholders
TINY = 1.0e-8


def _turn(v: np.ndarray, angle: float) -> np.ndarray:
    """Rotate unit vectors about the z axis (a longitude shift)."""
    c, s = math.cos(angle), math.sin(angle)
    out = v.copy()
    out[..., 0] = c * v[..., 0] - s * v[..., 1]
    out[..., 1] = s * v[..., 0] + c * v[..., 1]
    return out


def _lonlat(v: np.ndarray):
    return np.arctan2(v[..., 1], v[..., 0]), np.arcsin(np.clip(v[..., 2], -1.0, 1.0))


def _xyz(lon: np.ndarray, lat: np.ndarray) -> np.ndarray:
    return np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], axis=-1)


def _halo_fill(arrays, staggers, n: int, h: int) -> None:
    """Fill the edge halos of six-tile arrays (tile axis first) from the
    neighbouring tiles' compute points, as the reference's halo_update
    (one stagger) and vector_halo_update (an x, y pair) do over the cube.
    A pair (always lengths or length ratios here) is exchanged without signs,
    as the reference takes the absolute value after its vector update.  The three-tile corner blocks
    are left alone: no tile sits across a cube corner."""
    from .halo import build_map

    mp = build_map(n, tuple(staggers), h, h)
    src = [a.copy() for a in arrays]
    for c, a in enumerate(arrays):
        rows = mp[mp[:, 3] == c]
        for s, b in enumerate(src):
            r = rows[rows[:, 7] // 2 == s]
            v = b[r[:, 4], r[:, 5], r[:, 6]]
            a[r[:, 0], r[:, 1], r[:, 2]] = np.abs(v) if len(staggers) == 2 else v


def _fill_bgrid_x(q, n: int, h: int) -> None:
    """fill_corners_2d(gridtype "B", direction "x") (utils/corners.py:726-839)
    on corner-point arrays: each three-tile corner block takes the points of
    the west or east edge halo turned by a quarter."""
    e = h + n  # last corner index of the compute domain
    for i in range(1, h + 1):
        for j in range(1, h + 1):
            q[:, h - i, h - j] = q[:, h - j, h + i]
            q[:, h - i, e + j] = q[:, h - j, e - i]
            q[:, e + i, h - j] = q[:, e + j, h + i]
            q[:, e + i, e + j] = q[:, e + j, e - i]


def _fill_agrid_scalar(q, n: int, h: int, direction: str) -> None:
    """fill_corners_2d(gridtype "A") (utils/corners.py:770-818, 842-852) on
    cell arrays: direction "x" copies from the west/east halo, "y" from the
    south/north halo."""
    ie = h + n - 1  # last compute cell
    for i in range(1, h + 1):
        for j in range(1, h + 1):
            if direction == "x":
                q[:, h - i, h - j] = q[:, h - j, h + i - 1]
                q[:, h - i, ie + j] = q[:, h - j, ie - i + 1]
                q[:, ie + i, h - j] = q[:, ie + j, h + i - 1]
                q[:, ie + i, ie + j] = q[:, ie + j, ie - i + 1]
            else:
                q[:, h - j, h - i] = q[:, h + i - 1, h - j]
                q[:, h - j, ie + i] = q[:, h + i - 1, ie + j]
                q[:, ie + j, h - i] = q[:, ie - i + 1, h - j]
                q[:, ie + j, ie + i] = q[:, ie - i + 1, ie + j]


def _fill_agrid_pair(x, y, n: int, h: int) -> None:
    """fill_corners_agrid(vector=False) (utils/corners.py:855-890)."""
    ie = h + n - 1
    for i in range(1, h + 1):
        for j in range(1, h + 1):
            x[:, h - i, h - j] = y[:, h - j, h - 1 + i]
            y[:, h - j, h - i] = x[:, h - 1 + i, h - j]
            x[:, h - i, ie + j] = y[:, h - j, ie - i + 1]
            y[:, h - j, ie + i] = x[:, h - 1 + i, ie + j]
            x[:, ie + i, h - j] = y[:, ie + j, h - 1 + i]
            y[:, ie + j, h - i] = x[:, ie - i + 1, h - j]
            x[:, ie + i, ie + j] = y[:, ie + j, ie - i + 1]
            y[:, ie + j, ie + i] = x[:, ie - i + 1, ie + j]


def _fill_dgrid_pair(x, y, n: int, h: int) -> None:
    """fill_corners_dgrid(vector=False) (utils/corners.py:893-939): x on
    (cell, interface) points such as dx, y on (interface, cell) such as dy."""
    ie = h + n - 1
    for i in range(1, h + 1):
        for j in range(1, h + 1):
            x[:, h - i, h - j] = y[:, h - j, h - 1 + i]
            y[:, h - i, h - j] = x[:, h - 1 + j, h - i]
            x[:, h - i, ie + 1 + j] = y[:, h - j, ie + 1 - i]
            y[:, h - i, ie + j] = x[:, h - 1 + j, ie + 1 + i]
            x[:, ie + i, h - j] = y[:, ie + 1 + j, h - 1 + i]
            y[:, ie + 1 + i, h - j] = x[:, ie - j + 1, h - i]
            x[:, ie + i, ie + 1 + j] = y[:, ie + 1 + j, ie - i + 1]
            y[:, ie + 1 + i, ie + j] = x[:, ie - j + 1, ie + 1 + i]


def _fill_cgrid_pair(x, y, n: int, h: int) -> None:
    """fill_corners_cgrid(vector=False) (utils/corners.py:942-988): x on
    (interface, cell) points such as dxc, y on (cell, interface) such as dyc."""
    ie = h + n - 1
    for i in range(1, h + 1):
        for j in range(1, h + 1):
            x[:, h - i, h - j] = y[:, h - 1 + j, h - i]
            y[:, h - i, h - j] = x[:, h - j, h - 1 + i]
            x[:, h - i, ie + j] = y[:, h - 1 + j, ie + 1 + i]
            y[:, h - i, ie + 1 + j] = x[:, h - j, ie + 1 - i]
            x[:, ie + 1 + i, h - j] = y[:, ie + 1 - j, h - i]
            y[:, ie + i, h - j] = x[:, ie + 1 + j, h - 1 + i]
            x[:, ie + 1 + i, ie + j] = y[:, ie + 1 - j, ie + 1 + i]
            y[:, ie + i, ie + 1 + j] = x[:, ie + 1 + j, ie + 1 - i]


def grid_points(n: int, n_halo: int = 3) -> np.ndarray:
    """Corner unit vectors of the six tiles with their halos, shape
    (6, n+1+2h, n+1+2h, 3), as MetricTerms._init_dgrid builds them
    (generation.py:1218-1359): each tile's compute points, the longitude
    shift, the edge halos from the neighbouring tiles' points and the
    three-tile corner blocks by the B-grid "x" corner fill."""
    h = n_halo
    P = _turn(tile_corners(n, h), LON_SHIFT)
    from .halo import CORNER

    _halo_fill([P], (CORNER,), n, h)
    _fill_bgrid_x(P, n, h)
    return P


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


def _scos(c, a, b):
    """cos of the angle at ``c`` between the great circles to a and b (the
    reference's spherical_cos, grid/gnomonic.py:692-700)."""
    p = np.cross(c, a)
    q = np.cross(c, b)
    return np.einsum("...i,...i->...", p, q) / np.sqrt(
        np.einsum("...i,...i->...", p, p) * np.einsum("...i,...i->...", q, q))


def _corner_blocks(a, h: int, value) -> None:
    """Set the four three-tile corner blocks of cell arrays (tile axis first)."""
    a[:, :h, :h] = value
    a[:, :h, -h:] = value
    a[:, -h:, :h] = value
    a[:, -h:, -h:] = value


def _supergrid_trig(P, C, ec1, ec2, h: int):
    """cos/sin of the nine supergrid angles of every cell (geometry.py:163-250):
    components 0-3 at the west, south, east and north edge midpoints
    (between the line to the centre and the edge), 4 at the centre (between
    ec1 and ec2), 5-8 at the south-west, south-east, north-east and
    north-west corners; then the edge-halo sines next to the cube corners
    taken from the tile's own cells."""
    sw, se = P[:, :-1, :-1], P[:, 1:, :-1]
    nw, ne = P[:, :-1, 1:], P[:, 1:, 1:]
    cs = np.empty(C.shape[:-1] + (9,))
    cs[..., 5] = _scos(sw, se, nw)
    cs[..., 6] = -_scos(se, sw, ne)
    cs[..., 7] = _scos(ne, se, nw)
    cs[..., 8] = -_scos(nw, sw, ne)
    cs[..., 0] = _scos(_normalize(sw + nw), C, nw)
    cs[..., 1] = _scos(_normalize(sw + se), se, C)
    cs[..., 2] = _scos(_normalize(se + ne), C, se)
    cs[..., 3] = _scos(_normalize(nw + ne), nw, C)
    cs[..., 4] = np.sum(ec1 * ec2, axis=-1)
    cs[np.abs(1.0 - cs) < 1e-15] = 1.0
    sn = np.minimum(np.sqrt(np.maximum(1.0 - cs ** 2, 0.0)), 1.0)
    # geometry.py:226-245 (nc = cells per storage row)
    nc = C.shape[1]
    for k in range(h):
        sn[:, h - 1, k, 2] = sn[:, k, h, 1]
        sn[:, k, h - 1, 3] = sn[:, h, k, 0]
        sn[:, h - 1, nc - h + k, 2] = sn[:, h - 1 - k, nc - h - 1, 3]
        sn[:, k, nc - h, 1] = sn[:, h, nc - h - 2 + k, 0]
        sn[:, nc - h, k, 0] = sn[:, nc - 1 - k, h, 1]
        sn[:, nc - h + k, h - 1, 3] = sn[:, nc - h - 1, h - 1 - k, 2]
        sn[:, nc - h, nc - h + k, 0] = sn[:, nc - h + k, nc - h - 1, 3]
        sn[:, nc - h + k, nc - h, 1] = sn[:, nc - h - 1, nc - h + k, 2]
    return cs, sn


def _supergrid_corner_fix(cs, sn, h: int) -> None:
    """supergrid_corner_fix (geometry.py:430-511): the three-tile corner
    blocks get the placeholders, then the edge-halo angle terms beside each
    cube corner are the tile's own terms turned by a quarter.  Written as
    the south-west rule applied to views flipped so that each corner sits
    south-west, with the edge components renamed to match."""
    _corner_blocks(sn, h, TINY)
    _corner_blocks(cs, h, BIG)
    # (view flips, component of the south-west rule's 1 -> 2 and 0 -> 3 copies)
    for flip, (a_in, a_out, b_in, b_out) in (
            ((slice(None), slice(None)), (1, 2, 0, 3)),
            ((slice(None), slice(None, None, -1)), (3, 2, 0, 1)),
            ((slice(None, None, -1), slice(None)), (1, 0, 2, 3)),
            ((slice(None, None, -1), slice(None, None, -1)), (3, 0, 2, 1))):
        for arr in (sn, cs):
            v = arr[(slice(None),) + flip]
            v[:, h - 1, :h, a_out] = v[:, :h, h, a_in]
            v[:, :h, h - 1, b_out] = v[:, h, :h, b_in]


def tile_metrics(n: int, n_halo: int = 3, radius: float = RADIUS) -> Dict[str, np.ndarray]:
    """FV3 metric terms of the six tiles, halos included, shape (6, n+1+2h, n+1+2h).

    Follows MetricTerms (fv3core/grid/generation.py) step by step: the grid
    with its halo from the neighbouring tiles (:func:`grid_points`); cell
    centres from the corners with the neighbours' centres in the edge halos
    and the A-grid corner fills (longitude "x", latitude "y",
    generation.py:1361-1382); dx, dy, dxa, dya, dxc, dyc, area and area_c on
    the compute domain (the C-grid terms with the tile-border rules of
    gnomonic.py:418-590) and the halos from the neighbours plus the
    reference's corner fill of each; the supergrid angles, cosa/sina and
    their reciprocals (geometry.py:163-511, with the edge rules and the
    1e8 / 1e-8 placeholders where the reference leaves them); divg and del6
    (geometry.py:514-585) with their halos from the neighbours; the
    cell-centre transforms a11..a22 (geometry.py:588-601) and the a2b_ord4
    edge weights (geometry.py:604-715).

    Names are the reference's: dx, dy (D-grid edge lengths), dxc, dyc
    (C-grid), dxa, dya (A-grid), area, area_c and their reciprocals,
    cos_sg1..9 / sin_sg1..9, cosa, sina, rsina, cosa_u, cosa_v, sina_u,
    sina_v, rsin_u, rsin_v, cosa_s, rsin2, del6_u, del6_v, divg_u, divg_v,
    a11..a22, ec1, ec2, vlon, vlat, edge_w/e/s/n, fC (Coriolis at corners),
    f0 (centres), bgrid1/2 and agrid1/2 (longitude, latitude), grid3 and
    agrid3 (unit vectors), da_min/da_max/da_min_c/da_max_c.  Cell arrays
    occupy [:n+2h, :n+2h] of the (n+1+2h)^2 storage; the last row and column
    of a cell array hold 1 for lengths and areas and 0 / 1 for cosines /
    sines (the reference leaves zeros there; no stencil reads them).
    """
    from .halo import CENTER, CORNER, X_FACE, Y_FACE

    h = n_halo
    ng = n + 1 + 2 * h  # corner points per storage row
    nc = ng - 1  # cells per storage row
    e = h + n  # last compute corner index
    shp = (6, ng, ng)
    R2 = radius ** 2
    P = grid_points(n, h)
    blon, blat = _lonlat(P)

    # A grid: centres, the neighbours' centres, the corner fills
    C = _normalize(P[:, :-1, :-1] + P[:, 1:, :-1] + P[:, 1:, 1:] + P[:, :-1, 1:])
    _halo_fill([C], (CENTER,), n, h)
    alon, alat = _lonlat(C)
    _fill_agrid_scalar(alon, n, h, "x")
    _fill_agrid_scalar(alat, n, h, "y")
    C = _xyz(alon, alat)

    def cells(a, fill):
        out = np.full(shp, fill, dtype=np.float64)
        out[:, :a.shape[1], :a.shape[2]] = a
        return out

    m: Dict[str, np.ndarray] = {}
    # D grid (generation.py:1384-1416)
    dx = np.ones(shp)
    dy = np.ones(shp)
    dx[:, h:e, h:e + 1] = radius * _gc_dist(P[:, h:e, h:e + 1], P[:, h + 1:e + 1, h:e + 1])
    dy[:, h:e + 1, h:e] = radius * _gc_dist(P[:, h:e + 1, h:e], P[:, h:e + 1, h + 1:e + 1])
    _halo_fill([dx, dy], (Y_FACE, X_FACE), n, h)
    _fill_dgrid_pair(dx, dy, n, h)
    m["dx"], m["dy"] = dx, dy
    # A grid lengths between opposite edge midpoints (generation.py:1418-1452)
    mw = _normalize(P[:, :, :-1] + P[:, :, 1:])  # west-edge midpoint of cell (i, j) at [i, j]
    ms = _normalize(P[:, :-1, :] + P[:, 1:, :])  # south-edge midpoint at [i, j]
    dxa_t = radius * _gc_dist(mw[:, :-1], mw[:, 1:])
    dya_t = radius * _gc_dist(ms[:, :, :-1], ms[:, :, 1:])
    _fill_agrid_pair(dxa_t, dya_t, n, h)
    dxa, dya = cells(dxa_t, 1.0), cells(dya_t, 1.0)
    _halo_fill([dxa, dya], (CENTER, CENTER), n, h)
    m["dxa"], m["dya"] = dxa, dya
    # C grid (generation.py:1454-1516, gnomonic.py:546-590): between centres,
    # and at the tile edges twice the distance from the edge midpoint to the
    # tile's own centre
    t = radius * _gc_dist(C[:, :-1, :], C[:, 1:, :])
    dxc = np.ones(shp)
    dxc[:, 1:-1, :-1] = t
    dxc[:, 0, :-1] = t[:, 0]
    dxc[:, -1, :-1] = t[:, -1]
    t = radius * _gc_dist(C[:, :, :-1], C[:, :, 1:])
    dyc = np.ones(shp)
    dyc[:, :-1, 1:-1] = t
    dyc[:, :-1, 0] = t[:, :, 0]
    dyc[:, :-1, -1] = t[:, :, -1]
    wmid = _normalize(P[:, :, h:e] + P[:, :, h + 1:e + 1])  # edge midpoints along j
    smid = _normalize(P[:, h:e, :] + P[:, h + 1:e + 1, :])  # edge midpoints along i
    dxc[:, h, h:e] = 2.0 * radius * _gc_dist(wmid[:, h], C[:, h, h:e])
    dxc[:, e, h:e] = 2.0 * radius * _gc_dist(wmid[:, e], C[:, e - 1, h:e])
    dyc[:, h:e, h] = 2.0 * radius * _gc_dist(smid[:, :, h], C[:, h:e, h])
    dyc[:, h:e, e] = 2.0 * radius * _gc_dist(smid[:, :, e], C[:, h:e, e - 1])
    _halo_fill([dxc, dyc], (X_FACE, Y_FACE), n, h)
    _fill_cgrid_pair(dxc, dyc, n, h)
    m["dxc"], m["dyc"] = dxc, dyc
    # cell areas (generation.py:1518-1529): the corner blocks keep the
    # reference's -1e8 placeholder
    area = np.full(shp, -BIG)
    area[:, h:e, h:e] = R2 * _sph_area(P[:, h:e, h:e], P[:, h + 1:e + 1, h:e],
                                       P[:, h + 1:e + 1, h + 1:e + 1], P[:, h:e, h + 1:e + 1])
    area[:, nc, :] = 1.0
    area[:, :, nc] = 1.0
    _halo_fill([area], (CENTER,), n, h)
    m["area"] = area
    # dual-cell areas around the corners (generation.py:1531-1570): four
    # centres inside, twice the tile's own half on the tile edges
    # (gnomonic.py:418-510; the corner points take the last edge written:
    # west, north, east, south)
    area_c = np.zeros(shp)
    area_c[:, h:e + 1, h:e + 1] = R2 * _sph_area(
        C[:, h - 1:e, h - 1:e], C[:, h:e + 1, h - 1:e], C[:, h:e + 1, h:e + 1], C[:, h - 1:e, h:e + 1])
    jj = np.arange(h, e + 1)

    def half(line_lo, line_hi, c_lo, c_hi):
        return 2.0 * R2 * _sph_area(line_lo, c_lo, c_hi, line_hi)

    wm = _normalize(P[:, h, h - 1:e + 1] + P[:, h, h:e + 2])  # west edge segments j-1 .. e
    area_c[:, h, h:e + 1] = half(wm[:, jj - h], wm[:, jj - h + 1], C[:, h, jj - 1], C[:, h, jj])
    nm = _normalize(P[:, h - 1:e + 1, e] + P[:, h:e + 2, e])
    area_c[:, h:e + 1, e] = half(nm[:, jj - h], nm[:, jj - h + 1], C[:, jj - 1, e - 1],
                                 C[:, jj, e - 1])
    em = _normalize(P[:, e, h - 1:e + 1] + P[:, e, h:e + 2])
    area_c[:, e, h:e + 1] = half(em[:, jj - h], em[:, jj - h + 1], C[:, e - 1, jj - 1],
                                 C[:, e - 1, jj])
    sm = _normalize(P[:, h - 1:e + 1, h] + P[:, h:e + 2, h])
    area_c[:, h:e + 1, h] = half(sm[:, jj - h], sm[:, jj - h + 1], C[:, jj - 1, h], C[:, jj, h])
    _halo_fill([area_c], (CORNER,), n, h)
    _fill_bgrid_x(area_c, n, h)
    m["area_c"] = area_c

    # cell-centre unit vectors along the grid lines (geometry.py:14-74)
    Cp = _normalize(P[:, :-1, :-1] + P[:, 1:, :-1] + P[:, :-1, 1:] + P[:, 1:, 1:])
    pw = _normalize(P[:, :-1, :-1] + P[:, :-1, 1:])
    pe = _normalize(P[:, 1:, :-1] + P[:, 1:, 1:])
    ps = _normalize(P[:, :-1, :-1] + P[:, 1:, :-1])
    pn = _normalize(P[:, :-1, 1:] + P[:, 1:, 1:])
    ec1 = _normalize(np.cross(Cp, np.cross(pe, pw)))
    ec2 = _normalize(np.cross(Cp, np.cross(pn, ps)))
    _corner_blocks(ec1, h, BIG)
    _corner_blocks(ec2, h, BIG)

    # supergrid angles and the trigonometric terms (generation.py:1708-1814)
    cs, sn = _supergrid_trig(P, C, ec1, ec2, h)
    cosa = np.full(shp, BIG)
    sina = np.full(shp, BIG)
    cosa[:, h:e + 1, h:e + 1] = 0.5 * (cs[:, h - 1:e, h - 1:e, 7] + cs[:, h:e + 1, h:e + 1, 5])
    sina[:, h:e + 1, h:e + 1] = 0.5 * (sn[:, h - 1:e, h - 1:e, 7] + sn[:, h:e + 1, h:e + 1, 5])
    cosa_u = np.zeros(shp)
    sina_u = np.ones(shp)
    rsin_u = np.ones(shp)
    cosa_u[:, :, :nc] = BIG
    sina_u[:, :, :nc] = BIG
    rsin_u[:, :, :nc] = BIG
    cosa_u[:, 1:-1, :nc] = 0.5 * (cs[:, :-1, :, 2] + cs[:, 1:, :, 0])
    sina_u[:, 1:-1, :nc] = 0.5 * (sn[:, :-1, :, 2] + sn[:, 1:, :, 0])
    rsin_u[:, 1:-1, :nc] = 1.0 / np.maximum(sina_u[:, 1:-1, :nc] ** 2, TINY)
    cosa_v = np.zeros(shp)
    sina_v = np.ones(shp)
    rsin_v = np.ones(shp)
    cosa_v[:, :nc, :] = BIG
    sina_v[:, :nc, :] = BIG
    rsin_v[:, :nc, :] = BIG
    cosa_v[:, :nc, 1:-1] = 0.5 * (cs[:, :, :-1, 3] + cs[:, :, 1:, 1])
    sina_v[:, :nc, 1:-1] = 0.5 * (sn[:, :, :-1, 3] + sn[:, :, 1:, 1])
    rsin_v[:, :nc, 1:-1] = 1.0 / np.maximum(sina_v[:, :nc, 1:-1] ** 2, TINY)
    cosa_s = cs[..., 4].copy()
    _corner_blocks(cosa_s, h, BIG)
    rsin2 = 1.0 / np.maximum(sn[..., 4] ** 2, TINY)
    rsina = np.zeros(shp)
    rsina[:, h:e + 1, h:e + 1] = 1.0 / np.maximum(sina[:, h:e + 1, h:e + 1] ** 2, TINY)
    rsina[:, h, h:e + 1] = BIG
    rsina[:, e, h:e + 1] = BIG
    rsina[:, h:e + 1, h] = BIG
    rsina[:, h:e + 1, e] = BIG

    def edge_recip(s):
        s = np.where(np.abs(s) < TINY, TINY * np.sign(s), s)
        return 1.0 / s

    # on the tile edges the reference keeps 1 / sin, not 1 / sin^2
    rsin_u[:, h, :nc] = edge_recip(sina_u[:, h, :nc])
    rsin_u[:, e, :nc] = edge_recip(sina_u[:, e, :nc])
    rsin_v[:, :nc, h] = edge_recip(sina_v[:, :nc, h])
    rsin_v[:, :nc, e] = edge_recip(sina_v[:, :nc, e])
    _supergrid_corner_fix(cs, sn, h)
    for k in range(9):
        m[f"cos_sg{k + 1}"] = cells(cs[..., k], 0.0)
        m[f"sin_sg{k + 1}"] = cells(sn[..., k], 1.0)
    m["cosa"], m["sina"], m["rsina"] = cosa, sina, rsina
    m["cosa_u"], m["sina_u"], m["rsin_u"] = cosa_u, sina_u, rsin_u
    m["cosa_v"], m["sina_v"], m["rsin_v"] = cosa_v, sina_v, rsin_v
    m["cosa_s"] = cells(cosa_s, 0.0)
    m["rsin2"] = cells(rsin2, 1.0)
    m["rarea"] = 1.0 / area
    m["rarea_c"] = 1.0 / np.where(area_c == 0.0, 1.0, area_c)
    for r, b in (("rdx", "dx"), ("rdy", "dy"), ("rdxc", "dxc"), ("rdyc", "dyc"),
                 ("rdxa", "dxa"), ("rdya", "dya")):
        m[r] = 1.0 / m[b]

    # damping coefficients (geometry.py:514-585, generation.py:1930-1977)
    s1, s2, s3, s4 = (m[f"sin_sg{k}"] for k in (1, 2, 3, 4))
    divg_u = sina_v * dyc / dx
    del6_u = sina_v * dx / dyc
    divg_v = sina_u * dxc / dy
    del6_v = sina_u * dy / dxc
    for jl, jc_in, jc_out in ((h, h, h - 1), (e, e, e - 1)):
        f = 0.5 * (s2[:, :nc, jc_in] + s4[:, :nc, jc_out])
        divg_u[:, :nc, jl] = f * dyc[:, :nc, jl] / dx[:, :nc, jl]
        del6_u[:, :nc, jl] = f * dx[:, :nc, jl] / dyc[:, :nc, jl]
    for il, ic_in, ic_out in ((h, h, h - 1), (e, e, e - 1)):
        f = 0.5 * (s1[:, ic_in, :nc] + s3[:, ic_out, :nc])
        divg_v[:, il, :nc] = f * dxc[:, il, :nc] / dy[:, il, :nc]
        del6_v[:, il, :nc] = f * dy[:, il, :nc] / dxc[:, il, :nc]
    _halo_fill([divg_v, divg_u], (X_FACE, Y_FACE), n, h)
    _halo_fill([del6_v, del6_u], (X_FACE, Y_FACE), n, h)
    m["divg_u"], m["divg_v"], m["del6_u"], m["del6_v"] = divg_u, divg_v, del6_u, del6_v

    # Coriolis on corners and centres, longitudes and latitudes, unit vectors
    m["fC"] = 2.0 * OMEGA * np.sin(blat)
    m["f0"] = cells(2.0 * OMEGA * np.sin(alat), 0.0)
    m["grid3"] = P
    m["agrid3"] = np.zeros(shp + (3,))
    m["agrid3"][:, :nc, :nc] = C
    m["bgrid1"], m["bgrid2"] = blon, blat
    m["agrid1"], m["agrid2"] = cells(alon, 0.0), cells(alat, 0.0)
    m["ec1"] = np.zeros(shp + (3,))
    m["ec2"] = np.zeros(shp + (3,))
    m["ec1"][:, :nc, :nc] = ec1
    m["ec2"][:, :nc, :nc] = ec2
    # cell-centre transform from the covariant cube winds to east/north
    # winds (geometry.py:588-601, :888-905): a = 0.5 * inverse of
    # [[ec1.vlon, ec1.vlat], [ec2.vlon, ec2.vlat]] over sin_sg5
    zero = np.zeros_like(alon)
    vlon = np.stack([-np.sin(alon), np.cos(alon), zero], axis=-1)
    vlat = np.stack([-np.sin(alat) * np.cos(alon), -np.sin(alat) * np.sin(alon), np.cos(alat)],
                    axis=-1)
    m["vlon"] = np.zeros(shp + (3,))
    m["vlat"] = np.zeros(shp + (3,))
    m["vlon"][:, :nc, :nc] = vlon
    m["vlat"][:, :nc, :nc] = vlat
    z11 = np.sum(ec1 * vlon, axis=-1)
    z12 = np.sum(ec1 * vlat, axis=-1)
    z21 = np.sum(ec2 * vlon, axis=-1)
    z22 = np.sum(ec2 * vlat, axis=-1)
    s5 = sn[..., 4]
    m["a11"] = cells(0.5 * z22 / s5, 0.0)
    m["a12"] = cells(-0.5 * z12 / s5, 0.0)
    m["a21"] = cells(-0.5 * z21 / s5, 0.0)
    m["a22"] = cells(0.5 * z11 / s5, 0.0)
    # a2b_ord4 edge weights (geometry.py:604-715): at a corner point on a
    # tile edge, the weight of the along-edge neighbour behind it,
    # d2 / (d1 + d2) with d1, d2 the distances to the midpoints of the
    # A-grid pairs straddling the edge on either side.  Stored per IJ point.
    for name in ("edge_w", "edge_e", "edge_s", "edge_n"):
        m[name] = np.full(shp, 0.5)
    k = np.arange(h + 1, e)
    for name, ia, ib, along_i in (("edge_w", h - 1, h, False), ("edge_e", e, e - 1, False),
                                  ("edge_s", h - 1, h, True), ("edge_n", e, e - 1, True)):
        ip = h if ia < ib else e  # corner-point index across the edge
        if along_i:
            mid0 = _normalize(C[:, k - 1, ia] + C[:, k - 1, ib])
            mid1 = _normalize(C[:, k, ia] + C[:, k, ib])
            pt = P[:, k, ip]
        else:
            mid0 = _normalize(C[:, ia, k - 1] + C[:, ib, k - 1])
            mid1 = _normalize(C[:, ia, k] + C[:, ib, k])
            pt = P[:, ip, k]
        d1 = _gc_dist(mid0, pt)
        d2 = _gc_dist(mid1, pt)
        w = d2 / (d1 + d2)  # (6, n - 1)
        if along_i:
            m[name][:, k, :] = w[:, :, None]
        else:
            m[name][:, :, k] = w[:, None, :]
    comp = (slice(None), slice(h, e), slice(h, e))
    m["da_min"] = float(np.min(area[comp]))
    m["da_max"] = float(np.max(area[comp]))
    m["da_min_c"] = float(np.min(area_c[:, h:e + 1, h:e + 1]))
    m["da_max_c"] = float(np.max(area_c[:, h:e + 1, h:e + 1]))
    return m
