"""CPU oracle for PPM transport: xppm/yppm, CopyCornersXY, fv_tp_2d, DelnFlux.

Test infrastructure only (see oracle/common.py).  Every function follows the
reference statement by statement; the citations are to
``fv3core/stencils/*.py`` and ``fv3core/utils/corners.py`` of the reference.

Fields are numpy arrays indexed ``[i, j, k]`` in the reference's storage
index space (compute origin ``(n_halo, n_halo, 0)``).  IJ metric fields are
``[i, j]``.  Tile-edge anchors are absolute storage indices, ``None`` when the
tile has no edge there (the reference's int16 sentinel).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from .common import box, ij3, imask, restrict, sh, sign, transpose_ij

# fv3core/stencils/ppm.py:8-18
C1 = -2.0 / 14.0
C2 = 11.0 / 14.0
C3 = 5.0 / 14.0
P1 = 7.0 / 12.0
P2 = -1.0 / 12.0
S11 = 11.0 / 14.0
S14 = 4.0 / 7.0
S15 = 3.0 / 14.0


@dataclass(frozen=True)
class Tile:
    """Index facts of one tile: compute domain and edge flags."""

    nx: int
    ny: int
    n_halo: int = 3
    south: bool = True
    north: bool = True
    west: bool = True
    east: bool = True

    @property
    def isc(self):
        return self.n_halo

    @property
    def jsc(self):
        return self.n_halo

    @property
    def iec(self):
        return self.n_halo + self.nx - 1

    @property
    def jec(self):
        return self.n_halo + self.ny - 1

    @property
    def isd(self):
        return 0

    @property
    def jsd(self):
        return 0

    @property
    def ied(self):
        return self.iec + self.n_halo

    @property
    def jed(self):
        return self.jec + self.n_halo

    @property
    def i_start(self):
        return self.isc if self.west else None

    @property
    def i_end(self):
        return self.iec if self.east else None

    @property
    def j_start(self):
        return self.jsc if self.south else None

    @property
    def j_end(self):
        return self.jec if self.north else None

    def transposed(self) -> "Tile":
        return Tile(self.ny, self.nx, self.n_halo, self.west, self.east, self.south, self.north)


def _add(a, b):
    return None if a is None else a + b


# ---------------------------------------------------------------------------
# 1-D PPM in x (xppm.py).  yppm is the same code with i and j exchanged
# (the reference generates yppm.py from xppm.py by substitution,
# convert_xppm_yppm.sh:9-57), so yppm here is xppm on transposed arrays.
# ---------------------------------------------------------------------------


def _pert_ppm_standard_constraint(a0, al, ar):
    """ppm.py:21-34."""
    da1 = al - ar
    da2 = da1 * da1
    a6da = 3.0 * (al + ar) * da1
    neg = al * ar < 0.0
    ar_new = np.where(neg & (a6da < -da2), -2.0 * al, ar)
    al_new = np.where(neg & ~(a6da < -da2) & (a6da > da2), -2.0 * ar, al)
    al_new = np.where(neg, al_new, 0.0)
    ar_new = np.where(neg, ar_new, 0.0)
    return al_new, ar_new


def _compute_al(q, dxa, i_start, i_end):
    """xppm.py:147-181 (hord < 8)."""
    al = P1 * (sh(q, -1) + q) + P2 * (sh(q, -2) + sh(q, 1))
    s = q.shape
    m = imask(s, _add(i_start, -1)) | imask(s, i_end)
    al = np.where(m, C1 * sh(q, -2) + C2 * sh(q, -1) + C3 * q, al)
    m = imask(s, i_start) | imask(s, _add(i_end, 1))
    edge = 0.5 * (
        ((2.0 * sh(dxa, -1) + sh(dxa, -2)) * sh(q, -1) - sh(dxa, -1) * sh(q, -2))
        / (sh(dxa, -2) + sh(dxa, -1))
        + ((2.0 * dxa + sh(dxa, 1)) * q - dxa * sh(q, 1)) / (dxa + sh(dxa, 1))
    )
    al = np.where(m, edge, al)
    m = imask(s, _add(i_start, 1)) | imask(s, _add(i_end, 2))
    al = np.where(m, C3 * sh(q, -1) + C2 * q + C1 * sh(q, 1), al)
    return al


def _xt_dxa_edge_0(q, dxa):
    """xppm.py:105-113, 125-133 (xt_minmax is always True)."""
    xt = 0.5 * (
        ((2.0 * dxa + sh(dxa, -1)) * q - dxa * sh(q, -1)) / (sh(dxa, -1) + dxa)
        + ((2.0 * sh(dxa, 1) + sh(dxa, 2)) * sh(q, 1) - sh(dxa, 1) * sh(q, 2))
        / (sh(dxa, 1) + sh(dxa, 2))
    )
    minq = np.minimum(np.minimum(np.minimum(sh(q, -1), q), sh(q, 1)), sh(q, 2))
    maxq = np.maximum(np.maximum(np.maximum(sh(q, -1), q), sh(q, 1)), sh(q, 2))
    return np.minimum(np.maximum(xt, minq), maxq)


def _xt_dxa_edge_1(q, dxa):
    """xppm.py:115-122, 136-144."""
    xt = 0.5 * (
        ((2.0 * sh(dxa, -1) + sh(dxa, -2)) * sh(q, -1) - sh(dxa, -1) * sh(q, -2))
        / (sh(dxa, -2) + sh(dxa, -1))
        + ((2.0 * dxa + sh(dxa, 1)) * q - dxa * sh(q, 1)) / (dxa + sh(dxa, 1))
    )
    minq = np.minimum(np.minimum(np.minimum(sh(q, -2), sh(q, -1)), q), sh(q, 1))
    maxq = np.maximum(np.maximum(np.maximum(sh(q, -2), sh(q, -1)), q), sh(q, 1))
    return np.minimum(np.maximum(xt, minq), maxq)


def _dm_ord8(q):
    """xppm.py:84-89 (dm_iord8plus)."""
    xt = 0.25 * (sh(q, 1) - sh(q, -1))
    dqr = np.maximum(np.maximum(q, sh(q, -1)), sh(q, 1)) - q
    dql = q - np.minimum(np.minimum(q, sh(q, -1)), sh(q, 1))
    return sign(np.minimum(np.minimum(np.abs(xt), dqr), dql), xt)


def _dm_left(q):
    """dm of cell i-1 as written inline in bl_br_edges (xppm.py:193-197)."""
    xt = 0.25 * (q - sh(q, -2))
    dqr = np.maximum(np.maximum(sh(q, -1), sh(q, -2)), q) - sh(q, -1)
    dql = sh(q, -1) - np.minimum(np.minimum(sh(q, -1), sh(q, -2)), q)
    return sign(np.minimum(np.minimum(np.abs(xt), dqr), dql), xt)


def _dm_right(q):
    """dm of cell i+1 as written inline in bl_br_edges (xppm.py:205-209)."""
    xt = 0.25 * (sh(q, 2) - q)
    dqr = np.maximum(np.maximum(sh(q, 1), q), sh(q, 2)) - sh(q, 1)
    dql = sh(q, 1) - np.minimum(np.minimum(sh(q, 1), q), sh(q, 2))
    return sign(np.minimum(np.minimum(np.abs(xt), dqr), dql), xt)


def _blbr_ord8(q, dxa, i_start, i_end):
    """xppm.py:92-102, 184-266 (compute_blbr_ord8plus + bl_br_edges)."""
    s = q.shape
    dm = _dm_ord8(q)
    al = 0.5 * (sh(q, -1) + q) + 1.0 / 3.0 * (sh(dm, -1) - dm)
    xt = 2.0 * dm
    bl = -1.0 * sign(np.minimum(np.abs(xt), np.abs(al - q)), xt)
    br = sign(np.minimum(np.abs(xt), np.abs(sh(al, 1) - q)), xt)
    if i_start is None and i_end is None:
        return bl, br
    al_ip1 = sh(al, 1)
    xt_bl = np.full(s, np.nan)
    xt_br = np.full(s, np.nan)
    dm_l = _dm_left(q)
    dm_r = _dm_right(q)
    # region[i_start - 1, :]
    m = imask(s, _add(i_start, -1))
    xt_bl = np.where(m, S14 * dm_l + S11 * (sh(q, -1) - q) + q, xt_bl)
    xt_br = np.where(m, _xt_dxa_edge_0(q, dxa), xt_br)
    # region[i_start, :]
    m = imask(s, i_start)
    xt_bl = np.where(m, _xt_dxa_edge_1(q, dxa), xt_bl)
    xt_br = np.where(m, S15 * q + S11 * sh(q, 1) - S14 * dm_r, xt_br)
    # region[i_start + 1, :]
    m = imask(s, _add(i_start, 1))
    xt_bl = np.where(m, S15 * sh(q, -1) + S11 * q - S14 * dm, xt_bl)
    xt_br = np.where(m, al_ip1, xt_br)
    # region[i_end - 1, :]
    m = imask(s, _add(i_end, -1))
    xt_bl = np.where(m, al, xt_bl)
    xt_br = np.where(m, S15 * sh(q, 1) + S11 * q + S14 * dm, xt_br)
    # region[i_end, :]
    m = imask(s, i_end)
    xt_bl = np.where(m, S15 * q + S11 * sh(q, -1) + S14 * dm_l, xt_bl)
    xt_br = np.where(m, _xt_dxa_edge_0(q, dxa), xt_br)
    # region[i_end + 1, :]
    m = imask(s, _add(i_end, 1))
    xt_bl = np.where(m, _xt_dxa_edge_1(q, dxa), xt_bl)
    xt_br = np.where(m, S11 * (sh(q, 1) - q) - S14 * dm_r + q, xt_br)
    edge = np.zeros(s, dtype=bool)
    for off in (-1, 0, 1):
        edge |= imask(s, _add(i_start, off)) | imask(s, _add(i_end, off))
    bl = np.where(edge, xt_bl - q, bl)
    br = np.where(edge, xt_br - q, br)
    bl_c, br_c = _pert_ppm_standard_constraint(q, bl, br)
    bl = np.where(edge, bl_c, bl)
    br = np.where(edge, br_c, br)
    return bl, br


def _fx1(c, bl, br):
    """xppm.py:31-44 (fx1_fn) combined with the upwind choice of apply_flux."""
    b0 = bl + br
    pos = (1.0 - c) * (sh(br, -1) - c * sh(b0, -1))
    neg = (1.0 + c) * (bl + c * b0)
    return np.where(c > 0.0, pos, neg)


def xppm(q, c, dxa, iord: int, i_start=None, i_end=None):
    """Mean q advected through x-interfaces (xppm.py:269-280).

    ``q``, ``c``: (I, J, K); ``dxa``: (I, J).  Returns an (I, J, K) array whose
    entry ``[i, j, k]`` belongs to the x-interface between cells ``i-1`` and
    ``i``.  Entries whose stencil leaves the array are NaN.
    """
    dxa = ij3(dxa)
    mord = abs(iord)
    if mord < 8:
        if iord < 0:
            raise NotImplementedError("negative iord is rejected by the reference")
        al = _compute_al(q, dxa, i_start, i_end)
        bl = al - q
        br = sh(al, 1) - q
        b0 = bl + br
        if mord == 5:
            smt5 = bl * br < 0
        else:
            smt5 = (3.0 * np.abs(b0)) < np.abs(bl - br)
        smt5_m1 = np.zeros_like(smt5)
        smt5_m1[1:] = smt5[:-1]
        mask = np.where(smt5_m1 | smt5, 1.0, 0.0)
        fx1 = _fx1(c, bl, br)
        return np.where(c > 0.0, sh(q, -1) + fx1 * mask, q + fx1 * mask)
    if iord != 8:
        raise NotImplementedError("hord > 8 other than 8 is rejected (xppm.py:256)")
    bl, br = _blbr_ord8(q, dxa, i_start, i_end)
    fx1 = _fx1(c, bl, br)
    return np.where(c > 0.0, sh(q, -1) + fx1, q + fx1)


def yppm(q, c, dya, jord: int, j_start=None, j_end=None):
    """Mean q advected through y-interfaces (yppm.py; transposed xppm)."""
    out = xppm(transpose_ij(q), transpose_ij(c), transpose_ij(dya), jord, j_start, j_end)
    return transpose_ij(out)


# ---------------------------------------------------------------------------
# Corner copies (fv3core/utils/corners.py:305-540).  The table below is the
# reference's list of (destination regions, source offset) pairs.  Each row:
# ((dst anchor i, dst anchor j), (dst anchor i, dst anchor j), (di, dj)),
# anchors written as ("s"|"e", offset) meaning i_start/i_end + offset.
# ---------------------------------------------------------------------------

_X_TABLE = [
    ((("s", -3), ("s", -3)), (("e", 3), ("s", -3)), (0, 5)),
    ((("s", -2), ("s", -3)), (("e", 3), ("s", -2)), (-1, 4)),
    ((("s", -1), ("s", -3)), (("e", 3), ("s", -1)), (-2, 3)),
    ((("s", -3), ("s", -2)), (("e", 2), ("s", -3)), (1, 4)),
    ((("s", -2), ("s", -2)), (("e", 2), ("s", -2)), (0, 3)),
    ((("s", -1), ("s", -2)), (("e", 2), ("s", -1)), (-1, 2)),
    ((("s", -3), ("s", -1)), (("e", 1), ("s", -3)), (2, 3)),
    ((("s", -2), ("s", -1)), (("e", 1), ("s", -2)), (1, 2)),
    ((("s", -1), ("s", -1)), (("e", 1), ("s", -1)), (0, 1)),
    ((("s", -3), ("e", 1)), (("e", 1), ("e", 3)), (2, -3)),
    ((("s", -2), ("e", 1)), (("e", 1), ("e", 2)), (1, -2)),
    ((("s", -1), ("e", 1)), (("e", 1), ("e", 1)), (0, -1)),
    ((("s", -3), ("e", 2)), (("e", 2), ("e", 3)), (1, -4)),
    ((("s", -2), ("e", 2)), (("e", 2), ("e", 2)), (0, -3)),
    ((("s", -1), ("e", 2)), (("e", 2), ("e", 1)), (-1, -2)),
    ((("s", -3), ("e", 3)), (("e", 3), ("e", 3)), (0, -5)),
    ((("s", -2), ("e", 3)), (("e", 3), ("e", 2)), (-1, -4)),
    ((("s", -1), ("e", 3)), (("e", 3), ("e", 1)), (-2, -3)),
]

_Y_TABLE = [
    ((("s", -3), ("s", -3)), (("s", -3), ("e", 3)), (5, 0)),
    ((("s", -2), ("s", -3)), (("s", -3), ("e", 2)), (4, 1)),
    ((("s", -1), ("s", -3)), (("s", -3), ("e", 1)), (3, 2)),
    ((("s", -3), ("s", -2)), (("s", -2), ("e", 3)), (4, -1)),
    ((("s", -2), ("s", -2)), (("s", -2), ("e", 2)), (3, 0)),
    ((("s", -1), ("s", -2)), (("s", -2), ("e", 1)), (2, 1)),
    ((("s", -3), ("s", -1)), (("s", -1), ("e", 3)), (3, -2)),
    ((("s", -2), ("s", -1)), (("s", -1), ("e", 2)), (2, -1)),
    ((("s", -1), ("s", -1)), (("s", -1), ("e", 1)), (1, 0)),
    ((("e", 1), ("s", -3)), (("e", 3), ("e", 1)), (-3, 2)),
    ((("e", 2), ("s", -3)), (("e", 3), ("e", 2)), (-4, 1)),
    ((("e", 3), ("s", -3)), (("e", 3), ("e", 3)), (-5, 0)),
    ((("e", 1), ("s", -2)), (("e", 2), ("e", 1)), (-2, 1)),
    ((("e", 2), ("s", -2)), (("e", 2), ("e", 2)), (-3, 0)),
    ((("e", 3), ("s", -2)), (("e", 2), ("e", 3)), (-4, -1)),
    ((("e", 1), ("s", -1)), (("e", 1), ("e", 1)), (-1, 0)),
    ((("e", 2), ("s", -1)), (("e", 1), ("e", 2)), (-2, -1)),
    ((("e", 3), ("s", -1)), (("e", 1), ("e", 3)), (-3, -2)),
]


def _anchor(tile: Tile, axis: str, a):
    kind, off = a
    if axis == "i":
        base = tile.i_start if kind == "s" else tile.i_end
    else:
        base = tile.j_start if kind == "s" else tile.j_end
    return None if base is None else base + off


def _apply_corner_table(q_in, q_out, tile: Tile, table):
    """Scalar loop over the region table: q_out[dst] = q_in[dst + offset]."""
    for r1, r2, (di, dj) in table:
        for (ai, aj) in (r1, r2):
            i = _anchor(tile, "i", ai)
            j = _anchor(tile, "j", aj)
            if i is None or j is None:
                continue
            q_out[i, j, ...] = q_in[i + di, j + dj, ...]
    return q_out


def copy_corners_x(q: np.ndarray, tile: Tile) -> np.ndarray:
    """copy_corners_x_stencil_defn (corners.py:305-362): returns a new array."""
    return _apply_corner_table(q, q.copy(), tile, _X_TABLE)


def copy_corners_y(q: np.ndarray, tile: Tile) -> np.ndarray:
    """copy_corners_y_stencil_defn (corners.py:365-422): returns a new array."""
    return _apply_corner_table(q, q.copy(), tile, _Y_TABLE)


def copy_corners_xy(q: np.ndarray, tile: Tile):
    """copy_corners_xy_stencil_defn (corners.py:425-540) -> (q_x, q_y)."""
    return copy_corners_x(q, tile), copy_corners_y(q, tile)


# ---------------------------------------------------------------------------
# fv_tp_2d (fvtp2d.py:151-362)
# ---------------------------------------------------------------------------


def fvtp2d(
    q_x,
    q_y,
    crx,
    cry,
    x_area_flux,
    y_area_flux,
    area,
    dxa,
    dya,
    hord: int,
    tile: Tile,
    x_mass_flux=None,
    y_mass_flux=None,
    check_domains: bool = True,
):
    """Fluxes of q through x- and y-interfaces (fvtp2d.py:249-360, no damping).

    Returns ``(fx, fy)`` arrays with NaN outside their valid domains:
    fx on i in [isc, iec+1], j in [jsc, jec]; fy on i in [isc, iec], j in
    [jsc, jec+1]; all k of the inputs.
    """
    t = tile
    ord_outer = hord
    ord_inner = 8 if hord == 10 else hord
    area3 = ij3(area)
    # y-inner: origin compute + (-n_halo, 0), domain compute + (1 + 2h, 1)
    ymean = yppm(q_y, cry, dya, ord_inner, t.j_start, t.j_end)
    if check_domains:
        ymean = restrict(ymean, t.isc - t.n_halo, t.iec + 1 + t.n_halo, t.jsc, t.jec + 1)
    # q_i_stencil (fvtp2d.py:33-47): origin full + (0, 3), domain full + (0, -3)
    fyy = y_area_flux * ymean
    q_adv_y = (q_y * area3 + fyy - sh(fyy, 0, 1)) / (
        area3 + y_area_flux - sh(y_area_flux, 0, 1)
    )
    if check_domains:
        q_adv_y = restrict(q_adv_y, t.isd, t.ied, t.jsd + 3, t.jed)
    # x-outer
    xouter = xppm(q_adv_y, crx, dxa, ord_outer, t.i_start, t.i_end)
    # x-inner: origin compute + (0, -h), domain compute + (1, 1 + 2h)
    xmean = xppm(q_x, crx, dxa, ord_inner, t.i_start, t.i_end)
    if check_domains:
        xmean = restrict(xmean, t.isc, t.iec + 1, t.jsc - t.n_halo, t.jec + 1 + t.n_halo)
    # q_j_stencil (fvtp2d.py:50-60)
    fx1 = x_area_flux * xmean
    area_with_x_flux = area3 + x_area_flux - sh(x_area_flux, 1, 0)
    q_adv_x = (q_x * area3 + fx1 - sh(fx1, 1, 0)) / area_with_x_flux
    if check_domains:
        q_adv_x = restrict(q_adv_x, t.isd + 3, t.ied, t.jsd, t.jed)
    youter = yppm(q_adv_x, cry, dya, ord_outer, t.j_start, t.j_end)
    x_unit = x_area_flux if x_mass_flux is None else x_mass_flux
    y_unit = y_area_flux if y_mass_flux is None else y_mass_flux
    # final_fluxes (fvtp2d.py:63-92)
    fx = 0.5 * (xouter + xmean) * x_unit
    fy = 0.5 * (youter + ymean) * y_unit
    fx = restrict(fx, t.isc, t.iec + 1, t.jsc, t.jec)
    fy = restrict(fy, t.isc, t.iec, t.jsc, t.jec + 1)
    return fx, fy


# ---------------------------------------------------------------------------
# DelnFlux / DelnFluxNoSG (delnflux.py:901-1207)
# ---------------------------------------------------------------------------


def k_band(nk: int) -> np.ndarray:
    """Band index of each level: levels 0, 1, 2 and 3.. (delnflux.py intervals)."""
    return np.minimum(np.arange(nk), 3)


def damp_coefficient(nord_col, damp_c_col, da_min: float) -> np.ndarray:
    """calc_damp (delnflux.py:24-26): (damp_c * da_min) ** (nord + 1)."""
    return (np.asarray(damp_c_col, dtype=np.float64) * da_min) ** (
        np.asarray(nord_col, dtype=np.float64) + 1
    )


def _corner_copy_bands(q, tile: Tile, table, nord4):
    """copy_corners_{x,y}_nord: copy only on levels whose band has nord > 0."""
    out = q.copy()
    band = k_band(q.shape[2])
    active = np.array([nord4[b] > 0 for b in band])
    if active.any():
        copied = _apply_corner_table(q, q.copy(), tile, table)
        out[:, :, active] = copied[:, :, active]
    return out


def delnflux_nosg(q, del6_u, del6_v, rarea, nord4: Sequence[int], damp, tile: Tile,
                  mass=None):
    """DelnFluxNoSG.__call__ (delnflux.py:1156-1207) -> (fx2, fy2).

    ``nord4`` are the four band values nord[0..3]; ``damp`` is a per-level
    column.  Returns arrays with NaN outside the domains the reference writes.
    """
    t = tile
    nk = q.shape[2]
    nord4 = [int(v) for v in nord4]
    if max(nord4) > 3:
        raise ValueError("nord must be less than 3")
    nmax = max(nord4)
    band = k_band(nk)
    bandmax = np.array([nord4[b] for b in band])  # nord per level
    pos = bandmax > 0
    shape = q.shape
    damp = np.asarray(damp, dtype=np.float64)[:nk]
    # preamble: d2 on [isc-1-nmax, iec+1+nmax]; nord==0 bands only [isc-1, iec+1]
    full_box = box(shape[:2], t.isc - 1 - nmax, t.iec + 1 + nmax, t.jsc - 1 - nmax, t.jec + 1 + nmax)
    small_box = box(shape[:2], t.isc - 1, t.iec + 1, t.jsc - 1, t.jec + 1)
    mask = np.where(pos[None, None, :], full_box[:, :, None], small_box[:, :, None])
    if mass is None:
        val = damp[None, None, :] * q
    else:
        val = q.copy()
    d2 = np.where(mask, val, np.nan)
    d2 = _corner_copy_bands(d2, t, _X_TABLE, nord4)
    # fx: origin (isc-nmax, jsc-nmax), domain (nx+1+2nmax, ny+2nmax); nord==0
    # bands only [isc, iec+1] x [jsc, jec]
    fx_full = box(shape[:2], t.isc - nmax, t.iec + 1 + nmax, t.jsc - nmax, t.jec + nmax)
    fx_small = box(shape[:2], t.isc, t.iec + 1, t.jsc, t.jec)
    fxm = np.where(pos[None, None, :], fx_full[:, :, None], fx_small[:, :, None])
    fx2 = np.where(fxm, ij3(del6_v) * (sh(d2, -1, 0) - d2), np.nan)
    d2 = _corner_copy_bands(d2, t, _Y_TABLE, nord4)
    fy_full = box(shape[:2], t.isc - nmax, t.iec + nmax, t.jsc - nmax, t.jec + 1 + nmax)
    fy_small = box(shape[:2], t.isc, t.iec, t.jsc, t.jec + 1)
    fym = np.where(pos[None, None, :], fy_full[:, :, None], fy_small[:, :, None])
    fy2 = np.where(fym, ij3(del6_u) * (sh(d2, 0, -1) - d2), np.nan)
    for n in range(nmax):
        nt = nmax - 1 - n
        dbox = box(shape[:2], t.isc - nt - 1, t.iec + nt + 1, t.jsc - nt - 1, t.jec + nt + 1)
        m3 = dbox[:, :, None] & pos[None, None, :]
        d2new = (fx2 - sh(fx2, 1, 0) + fy2 - sh(fy2, 0, 1)) * ij3(rarea)
        d2 = np.where(m3, d2new, d2)
        d2 = _corner_copy_bands(d2, t, _X_TABLE, nord4)
        fbox = box(shape[:2], t.isc - nt, t.iec + 1 + nt, t.jsc - nt, t.jec + nt)
        m3 = fbox[:, :, None] & pos[None, None, :]
        fx2 = np.where(m3, -ij3(del6_v) * (sh(d2, -1, 0) - d2), fx2)
        d2 = _corner_copy_bands(d2, t, _Y_TABLE, nord4)
        fbox = box(shape[:2], t.isc - nt, t.iec + nt, t.jsc - nt, t.jec + 1 + nt)
        m3 = fbox[:, :, None] & pos[None, None, :]
        fy2 = np.where(m3, -ij3(del6_u) * (sh(d2, 0, -1) - d2), fy2)
    return fx2, fy2


def delnflux(q, fx, fy, del6_u, del6_v, rarea, nord_col, damp_c_col, da_min: float,
             tile: Tile, mass=None):
    """DelnFlux.__call__ (delnflux.py:955-998): adds damping fluxes to fx, fy.

    ``nord_col``/``damp_c_col`` are the reference's K columns; the damping is
    applied on levels ``[0, nk)`` with ``nk = q.shape[2]`` (the caller passes
    the nz levels the reference's grid domain has).  Returns new (fx, fy).
    """
    damp_c_col = np.asarray(damp_c_col, dtype=np.float64)
    if (damp_c_col <= 1e-4).all():
        return fx, fy
    if (damp_c_col[:-1] <= 1e-4).any():
        raise NotImplementedError(
            "damp_c currently must be always greater than 10^-4 for delnflux"
        )
    nk = q.shape[2]
    nord4 = [int(nord_col[b]) for b in range(4)]
    damp = damp_coefficient(nord_col, damp_c_col, da_min)[:nk]
    fx2, fy2 = delnflux_nosg(q, del6_u, del6_v, rarea, nord4, damp, tile, mass)
    t = tile
    fx = fx.copy()
    fy = fy.copy()
    # compute dims [X_INTERFACE, Y_INTERFACE, Z]: [isc, iec+1] x [jsc, jec+1]
    sl = (slice(t.isc, t.iec + 2), slice(t.jsc, t.jec + 2), slice(0, nk))
    if mass is None:
        fx[sl] = fx[sl] + fx2[sl]
        fy[sl] = fy[sl] + fy2[sl]
    else:
        d3 = damp[None, None, :]
        fx[sl] = (fx + 0.5 * d3 * (sh(mass, -1, 0) + mass) * fx2)[sl]
        fy[sl] = (fy + 0.5 * d3 * (sh(mass, 0, -1) + mass) * fy2)[sl]
    return fx, fy
