"""CPU oracle for the D-grid step: FiniteVolumeFluxPrep, xtp_u/ytp_v,
AGrid2BGridFourthOrder, DivergenceDamping and d_sw.

Test infrastructure only (see oracle/common.py).  Follows
fv3core/stencils/fxadv.py:1-520, xtp_u.py, ytp_v.py, a2b_ord4.py:1-695,
divergence_damping.py:1-490 and d_sw.py:1-1131 statement by statement; the
corner fills of fv3core/utils/corners.py:543-1100 are built from their
rotation rule (``corner_fill_entries``) and pinned against the reference's
tables by tests/test_corner_tables.py.  Arrays are ``[i, j, k]``; metric
terms ``g[name]`` are ``[i, j]``; ``t`` is an :class:`oracle.fvtp2d.Tile`.
Temporaries the reference allocates zero-filled start at zero here too.
"""
from __future__ import annotations

import math

import numpy as np

from .common import box, sh, sign
from .csw import contravariant, put, region
from .fvtp2d import (
    C1, C2, C3, P1, P2, S11, S14, S15, Tile, _compute_al, _dm_left, _dm_ord8, _dm_right,
    _pert_ppm_standard_constraint, _xt_dxa_edge_0, _xt_dxa_edge_1, copy_corners_xy,
    delnflux, delnflux_nosg, fvtp2d, transpose_ij,
)

ALL = (-10**6, 10**6)
DCON_THRESHOLD = 1e-5
RADIUS = 6.3712e6


def _g(g, name):
    return g[name][:, :, None]


def _off(a, d):
    return None if a is None else a + d


def _rng(a, d0, d1):
    """Inclusive range (a + d0, a + d1), None when the anchor is absent."""
    return None if a is None else (a + d0, a + d1)


def _or(*masks):
    out = masks[0].copy()
    for m in masks[1:]:
        out |= m
    return out


# ------------------------------------------------------------ corner fills
def corner_fill_entries(kind, t: Tile, depth=3):
    """Destination/source pairs of the B-grid and D-grid corner fills.

    ``kind``: "bgrid_x", "bgrid_y" (a corner-point field) or "dgrid" (the
    vector pair x = field on (i+1/2, j) points, y = field on (i, j+1/2)
    points).  Every destination in the depth x depth halo block of a tile
    corner takes the value at its image under the quarter turn about the
    corner point that keeps the sign of its x (x fills) or y (y fills)
    offset, i.e. the neighbouring x- or y-halo.  Returns a list of
    (dst_field, di, dj, src_field, si, sj, corner) in absolute indices.
    """
    anchors = dict(sw=(t.i_start, t.j_start), se=(t.i_end, t.j_start),
                   nw=(t.i_start, t.j_end), ne=(t.i_end, t.j_end))
    out = []
    for corner, (ai, aj) in anchors.items():
        if ai is None or aj is None:
            continue
        sx = -1 if corner in ("sw", "nw") else 1
        sy = -1 if corner in ("sw", "se") else 1
        # corner point in continuous (storage index) coordinates
        cx = ai if sx < 0 else ai + 1
        cy = aj if sy < 0 else aj + 1
        comps = {"bgrid_x": [("q", (0.0, 0.0), "x")], "bgrid_y": [("q", (0.0, 0.0), "y")],
                 "dgrid": [("x", (0.5, 0.0), "x"), ("y", (0.0, 0.5), "y")]}[kind]
        stag = {"q": (0.0, 0.0), "x": (0.5, 0.0), "y": (0.0, 0.5)}
        for name, (ox, oy), keep in comps:
            for a in range(1, depth + 1):
                for b in range(1, depth + 1):
                    # destination position relative to the corner point
                    px = sx * (a - 0.5) if ox == 0.5 else sx * a
                    py = sy * (b - 0.5) if oy == 0.5 else sy * b
                    # the two quarter turns: (y, -x) and (-y, x)
                    cand = [(py, -px), (-py, px)]
                    if keep == "x":
                        qx, qy = [c for c in cand if np.sign(c[0]) == np.sign(px)][0]
                    else:
                        qx, qy = [c for c in cand if np.sign(c[1]) == np.sign(py)][0]
                    src = {"q": "q", "x": "y", "y": "x"}[name]
                    sox, soy = stag[src]
                    di, dj = cx + px - ox, cy + py - oy
                    si, sj = cx + qx - sox, cy + qy - soy
                    out.append((name, int(round(di)), int(round(dj)), src, int(round(si)),
                                int(round(sj)), corner))
    return out


def fill_bgrid(q, t, direction):
    """FillCornersBGrid (corners.py:543-723): in place."""
    src = q.copy()
    for _, di, dj, _, si, sj, _ in corner_fill_entries("bgrid_" + direction, t):
        q[di, dj] = src[si, sj]
    return q


def fill_dgrid(x, y, t, mysign):
    """fill_corners_dgrid_defn (corners.py:993-1100): x, y in place; the
    south-west and north-east corners take ``mysign``, the others +1."""
    xs, ys = x.copy(), y.copy()
    fields = {"x": x, "y": y}
    srcs = {"x": xs, "y": ys}
    for name, di, dj, src, si, sj, corner in corner_fill_entries("dgrid", t):
        mult = mysign if corner in ("sw", "ne") else 1.0
        fields[name][di, dj] = mult * srcs[src][si, sj]
    return x, y


# ----------------------------------------------------------------- fxadv
def fxadv(uc, vc, crx, cry, xfx, yfx, uc_contra, vc_contra, dt, g, t):
    """FiniteVolumeFluxPrep.__call__ (fxadv.py:488-589); arrays in place."""
    s = uc.shape
    full = (t.isd, t.ied, t.jsd, t.jed)
    cosa_u, cosa_v = _g(g, "cosa_u"), _g(g, "cosa_v")
    rsin_u, rsin_v = _g(g, "rsin_u"), _g(g, "rsin_v")
    s1, s2, s3, s4 = (_g(g, f"sin_sg{k}") for k in (1, 2, 3, 4))
    # main_uc_contra
    utmp = uc_contra.copy()
    v = 0.25 * (sh(vc, -1, 0) + vc + sh(vc, -1, 1) + sh(vc, 0, 1))
    put(uc_contra, contravariant(uc, v, cosa_u, rsin_u), region(s, t, (t.isc - 1, t.iec + 2), ALL, full))
    m = _or(region(s, t, ALL, _rng(t.j_start, -1, 0), full), region(s, t, ALL, _rng(t.j_end, 0, 1), full))
    put(uc_contra, utmp, m)
    # uc_contra_y_edge
    m = _or(region(s, t, t.i_start, ALL, full), region(s, t, _off(t.i_end, 1), ALL, full))
    put(uc_contra, np.where(uc > 0, uc / sh(s3, -1, 0), uc / s1), m)
    # main_vc_contra
    vtmp = vc_contra.copy()
    u = 0.25 * (sh(uc, 0, -1) + sh(uc, 1, -1) + uc + sh(uc, 1, 0))
    put(vc_contra, contravariant(vc, u, cosa_v, rsin_v), region(s, t, ALL, (t.jsc - 1, t.jec + 2), full))
    m = _or(region(s, t, ALL, t.j_start, full), region(s, t, ALL, _off(t.j_end, 1), full))
    put(vc_contra, vtmp, m)
    # vc_contra_y_edge
    vtmp = vc_contra.copy()
    jr = (t.jsc, t.jec + 1)
    m = _or(region(s, t, _rng(t.i_start, -1, 0), jr, full), region(s, t, _rng(t.i_end, 0, 1), jr, full))
    uct = uc_contra
    u_contra = 0.25 * (sh(uct, 0, -1) + sh(uct, 1, -1) + uct + sh(uct, 1, 0))
    put(vc_contra, contravariant(vc, u_contra, cosa_v, 1.0), m)
    m = np.zeros(s[:2], dtype=bool)
    for ii in (_rng(t.i_start, -1, 0), _rng(t.i_end, 0, 1)):
        for jj in (_rng(t.j_start, 0, 1), _rng(t.j_end, 0, 1)):
            m |= region(s, t, ii, jj, full)
    put(vc_contra, vtmp, m)
    # vc_contra_x_edge
    m = _or(region(s, t, ALL, t.j_start, full), region(s, t, ALL, _off(t.j_end, 1), full))
    put(vc_contra, np.where(vc > 0, vc / sh(s4, 0, -1), vc / s2), m)
    # uc_contra_x_edge
    utmp = uc_contra.copy()
    ir = (t.isc, t.iec + 1)
    m = _or(region(s, t, ir, _rng(t.j_start, -1, 0), full), region(s, t, ir, _rng(t.j_end, 0, 1), full))
    vct = vc_contra
    v_contra = 0.25 * (sh(vct, -1, 0) + vct + sh(vct, -1, 1) + sh(vct, 0, 1))
    put(uc_contra, contravariant(uc, v_contra, cosa_u, 1.0), m)
    m = np.zeros(s[:2], dtype=bool)
    for ii in (_rng(t.i_start, 0, 1), _rng(t.i_end, 0, 1)):
        for jj in (_rng(t.j_start, -1, 0), _rng(t.j_end, 0, 1)):
            m |= region(s, t, ii, jj, full)
    put(uc_contra, utmp, m)
    # uc_contra_corners (domain full + (1, 1) origin)
    cdom = (t.isd + 1, t.ied, t.jsd + 1, t.jed)
    ucc = uc_contra.copy()
    vct = vc_contra

    def uc_corner(cv_off, vterms, uterms):
        cvv = sh(cosa_v, *cv_off)
        damp = 1.0 / (1.0 - 0.0625 * cosa_u * cvv)
        a, b, c, d = vterms
        e, f, h = uterms
        return (uc - 0.25 * cosa_u * (a + b + c + d - 0.25 * cvv * (e + f + h))) * damp

    val = uc_corner((-1, 0), (sh(vct, -1, 1), sh(vct, 0, 1), vct, sh(vc, -1, 0)),
                    (sh(ucc, -1, 0), sh(ucc, -1, -1), sh(ucc, 0, -1)))
    put(uc_contra, val, _or(region(s, t, _off(t.i_start, 1), _off(t.j_start, -1), cdom),
                            region(s, t, _off(t.i_start, 1), t.j_end, cdom)))
    val = uc_corner((-1, 1), (sh(vct, -1, 0), vct, sh(vct, 0, 1), sh(vc, -1, 1)),
                    (sh(ucc, -1, 0), sh(ucc, -1, 1), sh(ucc, 0, 1)))
    put(uc_contra, val, _or(region(s, t, _off(t.i_start, 1), t.j_start, cdom),
                            region(s, t, _off(t.i_start, 1), _off(t.j_end, 1), cdom)))
    val = uc_corner((0, 0), (sh(vct, 0, 1), sh(vct, -1, 1), sh(vct, -1, 0), vc),
                    (sh(ucc, 1, 0), sh(ucc, 1, -1), sh(ucc, 0, -1)))
    put(uc_contra, val, _or(region(s, t, t.i_end, _off(t.j_start, -1), cdom),
                            region(s, t, t.i_end, t.j_end, cdom)))
    val = uc_corner((0, 1), (vct, sh(vct, -1, 0), sh(vct, -1, 1), sh(vc, 0, 1)),
                    (sh(ucc, 1, 0), sh(ucc, 1, 1), sh(ucc, 0, 1)))
    put(uc_contra, val, _or(region(s, t, t.i_end, t.j_start, cdom),
                            region(s, t, t.i_end, _off(t.j_end, 1), cdom)))
    # vc_contra_corners
    vcc = vc_contra.copy()
    ut = uc_contra

    def vc_corner(cu_off, uterms, vterms):
        cuu = sh(cosa_u, *cu_off)
        damp = 1.0 / (1.0 - 0.0625 * cuu * cosa_v)
        a, b, c, d = uterms
        e, f, h = vterms
        return (vc - 0.25 * cosa_v * (a + b + c + d - 0.25 * cuu * (e + f + h))) * damp

    val = vc_corner((0, -1), (sh(ut, 1, -1), sh(ut, 1, 0), ut, sh(uc, 0, -1)),
                    (sh(vcc, 0, -1), sh(vcc, -1, -1), sh(vcc, -1, 0)))
    put(vc_contra, val, _or(region(s, t, _off(t.i_start, -1), _off(t.j_start, 1), cdom),
                            region(s, t, t.i_end, _off(t.j_start, 1), cdom)))
    val = vc_corner((1, -1), (sh(ut, 0, -1), ut, sh(ut, 1, 0), sh(uc, 1, -1)),
                    (sh(vcc, 0, -1), sh(vcc, 1, -1), sh(vcc, 1, 0)))
    put(vc_contra, val, _or(region(s, t, t.i_start, _off(t.j_start, 1), cdom),
                            region(s, t, _off(t.i_end, 1), _off(t.j_start, 1), cdom)))
    val = vc_corner((1, 0), (ut, sh(ut, 0, -1), sh(ut, 1, -1), sh(uc, 1, 0)),
                    (sh(vcc, 0, 1), sh(vcc, 1, 1), sh(vcc, 1, 0)))
    put(vc_contra, val, _or(region(s, t, _off(t.i_end, 1), t.j_end, cdom),
                            region(s, t, t.i_start, t.j_end, cdom)))
    val = vc_corner((0, 0), (sh(ut, 1, 0), sh(ut, 1, -1), sh(ut, 0, -1), uc),
                    (sh(vcc, 0, 1), sh(vcc, -1, 1), sh(vcc, -1, 0)))
    put(vc_contra, val, _or(region(s, t, t.i_end, t.j_end, cdom),
                            region(s, t, _off(t.i_start, -1), t.j_end, cdom)))
    # fxadv_fluxes_stencil
    rdxa, rdya = _g(g, "rdxa"), _g(g, "rdya")
    dx, dy = _g(g, "dx"), _g(g, "dy")
    prod = dt * uc_contra
    m = region(s, t, (t.isc, t.iec + 1), ALL, full)
    put(crx, np.where(prod > 0, prod * sh(rdxa, -1, 0), prod * rdxa), m)
    put(xfx, np.where(prod > 0, dy * prod * sh(s3, -1, 0), dy * prod * s1), m)
    prod = dt * vc_contra
    m = region(s, t, ALL, (t.jsc, t.jec + 1), full)
    put(cry, np.where(prod > 0, prod * sh(rdya, 0, -1), prod * rdya), m)
    put(yfx, np.where(prod > 0, dx * prod * sh(s4, 0, -1), dx * prod * s2), m)


# ------------------------------------------------------------- xtp_u / ytp_v
def _bl_br_xtp(u, dx, dxa, iord, t: Tile):
    """get_bl_br of xtp_u.py:8-47 along the first axis; ``t`` gives anchors."""
    s = u.shape
    dx3, dxa3 = dx[:, :, None], dxa[:, :, None]
    if iord < 8:
        al = _compute_al(u, dx3, t.i_start, t.i_end)
        bl = al - u
        br = sh(al, 1, 0) - u
    else:
        if iord != 8:
            raise NotImplementedError("xtp_u supports iord 5, 6, 7 and 8")
        dm = _dm_ord8(u)
        al = 0.5 * (sh(u, -1) + u) + 1.0 / 3.0 * (sh(dm, -1) - dm)
        xt = 2.0 * dm
        bl = -1.0 * sign(np.minimum(np.abs(xt), np.abs(al - u)), xt)
        br = sign(np.minimum(np.abs(xt), np.abs(sh(al, 1) - u)), xt)
        bl, br = _bl_br_edges(bl, br, u, dxa3, al, dm, t.i_start, t.i_end)
        m = region(s, t, _off(t.i_start, 1), ALL) | region(s, t, _off(t.i_end, -1), ALL)
        bl_c, br_c = _pert_ppm_standard_constraint(u, bl, br)
        bl = np.where(m[:, :, None], bl_c, bl)
        br = np.where(m[:, :, None], br_c, br)
    m = np.zeros(s[:2], dtype=bool)
    for ii in (_rng(t.i_start, -1, 0), _rng(t.i_end, 0, 1)):
        for jj in (t.j_start, _off(t.j_end, 1)):
            m |= region(s, t, ii, jj)
    bl = np.where(m[:, :, None], 0.0, bl)
    br = np.where(m[:, :, None], 0.0, br)
    return bl, br


def _bl_br_edges(bl, br, q, dxa, al, dm, i_start, i_end):
    """bl_br_edges (xppm.py:184-246) without the standard constraint."""
    from .common import imask
    s = q.shape
    if i_start is None and i_end is None:
        return bl, br
    xt_bl = np.full(s, np.nan)
    xt_br = np.full(s, np.nan)
    dm_l, dm_r = _dm_left(q), _dm_right(q)
    m = imask(s, _off(i_start, -1))
    xt_bl = np.where(m, S14 * dm_l + S11 * (sh(q, -1) - q) + q, xt_bl)
    xt_br = np.where(m, _xt_dxa_edge_0(q, dxa), xt_br)
    m = imask(s, i_start)
    xt_bl = np.where(m, _xt_dxa_edge_1(q, dxa), xt_bl)
    xt_br = np.where(m, S15 * q + S11 * sh(q, 1) - S14 * dm_r, xt_br)
    m = imask(s, _off(i_start, 1))
    xt_bl = np.where(m, S15 * sh(q, -1) + S11 * q - S14 * dm, xt_bl)
    xt_br = np.where(m, sh(al, 1), xt_br)
    m = imask(s, _off(i_end, -1))
    xt_bl = np.where(m, al, xt_bl)
    xt_br = np.where(m, S15 * sh(q, 1) + S11 * q + S14 * dm, xt_br)
    m = imask(s, i_end)
    xt_bl = np.where(m, S15 * q + S11 * sh(q, -1) + S14 * dm_l, xt_bl)
    xt_br = np.where(m, _xt_dxa_edge_0(q, dxa), xt_br)
    m = imask(s, _off(i_end, 1))
    xt_bl = np.where(m, _xt_dxa_edge_1(q, dxa), xt_bl)
    xt_br = np.where(m, S11 * (sh(q, 1) - q) - S14 * dm_r + q, xt_br)
    edge = np.zeros(s, dtype=bool)
    for off in (-1, 0, 1):
        edge |= imask(s, _off(i_start, off)) | imask(s, _off(i_end, off))
    return np.where(edge, xt_bl - q, bl), np.where(edge, xt_br - q, br)


def advect_u_along_x(u, ub_contra, rdx, dx, dxa, dt, iord, t: Tile):
    """advect_u_along_x (xtp_u.py:50-91)."""
    bl, br = _bl_br_xtp(u, dx, dxa, iord, t)
    b0 = bl + br
    rdx3 = rdx[:, :, None]
    cfl = np.where(ub_contra > 0, ub_contra * dt * sh(rdx3, -1, 0), ub_contra * dt * rdx3)
    fx0 = np.where(cfl > 0.0, (1.0 - cfl) * (sh(br, -1) - cfl * sh(b0, -1)),
                   (1.0 + cfl) * (bl + cfl * b0))
    if iord < 8:
        if iord == 5:
            smt5 = bl * br < 0
        else:
            smt5 = (3.0 * np.abs(b0)) < np.abs(bl - br)
        smt5_m1 = np.zeros_like(smt5)
        smt5_m1[1:] = smt5[:-1]
        mask = np.where(smt5_m1 | smt5, 1.0, 0.0)
    else:
        mask = 1.0
    return np.where(ub_contra > 0.0, sh(u, -1) + fx0 * mask, u + fx0 * mask)


def advect_v_along_y(v, vb_contra, rdy, dy, dya, dt, jord, t: Tile):
    """advect_v_along_y (ytp_v.py:50-91): xtp_u on transposed arrays."""
    tt = t.transposed()
    out = advect_u_along_x(transpose_ij(v), transpose_ij(vb_contra), rdy.T, dy.T, dya.T, dt,
                           jord, tt)
    return transpose_ij(out)


def interpolate_uc_vc_to_cell_corners(uc, vc, cosa, rsina, uc_contra, vc_contra, t, s):
    """d_sw.py:564-603."""
    ub_cov = 0.5 * (sh(uc, 0, -1) + uc)
    vb_cov = 0.5 * (sh(vc, -1, 0) + vc)
    ub = contravariant(ub_cov, vb_cov, cosa, rsina)
    vb = contravariant(vb_cov, ub_cov, cosa, rsina)
    rows = _or(region(s, t, ALL, t.j_start), region(s, t, ALL, _off(t.j_end, 1)))[:, :, None]
    cols = _or(region(s, t, t.i_start, ALL), region(s, t, _off(t.i_end, 1), ALL))[:, :, None]
    ub = np.where(rows, 0.25 * (-sh(uc_contra, 0, -2) + 3.0 * (sh(uc_contra, 0, -1) + uc_contra)
                                - sh(uc_contra, 0, 1)), ub)
    ub = np.where(cols, 0.5 * (sh(uc_contra, 0, -1) + uc_contra), ub)
    vb = np.where(cols, 0.25 * (-sh(vc_contra, -2, 0) + 3.0 * (sh(vc_contra, -1, 0) + vc_contra)
                                - sh(vc_contra, 1, 0)), vb)
    vb = np.where(rows, 0.5 * (sh(vc_contra, -1, 0) + vc_contra), vb)
    return ub, vb


def _corner_ke(u, v, ut, vt, dt, io1, jo1, io2, vsign):
    """corner_ke (d_sw.py:112-133)."""
    dt6 = dt / 6.0
    return dt6 * (
        (ut + sh(ut, 0, -1)) * ((io1 + 1) * u - (io1 * sh(u, -1, 0)))
        + (vt + sh(vt, -1, 0)) * ((jo1 + 1) * v - (jo1 * sh(v, 0, -1)))
        + (((jo1 + 1) * ut - (jo1 * sh(ut, 0, -1)))
           + vsign * ((io1 + 1) * vt - (io1 * sh(vt, -1, 0))))
        * ((io2 + 1) * u - (io2 * sh(u, -1, 0)))
    )


def kinetic_energy(uc, vc, u, v, uc_contra, vc_contra, dt, hord_mt, g, t):
    """kinetic_energy_update_part_1/2 (d_sw.py:196-237) -> ke on the corners."""
    s = u.shape
    dom = box(s[:2], t.isc, t.iec + 1, t.jsc, t.jec + 1)
    ub, vb = interpolate_uc_vc_to_cell_corners(uc, vc, _g(g, "cosa"), _g(g, "rsina"),
                                               uc_contra, vc_contra, t, s)
    adv_v = advect_v_along_y(v, vb, g["rdy"], g["dy"], g["dya"], dt, hord_mt, t)
    adv_u = advect_u_along_x(u, ub, g["rdx"], g["dx"], g["dxa"], dt, hord_mt, t)
    ke = 0.5 * dt * (ub * adv_u + vb * adv_v)
    for (ai, aj, args) in ((t.i_start, t.j_start, (0, 0, -1, 1)),
                           (_off(t.i_end, 1), t.j_start, (-1, 0, 0, -1)),
                           (_off(t.i_end, 1), _off(t.j_end, 1), (-1, -1, 0, 1)),
                           (t.i_start, _off(t.j_end, 1), (0, -1, -1, -1))):
        m = region(s, t, ai, aj)
        ke = np.where(m[:, :, None], _corner_ke(u, v, uc_contra, vc_contra, dt, *args), ke)
    out = np.full(s, np.nan)
    put(out, ke, dom)
    return out, ub, vb


# ---------------------------------------------------------------- a2b_ord4
A1, A2 = 9.0 / 16.0, -1.0 / 16.0
B1, B2 = 7.0 / 12.0, -1.0 / 12.0
CC1, CC2 = 2.0 / 3.0, -1.0 / 6.0


def _gc_dist(p1a, p1b, p2a, p2b):
    """great_circle_dist (a2b_ord4.py:37-41)."""
    tb = np.sin((p1b - p2b) / 2.0) ** 2.0
    ta = np.sin((p1a - p2a) / 2.0) ** 2.0
    return np.arcsin(np.sqrt(tb + np.cos(p1b) * np.cos(p2b) * ta)) * 2.0


def _extrap(p0, p1, p2, qa, qb):
    x1 = _gc_dist(p1[0], p1[1], p0[0], p0[1])
    x2 = _gc_dist(p2[0], p2[1], p0[0], p0[1])
    return qa + x1 / (x2 - x1) * (qa - qb)


def _corner_value(qin, g, i, j, kind):
    """The _sw/_nw/_ne/_se_corner stencils (a2b_ord4.py:44-268) at (i, j)."""
    a1, a2, b1, b2 = g["agrid1"], g["agrid2"], g["bgrid1"], g["bgrid2"]
    p0 = (b1[i, j], b2[i, j])

    def ec(o1, o2):
        (di1, dj1), (di2, dj2) = o1, o2
        return _extrap(p0, (a1[i + di1, j + dj1], a2[i + di1, j + dj1]),
                       (a1[i + di2, j + dj2], a2[i + di2, j + dj2]),
                       qin[i + di1, j + dj1], qin[i + di2, j + dj2])

    s00, sm10, s0m1, sm1m1 = ((0, 0), (1, 1)), ((-1, 0), (-2, 1)), ((0, -1), (1, -2)), \
        ((-1, -1), (-2, -2))
    order = {"sw": (s00, sm10, s0m1), "nw": (sm10, sm1m1, s00),
             "ne": (sm1m1, s0m1, sm10), "se": (s0m1, sm1m1, s00)}[kind]
    e1, e2, e3 = (ec(*o) for o in order)
    return (e1 + e2 + e3) * (1.0 / 3.0)


def _qx_edge_west(q, dxa):
    g_in = sh(dxa, 1) / dxa
    g_ou = sh(dxa, -2) / sh(dxa, -1)
    return 0.5 * (((2.0 + g_in) * q - sh(q, 1)) / (1.0 + g_in)
                  + ((2.0 + g_ou) * sh(q, -1) - sh(q, -2)) / (1.0 + g_ou))


def _qx_edge_west2(q, dxa):
    g_in = dxa / sh(dxa, -1)
    g_ou = sh(dxa, -3) / sh(dxa, -2)
    qxleft = 0.5 * (((2.0 + g_in) * sh(q, -1) - q) / (1.0 + g_in)
                    + ((2.0 + g_ou) * sh(q, -2) - sh(q, -3)) / (1.0 + g_ou))
    qxright = B2 * (sh(q, -1) + sh(q, 2)) + B1 * (q + sh(q, 1))
    return (3.0 * (g_in * sh(q, -1) + q) - (g_in * qxleft + qxright)) / (2.0 + 2.0 * g_in)


def _qx_edge_east(q, dxa):
    g_in = sh(dxa, -2) / sh(dxa, -1)
    g_ou = sh(dxa, 1) / dxa
    return 0.5 * (((2.0 + g_in) * sh(q, -1) - sh(q, -2)) / (1.0 + g_in)
                  + ((2.0 + g_ou) * q - sh(q, 1)) / (1.0 + g_ou))


def _qx_edge_east2(q, dxa):
    g_in = sh(dxa, -1) / dxa
    g_ou = sh(dxa, 2) / sh(dxa, 1)
    qxright = 0.5 * (((2.0 + g_in) * q - sh(q, -1)) / (1.0 + g_in)
                     + ((2.0 + g_ou) * sh(q, 1) - sh(q, 2)) / (1.0 + g_ou))
    qxleft = B2 * (sh(q, -3) + q) + B1 * (sh(q, -2) + sh(q, -1))
    return (3.0 * (sh(q, -1) + g_in * q) - (g_in * qxright + qxleft)) / (2.0 + 2.0 * g_in)


def _ppm_volume_mean_x(qin, dxa, t):
    """ppm_volume_mean_x (a2b_ord4.py:300-318) on all points (edges by region)."""
    s = qin.shape
    dxa3 = dxa[:, :, None]
    qx = B2 * (sh(qin, -2) + sh(qin, 1)) + B1 * (sh(qin, -1) + qin)
    for a, f in ((t.i_start, _qx_edge_west), (_off(t.i_start, 1), _qx_edge_west2),
                 (_off(t.i_end, 1), _qx_edge_east), (t.i_end, _qx_edge_east2)):
        m = region(s, t, a, ALL)[:, :, None]
        qx = np.where(m, f(qin, dxa3), qx)
    return qx


def a2b_ord4(qin, qout, g, t, replace=False, tmp_edges=None):
    """AGrid2BGridFourthOrder.__call__ (a2b_ord4.py:617-695): qout (and qin
    when ``replace``) in place; returns the edge temporary."""
    s = qin.shape
    edges = np.zeros(s) if tmp_edges is None else tmp_edges
    # corner stencils, applied on every tile as in the reference
    for (i, j, kind) in ((t.isc, t.jsc, "sw"), (t.iec + 1, t.jsc, "nw"),
                         (t.iec + 1, t.jec + 1, "ne"), (t.isc, t.jec + 1, "se")):
        val = _corner_value(qin, g, i, j, kind)
        qout[i, j] = val
        edges[i, j] = val
    dxa, dya = g["dxa"][:, :, None], g["dya"][:, :, None]
    js2 = t.jsc + 1 if t.south else t.jsc
    je1 = t.jec if t.north else t.jec + 1
    is2 = t.isc + 1 if t.west else t.isc
    ie1 = t.iec if t.east else t.iec + 1
    # qout_x_edge (west/east): q2 at j and j - 1
    q2 = (sh(qin, -1, 0) * dxa + qin * sh(dxa, -1, 0)) / (sh(dxa, -1, 0) + dxa)
    for flag, i, ew in ((t.west, t.isc, "edge_w"), (t.east, t.iec + 1, "edge_e")):
        if not flag:
            continue
        w = g[ew][:, :, None]
        val = w * sh(q2, 0, -1) + (1.0 - w) * q2
        m = box(s[:2], i, i, js2, je1)
        put(qout, val, m)
        put(edges, val, m)
    q1 = (sh(qin, 0, -1) * dya + qin * sh(dya, 0, -1)) / (sh(dya, 0, -1) + dya)
    for flag, j, es in ((t.south, t.jsc, "edge_s"), (t.north, t.jec + 1, "edge_n")):
        if not flag:
            continue
        w = g[es][:, :, None]
        val = w * sh(q1, -1, 0) + (1.0 - w) * q1
        m = box(s[:2], is2, ie1, j, j)
        put(qout, val, m)
        put(edges, val, m)
    # volume means: qx on [isc, iec+1] x [jsc-2, jec+2]; qy on [isc-2, iec+2] x [jsc, jec+1]
    qx = np.full(s, np.nan)
    put(qx, _ppm_volume_mean_x(qin, g["dxa"], t), box(s[:2], t.isc, t.iec + 1, t.jsc - 2, t.jec + 2))
    qyt = _ppm_volume_mean_x(transpose_ij(qin), g["dya"].T, t.transposed())
    qy = np.full(s, np.nan)
    put(qy, transpose_ij(qyt), box(s[:2], t.isc - 2, t.iec + 2, t.jsc, t.jec + 1))
    # a2b_interpolation on the corners minus the tile-edge lines
    i0 = t.isc + 1 if t.west else t.isc
    i1 = t.iec if t.east else t.iec + 1
    j0 = t.jsc + 1 if t.south else t.jsc
    j1 = t.jec if t.north else t.jec + 1
    dom = (i0, i1, j0, j1)
    qxx = A2 * (sh(qx, 0, -2) + sh(qx, 0, 1)) + A1 * (sh(qx, 0, -1) + qx)
    qyy = A2 * (sh(qy, -2, 0) + sh(qy, 1, 0)) + A1 * (sh(qy, -1, 0) + qy)
    m = region(s, t, ALL, _off(t.j_start, 1), dom)[:, :, None]
    upper = A2 * (sh(qx, 0, -1) + sh(qx, 0, 2)) + A1 * (qx + sh(qx, 0, 1))
    qxx = np.where(m, CC1 * (sh(qx, 0, -1) + qx) + CC2 * (sh(edges, 0, -1) + upper), qxx)
    m = region(s, t, ALL, t.j_end, dom)[:, :, None]
    lower = A2 * (sh(qx, 0, -3) + qx) + A1 * (sh(qx, 0, -2) + sh(qx, 0, -1))
    qxx = np.where(m, CC1 * (sh(qx, 0, -1) + qx) + CC2 * (sh(edges, 0, 1) + lower), qxx)
    m = region(s, t, _off(t.i_start, 1), ALL, dom)[:, :, None]
    right = A2 * (sh(qy, -1, 0) + sh(qy, 2, 0)) + A1 * (qy + sh(qy, 1, 0))
    qyy = np.where(m, CC1 * (sh(qy, -1, 0) + qy) + CC2 * (sh(edges, -1, 0) + right), qyy)
    m = region(s, t, t.i_end, ALL, dom)[:, :, None]
    left = A2 * (sh(qy, -3, 0) + qy) + A1 * (sh(qy, -2, 0) + sh(qy, -1, 0))
    qyy = np.where(m, CC1 * (sh(qy, -1, 0) + qy) + CC2 * (sh(edges, 1, 0) + left), qyy)
    put(qout, 0.5 * (qxx + qyy), box(s[:2], *dom))
    if replace:
        m = box(s[:2], t.isc, t.iec + 1, t.jsc, t.jec + 1)
        put(qin, qout, m)
    return edges


# ---------------------------------------------------------- divergence damping
def _damp_tmp(q, da_min_c, d2_bg, dddmp):
    mintmp = np.minimum(0.2, dddmp * np.abs(q))
    return da_min_c * np.maximum(d2_bg, mintmp)


def divergence_damping(u, v, va, ptc, vort, ua, divg_d, vc, uc, delpc, ke, wk, dt, g, t,
                       nord_col, d2_bg_col, dddmp, d4_bg, da_min_c, nord):
    """DivergenceDamping.__call__ (divergence_damping.py:319-441); in place.

    ``ptc`` receives u_contra_dyc and ``vort`` v_contra_dxc, as in the
    reference.  All arrays have the levels of the call.
    """
    s = u.shape
    nk = s[2]
    nonzero_k = 0
    nonzero_nord = int(nord)
    for k in range(len(nord_col)):
        if nord_col[k] > 0:
            nonzero_k = k
            nonzero_nord = int(nord_col[k])
            break
    dd8 = (da_min_c * d4_bg) ** (nonzero_nord + 1)
    comp = box(s[:2], t.isc, t.iec + 1, t.jsc, t.jec + 1)
    d2_bg = np.asarray(d2_bg_col, dtype=np.float64)[None, None, :nk]
    lo = slice(0, nonzero_k)
    hi = slice(nonzero_k, nk)
    if nonzero_k > 0:
        # get_delpc (divergence_damping.py:31-93) on the low levels
        cosa_u, sina_u = _g(g, "cosa_u"), _g(g, "sina_u")
        cosa_v, sina_v = _g(g, "cosa_v"), _g(g, "sina_v")
        s1, s2, s3, s4 = (_g(g, f"sin_sg{k}") for k in (1, 2, 3, 4))
        dxc, dyc = _g(g, "dxc"), _g(g, "dyc")
        ul, vl, ual, val_, ucl, vcl = (a[..., lo] for a in (u, v, ua, va, uc, vc))
        vc_from_va = 0.5 * (sh(val_, 0, -1) + val_)
        u_contra = contravariant(ul, vc_from_va, cosa_v, sina_v)
        m = _or(region(s, t, ALL, t.j_start), region(s, t, ALL, _off(t.j_end, 1)))[:, :, None]
        u_contra = np.where(m, np.where(vcl > 0, ul * sh(s4, 0, -1), ul * s2), u_contra)
        u_contra_dyc = u_contra * dyc
        uc_from_ua = 0.5 * (sh(ual, -1, 0) + ual)
        v_contra = contravariant(vl, uc_from_ua, cosa_u, sina_u)
        m = _or(region(s, t, t.i_start, ALL), region(s, t, _off(t.i_end, 1), ALL))[:, :, None]
        v_contra = np.where(m, np.where(ucl > 0, vl * sh(s3, -1, 0), vl * s1), v_contra)
        v_contra_dxc = v_contra * dxc
        m = _or(region(s, t, t.i_start, _off(t.j_end, 1)), region(s, t, _off(t.i_end, 1), _off(t.j_end, 1)),
                region(s, t, t.i_start, _off(t.j_start, -1)),
                region(s, t, _off(t.i_end, 1), _off(t.j_start, -1)))[:, :, None]
        v_contra_dxc = np.where(m, 0.0, v_contra_dxc)
        dpc = (sh(v_contra_dxc, 0, -1) - v_contra_dxc + sh(u_contra_dyc, -1, 0) - u_contra_dyc)
        dpc = _g(g, "rarea_c") * dpc
        # written fields: delpc on the compute corners; the contravariant
        # temporaries on the extent the delpc statement reads
        tmp = delpc[..., lo]
        put(tmp, dpc, comp)
        delpc[..., lo] = tmp
        tmp = ptc[..., lo]
        put(tmp, u_contra_dyc, box(s[:2], t.isc - 1, t.iec + 1, t.jsc, t.jec + 1))
        ptc[..., lo] = tmp
        # damping (divergence_damping.py:96-107)
        delpcdt = dpc * dt
        damp = _damp_tmp(delpcdt, da_min_c, d2_bg[..., lo], dddmp)
        vv = damp * dpc
        tmp = vort[..., lo]
        put(tmp, v_contra_dxc, box(s[:2], t.isc, t.iec + 1, t.jsc - 1, t.jec + 1))
        put(tmp, vv, comp)
        vort[..., lo] = tmp
        tmp = ke[..., lo]
        put(tmp, tmp + vv, comp)
        ke[..., lo] = tmp
    # high levels
    dh = divg_d[..., hi]
    tmp = delpc[..., hi]
    put(tmp, dh, comp)
    delpc[..., hi] = tmp
    vch, uch = vc[..., hi], uc[..., hi]
    grid_has_corner = any(a is not None for a in (t.i_start, t.i_end)) and any(
        a is not None for a in (t.j_start, t.j_end))
    for n in range(nonzero_nord):
        nt = nonzero_nord - (n + 1)
        fillc = (n + 1 != nonzero_nord) and grid_has_corner
        if fillc:
            fill_bgrid(dh, t, "x")
        m = box(s[:2], t.isc - nt - 1, t.iec + nt + 1, t.jsc - nt, t.jec + nt + 1)
        put(vch, (sh(dh, 1, 0) - dh) * _g(g, "divg_u"), m)
        if fillc:
            fill_bgrid(dh, t, "y")
        m = box(s[:2], t.isc - nt, t.iec + nt + 1, t.jsc - nt - 1, t.jec + nt + 1)
        put(uch, (sh(dh, 0, 1) - dh) * _g(g, "divg_v"), m)
        if fillc:
            fill_dgrid(vch, uch, t, -1.0)
        dom = (t.isc - nt, t.iec + nt + 1, t.jsc - nt, t.jec + nt + 1)
        val = sh(uch, 0, -1) - uch + sh(vch, -1, 0) - vch
        m = _or(region(s, t, t.i_start, t.j_start, dom), region(s, t, _off(t.i_end, 1), t.j_start, dom))
        val = np.where(m[:, :, None], sh(vch, -1, 0) - vch - uch, val)
        m = _or(region(s, t, t.i_start, _off(t.j_end, 1), dom),
                region(s, t, _off(t.i_end, 1), _off(t.j_end, 1), dom))
        val = np.where(m[:, :, None], sh(uch, 0, -1) + sh(vch, -1, 0) - vch, val)
        val = val * _g(g, "rarea_c")
        put(dh, val, box(s[:2], *dom))
    divg_d[..., hi] = dh
    vc[..., hi] = vch
    uc[..., hi] = uch
    # _vorticity_calc
    vh = vort[..., hi]
    if dddmp < 1e-5:
        put(vh, 0.0, box(s[:2], t.isd, t.ied + 1, t.jsd, t.jed + 1))
    else:
        wkh = wk[..., hi]
        a2b_ord4(wkh, vh, g, t, replace=False)
        dph = delpc[..., hi]
        put(vh, abs(dt) * (dph ** 2.0 + vh ** 2.0) ** 0.5, comp)
    # damping_nord_highorder_stencil
    damp = _damp_tmp(vh, da_min_c, d2_bg[..., hi], dddmp)
    newv = damp * delpc[..., hi] + dd8 * dh
    put(vh, newv, comp)
    vort[..., hi] = vh
    kh = ke[..., hi]
    put(kh, kh + newv, comp)
    ke[..., hi] = kh


# ---------------------------------------------------------------------- d_sw
def column_namelist(cfg, npz):
    """get_column_namelist (d_sw.py:512-563) from a dict of namelist values."""
    col = {n: np.zeros(npz + 1) for n in ("ke_bg", "d_con", "nord", "nord_v", "nord_w",
                                           "nord_t", "damp_vt", "damp_w", "damp_t", "d2_divg")}
    for n in ("ke_bg", "d_con", "nord"):
        col[n][:] = cfg[n]
    col["d2_divg"][:] = min(0.2, cfg["d2_bg"])
    col["nord_v"][:] = min(2, col["nord"][0])
    col["nord_w"][:] = col["nord_v"][0]
    col["nord_t"][:] = col["nord_v"][0]
    col["damp_vt"][:] = cfg["vtdm4"] if cfg["do_vort_damp"] else 0
    col["damp_w"][:] = col["damp_vt"][0]
    col["damp_t"][:] = col["damp_vt"][0]

    def set_low(k):
        for n in ("nord", "nord_w", "d_con"):
            col[n][k] = 0
        col["damp_w"][k] = col["d2_divg"][k]

    def low(k):
        set_low(k)
        if cfg["do_vort_damp"]:
            col["nord_v"][k] = 0
            col["damp_vt"][k] = 0.5 * col["d2_divg"][k]

    if npz == 1 or cfg["n_sponge"] < 0:
        col["d2_divg"][0] = cfg["d2_bg"]
    else:
        col["d2_divg"][0] = max(0.01, cfg["d2_bg"], cfg["d2_bg_k1"])
        low(0)
        if cfg["d2_bg_k2"] > 0.01:
            col["d2_divg"][1] = max(cfg["d2_bg"], cfg["d2_bg_k2"])
            low(1)
        if cfg["d2_bg_k2"] > 0.05:
            col["d2_divg"][2] = max(cfg["d2_bg"], 0.2 * cfg["d2_bg_k2"])
            set_low(2)
    return col


def _flux_increment(gx, gy, rarea):
    return (gx - sh(gx, 1, 0) + gy - sh(gy, 0, 1)) * rarea


def d_sw(state, dt, g, t, cfg, col, damp):
    """DGridShallowWaterLagrangianDynamics.__call__ (d_sw.py:843-1131).

    ``state``: dict of the call's arrays (delpc, delp, ptc, pt, u, v, w, uc,
    vc, ua, va, divgd, mfx, mfy, cx, cy, crx, cry, xfx, yfx, q_con, zh,
    heat_source, diss_est), nz levels each (zh: nz + 1); copied.  ``cfg``:
    namelist dict (hord_*, dddmp, d4_bg, nord, d_con, do_f3d, hydrostatic);
    ``col``: column_namelist; ``damp``: dict with del6_u, del6_v, da_min,
    da_min_c.  Returns the dict of arrays after the call.
    """
    st = {k: np.array(v, dtype=np.float64, copy=True) for k, v in state.items()}
    nz = st["delp"].shape[2]
    s = st["delp"].shape
    s2 = s[:2]
    comp = box(s2, t.isc, t.iec, t.jsc, t.jec)
    full = box(s2, t.isd, t.ied, t.jsd, t.jed)
    rarea = _g(g, "rarea")
    z = np.zeros(s)
    uc_contra, vc_contra = np.zeros(s), np.zeros(s)
    fxadv(st["uc"], st["vc"], st["crx"], st["cry"], st["xfx"], st["yfx"], uc_contra, vc_contra,
          dt, g, t)
    del6_u, del6_v = damp["del6_u"], damp["del6_v"]

    def transport(q, hord, nord_name, damp_name, mfx=None, mfy=None, mass=None):
        qx, qy = copy_corners_xy(q, t)
        q[...] = qx
        fx, fy = fvtp2d(qx, qy, st["crx"], st["cry"], st["xfx"], st["yfx"], g["area"], g["dxa"],
                        g["dya"], hord, t, x_mass_flux=mfx, y_mass_flux=mfy, check_domains=False)
        if nord_name is not None:
            fx, fy = delnflux(q, fx, fy, del6_u, del6_v, g["rarea"], col[nord_name],
                              col[damp_name], damp["da_min"], t, mass=mass)
        # the component's flux temporaries start at zero outside what is written
        sx = box(s2, t.isc, t.iec + 1, t.jsc, t.jec)
        sy = box(s2, t.isc, t.iec, t.jsc, t.jec + 1)
        ox, oy = np.zeros(s), np.zeros(s)
        put(ox, fx, sx)
        put(oy, fy, sy)
        return ox, oy

    fx, fy = transport(st["delp"], cfg["hord_dp"], "nord_v", "damp_vt")
    # flux_capacitor on the full domain
    put(st["cx"], st["cx"] + st["crx"], full)
    put(st["cy"], st["cy"] + st["cry"], full)
    put(st["mfx"], st["mfx"] + fx, full)
    put(st["mfy"], st["mfy"] + fy, full)
    damp_w_col = col["damp_w"][None, None, :nz]
    heat_s = np.zeros(s)
    dw = np.zeros(s)
    if not cfg["hydrostatic"]:
        dampw = (col["damp_w"][:nz] * damp["da_min_c"]) ** (col["nord_w"][:nz] + 1)
        nord4 = [int(col["nord_w"][b]) for b in range(4)]
        fx2, fy2 = delnflux_nosg(st["w"], del6_u, del6_v, g["rarea"], nord4, dampw, t)
        # heat_diss (d_sw.py:271-285) on the compute domain
        dwv = (fx2 - sh(fx2, 1, 0) + fy2 - sh(fy2, 0, 1)) * rarea
        dd8 = col["ke_bg"][None, None, :nz] * abs(dt)
        hs = dd8 - dwv * (st["w"] + 0.5 * dwv)
        on = np.broadcast_to(damp_w_col > 1e-5, s) & comp[:, :, None]
        dw[on] = dwv[on]
        heat_s[on] = hs[on]
        st["diss_est"][on] = (st["diss_est"] + hs)[on]
        gx, gy = transport(st["w"], cfg["hord_vt"], None, None, mfx=fx, mfy=fy)
        put(st["w"], st["w"] * st["delp"] + _flux_increment(gx, gy, rarea), comp)
    gx, gy = transport(st["q_con"], cfg["hord_dp"], "nord_t", "damp_t", mfx=fx, mfy=fy,
                       mass=st["delp"])
    put(st["q_con"], st["q_con"] * st["delp"] + _flux_increment(gx, gy, rarea), comp)
    gx, gy = transport(st["pt"], cfg["hord_tm"], "nord_v", "damp_vt", mfx=fx, mfy=fy,
                       mass=st["delp"])
    # apply_pt_delp_fluxes on the compute domain
    ptn = st["pt"] * st["delp"] + _flux_increment(gx, gy, rarea)
    dpn = st["delp"] + _flux_increment(fx, fy, rarea)
    put(st["pt"], ptn / dpn, comp)
    put(st["delp"], dpn, comp)
    ke, _, _ = kinetic_energy(st["uc"], st["vc"], st["u"], st["v"], uc_contra, vc_contra, dt,
                              cfg["hord_mt"], g, t)
    # compute_vorticity on the full domain
    dx, dy = _g(g, "dx"), _g(g, "dy")
    u, v = st["u"], st["v"]
    wkv = (u - sh(u, 0, 1) * sh(dx, 0, 1) / dx) * (rarea * dx) + (
        sh(v, 1, 0) * sh(dy, 1, 0) / dy - v) * (rarea * dy)
    wk = np.zeros(s)
    put(wk, wkv, full)
    # adjust_w_and_qcon on the compute domain
    w_new = st["w"] / st["delp"]
    w_new = np.where(damp_w_col > 1e-5, w_new + dw, w_new)
    put(st["w"], w_new, comp)
    put(st["q_con"], st["q_con"] / st["delp"], comp)
    vort = np.zeros(s)
    divergence_damping(u, v, st["va"], st["ptc"], vort, st["ua"], st["divgd"], st["vc"],
                       st["uc"], st["delpc"], ke, wk, dt, g, t, col["nord"], col["d2_divg"],
                       cfg["dddmp"], cfg["d4_bg"], damp["da_min_c"], cfg["nord"])
    # vort_differencing
    dcon = col["d_con"][None, None, :nz] > DCON_THRESHOLD
    vxd, vyd = np.zeros(s), np.zeros(s)
    vx = vort - sh(vort, 1, 0)
    vy = vort - sh(vort, 0, 1)
    mx = box(s2, t.isc, t.iec, t.jsc, t.jec + 1)[:, :, None] & dcon
    my = box(s2, t.isc, t.iec + 1, t.jsc, t.jec)[:, :, None] & dcon
    vxd[mx] = vx[mx]
    vyd[my] = vy[my]
    # compute_vort on the full domain
    put(vort, wk + _g(g, "f0"), full)
    fxv, fyv = transport(vort, cfg["hord_vt"], None, None)
    # u_and_v_from_ke
    put(st["u"], st["u"] * dx + ke - sh(ke, 1, 0) + fyv, box(s2, t.isc, t.iec, t.jsc, t.jec + 1))
    put(st["v"], st["v"] * dy + ke - sh(ke, 0, 1) - fxv, box(s2, t.isc, t.iec + 1, t.jsc, t.jec))
    # delnflux_nosg_v on wk
    dampv = (col["damp_vt"][:nz] * damp["da_min_c"]) ** (col["nord_v"][:nz] + 1)
    nord4 = [int(col["nord_v"][b]) for b in range(4)]
    ut, vt = delnflux_nosg(wk, del6_u, del6_v, g["rarea"], nord4, dampv, t)
    # heat_source_from_vorticity_damping
    rdx, rdy = _g(g, "rdx"), _g(g, "rdy")
    u, v = st["u"], st["v"]
    ubt = (vxd + vt) * rdx
    fy_ = u * rdx
    gy = fy_ * ubt
    vbt = (vyd - ut) * rdy
    fx_ = v * rdy
    gx = fx_ * vbt
    kef = col["d_con"][None, None, :nz]
    u2 = fy_ + sh(fy_, 0, 1)
    du2 = ubt + sh(ubt, 0, 1)
    v2 = fx_ + sh(fx_, 1, 0)
    dv2 = vbt + sh(vbt, 1, 0)
    dampterm = _g(g, "rsin2") * (
        (ubt * ubt + sh(ubt, 0, 1) * sh(ubt, 0, 1) + vbt * vbt + sh(vbt, 1, 0) * sh(vbt, 1, 0))
        + 2.0 * (gy + sh(gy, 0, 1) + gx + sh(gx, 1, 0))
        - _g(g, "cosa_s") * (u2 * dv2 + v2 * du2 + du2 * dv2))
    hs_new = st["delp"] * (heat_s - 0.25 * kef * dampterm)
    use = np.broadcast_to(kef > DCON_THRESHOLD, s) | cfg.get("do_skeb", False)
    heat_s = np.where(use, hs_new, heat_s)
    if cfg["d_con"] > DCON_THRESHOLD or cfg.get("do_skeb", False):
        put(st["heat_source"], st["heat_source"] + heat_s, comp)
        if cfg.get("do_skeb", False):
            put(st["diss_est"], st["diss_est"] - dampterm, comp)
    # update_u_and_v
    on = col["damp_vt"][None, None, :nz] > 1e-5
    mu = box(s2, t.isc, t.iec, t.jsc, t.jec + 1)[:, :, None] & on
    mv = box(s2, t.isc, t.iec + 1, t.jsc, t.jec)[:, :, None] & on
    st["u"][mu] = (st["u"] + vt)[mu]
    st["v"][mv] = (st["v"] - ut)[mv]
    return st
