"""CPU oracle for the C-grid half step: d2a2c_vect and c_sw.

Test infrastructure only (see oracle/common.py).  Follows
fv3core/stencils/d2a2c_vect.py:1-651 and fv3core/stencils/c_sw.py:1-659
statement by statement, with the corner fills of fv3core/utils/corners.py:
127-300.  Fields are ``[i, j, k]`` arrays over the storage index space,
metric terms ``[i, j]``; ``g`` is a dict of metric arrays and ``t`` the
:class:`oracle.fvtp2d.Tile` of the call.
"""
from __future__ import annotations

import numpy as np

from .common import box, sh

# fv3core/stencils/a2b_ord4.py:26-34 (Lagrange weights)
A1 = 9.0 / 16.0
A2 = -1.0 / 16.0
# fv3core/stencils/d2a2c_vect.py:12-15
VC1 = -2.0 / 14.0
VC2 = 11.0 / 14.0
VC3 = 5.0 / 14.0
OFFSET = 2
BIG_NUMBER = 1e30


def _g(g, name):
    return g[name][:, :, None]


def put(dst, val, mask2):
    """``dst[mask] = val[mask]`` for a 2-D mask over the first two axes."""
    dst[mask2] = np.broadcast_to(val, dst.shape)[mask2]


def region(shape, t, i, j, dom=None):
    """Mask of ``region[i, j]`` restricted to the stencil domain ``dom``.

    ``i``/``j``: an absolute index, an inclusive ``(lo, hi)`` pair, or None
    when the anchor it derives from is absent (tile not on that edge).
    """
    if i is None or j is None:
        return np.zeros(shape[:2], dtype=bool)
    ilo, ihi = (i, i) if isinstance(i, (int, np.integer)) else i
    jlo, jhi = (j, j) if isinstance(j, (int, np.integer)) else j
    m = box(shape[:2], ilo, ihi, jlo, jhi)
    if dom is not None:
        m &= box(shape[:2], *dom)
    return m


def _off(a, d):
    return None if a is None else a + d


# --------------------------------------------------------------- corner fills
# Each entry: (corner, anchor_i, di, anchor_j, dj, src_di, src_dj); the
# destination is (anchor_i + di, anchor_j + dj) and the source sits at
# (src_di, src_dj) from it (corners.py:127-300).
_FILL2_X = [
    ("sw", "is", -1, "js", -1, 0, 1), ("sw", "is", -2, "js", -1, 1, 2),
    ("se", "ie", 1, "js", -1, 0, 1), ("se", "ie", 2, "js", -1, -1, 2),
    ("nw", "is", -1, "je", 1, 0, -1), ("nw", "is", -2, "je", 1, 1, -2),
    ("ne", "ie", 1, "je", 1, 0, -1), ("ne", "ie", 2, "je", 1, -1, -2),
]
_FILL3_X = _FILL2_X + [
    ("sw", "is", -3, "js", -1, 2, 3), ("se", "ie", 3, "js", -1, -2, 3),
    ("nw", "is", -3, "je", 1, 2, -3), ("ne", "ie", 3, "je", 1, -2, -3),
]
_FILL2_Y = [
    ("sw", "is", -1, "js", -1, 1, 0), ("sw", "is", -1, "js", -2, 2, 1),
    ("se", "ie", 1, "js", -1, -1, 0), ("se", "ie", 1, "js", -2, -2, 1),
    ("nw", "is", -1, "je", 1, 1, 0), ("nw", "is", -1, "je", 2, 2, -1),
    ("ne", "ie", 1, "je", 1, -1, 0), ("ne", "ie", 1, "je", 2, -2, -1),
]
_FILL3_Y = _FILL2_Y + [
    ("sw", "is", -1, "js", -3, 3, 2), ("se", "ie", 1, "js", -3, -3, 2),
    ("nw", "is", -1, "je", 3, 3, -2), ("ne", "ie", 1, "je", 3, -3, -2),
]
FILL_TABLES = {"2x": _FILL2_X, "3x": _FILL3_X, "2y": _FILL2_Y, "3y": _FILL3_Y}


def fill_corners(q, qc, t, kind, mult=(1.0, 1.0, 1.0, 1.0), dom=None):
    """fill_corners_{2,3}cells_mult_{x,y}: ``q`` updated in place from ``qc``.

    ``mult`` = (sw, se, nw, ne).  Statements run in the reference's order,
    so a source written by an earlier statement is read after the write.
    """
    anchors = {"is": t.i_start, "ie": t.i_end, "js": t.j_start, "je": t.j_end}
    mults = dict(zip(("sw", "se", "nw", "ne"), mult))
    for corner, ai, di, aj, dj, sdi, sdj in FILL_TABLES[kind]:
        if anchors[ai] is None or anchors[aj] is None:
            continue
        i, j = anchors[ai] + di, anchors[aj] + dj
        if dom is not None and not (dom[0] <= i <= dom[1] and dom[2] <= j <= dom[3]):
            continue
        q[i, j] = mults[corner] * qc[i + sdi, j + sdj]
    return q


# ------------------------------------------------------------------ d2a2c_vect
def contravariant(v1, v2, cosa, rsin2):
    """d2a2c_vect.py:142-164."""
    return (v1 - v2 * cosa) * rsin2


def _lagrange_x(q):
    return A2 * (sh(q, -2, 0) + sh(q, 1, 0)) + A1 * (sh(q, -1, 0) + q)


def _lagrange_y(q):
    return A2 * (sh(q, 0, -2) + sh(q, 0, 1)) + A1 * (sh(q, 0, -1) + q)


def _vol_x(u):
    return VC1 * sh(u, -2, 0) + VC2 * sh(u, -1, 0) + VC3 * u


def _vol_x_rev(u):
    return VC1 * sh(u, 1, 0) + VC2 * u + VC3 * sh(u, -1, 0)


def _vol_y(v):
    return VC1 * sh(v, 0, -2) + VC2 * sh(v, 0, -1) + VC3 * v


def _vol_y_rev(v):
    return VC1 * sh(v, 0, 1) + VC2 * v + VC3 * sh(v, 0, -1)


def _edge4_x(ua, dxa):
    """edge_interpolate4_x (d2a2c_vect.py:254-260), including its use of t1 in n2."""
    t1 = sh(dxa, -2, 0) + sh(dxa, -1, 0)
    t2 = dxa + sh(dxa, 1, 0)
    n1 = (t1 + sh(dxa, -1, 0)) * sh(ua, -1, 0) - sh(dxa, -1, 0) * sh(ua, -2, 0)
    n2 = (t1 + dxa) * ua - dxa * sh(ua, 1, 0)
    return 0.5 * (n1 / t1 + n2 / t2)


def _edge4_y(va, dya):
    t1 = sh(dya, 0, -2) + sh(dya, 0, -1)
    t2 = dya + sh(dya, 0, 1)
    n1 = (t1 + sh(dya, 0, -1)) * sh(va, 0, -1) - sh(dya, 0, -1) * sh(va, 0, -2)
    n2 = (t1 + dya) * va - dya * sh(va, 0, 1)
    return 0.5 * (n1 / t1 + n2 / t2)


def d2a2c_bounds(t, nested=False, dord4=True):
    """Index bounds of DGrid2AGrid2CGridVectors.__init__ (d2a2c_vect.py:374-470)."""
    nx, ny = t.iec + 1, t.jec + 1
    npt = 4 if not nested else 0
    if npt > t.nx - 1 or npt > t.ny - 1:
        npt = 0
    b = dict(npt=npt)
    b["js1"] = npt + OFFSET if t.south else t.jsc - 1
    b["je1"] = ny - npt if t.north else t.jec + 1
    b["is1"] = npt + OFFSET if t.west else t.isd
    b["ie1"] = nx - npt if t.east else t.ied
    b["is2"] = npt + OFFSET if t.west else t.isc - 1
    b["ie2"] = nx - npt if t.east else t.iec + 1
    b["js2"] = npt + OFFSET if t.south else t.jsd
    b["je2"] = ny - npt if t.north else t.jed
    b["ifirst"] = t.isc + 2 if t.west else t.isc - 1
    b["ilast"] = t.iec - 1 if t.east else t.iec + 2
    b["jfirst"] = t.jsc + 2 if t.south else t.jsc - 1
    b["jlast"] = t.jec - 1 if t.north else t.jec + 2
    b["avg_offset"] = -1 if npt == 0 else 3
    b["id"] = 1 if dord4 else 0
    return b


def d2a2c_vect(uc, vc, u, v, ua, va, utc, vtc, g, t):
    """DGrid2AGrid2CGridVectors.__call__ (d2a2c_vect.py:573-651), in place.

    Returns the temporaries (utmp, vtmp) for inspection.
    """
    shape = u.shape
    b = d2a2c_bounds(t)
    full = (t.isd, t.ied, t.jsd, t.jed)
    # set_tmps on the full domain (compute_halos 3)
    utmp = np.full(shape, np.nan)
    vtmp = np.full(shape, np.nan)
    put(utmp, BIG_NUMBER, box(shape[:2], *full))
    put(vtmp, BIG_NUMBER, box(shape[:2], *full))
    # lagrange_interpolation_y_p1 / x_p1 (d2a2c_vect.py:117-138)
    put(utmp, A2 * (sh(u, 0, -1) + sh(u, 0, 2)) + A1 * (u + sh(u, 0, 1)),
        box(shape[:2], b["is1"], b["ie1"], b["js1"], b["je1"]))
    put(vtmp, A2 * (sh(v, -1, 0) + sh(v, 2, 0)) + A1 * (v + sh(v, 1, 0)),
        box(shape[:2], b["is2"], b["ie2"], b["js2"], b["je2"]))
    # avg_box (d2a2c_vect.py:141-163): bands of width D2A2C_AVG_OFFSET at the edges
    off = b["avg_offset"]
    m = np.zeros(shape[:2], dtype=bool)
    if t.south:
        m |= region(shape, t, (-10**6, 10**6), (-10**6, t.jsc + off - 1))
    if t.north:
        m |= region(shape, t, (-10**6, 10**6), (t.jec - off + 1, 10**6))
    if t.west:
        m |= region(shape, t, (-10**6, t.isc + off - 1), (-10**6, 10**6))
    if t.east:
        m |= region(shape, t, (t.iec - off + 1, 10**6), (-10**6, 10**6))
    m &= box(shape[:2], *full)
    put(utmp, 0.5 * (u + sh(u, 0, 1)), m)
    put(vtmp, 0.5 * (v + sh(v, 1, 0)), m)
    # contravariant_components on compute + (1 + id) halo
    h = 1 + b["id"]
    m = box(shape[:2], t.isc - h, t.iec + h, t.jsc - h, t.jec + h)
    cosa_s, rsin2 = _g(g, "cosa_s"), _g(g, "rsin2")
    put(ua, contravariant(utmp, vtmp, cosa_s, rsin2), m)
    put(va, contravariant(vtmp, utmp, cosa_s, rsin2), m)
    edges = (t.isc - 3, t.iec + 3, t.jsc - 3, t.jec + 3)
    mults = (-1.0, 1.0, 1.0, -1.0)  # sw, se, nw, ne
    # fill_corners_x
    fill_corners(utmp, vtmp, t, "3x", mults, edges)
    fill_corners(ua, va, t, "2x", mults, edges)
    # ut_main
    cosa_u, rsin_u = _g(g, "cosa_u"), _g(g, "rsin_u")
    m = box(shape[:2], b["ifirst"], b["ilast"], t.jsc - 1, t.jec + 1)
    ucv = _lagrange_x(utmp)
    put(uc, ucv, m)
    put(utc, contravariant(uc, v, cosa_u, rsin_u), m)
    # east_west_edges (d2a2c_vect.py:43-98)
    jr = (t.jsc - 1, t.jec + 1)
    s1, s3 = _g(g, "sin_sg1"), _g(g, "sin_sg3")
    dxa = g["dxa"]

    def edge_uc(i):
        mm = region(shape, t, i, jr, edges)
        put(utc, _edge4_x(ua, dxa[:, :, None]), mm)
        put(uc, np.where(utc > 0, utc * sh(s3, -1, 0), utc * s1), mm)

    for a, east in ((t.i_start, False), (t.i_end, True)):
        if a is None:
            continue
        first = a - 1 if not east else a        # vol_conserv (forward)
        mid = a if not east else a + 1          # edge interpolation
        last = a + 1 if not east else a + 2     # vol_conserv (reverse)
        put(uc, _vol_x(utmp), region(shape, t, first, jr, edges))
        edge_uc(mid)
        put(uc, _vol_x_rev(utmp), region(shape, t, last, jr, edges))
        put(utc, contravariant(uc, v, cosa_u, rsin_u), region(shape, t, first, jr, edges))
        put(utc, contravariant(uc, v, cosa_u, rsin_u), region(shape, t, last, jr, edges))
    # fill_corners_y
    fill_corners(vtmp, utmp, t, "3y", mults, edges)
    fill_corners(va, ua, t, "2y", mults, edges)
    # north_south_edges (d2a2c_vect.py:101-114 in the source numbering)
    cosa_v, rsin_v = _g(g, "cosa_v"), _g(g, "rsin_v")
    s2, s4 = _g(g, "sin_sg2"), _g(g, "sin_sg4")
    dya = g["dya"]
    ir = (t.isc - 1, t.iec + 1)
    mm = region(shape, t, ir, (t.jsc - 1, t.jec + 2), edges)
    put(vc, _lagrange_y(vtmp), mm)
    put(vtc, contravariant(vc, u, cosa_v, rsin_v), mm)
    for a, north in ((t.j_start, False), (t.j_end, True)):
        if a is None:
            continue
        first = a - 1 if not north else a
        mid = a if not north else a + 1
        last = a + 1 if not north else a + 2
        mm = region(shape, t, ir, first, edges)
        put(vc, _vol_y(vtmp), mm)
        put(vtc, contravariant(vc, u, cosa_v, rsin_v), mm)
        mm = region(shape, t, ir, mid, edges)
        put(vtc, _edge4_y(va, dya[:, :, None]), mm)
        put(vc, np.where(vtc > 0, vtc * sh(s4, 0, -1), vtc * s2), mm)
        mm = region(shape, t, ir, last, edges)
        put(vc, _vol_y_rev(vtmp), mm)
        put(vtc, contravariant(vc, u, cosa_v, rsin_v), mm)
    # vt_main
    m = box(shape[:2], t.isc - 1, t.iec + 1, b["jfirst"], b["jlast"])
    put(vc, _lagrange_y(vtmp), m)
    put(vtc, contravariant(vc, u, cosa_v, rsin_v), m)
    return utmp, vtmp


# ------------------------------------------------------------------------ c_sw
def divergence_corner(u, v, ua, va, g, t):
    """divergence_corner (c_sw.py:172-289) on the compute corner points."""
    dxc, dyc = _g(g, "dxc"), _g(g, "dyc")
    s1, s2, s3, s4 = (_g(g, f"sin_sg{k}") for k in (1, 2, 3, 4))
    c1, c2, c3, c4 = (_g(g, f"cos_sg{k}") for k in (1, 2, 3, 4))
    rarea_c = _g(g, "rarea_c")
    shape = u.shape
    dom = (t.isc, t.iec + 1, t.jsc, t.jec + 1)
    out = np.full(shape, np.nan)
    uf = (u - 0.25 * (sh(va, 0, -1) + va) * (sh(c4, 0, -1) + c2)) * dyc * 0.5 * (sh(s4, 0, -1) + s2)
    vf = (v - 0.25 * (sh(ua, -1, 0) + ua) * (sh(c3, -1, 0) + c1)) * dxc * 0.5 * (sh(s3, -1, 0) + s1)
    put(out, (sh(vf, 0, -1) - vf + sh(uf, -1, 0) - uf) * rarea_c, box(shape[:2], *dom))
    vf0 = v * dxc * 0.5 * (sh(s3, -1, 0) + s1)
    vf1 = sh(v, 0, -1) * sh(dxc, 0, -1) * 0.5 * (sh(s3, -1, -1) + sh(s1, 0, -1))
    uf1 = ((sh(u, -1, 0) - 0.25 * (sh(va, -1, -1) + sh(va, -1, 0)) * (sh(c4, -1, -1) + sh(c2, -1, 0)))
           * sh(dyc, -1, 0) * 0.5 * (sh(s4, -1, -1) + sh(s2, -1, 0)))
    ie = _off(t.i_end, 1)
    je = _off(t.j_end, 1)
    mm = region(shape, t, t.i_start, (-10**6, 10**6), dom) | region(shape, t, ie, (-10**6, 10**6), dom)
    put(out, (vf1 - vf0 + uf1 - uf) * rarea_c, mm)
    uf0 = u * dyc * 0.5 * (sh(s4, 0, -1) + s2)
    uf1 = sh(u, -1, 0) * sh(dyc, -1, 0) * 0.5 * (sh(s4, -1, -1) + sh(s2, -1, 0))
    vf1 = ((sh(v, 0, -1) - 0.25 * (sh(ua, -1, -1) + sh(ua, 0, -1)) * (sh(c3, -1, -1) + sh(c1, 0, -1)))
           * sh(dxc, 0, -1) * 0.5 * (sh(s3, -1, -1) + sh(s1, 0, -1)))
    mm = region(shape, t, (-10**6, 10**6), t.j_start, dom) | region(shape, t, (-10**6, 10**6), je, dom)
    put(out, (vf1 - vf + uf1 - uf0) * rarea_c, mm)
    uf1 = sh(u, -1, 0) * sh(dyc, -1, 0) * 0.5 * (sh(s4, -1, -1) + sh(s2, -1, 0))
    vf0 = v * dxc * 0.5 * (sh(s3, -1, 0) + s1)
    uf0 = u * dyc * 0.5 * (sh(s4, 0, -1) + s2)
    mm = region(shape, t, t.i_start, t.j_start, dom) | region(shape, t, ie, t.j_start, dom)
    put(out, (-vf0 + uf1 - uf0) * rarea_c, mm)
    vf1 = sh(v, 0, -1) * sh(dxc, 0, -1) * 0.5 * (sh(s3, -1, -1) + sh(s1, 0, -1))
    mm = region(shape, t, ie, je, dom) | region(shape, t, t.i_start, je, dom)
    put(out, (vf1 + uf1 - uf0) * rarea_c, mm)
    return out


def c_sw(delp, pt, u, v, w, uc, vc, ua, va, ut, vt, divgd, omga, dt2, g, t, nord):
    """CGridShallowWaterDynamics.__call__ (c_sw.py:534-659).

    Arguments are copied; returns a dict of every field after the call
    (including the in-place corner fills of delp, pt and w) plus delpc, ptc.
    """
    cp = {k: np.array(a, dtype=np.float64, copy=True) for k, a in dict(
        delp=delp, pt=pt, u=u, v=v, w=w, uc=uc, vc=vc, ua=ua, va=va, ut=ut, vt=vt,
        divgd=divgd, omga=omga).items()}
    delp, pt, u, v, w = cp["delp"], cp["pt"], cp["u"], cp["v"], cp["w"]
    uc, vc, ua, va, ut, vt = cp["uc"], cp["vc"], cp["ua"], cp["va"], cp["ut"], cp["vt"]
    divgd, omga = cp["divgd"], cp["omga"]
    shape = delp.shape
    s2 = shape[:2]
    delpc = np.zeros(shape)
    ptc = np.zeros(shape)
    d2a2c_vect(uc, vc, u, v, ua, va, ut, vt, g, t)
    if nord > 0:
        put(divgd, divergence_corner(u, v, ua, va, g, t), box(s2, t.isc, t.iec + 1, t.jsc, t.jec + 1))
    s1, s2g, s3, s4 = (_g(g, f"sin_sg{k}") for k in (1, 2, 3, 4))
    c1, c2, c3, c4 = (_g(g, f"cos_sg{k}") for k in (1, 2, 3, 4))
    dx, dy = _g(g, "dx"), _g(g, "dy")
    # geoadjust_ut / geoadjust_vt (c_sw.py:18-37)
    m = box(s2, t.isc - 1, t.iec + 2, t.jsc - 1, t.jec + 1)
    put(ut, np.where(ut > 0, dt2 * ut * dy * sh(s3, -1, 0), dt2 * ut * dy * s1), m)
    m = box(s2, t.isc - 1, t.iec + 1, t.jsc - 1, t.jec + 2)
    put(vt, np.where(vt > 0, dt2 * vt * dx * sh(s4, 0, -1), dt2 * vt * dx * s2g), m)
    full = (t.isc - 3, t.iec + 3, t.jsc - 3, t.jec + 3)
    for q in (delp, pt, w):
        fill_corners(q, q, t, "2x", dom=full)
    # compute_nonhydro_fluxes_x (c_sw.py:57-72)
    m = box(s2, t.isc - 1, t.iec + 2, t.jsc - 1, t.jec + 1)
    up = ut > 0.0
    fx1 = np.full(shape, np.nan)
    fx = np.full(shape, np.nan)
    fx2 = np.full(shape, np.nan)
    put(fx1, ut * np.where(up, sh(delp, -1, 0), delp), m)
    put(fx, fx1 * np.where(up, sh(pt, -1, 0), pt), m)
    put(fx2, fx1 * np.where(up, sh(w, -1, 0), w), m)
    for q in (delp, pt, w):
        fill_corners(q, q, t, "2y", dom=full)
    # transportdelp_update_vorticity_and_kineticenergy (c_sw.py:75-169)
    dom = (t.isc - 1, t.iec + 1, t.jsc - 1, t.jec + 1)
    m = box(s2, *dom)
    rarea = _g(g, "rarea")
    vp = vt > 0.0
    fy1 = vt * np.where(vp, sh(delp, 0, -1), delp)
    fy = fy1 * np.where(vp, sh(pt, 0, -1), pt)
    fy2 = fy1 * np.where(vp, sh(w, 0, -1), w)
    dpc = delp + (fx1 - sh(fx1, 1, 0) + fy1 - sh(fy1, 0, 1)) * rarea
    put(delpc, dpc, m)
    put(ptc, (pt * delp + (fx - sh(fx, 1, 0) + fy - sh(fy, 0, 1)) * rarea) / dpc, m)
    put(omga, (w * delp + (fx2 - sh(fx2, 1, 0) + fy2 - sh(fy2, 0, 1)) * rarea) / dpc, m)
    ke = np.where(ua > 0.0, uc, sh(uc, 1, 0))
    vort = np.where(va > 0.0, vc, sh(vc, 0, 1))
    allr = (-10**6, 10**6)
    mm = region(shape, t, allr, _off(t.j_start, -1), dom) | region(shape, t, allr, t.j_end, dom)
    put(vort, np.where(va <= 0.0, vort * s4 + sh(u, 0, 1) * c4, vort), mm)
    mm = region(shape, t, allr, t.j_start, dom) | region(shape, t, allr, _off(t.j_end, 1), dom)
    put(vort, np.where(va > 0.0, vort * s2g + u * c2, vort), mm)
    mm = region(shape, t, t.i_end, allr, dom) | region(shape, t, _off(t.i_start, -1), allr, dom)
    put(ke, np.where(ua <= 0.0, ke * s3 + sh(v, 1, 0) * c3, ke), mm)
    mm = region(shape, t, _off(t.i_end, 1), allr, dom) | region(shape, t, t.i_start, allr, dom)
    put(ke, np.where(ua > 0.0, ke * s1 + v * c1, ke), mm)
    ke = 0.5 * dt2 * (ua * ke + va * vort)
    ke_out = np.full(shape, np.nan)
    put(ke_out, ke, m)
    # circulation_cgrid + absolute_vorticity (c_sw.py:292-322, 40-42)
    dom = (t.isc, t.iec + 1, t.jsc, t.jec + 1)
    dxc, dyc = _g(g, "dxc"), _g(g, "dyc")
    fxc = dxc * uc
    fyc = dyc * vc
    fxc1 = sh(dxc, 0, -1) * sh(uc, 0, -1)
    fyc1 = sh(dyc, -1, 0) * sh(vc, -1, 0)
    vc_ = np.full(shape, np.nan)
    put(vc_, fxc1 - fxc - fyc1 + fyc, box(s2, *dom))
    ie1, je1 = _off(t.i_end, 1), _off(t.j_end, 1)
    mm = region(shape, t, t.i_start, t.j_start, dom) | region(shape, t, t.i_start, je1, dom)
    put(vc_, fxc1 - fxc + fyc, mm)
    mm = region(shape, t, ie1, t.j_start, dom) | region(shape, t, ie1, je1, dom)
    put(vc_, fxc1 - fxc - fyc1, mm)
    vort_c = np.full(shape, np.nan)
    put(vort_c, _g(g, "fC") + _g(g, "rarea_c") * vc_, box(s2, *dom))
    # update_y_velocity then update_x_velocity (c_sw.py:325-370)
    dom = (t.isc, t.iec, t.jsc, t.jec + 1)
    tmp = dt2 * (u - vc * _g(g, "cosa_v")) / _g(g, "sina_v")
    allr = (-10**6, 10**6)
    mm = region(shape, t, allr, t.j_start, dom) | region(shape, t, allr, je1, dom)
    tmp = np.where(mm[:, :, None], dt2 * u, tmp)
    flux = np.where(tmp > 0.0, vort_c, sh(vort_c, 1, 0))
    vcn = vc - tmp * flux + _g(g, "rdyc") * (sh(ke_out, 0, -1) - ke_out)
    put(vc, vcn, box(s2, *dom))
    dom = (t.isc, t.iec + 1, t.jsc, t.jec)
    tmp = dt2 * (v - uc * _g(g, "cosa_u")) / _g(g, "sina_u")
    mm = region(shape, t, t.i_start, allr, dom) | region(shape, t, ie1, allr, dom)
    tmp = np.where(mm[:, :, None], dt2 * v, tmp)
    flux = np.where(tmp > 0.0, vort_c, sh(vort_c, 0, 1))
    ucn = uc + tmp * flux + _g(g, "rdxc") * (sh(ke_out, -1, 0) - ke_out)
    put(uc, ucn, box(s2, *dom))
    cp.update(delpc=delpc, ptc=ptc, ke=ke_out, vort=vort_c)
    return cp
