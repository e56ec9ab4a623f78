"""CPU oracle for the rest of the acoustic step and its driver.

Test infrastructure only (see oracle/common.py).  Follows
fv3core/stencils/updatedzc.py, updatedzd.py, nh_p_grad.py, pk3_halo.py,
pe_halo.py, ray_fast.py, del2cubed.py, temperature_adjust.py and
dyn_core.py statement by statement.  Arrays are ``[i, j, k]``; K columns are
1-D; ``g`` holds the metric terms of one tile and ``t`` its Tile.
"""
from __future__ import annotations

import math

import numpy as np

from .common import box, finite_on, sh, shk, sign
from .csw import fill_corners, put, region
from .dsw import a2b_ord4
from .fvtp2d import copy_corners_x, copy_corners_y, copy_corners_xy, delnflux_nosg, fvtp2d

# fv3core/utils/global_constants.py
GRAV = 9.80665
RDGAS = 287.05
CP_AIR = 1004.6
KAPPA = RDGAS / CP_AIR
DZ_MIN = 2.0
CV_AIR = CP_AIR - RDGAS
RDG = -RDGAS / GRAV
CNST_0P20 = 0.2
SDAY = 86400.0


def _g(g, name):
    return g[name][:, :, None]


def _k(a, nk):
    return np.asarray(a, dtype=np.float64)[None, None, :nk]


# ----------------------------------------------------------------- updatedzc
def update_dz_c(dp_ref, zs, area, ut, vt, gz, ws, dt, t):
    """UpdateGeopotentialHeightOnCGrid.__call__ (updatedzc.py:122-163).

    gz has nz + 1 levels and is updated in place on [isc-1, iec+1]^2; ws (IJ)
    gets the surface value.  dp_ref is the K column of layer thicknesses.
    """
    s = gz.shape
    nk = s[2]  # nz + 1
    gz_x = gz.copy()
    gz_y = gz.copy()
    dom = (t.isd, t.ied, t.jsd, t.jed)
    fill_corners(gz_x, gz_x, t, "2x", dom=dom)
    fill_corners(gz_y, gz_y, t, "2y", dom=dom)
    dp0 = np.asarray(dp_ref, dtype=np.float64)

    def interfaces(vel):
        out = np.full(s, np.nan)
        # top: vel + (vel - vel[k+1]) * dp0 / (dp0 + dp0[1])
        ratio = dp0[0] / (dp0[0] + dp0[1])
        out[..., 0] = vel[..., 0] + (vel[..., 0] - vel[..., 1]) * ratio
        for k in range(1, nk - 1):
            ir = 1.0 / (dp0[k - 1] + dp0[k])
            out[..., k] = (dp0[k] * vel[..., k - 1] + dp0[k - 1] * vel[..., k]) * ir
        k = nk - 1
        ratio = dp0[k - 1] / (dp0[k - 2] + dp0[k - 1])
        out[..., k] = vel[..., k - 1] + (vel[..., k - 1] - vel[..., k - 2]) * ratio
        return out

    xfx = interfaces(ut)
    yfx = interfaces(vt)
    fx = xfx * np.where(xfx > 0.0, sh(gz_x, -1, 0), gz_x)
    fy = yfx * np.where(yfx > 0.0, sh(gz_y, 0, -1), gz_y)
    a3 = area[:, :, None]
    new = (gz_y * a3 + fx - sh(fx, 1, 0) + fy - sh(fy, 0, 1)) / (
        a3 + xfx - sh(xfx, 1, 0) + yfx - sh(yfx, 0, 1))
    m = box(s[:2], t.isc - 1, t.iec + 1, t.jsc - 1, t.jec + 1)
    put(gz, new, m)
    rdt = 1.0 / dt
    ws[m] = ((zs - gz[..., nk - 1]) * rdt)[m]
    for k in range(nk - 2, -1, -1):
        kp1 = gz[..., k + 1] + DZ_MIN
        gz[..., k] = np.where(m, np.where(gz[..., k] > kp1, gz[..., k], kp1), gz[..., k])


# ----------------------------------------------------------------- updatedzd
def spline_constants(dp0, nk):
    """cubic_spline_interpolation_constants (updatedzd.py:121-145) on nk levels."""
    gk, beta, gamma = np.zeros(nk), np.zeros(nk), np.zeros(nk)
    gk[0] = dp0[1] / dp0[0]
    beta[0] = gk[0] * (gk[0] + 0.5)
    gamma[0] = (1.0 + gk[0] * (gk[0] + 1.5)) / beta[0]
    for k in range(1, nk - 1):
        gk[k] = dp0[k - 1] / dp0[k]
        beta[k] = 2.0 + 2.0 * gk[k] - gamma[k - 1]
        gamma[k] = gk[k] / beta[k]
    return gk, beta, gamma


def spline_to_interfaces(qc, gk, beta, gamma):
    """cubic_spline_interpolation_from_layer_center_to_interfaces (updatedzd.py:148-187)."""
    nk = qc.shape[2]
    qi = np.zeros(qc.shape)
    xt1 = 2.0 * gk[0] * (gk[0] + 1.0)
    qi[..., 0] = (xt1 * qc[..., 0] + qc[..., 1]) / beta[0]
    for k in range(1, nk - 1):
        qi[..., k] = (3.0 * (qc[..., k - 1] + gk[k] * qc[..., k]) - qi[..., k - 1]) / beta[k]
    k = nk - 1
    a_bot = 1.0 + gk[k - 1] * (gk[k - 1] + 1.5)
    xt1 = 2.0 * gk[k - 1] * (gk[k - 1] + 1.0)
    xt2 = gk[k - 1] * (gk[k - 1] + 0.5) - a_bot * gamma[k - 1]
    qi[..., k] = (xt1 * qc[..., k - 1] + qc[..., k - 2] - a_bot * qi[..., k - 1]) / xt2
    for k in range(nk - 2, -1, -1):
        qi[..., k] = qi[..., k] - gamma[k] * qi[..., k + 1]
    return qi


def update_dz_d(zs, height, crx, cry, xfx, yfx, ws, dt, g, t, hord_tm, nord_v, damp_vt,
                dp0):
    """UpdateHeightOnDGrid.__call__ (updatedzd.py:323-393); height (nz + 1
    levels) and ws (IJ) in place.  The inputs have nz + 1 levels."""
    s = height.shape
    nk = s[2]
    gk, beta, gamma = spline_constants(dp0, nk)
    crx_i = spline_to_interfaces(crx, gk, beta, gamma)
    xfx_i = spline_to_interfaces(xfx, gk, beta, gamma)
    cry_i = spline_to_interfaces(cry, gk, beta, gamma)
    yfx_i = spline_to_interfaces(yfx, gk, beta, gamma)
    qx, qy = copy_corners_xy(height, t)
    height[...] = qx
    fx, fy = fvtp2d(qx, qy, crx_i, cry_i, xfx_i, yfx_i, g["area"], g["dxa"], g["dya"], hord_tm,
                    t, check_domains=False)
    nord4 = [int(nord_v[b]) for b in range(4)]
    gxd, gyd = delnflux_nosg(height, g["del6_u"], g["del6_v"], g["rarea"], nord4,
                             np.asarray(damp_vt, dtype=np.float64)[:nk], t)
    a3 = g["area"][:, :, None]
    area_after = (a3 + xfx_i - sh(xfx_i, 1, 0)) + (a3 + yfx_i - sh(yfx_i, 0, 1)) - a3
    adv = (height * a3 + fx - sh(fx, 1, 0) + fy - sh(fy, 0, 1)) / area_after
    new = adv + (gxd - sh(gxd, 1, 0) + gyd - sh(gyd, 0, 1)) / a3
    m = box(s[:2], t.isc, t.iec, t.jsc, t.jec)
    put(height, new, m)
    ws[m] = ((zs - height[..., nk - 1]) / dt)[m]
    for k in range(nk - 2, -1, -1):
        other = height[..., k + 1] + DZ_MIN
        height[..., k] = np.where(m, np.where(height[..., k] > other, height[..., k], other),
                                  height[..., k])


# ------------------------------------------------------------------ nh_p_grad
def nh_p_grad(u, v, pp, gz, pk3, delp, dt, ptop, akap, g, t):
    """NonHydrostaticPressureGradient.__call__ (nh_p_grad.py:147-206); in place.

    pp, gz, pk3 have nz + 1 levels, u, v, delp nz (and more storage)."""
    nz = delp.shape[2]
    ptk = ptop ** akap
    wk1 = np.zeros(pp.shape)
    # a2b_k1 on levels 1..nz of pp and pk3 (replace), a2b_kbuffer on gz 0..nz
    a2b_ord4(pp[..., 1:nz + 1], wk1[..., 1:nz + 1], g, t, replace=True)
    a2b_ord4(pk3[..., 1:nz + 1], wk1[..., 1:nz + 1], g, t, replace=True)
    a2b_ord4(gz[..., :nz + 1], wk1[..., :nz + 1], g, t, replace=True)
    a2b_ord4(delp[..., :nz], wk1[..., :nz], g, t, replace=False)
    s = u.shape
    m = box(s[:2], t.isc, t.iec + 1, t.jsc, t.jec + 1)
    pp[..., 0][m] = 0.0
    pk3[..., 0][m] = ptk
    wk = np.full(pp.shape, np.nan)
    wk[..., :nz] = pk3[..., 1:nz + 1] - pk3[..., :nz]
    K = slice(0, nz)
    Kp = slice(1, nz + 1)
    rdx, rdy = _g(g, "rdx"), _g(g, "rdy")
    w, w1 = wk[..., K], wk1[..., K]
    gz0, gz1 = gz[..., K], gz[..., Kp]
    pk0, pk1 = pk3[..., K], pk3[..., Kp]
    pp0, pp1 = pp[..., K], pp[..., Kp]
    mu = box(s[:2], t.isc, t.iec, t.jsc, t.jec + 1)
    mv = box(s[:2], t.isc, t.iec + 1, t.jsc, t.jec)
    # a2b leaves wk1 at zero outside the B-grid box, so the divisions below
    # are only defined on the boxes they are written to
    with np.errstate(divide="ignore", invalid="ignore"):
        du = dt / (w + sh(w, 1, 0)) * ((gz1 - sh(gz0, 1, 0)) * (sh(pk1, 1, 0) - pk0)
                                       + (gz0 - sh(gz1, 1, 0)) * (pk1 - sh(pk0, 1, 0)))
        un = (u[..., K] + du + dt / (w1 + sh(w1, 1, 0)) * (
            (gz1 - sh(gz0, 1, 0)) * (sh(pp1, 1, 0) - pp0)
            + (gz0 - sh(gz1, 1, 0)) * (pp1 - sh(pp0, 1, 0))
        )) * rdx
        dv = dt / (w + sh(w, 0, 1)) * ((gz1 - sh(gz0, 0, 1)) * (sh(pk1, 0, 1) - pk0)
                                       + (gz0 - sh(gz1, 0, 1)) * (pk1 - sh(pk0, 0, 1)))
        vn = (v[..., K] + dv + dt / (w1 + sh(w1, 0, 1)) * (
            (gz1 - sh(gz0, 0, 1)) * (sh(pp1, 0, 1) - pp0)
            + (gz0 - sh(gz1, 0, 1)) * (pp1 - sh(pp0, 0, 1))
        )) * rdy
    uk = u[..., K]
    put(uk, finite_on(un, mu, "nh_p_grad u"), mu)
    vk = v[..., K]
    put(vk, finite_on(vn, mv, "nh_p_grad v"), mv)


# ------------------------------------------------------------ halo pressure
def _ring(shape, t, width):
    """Horizontal ring of the given width around the compute domain."""
    m = box(shape[:2], t.isc - width, t.iec + width, t.jsc - width, t.jec + width)
    m &= ~box(shape[:2], t.isc, t.iec, t.jsc, t.jec)
    return m


def pk3_halo(pk3, delp, ptop, akap, t):
    """PK3Halo.__call__ (pk3_halo.py:11-68): pk3 in place on the 2-wide ring."""
    s = pk3.shape
    m = _ring(s, t, 2)
    pe = np.full(s[:2], ptop)
    for k in range(1, s[2]):
        pe = pe + delp[..., k - 1]
        pk3[..., k][m] = (pe ** akap)[m]


def pe_halo(pe, delp, ptop, t):
    """edge_pe (pe_halo.py:6-33): pe in place on the 1-wide ring."""
    m = _ring(pe.shape, t, 1)
    pe[..., 0][m] = ptop
    for k in range(1, pe.shape[2]):
        pe[..., k][m] = (pe[..., k - 1] + delp[..., k - 1])[m]


# ------------------------------------------------------------------ p_grad_c
def p_grad_c(uc, vc, delpc, pkc, gz, dt2, g, t, hydrostatic=False):
    """p_grad_c_stencil (dyn_core.py:101-145) on [isc, iec+1]^2; in place."""
    nz = delpc.shape[2]
    K, Kp = slice(0, nz), slice(1, nz + 1)
    wk = (pkc[..., Kp] - pkc[..., K]) if hydrostatic else delpc[..., K]
    gz0, gz1, pk0, pk1 = gz[..., K], gz[..., Kp], pkc[..., K], pkc[..., Kp]
    rdxc, rdyc = _g(g, "rdxc"), _g(g, "rdyc")
    m = box(uc.shape[:2], t.isc, t.iec + 1, t.jsc, t.jec + 1)
    # wk is only defined where delpc (pkc) was formed
    with np.errstate(divide="ignore", invalid="ignore"):
        un = uc[..., K] + dt2 * rdxc / (sh(wk, -1, 0) + wk) * (
            (sh(gz1, -1, 0) - gz0) * (pk1 - sh(pk0, -1, 0))
            + (sh(gz0, -1, 0) - gz1) * (sh(pk1, -1, 0) - pk0))
        vn = vc[..., K] + dt2 * rdyc / (sh(wk, 0, -1) + wk) * (
            (sh(gz1, 0, -1) - gz0) * (pk1 - sh(pk0, 0, -1))
            + (sh(gz0, 0, -1) - gz1) * (sh(pk1, 0, -1) - pk0))
    uk, vk = uc[..., K], vc[..., K]
    put(uk, finite_on(un, m, "p_grad_c uc"), m)
    put(vk, finite_on(vn, m, "p_grad_c vc"), m)


# ------------------------------------------------------------------ ray_fast
def rayleigh_columns(pfull, dp, dt, ptop, rf_cutoff, tau):
    """Per-level rf, the nudge mask and the column mass dm (ray_fast.py:25-75)."""
    pfull = np.asarray(pfull, dtype=np.float64)
    dp = np.asarray(dp, dtype=np.float64)
    nk = len(pfull)
    nudge = rf_cutoff + min(100.0, 10.0 * ptop)
    use_rf = pfull < rf_cutoff
    use_n = pfull < nudge
    rf = np.ones(nk)
    tau0 = tau * SDAY
    with np.errstate(all="ignore"):
        rfv = dt / tau0 * np.sin(0.5 * math.pi * np.log(rf_cutoff / pfull)
                                 / np.log(rf_cutoff / ptop)) ** 2
        rfv = 1.0 / (1.0 + rfv)
    rf = np.where(use_rf, rfv, rf)
    dm = np.zeros(nk)
    acc = 0.0
    for k in range(nk):
        if k == 0:
            acc = dp[0] if use_n[0] else 0.0
        elif use_n[k]:
            acc = acc + dp[k]
        dm[k] = acc
    for k in range(nk - 2, -1, -1):
        if use_n[k]:
            dm[k] = dm[k + 1]
    if not use_rf[0]:
        dm[0] = 0.0
    return rf, use_rf, use_n, dm


def ray_fast(u, v, w, dp, pfull, dt, ptop, rf_cutoff, tau, t, hydrostatic=False):
    """RayleighDamping.__call__ (ray_fast.py:130-192); u, v, w in place (nz levels)."""
    nz = len(pfull)
    rf, use_rf, use_n, dm = rayleigh_columns(pfull, dp, dt, ptop, rf_cutoff, tau)
    dp = np.asarray(dp, dtype=np.float64)
    s = u.shape

    def damp_component(q, mask):
        dmdir = np.zeros(s[:2])
        for k in range(nz):
            if use_rf[k]:
                dmdir = dmdir + (1.0 - rf[k]) * dp[k] * q[..., k]
                q[..., k] = np.where(mask, q[..., k] * rf[k], q[..., k])
        for k in range(nz):
            if use_n[k]:
                q[..., k] = np.where(mask, q[..., k] + dmdir / dm[k], q[..., k])

    mu = box(s[:2], t.isc, t.iec, t.jsc, t.jec + 1)
    mv = box(s[:2], t.isc, t.iec + 1, t.jsc, t.jec)
    damp_component(u, mu)
    damp_component(v, mv)
    if not hydrostatic:
        mw = box(s[:2], t.isc, t.iec, t.jsc, t.jec)
        for k in range(nz):
            if use_rf[k]:
                w[..., k] = np.where(mw, w[..., k] * rf[k], w[..., k])


# ----------------------------------------------------------------- del2cubed
def _corner_fill_avg(q, t):
    """corner_fill (del2cubed.py:30-62): 12 three-cell averages at the corners."""
    out = q.copy()
    third = 1.0 / 3.0

    def o(a, d):
        return None if a is None else a + d

    rows = [
        ((t.i_start, t.j_start), ((0, 0), (-1, 0), (0, -1))),
        ((o(t.i_start, -1), t.j_start), ((1, 0), (0, 0), (1, -1))),
        ((t.i_start, o(t.j_start, -1)), ((0, 1), (-1, 1), (0, 0))),
        ((t.i_end, t.j_start), ((0, 0), (1, 0), (0, -1))),
        ((o(t.i_end, 1), t.j_start), ((-1, 0), (0, 0), (-1, -1))),
        ((t.i_end, o(t.j_start, -1)), ((0, 1), (1, 1), (0, 0))),
        ((t.i_end, t.j_end), ((0, 0), (1, 0), (0, 1))),
        ((o(t.i_end, 1), t.j_end), ((-1, 0), (0, 0), (-1, 1))),
        ((t.i_end, o(t.j_end, 1)), ((0, -1), (1, -1), (0, 0))),
        ((t.i_start, t.j_end), ((0, 0), (-1, 0), (0, 1))),
        ((o(t.i_start, -1), t.j_end), ((1, 0), (0, 0), (1, 1))),
        ((t.i_start, o(t.j_end, 1)), ((0, -1), (-1, -1), (0, 0))),
    ]
    for (i, j), offs in rows:
        if i is None or j is None:
            continue
        a, b, c = (q[i + di, j + dj] for di, dj in offs)
        out[i, j] = (a + b + c) * third
    return out


def del2cubed(qdel, cd, g, t, nmax):
    """HyperdiffusionDamping.__call__ (del2cubed.py:78-181); qdel in place (nz levels)."""
    ntimes = min(3, nmax)
    s = qdel.shape
    rarea = _g(g, "rarea")
    for n in range(ntimes):
        nt = ntimes - (n + 1)
        full = box(s[:2], t.isd, t.ied, t.jsd, t.jed)
        q = np.full(s, np.nan)
        put(q, _corner_fill_avg(qdel, t), full)
        qx = copy_corners_x(q, t) if nt > 0 else q
        fx = np.full(s, np.nan)
        put(fx, _g(g, "del6_v") * (sh(qx, -1, 0) - qx), box(s[:2], t.isc - nt, t.iec + 1 + nt,
                                                           t.jsc - nt, t.jec + nt))
        qy = copy_corners_y(qx, t) if nt > 0 else qx
        fy = np.full(s, np.nan)
        put(fy, _g(g, "del6_u") * (sh(qy, 0, -1) - qy), box(s[:2], t.isc - nt, t.iec + nt,
                                                           t.jsc - nt, t.jec + 1 + nt))
        put(qdel, qy, full)
        m = box(s[:2], t.isc - nt, t.iec + nt, t.jsc - nt, t.jec + nt)
        put(qdel, qdel + cd * rarea * (fx - sh(fx, 1, 0) + fy - sh(fy, 0, 1)), m)


# -------------------------------------------------------- temperature_adjust
def pkz_tempadjust(delp, delz, cappa, heat_source, pt, pkz, delt_time_factor, t, nk):
    """compute_pkz_tempadjust (temperature_adjust.py:8-44) on [0, nk); in place."""
    m = box(delp.shape[:2], t.isc, t.iec, t.jsc, t.jec)
    K = slice(0, nk)
    dp, dz, cp, hs, ptk = (a[..., K] for a in (delp, delz, cappa, heat_source, pt))
    pk = (RDG * dp / dz * ptk) ** (cp / (1.0 - cp))
    dtmp = hs / (CV_AIR * dp)
    fac = np.array([0.1, 0.5] + [1.0] * max(0, nk - 2))[:nk][None, None, :]
    deltmin = sign(np.minimum(delt_time_factor * fac, np.abs(dtmp)), dtmp)
    ptn = ptk + deltmin / pk
    pkzk = pkz[..., K]
    put(pkzk, pk, m)
    ptv = pt[..., K]
    put(ptv, ptn, m)


# ------------------------------------------------------------ small stencils
def dp_ref_compute(ak, bk):
    """dp_ref_compute (dyn_core.py:66-77): the layer thickness column."""
    ak, bk = np.asarray(ak, dtype=np.float64), np.asarray(bk, dtype=np.float64)
    return ak[1:] - ak[:-1] + (bk[1:] - bk[:-1]) * 1.0e5


def set_gz(zs, delz, gz, t):
    """set_gz (dyn_core.py:80-85) on the compute domain; gz has nz + 1 levels."""
    m = box(gz.shape[:2], t.isc, t.iec, t.jsc, t.jec)
    nk = gz.shape[2]
    gz[..., nk - 1][m] = zs[m]
    for k in range(nk - 2, -1, -1):
        gz[..., k][m] = (gz[..., k + 1] - delz[..., k])[m]


def set_pem(delp, pem, ptop, t):
    """set_pem (dyn_core.py:88-93) on [isc-1, iec+1]^2, nz levels."""
    m = box(delp.shape[:2], t.isc - 1, t.iec + 1, t.jsc - 1, t.jec + 1)
    nz = delp.shape[2]
    pem[..., 0][m] = ptop
    for k in range(1, nz):
        pem[..., k][m] = (pem[..., k - 1] + delp[..., k])[m]


# ----------------------------------------------------------- the driver
def acoustic_dynamics(state, cfg, metrics, scalars, ak, bk, pfull, halo, sync, col, nz,
                      edges=True):
    """AcousticDynamics.__call__ (dyn_core.py:479-761) on a batch of tiles, in place.

    ``edges``: the tile-edge flags (True for cube tiles; False for a doubly
    periodic plane tile, whose ``halo`` wraps periodically).
    ``state``: dict of arrays [T, I, J, nz + 1] (IJ fields [T, I, J]),
    including the component temporaries (ut, vt, gz, zh, pem, pkc, pk3,
    heat_source, divgd, crx, xfx, cry, yfx, ws3, delpc, ptc) as the
    reference keeps them in the state.  ``cfg``: dict of namelist values;
    ``metrics``: per-tile metric dicts; ``scalars``: da_min, da_min_c;
    ``halo(kind, names, n_halo=3)`` refreshes the halos of state arrays;
    ``sync(u, v)`` synchronises the D-grid edge winds; ``col``: d_sw column
    namelist.
    """
    from . import csw as ocs
    from . import dsw as ods
    from . import riemann as ori
    from .fvtp2d import Tile

    n = state["delp"].shape[1] - 7
    ntiles = state["delp"].shape[0]
    t = Tile(n, n, 3, edges, edges, edges, edges)
    K = slice(0, nz)
    Ki = slice(0, nz + 1)
    dt = cfg["mdt"] / cfg["n_split"]
    dt2 = 0.5 * dt
    akap = KAPPA
    ptop = cfg["ptop"]
    dp_ref = np.zeros(nz + 1)
    dp_ref[:nz] = dp_ref_compute(ak, bk)[:nz]
    zs = state["phis"] * (1.0 / GRAV)
    halo("scalar", ["q_con", "cappa"])
    halo("scalar", ["delp", "pt"])
    halo("vector_d", ["u", "v"])
    # zero_data
    for name in ("mfxd", "mfyd", "cxd", "cyd"):
        state[name][:, t.isd:t.ied + 1, t.jsd:t.jed + 1, K] = 0.0
    if cfg["n_map"] == 1:
        state["diss_estd"][:, t.isc:t.iec + 1, t.jsc:t.jec + 1, K] = 0.0
    nk_heat = nz if (cfg["convert_ke"] or cfg["vtdm4"] > 1.0e-4) else 0
    dsw_cfg = dict(hord_dp=cfg["hord_dp"], hord_tm=cfg["hord_tm"], hord_mt=cfg["hord_mt"],
                   hord_vt=cfg["hord_vt"], dddmp=cfg["dddmp"], d4_bg=cfg["d4_bg"],
                   nord=cfg["nord"], d_con=cfg["d_con"], do_f3d=False, hydrostatic=False,
                   do_skeb=False)
    end_step = cfg["n_map"] == cfg["k_split"]
    S = state
    for it in range(cfg["n_split"]):
        remap_step = it == cfg["n_split"] - 1
        halo("scalar", ["w"])
        if it == 0:
            for ti in range(ntiles):
                set_gz(zs[ti], S["delz"][ti], S["gz"][ti][..., Ki], t)
            halo("scalar", ["gz"])
        if it == cfg["n_split"] - 1 and end_step:
            for ti in range(ntiles):
                set_pem(S["delp"][ti][..., K], S["pem"][ti][..., K], ptop, t)
        for ti in range(ntiles):
            g = metrics[ti]
            out = ocs.c_sw(*(S[k][ti][..., K] for k in ("delp", "pt", "u", "v", "w", "uc", "vc",
                                                         "ua", "va", "ut", "vt", "divgd",
                                                         "omga")), dt2, g, t, cfg["nord"])
            for k in ("delp", "pt", "u", "v", "w", "uc", "vc", "ua", "va", "ut", "vt", "divgd",
                      "omga"):
                S[k][ti][..., K] = out[k]
            S["delpc"][ti][..., K] = out["delpc"]
            S["ptc"][ti][..., K] = out["ptc"]
        if cfg["nord"] > 0:
            halo("corner", ["divgd"])
        for ti in range(ntiles):
            if it == 0:
                S["zh"][ti][t.isd:t.ied + 1, t.jsd:t.jed + 1, Ki] = \
                    S["gz"][ti][t.isd:t.ied + 1, t.jsd:t.jed + 1, Ki]
            else:
                S["gz"][ti][t.isd:t.ied + 1, t.jsd:t.jed + 1, Ki] = \
                    S["zh"][ti][t.isd:t.ied + 1, t.jsd:t.jed + 1, Ki]
        for ti in range(ntiles):
            g = metrics[ti]
            gz = S["gz"][ti][..., Ki]
            update_dz_c(dp_ref, zs[ti], g["area"], S["ut"][ti][..., Ki], S["vt"][ti][..., Ki],
                        gz, S["ws3"][ti], dt2, t)
            b = (slice(t.isc - 1, t.iec + 2), slice(t.jsc - 1, t.jec + 2))
            rgz, rpef = ori.riem_solver_c(
                dt2, S["cappa"][ti][b], ptop, S["phis"][ti][b], S["ws3"][ti][b],
                S["ptc"][ti][b], S["q_con"][ti][b], S["delpc"][ti][b], S["gz"][ti][b],
                S["omga"][ti][b], cfg["p_fac"])
            S["gz"][ti][b] = rgz
            S["pkc"][ti][b] = rpef
            uc, vc = S["uc"][ti][..., Ki], S["vc"][ti][..., Ki]
            p_grad_c(uc, vc, S["delpc"][ti][..., K], S["pkc"][ti][..., Ki], S["gz"][ti][..., Ki],
                     dt2, g, t)
        halo("vector_c", ["uc", "vc"])
        for ti in range(ntiles):
            g = metrics[ti]
            st = {"delpc": S["vt"][ti][..., K], "ptc": S["ptc"][ti][..., K]}
            for k in ("delp", "pt", "u", "v", "w", "uc", "vc", "ua", "va", "divgd", "q_con",
                      "zh", "crx", "cry", "xfx", "yfx", "heat_source"):
                st[k] = S[k][ti][..., K]
            st.update(mfx=S["mfxd"][ti][..., K], mfy=S["mfyd"][ti][..., K],
                      cx=S["cxd"][ti][..., K], cy=S["cyd"][ti][..., K],
                      diss_est=S["diss_estd"][ti][..., K])
            damp = dict(del6_u=g["del6_u"], del6_v=g["del6_v"], da_min=scalars["da_min"],
                        da_min_c=scalars["da_min_c"])
            out = ods.d_sw(st, dt, g, t, dsw_cfg, col, damp)
            back = dict(vt="delpc", ptc="ptc", mfxd="mfx", mfyd="mfy", cxd="cx", cyd="cy",
                        diss_estd="diss_est")
            for k in ("delp", "pt", "u", "v", "w", "uc", "vc", "divgd", "q_con", "crx", "cry",
                      "xfx", "yfx", "heat_source"):
                back[k] = k
            for sk, ok in back.items():
                S[sk][ti][..., K] = out[ok]
        halo("scalar", ["delp", "pt", "q_con"])
        for ti in range(ntiles):
            g = metrics[ti]
            zh = S["zh"][ti][..., Ki]
            update_dz_d(zs[ti], zh, S["crx"][ti][..., Ki], S["cry"][ti][..., Ki],
                        S["xfx"][ti][..., Ki], S["yfx"][ti][..., Ki], S["wsd"][ti], dt, g, t,
                        cfg["hord_tm"], col["nord_v"], col["damp_vt"], dp_ref)
            c = (slice(t.isc, t.iec + 1), slice(t.jsc, t.jec + 1))
            res = ori.riem_solver3(remap_step, dt, S["cappa"][ti][c], ptop, zs[ti][c],
                                   S["wsd"][ti][c], S["delz"][ti][c], S["q_con"][ti][c],
                                   S["delp"][ti][c], S["pt"][ti][c], S["zh"][ti][c],
                                   S["pe"][ti][c], S["pkc"][ti][c], S["pk3"][ti][c],
                                   S["pk"][ti][c], S["peln"][ti][c], S["w"][ti][c], cfg["p_fac"])
            for k, a in zip(("zh", "pe", "pkc", "pk3", "pk", "peln", "w", "delz"), res):
                S[k][ti][c] = a
        halo("scalar", ["zh"])
        halo("scalar", ["pkc"], n_halo=2)
        for ti in range(ntiles):
            if remap_step:
                pe_halo(S["pe"][ti][..., Ki], S["delp"][ti], ptop, t)
            pk3_halo(S["pk3"][ti][..., Ki], S["delp"][ti], ptop, akap, t)
            b = (slice(t.isc - 2, t.iec + 3), slice(t.jsc - 2, t.jec + 3), Ki)
            S["gz"][ti][b] = S["zh"][ti][b] * GRAV
            g = metrics[ti]
            nh_p_grad(S["u"][ti], S["v"][ti], S["pkc"][ti][..., Ki], S["gz"][ti][..., Ki],
                      S["pk3"][ti][..., Ki], S["delp"][ti][..., K], dt, ptop, akap, g, t)
            ray_fast(S["u"][ti][..., K], S["v"][ti][..., K], S["w"][ti][..., K], dp_ref[:nz],
                     pfull[:nz], dt, ptop, cfg["rf_cutoff"], cfg["tau"], t)
        if it != cfg["n_split"] - 1:
            halo("vector_d", ["u", "v"])
        else:
            sync()
    if nk_heat != 0 and cfg["d_con"] > 1.0e-5:
        halo("scalar", ["heat_source"])
        cd = CNST_0P20 * scalars["da_min"]
        for ti in range(ntiles):
            g = metrics[ti]
            hs = S["heat_source"][ti][..., K]
            del2cubed(hs, cd, g, t, min(3, cfg["nord"] + 1))
            S["heat_source"][ti][..., K] = hs
            pkz_tempadjust(S["delp"][ti], S["delz"][ti], S["cappa"][ti], S["heat_source"][ti],
                           S["pt"][ti], S["pkz"][ti], abs(dt * cfg["delt_max"]), t, nk_heat)
