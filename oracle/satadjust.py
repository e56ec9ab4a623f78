"""CPU oracle for the fast saturation adjustment (SatAdjust3d).

Test infrastructure only: imported by tests, never by the product path.
Follows fv3core/stencils/saturation_adjustment.py statement by statement:
the saturation-pressure "tables" (:43-168), which the stencil evaluates
from their closed forms at the table index instead of reading stored
tables, the microphysical helper functions (:171-535) and the satadjust
stencil (:554-896), with its SatAdjust3d driver (:899-1024).

Every function works on numpy arrays of the cells of the stencil's domain
(any shape; IJ fields broadcast); a gtscript ``if`` on a point becomes an
``np.where`` over the points, evaluated in the stencil's statement order.
"""
from __future__ import annotations

import math

import numpy as np

GRAV = 9.80665
RDGAS = 287.05
RVGAS = 461.50
HLV = 2.5e6
HLF = 3.3358e5
CP_AIR = 1004.6
CV_AIR = CP_AIR - RDGAS
CV_VAP = 3.0 * RVGAS
CP_VAP = 4.0 * RVGAS
C_ICE = 1972.0
C_LIQ = 4.1855e3
TICE = 273.16
ZVIR = RVGAS / RDGAS - 1
RDG = -RDGAS / GRAV

# saturation_adjustment.py:19-41
DC_VAP = CP_VAP - C_LIQ
DC_ICE = C_LIQ - C_ICE
LV0 = HLV - DC_VAP * TICE
LI00 = HLF - DC_ICE * TICE
LI2 = LV0 + LI00
D2ICE = DC_VAP + DC_ICE
E00 = 611.21
TMIN = TICE - 160.0
DELT = 0.1
T_WFR = TICE - 40.0
TICE0 = TICE - 0.01
LAT2 = (HLV + HLF) ** 2
QS_LENGTH = 2621


# ------------------------------------------------------------------ tables
def tem_lower(i):
    """:43-45"""
    return TMIN + DELT * i


def tem_upper(i):
    """:48-50"""
    return 253.16 + DELT * i


def q_table_oneline(dhc, lhc, tem):
    """:53-61"""
    return E00 * np.exp((dhc * np.log(tem / TICE) + (tem - TICE) / (tem * TICE) * lhc) / RVGAS)


def table_vapor_oneline(tem):
    return q_table_oneline(DC_VAP, LV0, tem)


def table_ice_oneline(tem):
    return q_table_oneline(D2ICE, LI2, tem)


def qs_table_fn(i):
    """:76-92 (ice below -20 C, water above 0 C, blended in between)."""
    i = np.asarray(i, dtype=np.float64)
    tem_l = tem_lower(i)
    tem_u = tem_upper(i - 1400)
    with np.errstate(all="ignore"):
        table = np.where(i < 1600, table_ice_oneline(tem_l), 0.0)
        table = np.where((i >= 1600) & (i < 1400 + 1221), table_vapor_oneline(tem_u), table)
        esupc = table_vapor_oneline(tem_u)
        wice = 0.05 * (TICE - tem_u)
        wh2o = 0.05 * (tem_u - 253.16)
        blend = wice * table + wh2o * esupc
    return np.where((i >= 1400) & (i < 1600), blend, table)


def qs_table2_fn(i):
    """:95-129 (ice below 0 C, water above, smoothed at indices 1599/1600)."""
    i = np.asarray(i, dtype=np.float64)
    tem0 = tem_lower(i)
    table2 = np.where(i < 1600, table_ice_oneline(tem0), table_vapor_oneline(tem0))
    # i == 1599
    t99 = table_ice_oneline(tem_lower(1599.0))
    tu = tem_upper(1599.0 - 1400)
    t99 = (0.05 * (TICE - tu)) * t99 + (0.05 * (tu - 253.16)) * table_vapor_oneline(tu)
    v99 = 0.25 * (table_ice_oneline(tem_lower(1598.0)) + 2.0 * t99 +
                  table_vapor_oneline(tem_lower(1600.0)))
    # i == 1600
    t00 = table_vapor_oneline(tem_upper(1600.0 - 1400))
    v00 = 0.25 * (table_ice_oneline(tem_lower(1599.0)) + 2.0 * t00 +
                  table_vapor_oneline(tem_lower(1601.0)))
    table2 = np.where(i == 1599, v99, table2)
    return np.where(i == 1600, v00, table2)


def qs_tablew_fn(i):
    """:132-135"""
    return table_vapor_oneline(tem_lower(np.asarray(i, dtype=np.float64)))


def des2_table(i):
    """:149-158: forward difference of table2, clamped at zero; the last
    entry repeats the one before."""
    i = np.asarray(i, dtype=np.float64)
    t = qs_table2_fn(i)
    des2 = np.maximum(0.0, qs_table2_fn(i + 1) - t)
    last = np.maximum(0.0, t - qs_table2_fn(i - 1))
    return np.where(i == QS_LENGTH - 1, last, des2)


def desw_table(i):
    """:161-168"""
    i = np.asarray(i, dtype=np.float64)
    t = qs_tablew_fn(i)
    desw = np.maximum(0.0, qs_tablew_fn(i + 1) - t)
    last = np.maximum(0.0, t - qs_table2_fn(i - 1))  # des_end reads table2 (:138-146)
    return np.where(i == QS_LENGTH - 1, last, desw)


# ---------------------------------------------------------------- helpers
def dim(a, b):
    """basic_operations.py:44-46"""
    return np.where(a - b > 0, a - b, 0.0)


def compute_cvm(mc_air, qv, c_vap, q_liq, q_sol):
    return mc_air + qv * c_vap + q_liq * C_LIQ + q_sol * C_ICE


def ap1_for_wqs2(ta):
    """:454-457"""
    ap1 = 10.0 * dim(ta, TMIN) + 1.0
    return np.minimum(ap1, QS_LENGTH) - 1


def wqs2(ta, den, water: bool):
    """wqs2_fn_w (water) / wqs2_fn_2 (ice below 0 C), :487-522."""
    ap1 = ap1_for_wqs2(ta)
    it = np.floor(ap1)
    it2 = np.floor(ap1 - 0.5)
    it2_p1 = it2 + 1
    if water:
        tab, des, des2, des_p1 = qs_tablew_fn(it), desw_table(it), desw_table(it2), \
            desw_table(it2_p1)
    else:
        tab, des, des2, des_p1 = qs_table2_fn(it), des2_table(it), des2_table(it2), \
            des2_table(it2_p1)
    es = tab + (ap1 - it) * des
    denom = RVGAS * ta * den
    wqsat = es / denom
    dqdt = 10.0 * (des2 + (ap1 - it2) * (des_p1 - des2))
    return wqsat, dqdt / denom


def wqs1(it, ap1, ta, den, water: bool):
    """wqs1_fn_w / wqs1_fn_2, :525-535."""
    if water:
        tab, des = qs_tablew_fn(it), desw_table(it)
    else:
        tab, des = qs_table2_fn(it), des2_table(it)
    return (tab + (ap1 - it) * des) / (RVGAS * ta * den)


def lat_i(pt1, cvm):
    """update_latent_heat_coefficient_i, :399-403."""
    lhi = LI00 + DC_ICE * pt1
    return lhi, lhi / cvm


def lat_l(pt1, cvm, lv00, d0_vap):
    """update_latent_heat_coefficient_l, :406-410."""
    lhl = lv00 + d0_vap * pt1
    return lhl, lhl / cvm


# ------------------------------------------------------------- satadjust
def satadjust(fields, area, hs, *, sdt, zvir, fac_i2s, do_qa, consv_te, c_air, c_vap, mdt,
              fac_r2g, fac_smlt, fac_l2r, fac_imlt, d0_vap, lv00, fac_v2l, fac_l2v, last_step,
              ext):
    """The satadjust stencil (:554-896) on the cells of its domain.

    fields: dict of arrays [..., nk] (k from kmp to nz - 1): qv, ql, qi, qr,
    qs, qg, qa, pt, dp, delz, te0, q_con, cappa, pkz, and peln (nk + 1
    levels, only read when ext["hydrostatic"]).  area, hs: arrays
    broadcasting against the cell arrays.  ext: the stencil's externals.
    Returns a new dict with the updated fields.
    """
    f = {k: np.array(v, dtype=np.float64, copy=True) for k, v in fields.items()}
    hydro = bool(ext["hydrostatic"])
    qv, ql, qi, qr, qs, qg = f["qv"], f["ql"], f["qi"], f["qr"], f["qs"], f["qg"]
    pt, dp, delz = f["pt"], f["dp"], f["delz"]
    if hydro:  # interval(1, None): delz = delz[k - 1] (:612-615)
        delz = delz.copy()
        delz[..., 1:] = f["delz"][..., :-1]
        f["delz"] = delz
    q_liq = ql + qr
    q_sol = qi + qs + qg
    qpz = q_liq + q_sol
    pt1 = pt / ((1.0 + zvir * qv) * (1.0 - qpz))
    t0 = pt1
    qpz = qpz + qv
    if hydro:
        peln = f["peln"]
        den = dp / ((peln[..., 1:] - peln[..., :-1]) * RDGAS * pt)
    else:
        den = -dp / (GRAV * delz)
    mc_air = (1.0 - qpz) * c_air
    cvm = compute_cvm(mc_air, qv, c_vap, q_liq, q_sol)
    lhi, icp2 = lat_i(pt1, cvm)
    te0 = f["te0"]
    if consv_te:
        te0 = -c_air * t0 if hydro else -cvm * t0
    # fix negative cloud ice with snow
    neg = qi < 0.0
    qs = np.where(neg, qs + qi, qs)
    qi = np.where(neg, 0.0, qi)

    # melt_cloud_ice (:186-200)
    c = (qi > 1.0e-8) & (pt1 > TICE)
    factmp = fac_imlt * (pt1 - TICE) / icp2
    sink = np.where(qi < factmp, qi, factmp)
    qi = np.where(c, qi - sink, qi)
    ql = np.where(c, ql + sink, ql)
    q_liq = np.where(c, q_liq + sink, q_liq)
    q_sol = np.where(c, q_sol - sink, q_sol)
    cvm = np.where(c, compute_cvm(mc_air, qv, c_vap, q_liq, q_sol), cvm)
    pt1 = np.where(c, pt1 + (-sink) * lhi / cvm, pt1)
    lhi, icp2 = lat_i(pt1, cvm)

    # fix_negative_snow (:209-219)
    c1 = qs < 0.0
    c2 = ~c1 & (qg < 0.0)
    tmp = np.minimum(-qg, np.maximum(qs, 0.0))
    qg_n = np.where(c1, qg + qs, np.where(c2, qg + tmp, qg))
    qs = np.where(c1, 0.0, np.where(c2, qs - tmp, qs))
    qg = qg_n
    # fix_negative_cloud_water (:222-233)
    c1 = ql < 0.0
    c2 = ~c1 & (qr < 0.0)
    t1 = np.minimum(-ql, np.maximum(qr, 0.0))
    t2 = np.minimum(-qr, np.maximum(ql, 0.0))
    ql_n = np.where(c1, ql + t1, np.where(c2, ql - t2, ql))
    qr = np.where(c1, qr - t1, np.where(c2, qr + t2, qr))
    ql = ql_n

    # complete_freezing (:236-247)
    dtmp = TICE - 48.0 - pt1
    c = (ql > 0.0) & (dtmp > 0.0)
    sink = np.minimum(ql, dtmp / icp2)
    ql, qi, q_liq, q_sol, cvm, pt1 = _freeze(c, sink, ql, qi, q_liq, q_sol, cvm, pt1, mc_air,
                                             qv, c_vap, lhi)

    wqsat, dq2dt = wqs2(pt1, den, True)
    lhl, lcp2 = lat_l(pt1, cvm, lv00, d0_vap)
    lhi, icp2 = lat_i(pt1, cvm)
    dimmin = np.minimum(1.0, dim(TICE, pt1) / 48.0)
    tcp3 = lcp2 + icp2 * dimmin
    dq0 = (qv - wqsat) / (1.0 + tcp3 * dq2dt)
    src_cond = np.minimum(ext["sat_adj0"] * dq0, np.maximum(ext["ql_gen"] - ql, fac_v2l * dq0))
    src = np.where(dq0 > 0, src_cond, _evaporation(wqsat, qv, ql, dq0, fac_l2v))
    qv, ql, q_liq, cvm, pt1 = _wqsat_correct(src, pt1, lhl, qv, ql, q_liq, q_sol, mc_air, c_vap)
    lhl, lcp2 = lat_l(pt1, cvm, lv00, d0_vap)
    lhi, icp2 = lat_i(pt1, cvm)
    dimmin = np.minimum(1.0, dim(TICE, pt1) / 48.0)
    tcp3 = lcp2 + icp2 * dimmin

    if last_step:
        wqsat, dq2dt = wqs2(pt1, den, True)
        dq0 = (qv - wqsat) / (1.0 + tcp3 * dq2dt)
        src = np.where(dq0 > 0, dq0, _evaporation(wqsat, qv, ql, dq0, fac_l2v))
        qv, ql, q_liq, cvm, pt1 = _wqsat_correct(src, pt1, lhl, qv, ql, q_liq, q_sol, mc_air,
                                                 c_vap)
        lhl, lcp2 = lat_l(pt1, cvm, lv00, d0_vap)
        lhi, icp2 = lat_i(pt1, cvm)

    # homogenous_freezing (:250-263)
    dtmp = T_WFR - pt1
    c = (ql > 0.0) & (dtmp > 0.0)
    sink = np.minimum(np.minimum(ql, dtmp / icp2), ql * dtmp * 0.125)
    ql, qi, q_liq, q_sol, cvm, pt1 = _freeze(c, sink, ql, qi, q_liq, q_sol, cvm, pt1, mc_air,
                                             qv, c_vap, lhi)
    lhi, icp2 = lat_i(pt1, cvm)
    exptc = np.exp(0.66 * (TICE0 - pt1))
    # heterogeneous_freezing, Bigg mechanism (:266-281)
    tc = TICE0 - pt1
    c = (ql > 0.0) & (tc > 0.0)
    sink = 3.3333e-10 * mdt * (exptc - 1.0) * den * ql ** 2
    sink = np.minimum(np.minimum(ql, sink), tc / icp2)
    ql, qi, q_liq, q_sol, cvm, pt1 = _freeze(c, sink, ql, qi, q_liq, q_sol, cvm, pt1, mc_air,
                                             qv, c_vap, lhi)
    lhi, icp2 = lat_i(pt1, cvm)

    # make_graupel (:284-298)
    dtmp = (TICE - 0.1) - pt1
    c = (qr > 1e-7) & (dtmp > 0.0)
    rainfac = (dtmp * 0.025) ** 2
    tmp = np.where(1.0 < rainfac, qr, rainfac * qr)
    sink = np.minimum(tmp, fac_r2g * dtmp / icp2)
    qr = np.where(c, qr - sink, qr)
    qg = np.where(c, qg + sink, qg)
    q_liq = np.where(c, q_liq - sink, q_liq)
    q_sol = np.where(c, q_sol + sink, q_sol)
    cvm = np.where(c, compute_cvm(mc_air, qv, c_vap, q_liq, q_sol), cvm)
    pt1 = np.where(c, pt1 + sink * lhi / cvm, pt1)
    lhi, icp2 = lat_i(pt1, cvm)

    # melt_snow (:301-321)
    dtmp = pt1 - (TICE + 0.1)
    dimqs = dim(ext["qs_mlt"], ql)
    c = (qs > 1e-7) & (dtmp > 0.0)
    snowfac = (dtmp * 0.1) ** 2
    tmp = np.where(1.0 < snowfac, qs, snowfac * qs)
    sink = np.minimum(tmp, fac_smlt * dtmp / icp2)
    tmp = np.minimum(sink, dimqs)
    qs = np.where(c, qs - sink, qs)
    ql = np.where(c, ql + tmp, ql)
    qr = np.where(c, qr + sink - tmp, qr)
    q_liq = np.where(c, q_liq + sink, q_liq)
    q_sol = np.where(c, q_sol - sink, q_sol)
    cvm = np.where(c, compute_cvm(mc_air, qv, c_vap, q_liq, q_sol), cvm)
    pt1 = np.where(c, pt1 - sink * lhi / cvm, pt1)

    # autoconversion_cloud_to_rain (:324-330)
    c = ql > ext["ql0_max"]
    sink = fac_l2r * (ql - ext["ql0_max"])
    qr = np.where(c, qr + sink, qr)
    ql = np.where(c, ql - sink, ql)

    iqs2, dqsdt = wqs2(pt1, den, False)
    with np.errstate(all="ignore"):
        expsubl = np.exp(0.875 * np.log(qi * den))
    lhl, lcp2 = lat_l(pt1, cvm, lv00, d0_vap)
    lhi, icp2 = lat_i(pt1, cvm)
    tcp2 = lcp2 + icp2
    adj_fac = 1.0 if last_step else ext["sat_adj0"]

    # sublimation (:333-396)
    t_sub, qi_gen, qi_lim = ext["t_sub"], ext["qi_gen"], ext["qi_lim"]
    cold = pt1 < t_sub
    mid = ~cold & (pt1 < TICE0)
    dq = qv - iqs2
    sink = adj_fac * dq / (1.0 + tcp2 * dqsdt)
    with np.errstate(all="ignore"):
        pidep = np.where(qi > 1.0e-8,
                         sdt * dq * 349138.78 * expsubl /
                         (iqs2 * den * LAT2 / (0.0243 * RVGAS * pt1 ** 2.0) + 4.42478e4), 0.0)
    tmp = TICE - pt1
    qi_crt = np.where(qi_lim < 0.1 * tmp, qi_gen * qi_lim / den, qi_gen * 0.1 * tmp / den)
    maxtmp = np.where(qi_crt - qi > pidep, qi_crt - qi, pidep)
    src_pos = np.where(sink < maxtmp, sink, maxtmp)
    src_pos = np.where(src_pos < tmp / tcp2, src_pos, tmp / tcp2)
    dimtmp = dim(pt1, t_sub)
    pidep_n = np.where(1.0 < dimtmp * 0.2, pidep, pidep * dimtmp * 0.2)
    src_neg = np.where(pidep_n > sink, pidep_n, sink)
    src_neg = np.where(src_neg > -qi, src_neg, -qi)
    src = np.where(cold, dim(qv, 1e-6), np.where(mid, np.where(dq > 0.0, src_pos, src_neg),
                                                 0.0))
    qv = qv - src
    qi = qi + src
    q_sol = q_sol + src
    cvm = compute_cvm(mc_air, qv, c_vap, q_liq, q_sol)
    pt1 = pt1 + src * (lhl + lhi) / cvm

    # virtual temperature, cappa (:791-797)
    q_con = q_liq + q_sol
    tmp = 1.0 + zvir * qv
    pt = pt1 * tmp * (1.0 - q_con)
    tmp = tmp * RDGAS
    cappa = tmp / (tmp + cvm)
    # fix negative graupel with available cloud ice
    c = qg < 0
    mintmp = np.minimum(-qg, np.maximum(0.0, qi))
    qg = np.where(c, qg + mintmp, qg)
    qi = np.where(c, qi - mintmp, qi)
    # autoconversion from cloud ice to snow
    qim = ext["qi0_max"] / den
    c = qi > qim
    sink = fac_i2s * (qi - qim)
    qi = np.where(c, qi - sink, qi)
    qs = np.where(c, qs + sink, qs)
    if consv_te:
        te0 = dp * (te0 + c_air * pt1) if hydro else dp * (te0 + cvm * pt1)
    cvm = mc_air + (qv + q_liq + q_sol) * c_vap
    lhl, lcp2 = lat_l(pt1, cvm, lv00, d0_vap)
    lhi, icp2 = lat_i(pt1, cvm)

    qa = f["qa"]
    if do_qa and last_step:
        if ext["rad_snow"]:
            q_sol = qi + qs + qg if ext["rad_graupel"] else qi + qs
        else:
            q_sol = qi
        q_liq = ql + qr if ext["rad_rain"] else ql
        q_cond = q_sol + q_liq
        tin = pt1 if ext["tintqs"] else pt1 - (lcp2 * q_cond + icp2 * q_sol)
        ap1 = ap1_for_wqs2(tin)
        it = np.floor(ap1)
        wqs1v = wqs1(it, ap1, tin, den, True)
        iqs1v = wqs1(it, ap1, tin, den, False)
        with np.errstate(all="ignore"):
            rqi = np.where(q_cond > 1e-6, q_sol / q_cond, (TICE - tin) / (TICE - T_WFR))
        qstar = np.where(tin < T_WFR, iqs1v,
                         np.where(tin >= TICE, wqs1v, rqi * iqs1v + (1.0 - rqi) * wqs1v))
        mindw = np.minimum(1.0, np.abs(hs) / (10.0 * GRAV))
        dw = ext["dw_ocean"] + (ext["dw_land"] - ext["dw_ocean"]) * mindw
        dbl_sqrt_area = dw * (area ** 0.5 / 100.0e3) ** 0.5
        hvar = np.minimum(0.2, np.maximum(0.01, dbl_sqrt_area))
        rh = qpz / qstar
        dq = hvar * qpz
        q_plus = qpz + dq
        q_minus = qpz - dq
        with np.errstate(all="ignore"):
            if ext["icloud_f"] == 2:
                qa_c = np.where(qpz > qstar, 1.0,
                                np.where((qstar < q_plus) & (q_cond > 1.0e-8),
                                         np.minimum(1.0, ((q_plus - qstar) / dq) ** 2), 0.0))
            else:
                if ext["icloud_f"] == 0:
                    part = (q_plus - qstar) / (dq + dq)
                else:
                    part = (q_plus - qstar) / (2.0 * dq * (1.0 - q_cond))
                inner = np.where(qstar < q_plus, part, 0.0)
                inner = np.where(q_cond > 1.0e-8, np.maximum(ext["cld_min"], inner), inner)
                inner = np.minimum(1.0, inner)
                qa_c = np.where(qstar < q_minus, 1.0, inner)
        qa = np.where((rh > 0.75) & (qpz > 1.0e-8), qa_c, 0.0)

    pkz = f["pkz"]
    if not hydro:  # compute_pkz_func (moist_cv.py:166-168)
        with np.errstate(all="ignore"):
            pkz = np.exp(cappa * np.log(RDG * dp / delz * pt))
    f.update(qv=qv, ql=ql, qi=qi, qr=qr, qs=qs, qg=qg, qa=qa, pt=pt, te0=te0, q_con=q_con,
             cappa=cappa, pkz=pkz)
    return f


def _freeze(c, sink, ql, qi, q_liq, q_sol, cvm, pt1, mc_air, qv, c_vap, lhi):
    """The liquid -> ice move shared by complete / homogeneous / Bigg freezing."""
    ql = np.where(c, ql - sink, ql)
    qi = np.where(c, qi + sink, qi)
    q_liq = np.where(c, q_liq - sink, q_liq)
    q_sol = np.where(c, q_sol + sink, q_sol)
    cvm = np.where(c, compute_cvm(mc_air, qv, c_vap, q_liq, q_sol), cvm)
    pt1 = np.where(c, pt1 + sink * lhi / cvm, pt1)
    return ql, qi, q_liq, q_sol, cvm, pt1


def _evaporation(wqsat, qv, ql, dq0, fac_l2v):
    """ql_evaporation (:437-441): src = -min(ql, factor * dq0)."""
    factor = -np.minimum(1, fac_l2v * 10.0 * (1.0 - qv / wqsat))
    return -np.minimum(ql, factor * dq0)


def _wqsat_correct(src, pt1, lhl, qv, ql, q_liq, q_sol, mc_air, c_vap):
    """:444-451"""
    qv = qv - src
    ql = ql + src
    q_liq = q_liq + src
    cvm = compute_cvm(mc_air, qv, c_vap, q_liq, q_sol)
    pt1 = pt1 + src * lhl / cvm
    return qv, ql, q_liq, cvm, pt1


def scalars(config, mdt):
    """The conversion factors SatAdjust3d.__call__ forms (:960-996)."""
    sdt = 0.5 * mdt
    fac_l2v = min(config.sat_adj0, 1.0 - math.exp(-sdt / config.tau_l2v))
    if config.hydrostatic:
        c_air, c_vap = CP_AIR, CP_VAP
    else:
        c_air, c_vap = CV_AIR, CV_VAP
    d0_vap = c_vap - C_LIQ
    return dict(
        sdt=sdt,
        fac_i2s=1.0 - math.exp(-mdt / config.tau_i2s),
        fac_v2l=1.0 - math.exp(-sdt / config.tau_v2l),
        fac_r2g=1.0 - math.exp(-mdt / config.tau_r2g),
        fac_l2r=1.0 - math.exp(-mdt / config.tau_l2r),
        fac_l2v=fac_l2v,
        fac_imlt=1.0 - math.exp(-sdt / config.tau_imlt),
        fac_smlt=1.0 - math.exp(-mdt / config.tau_smlt),
        c_air=c_air, c_vap=c_vap, d0_vap=d0_vap, lv00=HLV - d0_vap * TICE, mdt=mdt,
    )


def externals(config):
    """The stencil externals of SatAdjust3d (:905-925)."""
    return {k: getattr(config, k) for k in (
        "hydrostatic", "rad_snow", "rad_rain", "rad_graupel", "tintqs", "sat_adj0", "ql_gen",
        "qs_mlt", "ql0_max", "t_sub", "qi_gen", "qi_lim", "qi0_max", "dw_ocean", "dw_land",
        "icloud_f", "cld_min")}


def sat_adjust_3d(config, area, te, qvapor, qliquid, qice, qrain, qsnow, qgraupel, qcld, hs,
                  peln, delp, delz, q_con, pt, pkz, cappa, r_vir, mdt, fast_mp_consv, last_step,
                  kmp, box, nz):
    """SatAdjust3d.__call__ (:936-1024) on storage arrays [i, j, k] (area,
    hs: [i, j]): the stencil on box = (i0, i1, j0, j1) (inclusive) and levels
    [kmp, nz); returns the updated arrays keyed qv, ql, qi, qr, qs, qg, qa,
    pt, dp, delz, te0, q_con, cappa, pkz."""
    i0, i1, j0, j1 = box
    sl = (slice(i0, i1 + 1), slice(j0, j1 + 1))
    nzc = nz
    ks = slice(kmp, nzc)
    names = dict(qv=qvapor, ql=qliquid, qi=qice, qr=qrain, qs=qsnow, qg=qgraupel, qa=qcld, pt=pt,
                 dp=delp, delz=delz, te0=te, q_con=q_con, cappa=cappa, pkz=pkz)
    fields = {k: v[sl + (ks,)] for k, v in names.items()}
    fields["peln"] = peln[sl + (slice(kmp, nzc + 1),)]
    sc = scalars(config, mdt)
    out = satadjust(fields, np.asarray(area)[sl][..., None], np.asarray(hs)[sl][..., None],
                    sdt=sc["sdt"], zvir=r_vir,
                    fac_i2s=sc["fac_i2s"], do_qa=True, consv_te=fast_mp_consv,
                    c_air=sc["c_air"], c_vap=sc["c_vap"], mdt=mdt, fac_r2g=sc["fac_r2g"],
                    fac_smlt=sc["fac_smlt"], fac_l2r=sc["fac_l2r"], fac_imlt=sc["fac_imlt"],
                    d0_vap=sc["d0_vap"], lv00=sc["lv00"], fac_v2l=sc["fac_v2l"],
                    fac_l2v=sc["fac_l2v"], last_step=last_step, ext=externals(config))
    result = {}
    for k, arr in names.items():
        a = np.array(arr, copy=True)
        a[sl + (ks,)] = out[k]
        result[k] = a
    return result
