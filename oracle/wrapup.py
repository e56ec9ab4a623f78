"""CPU oracle for the step's wrap-up: AdjustNegativeTracerMixingRatio
(neg_adj3) and CubedToLatLon (c2l_ord).

Test infrastructure only: imported by tests, the smoke check and the CPU
baseline of the benchmark, never by the product path.  Follows
fv3core/stencils/neg_adj3.py:14-372 and c2l_ord.py:1-141 statement by
statement.  Column functions take arrays ``[..., nz]`` (the layer levels of
the compute-domain columns); a ``computation`` with several ``interval``
blocks runs them in the order they are written, each over its levels.

Two readings the stencil text leaves open are fixed as follows:
* ``fix_water_vapor_down`` writes the temporary ``dp_bot`` on the bottom
  level only and reads it on every level above (neg_adj3.py:255-271); a
  temporary holds nothing where it was not written, and the Fortran
  routine the stencil restates (the reference's own comment,
  neg_adj3.py:177-190) uses the bottom layer's thickness there, so the
  oracle uses ``dp`` of the bottom level on every level;
* in the same stencil the ``interval(0, 1)`` block zeroes a negative top
  level before the ``interval(1, 2)`` block reads it, so the top level's
  deficit is dropped, not moved down (blocks run in order).
"""
from __future__ import annotations

import numpy as np

from .common import sh

RDGAS = 287.05
RVGAS = 461.50
CP_AIR = 1004.6
CV_AIR = CP_AIR - RDGAS
CV_VAP = 3.0 * RVGAS
C_ICE = 1972.0
C_LIQ = 4.1855e3
HLV = 2.5e6
HLF = 3.3358e5
TICE = 273.16
DC_ICE = C_LIQ - C_ICE
LI0 = HLF - DC_ICE * TICE
# nonhydrostatic constants of AdjustNegativeTracerMixingRatio.__init__
# (neg_adj3.py:315-320)
D0_VAP = CV_VAP - C_LIQ
LV00 = HLV - D0_VAP * TICE


def _lt(a, b):
    """``a if a < b else b`` as the stencils write it."""
    return np.where(a < b, a, b)


def fix_negative_ice(qv, qi, qs, qg, qr, ql, pt, lcpk, icpk):
    """fix_negative_ice (neg_adj3.py:14-54), element-wise."""
    qsum = qi + qs
    pos = qsum > 0.0
    c1 = pos & (qi < 0.0)
    c2 = pos & ~(qi < 0.0) & (qs < 0.0)
    qi, qs = np.where(c1, 0.0, qi), np.where(c1, qsum, qs)
    qs, qi = np.where(c2, 0.0, qs), np.where(c2, qsum, qi)
    qi = np.where(~pos, 0.0, qi)
    qs = np.where(~pos, 0.0, qs)
    qg = np.where(~pos, qg + qsum, qg)
    m = qg < 0.0
    dq = _lt(qs, -qg)
    qs = np.where(m, qs - dq, qs)
    qg = np.where(m, qg + dq, qg)
    m2 = m & (qg < 0.0)
    dq = _lt(qi, -qg)
    qi = np.where(m2, qi - dq, qi)
    qg = np.where(m2, qg + dq, qg)
    # borrow from rain: the stencil takes the amount out of qliquid
    m = (qg < 0.0) & (qr > 0.0)
    dq = _lt(qr, -qg)
    qg = np.where(m, qg + dq, qg)
    ql = np.where(m, ql - dq, ql)
    pt = np.where(m, pt + dq * icpk, pt)
    m = (qg < 0.0) & (ql > 0.0)
    dq = _lt(ql, -qg)
    qg = np.where(m, qg + dq, qg)
    ql = np.where(m, ql - dq, ql)
    pt = np.where(m, pt + dq * icpk, pt)
    m = (qg < 0.0) & (qv > 0.0)
    dq = _lt(0.999 * qv, -qg)
    qg = np.where(m, qg + dq, qg)
    qv = np.where(m, qv - dq, qv)
    pt = np.where(m, pt + dq * (icpk + lcpk), pt)
    return qv, qi, qs, qg, qr, ql, pt


def fix_negative_liq(qv, qi, qs, qg, qr, ql, pt, lcpk, icpk):
    """fix_negative_liq (neg_adj3.py:57-97), element-wise."""
    qsum = ql + qr
    pos_qg = np.where(0.0 > qg, 0.0, qg)
    pos = qsum > 0.0
    c1 = pos & (qr < 0.0)
    c2 = pos & ~(qr < 0.0) & (ql < 0.0)
    qr, ql = np.where(c1, 0.0, qr), np.where(c1, qsum, ql)
    ql, qr = np.where(c2, 0.0, ql), np.where(c2, qsum, qr)
    neg = ~pos
    ql = np.where(neg, 0.0, ql)
    qrain_tmp = qsum
    dq = _lt(pos_qg, -qrain_tmp)
    qrain_tmp = qrain_tmp + dq
    qg = np.where(neg, qg - dq, qg)
    pt = np.where(neg, pt - dq * icpk, pt)
    m = neg & (qr < 0.0)
    dq = _lt(qi + qs, -qrain_tmp)
    qrain_tmp = np.where(m, qrain_tmp + dq, qrain_tmp)
    dq1 = _lt(dq, qs)
    qs = np.where(m, qs - dq1, qs)
    qi = np.where(m, qi + dq1 - dq, qi)
    pt = np.where(m, pt - dq * icpk, pt)
    qr = np.where(neg, qrain_tmp, qr)
    m = neg & (qr < 0.0) & (qv > 0.0)
    dq = _lt(0.999 * qv, -qr)
    qv = np.where(m, qv - dq, qv)
    qr = np.where(m, qr + dq, qr)
    pt = np.where(m, pt + dq * lcpk, pt)
    return qv, qi, qs, qg, qr, ql, pt


def fix_neg_water(pt, qv, ql, qr, qs, qi, qg):
    """fix_neg_water (neg_adj3.py:122-150): returns the new
    (pt, qv, ql, qr, qs, qi, qg)."""
    q_liq = np.where(0.0 > ql + qr, 0.0, ql + qr)
    q_sol = np.where(0.0 > qi + qs, 0.0, qi + qs)
    cpm = ((1.0 - (qv + q_liq + q_sol)) * CV_AIR + qv * CV_VAP + q_liq * C_LIQ
           + q_sol * C_ICE)
    lcpk = (LV00 + D0_VAP * pt) / cpm
    icpk = (LI0 + DC_ICE * pt) / cpm
    qv, qi, qs, qg, qr, ql, pt = fix_negative_ice(qv, qi, qs, qg, qr, ql, pt, lcpk, icpk)
    qv, qi, qs, qg, qr, ql, pt = fix_negative_liq(qv, qi, qs, qg, qr, ql, pt, lcpk, icpk)
    return pt, qv, ql, qr, qs, qi, qg


def fillq(q, dp):
    """fillq (neg_adj3.py:100-119) on columns: returns the new q."""
    q = np.array(q, dtype=np.float64, copy=True)
    nz = q.shape[-1]
    sum1 = np.zeros(q.shape[:-1])
    sum2 = np.zeros(q.shape[:-1])
    for k in range(nz):
        sum1 = np.where(q[..., k] > 0, sum1 + q[..., k] * dp[..., k], sum1)
    for k in range(nz - 1, -1, -1):
        m = (q[..., k] < 0.0) & (sum1 >= 0)
        dq = _lt(sum1, -q[..., k] * dp[..., k])
        sum1 = np.where(m, sum1 - dq, sum1)
        sum2 = np.where(m, sum2 + dq, sum2)
        q[..., k] = np.where(m, q[..., k] + dq / dp[..., k], q[..., k])
    for k in range(nz - 1, -1, -1):
        m = (q[..., k] > 0.0) & (sum1 >= 1e-12) & (sum2 > 0)
        dq = _lt(sum2, q[..., k] * dp[..., k])
        sum2 = np.where(m, sum2 - dq, sum2)
        q[..., k] = np.where(m, q[..., k] - dq / dp[..., k], q[..., k])
    return q


def fix_water_vapor_down(qv, dp):
    """fix_water_vapor_down (neg_adj3.py:193-276) on columns: returns the new qv."""
    qv = np.array(qv, dtype=np.float64, copy=True)
    nz = qv.shape[-1]
    upper_fix = np.zeros(qv.shape)
    lower_fix = np.zeros(qv.shape)
    qv[..., 0] = np.where(qv[..., 0] < 0.0, 0.0, qv[..., 0])
    m = qv[..., 0] < 0
    qv[..., 1] = np.where(m, qv[..., 1] + qv[..., 0] * dp[..., 0] / dp[..., 1], qv[..., 1])
    for k in range(1, nz - 1):
        dq = qv[..., k - 1] * dp[..., k - 1]
        m = lower_fix[..., k - 1] != 0
        qv[..., k] = np.where(m, qv[..., k] + lower_fix[..., k - 1] / dp[..., k], qv[..., k])
        m = (qv[..., k] < 0) & (qv[..., k - 1] > 0)
        dq = np.where(m, _lt(dq, -qv[..., k] * dp[..., k]), dq)
        upper_fix[..., k] = np.where(m, dq, upper_fix[..., k])
        qv[..., k] = np.where(m, qv[..., k] + dq / dp[..., k], qv[..., k])
        m = qv[..., k] < 0
        lower_fix[..., k] = np.where(m, qv[..., k] * dp[..., k], lower_fix[..., k])
        qv[..., k] = np.where(m, 0.0, qv[..., k])
    for k in range(0, nz - 2):
        m = upper_fix[..., k + 1] != 0
        qv[..., k] = np.where(m, qv[..., k] - upper_fix[..., k + 1] / dp[..., k], qv[..., k])
    kb = nz - 1
    m = lower_fix[..., kb - 1] > 0
    qv[..., kb] = np.where(m, qv[..., kb] + lower_fix[..., kb] / dp[..., kb], qv[..., kb])
    upper_fix[..., kb] = qv[..., kb]
    dp_bot = dp[..., kb]
    for k in range(nz - 2, -1, -1):
        dq = qv[..., k] * dp[..., k]
        m = (upper_fix[..., k + 1] < 0) & (qv[..., k] > 0)
        cap = -upper_fix[..., k + 1] * dp_bot
        dq = np.where(m & (dq >= cap), cap, dq)
        qv[..., k] = np.where(m, qv[..., k] - dq / dp[..., k], qv[..., k])
        upper_fix[..., k] = np.where(m, upper_fix[..., k + 1] + dq / dp_bot,
                                     upper_fix[..., k + 1])
    for k in range(1, nz):
        upper_fix[..., k] = upper_fix[..., k - 1]
    qv[..., kb] = upper_fix[..., kb]
    return qv


def fix_neg_cloud(dp, qcld):
    """fix_neg_cloud (neg_adj3.py:153-174) on columns: returns the new qcld."""
    q = np.array(qcld, dtype=np.float64, copy=True)
    nz = q.shape[-1]
    for k in range(1, nz - 1):
        m = q[..., k - 1] < 0.0
        q[..., k] = np.where(m, q[..., k] + q[..., k - 1] * dp[..., k - 1] / dp[..., k], q[..., k])
    q[..., 1:nz - 1] = np.where(q[..., 1:nz - 1] < 0.0, 0.0, q[..., 1:nz - 1])
    k = nz - 2
    m = (q[..., k + 1] < 0.0) & (q[..., k] > 0)
    dq = _lt(-q[..., k] * dp[..., k], q[..., k + 1] * dp[..., k + 1])
    q[..., k] = np.where(m, q[..., k] - dq / dp[..., k], q[..., k])
    k = nz - 1
    m = (q[..., k] < 0) & (q[..., k - 1] > 0.0)
    dq = _lt(-q[..., k] * dp[..., k], q[..., k - 1] * dp[..., k - 1])
    qk = q[..., k] + dq / dp[..., k]
    qk = np.where(0.0 > qk, 0.0, qk)
    q[..., k] = np.where(m, qk, q[..., k])
    return q


def neg_adj3(qvapor, qliquid, qrain, qsnow, qice, qgraupel, qcld, pt, delp):
    """AdjustNegativeTracerMixingRatio.__call__ (neg_adj3.py:335-372) on
    columns ``[..., nz]``: returns a dict of the new fields."""
    pt, qv, ql, qr, qs, qi, qg = fix_neg_water(pt, qvapor, qliquid, qrain, qsnow, qice,
                                               qgraupel)
    qg = fillq(qg, delp)
    qr = fillq(qr, delp)
    qv = fix_water_vapor_down(qv, delp)
    qc = fix_neg_cloud(delp, qcld)
    return dict(qvapor=qv, qliquid=ql, qrain=qr, qsnow=qs, qice=qi, qgraupel=qg, qcld=qc,
                pt=pt)


# ------------------------------------------------------------- CubedToLatLon
C1 = 1.125
C2 = -0.125


def c2l_ord4(u, v, dx, dy, a11, a12, a21, a22, isc, iec, jsc, jec):
    """ord4_transform (c2l_ord.py:43-72) on one tile ``[I, J, K]`` with the
    halos of u and v filled: returns (ua, va) on the compute domain (NaN
    elsewhere)."""
    dx3, dy3 = dx[..., None], dy[..., None]
    utmp = C2 * (sh(u, 0, -1) + sh(u, 0, 2)) + C1 * (u + sh(u, 0, 1))
    vtmp = C2 * (sh(v, -1, 0) + sh(v, 2, 0)) + C1 * (v + sh(v, 1, 0))
    ni, nj = u.shape[:2]
    ii = np.arange(ni)[:, None, None]
    jj = np.arange(nj)[None, :, None]
    vedge = 2.0 * ((v * dy3) + (sh(v, 1, 0) * sh(dy3, 1, 0))) / (dy3 + sh(dy3, 1, 0))
    uedge = 2.0 * (u * dx3 + sh(u, 0, 1) * sh(dx3, 0, 1)) / (dx3 + sh(dx3, 0, 1))
    sn = (jj == jsc) | (jj == jec)
    vtmp = np.where(sn, vedge, vtmp)
    utmp = np.where(sn, uedge, utmp)
    we = (ii == isc) | (ii == iec)
    uedge2 = 2.0 * ((u * dx3) + (sh(u, 0, 1) * sh(dx3, 0, 1))) / (dx3 + sh(dx3, 0, 1))
    utmp = np.where(we, uedge2, utmp)
    vtmp = np.where(we, vedge, vtmp)
    ua = a11[..., None] * utmp + a12[..., None] * vtmp
    va = a21[..., None] * utmp + a22[..., None] * vtmp
    out_u = np.full(u.shape, np.nan)
    out_v = np.full(u.shape, np.nan)
    c = (slice(isc, iec + 1), slice(jsc, jec + 1))
    out_u[c] = ua[c]
    out_v[c] = va[c]
    return out_u, out_v


def c2l_ord2(u, v, dx, dy, a11, a12, a21, a22, isc, iec, jsc, jec):
    """c2l_ord2 (c2l_ord.py:17-40) on the compute domain plus one halo cell:
    returns (ua, va), NaN outside that box."""
    dx3, dy3 = dx[..., None], dy[..., None]
    wu = u * dx3
    wv = v * dy3
    u1 = 2.0 * (wu + sh(wu, 0, 1)) / (dx3 + sh(dx3, 0, 1))
    v1 = 2.0 * (wv + sh(wv, 1, 0)) / (dy3 + sh(dy3, 1, 0))
    ua = a11[..., None] * u1 + a12[..., None] * v1
    va = a21[..., None] * u1 + a22[..., None] * v1
    out_u = np.full(u.shape, np.nan)
    out_v = np.full(u.shape, np.nan)
    c = (slice(isc - 1, iec + 2), slice(jsc - 1, jec + 2))
    out_u[c] = ua[c]
    out_v[c] = va[c]
    return out_u, out_v
