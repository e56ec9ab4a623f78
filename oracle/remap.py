"""CPU oracle for the vertical remap: RemapProfile (cs_profile) and MapSingle.

Test infrastructure only (see oracle/common.py).  Follows
fv3core/stencils/remap_profile.py:119-638 and map_single.py:12-173 of the
reference statement by statement, vectorised over the horizontal and looping
over k where the reference runs FORWARD/BACKWARD.

Arrays are ``[i, j, k]``; the caller passes the horizontal box
(``i1..i2, j1..j2`` inclusive) the reference computes on.  Interval notation:
the profile's forward/backward pass runs on ``km + 1`` levels, the
constraints on ``km`` levels (``km`` = number of layers).
"""
from __future__ import annotations

import numpy as np


def _third(x):
    # the reference writes "x * 1.0 / 3.0", i.e. (x * 1.0) / 3.0
    return (x * 1.0) / 3.0


def remap_profile(qs, a4_1, delp, kord: int, iv: int, qmin: float = 0.0):
    """cs_profile: returns (a4_1, a4_2, a4_3, a4_4) for columns ``[..., km]``.

    ``qs`` is an IJ array (bottom boundary value), ``a4_1`` and ``delp`` are
    ``[..., km]`` (layer means and layer thicknesses).
    """
    if kord > 10:
        raise NotImplementedError(f"kord {kord} not implemented.")
    kord = abs(kord)
    a4_1 = np.array(a4_1, dtype=np.float64)
    delp = np.asarray(delp, dtype=np.float64)
    qs = np.asarray(qs, dtype=np.float64)
    km = a4_1.shape[-1]
    hshape = a4_1.shape[:-1]
    q = np.zeros(hshape + (km + 1,))
    gam = np.zeros(hshape + (km + 1,))
    a4_2 = np.zeros(hshape + (km,))
    a4_3 = np.zeros(hshape + (km,))
    a4_4 = np.zeros(hshape + (km,))

    # ---- set_initial_vals (remap_profile.py:119-207), domain km+1 levels
    if iv == -2:
        q[..., 0] = 1.5 * a4_1[..., 0]
        gam[..., 1] = 0.5
        grid_ratio = delp[..., 0] / delp[..., 1]
        bet = 2.0 + grid_ratio + grid_ratio - gam[..., 1]
        q[..., 1] = (3.0 * (a4_1[..., 0] + a4_1[..., 1]) - q[..., 0]) / bet
        gr = np.zeros(hshape + (km + 1,))
        for k in range(2, km):  # interval(2, -1) over km+1 levels
            old_grid_ratio = delp[..., k - 2] / delp[..., k - 1]
            old_bet = 2.0 + old_grid_ratio + old_grid_ratio - gam[..., k - 1]
            gam[..., k] = old_grid_ratio / old_bet
            gr[..., k] = delp[..., k - 1] / delp[..., k]
        for k in range(2, km - 1):  # interval(2, -2)
            bet = 2.0 + gr[..., k] + gr[..., k] - gam[..., k]
            q[..., k] = (3.0 * (a4_1[..., k - 1] + a4_1[..., k]) - q[..., k - 1]) / bet
        k = km - 1  # interval(-2, -1)
        q[..., k] = (
            3.0 * (a4_1[..., k - 1] + a4_1[..., k]) - gr[..., k] * qs - q[..., k - 1]
        ) / (2.0 + gr[..., k] + gr[..., k] - gam[..., k])
        q[..., km] = qs  # interval(-1, None)
        q_bot_km = qs
        for k in range(km - 2, -1, -1):  # BACKWARD interval(0, -2)
            q[..., k] = q[..., k] - gam[..., k + 1] * q[..., k + 1]
    else:
        grid_ratio = delp[..., 1] / delp[..., 0]
        bet = grid_ratio * (grid_ratio + 0.5)
        q[..., 0] = ((grid_ratio + grid_ratio) * (grid_ratio + 1.0) * a4_1[..., 0]
                     + a4_1[..., 1]) / bet
        gam[..., 0] = (1.0 + grid_ratio * (grid_ratio + 1.5)) / bet
        for k in range(1, km):  # interval(1, -1) over km+1 levels
            d4 = delp[..., k - 1] / delp[..., k]
            bet = 2.0 + d4 + d4 - gam[..., k - 1]
            q[..., k] = (3.0 * (a4_1[..., k - 1] + d4 * a4_1[..., k]) - q[..., k - 1]) / bet
            gam[..., k] = d4 / bet
        d4 = delp[..., km - 2] / delp[..., km - 1]
        a_bot = 1.0 + d4 * (d4 + 1.5)
        q[..., km] = (2.0 * d4 * (d4 + 1.0) * a4_1[..., km - 1] + a4_1[..., km - 2]
                      - a_bot * q[..., km - 1]) / (d4 * (d4 + 0.5) - a_bot * gam[..., km - 1])
        q_bot_km = None
        for k in range(km - 1, -1, -1):  # BACKWARD interval(0, -1)
            q[..., k] = q[..., k] - gam[..., k] * q[..., k + 1]
    del q_bot_km

    # ---- apply_constraints (remap_profile.py:210-280), domain km levels
    tmp = np.zeros(hshape + (km,))
    tmp2 = np.zeros(hshape + (km,))
    g = np.zeros(hshape + (km + 1,))  # the reference reuses gam here
    for k in range(1, km):
        a0 = a4_1[..., k - 1]
        tmp[..., k] = np.where(a0 > a4_1[..., k], a0, a4_1[..., k])
        tmp2[..., k] = np.where(a0 < a4_1[..., k], a0, a4_1[..., k])
        g[..., k] = a4_1[..., k] - a0
    if iv == -2:
        # gam entries the constraints do not overwrite keep the profile's
        # values (levels 0 and km)
        g[..., 0] = gam[..., 0]
        g[..., km] = gam[..., km]
    else:
        g[..., 0] = gam[..., 0]
        g[..., km] = gam[..., km]
    qc = q.copy()

    def clamp_both(x, k):
        x = np.where(x >= tmp[..., k], tmp[..., k], x)
        x = np.where(x <= tmp2[..., k], tmp2[..., k], x)
        return x

    qc[..., 1] = clamp_both(qc[..., 1], 1)
    for k in range(2, km - 1):
        x = qc[..., k]
        both = (g[..., k - 1] * g[..., k + 1]) > 0
        locmax = ~both & (g[..., k - 1] > 0)
        locmin = ~both & ~(g[..., k - 1] > 0)
        xb = clamp_both(x, k)
        xmax = np.where(x <= tmp2[..., k], tmp2[..., k], x)
        xmin = np.where(x >= tmp[..., k], tmp[..., k], x)
        if iv == 0:
            xmin = np.where(xmin < 0.0, 0.0, xmin)
        qc[..., k] = np.where(both, xb, np.where(locmax, xmax, np.where(locmin, xmin, x)))
    qc[..., km - 1] = clamp_both(qc[..., km - 1], km - 1)
    a4_2 = qc[..., :km].copy()
    a4_3 = qc[..., 1:km + 1].copy()
    extm = np.zeros(hshape + (km,), dtype=bool)
    extm[..., 0] = (a4_2[..., 0] - a4_1[..., 0]) * (a4_3[..., 0] - a4_1[..., 0]) > 0.0
    for k in range(1, km - 1):
        extm[..., k] = g[..., k] * g[..., k + 1] < 0.0
    extm[..., km - 1] = ((a4_2[..., km - 1] - a4_1[..., km - 1])
                         * (a4_3[..., km - 1] - a4_1[..., km - 1]) > 0.0)
    ext5 = np.zeros(hshape + (km,), dtype=bool)
    ext6 = np.zeros(hshape + (km,), dtype=bool)
    if kord > 9:
        x0 = 2.0 * a4_1 - (a4_2 + a4_3)
        x1 = np.abs(a4_2 - a4_3)
        a4_4 = 3.0 * x0
        ext5 = np.abs(x0) > x1
        ext6 = np.abs(a4_4) > x1

    # ---- set_interpolation_coefficients (remap_profile.py:283-495)
    def a4_4_std(k):
        return 3.0 * (2.0 * a4_1[..., k] - (a4_2[..., k] + a4_3[..., k]))

    if iv == 0:
        a4_2[..., 0] = np.where(a4_2[..., 0] < 0.0, 0.0, a4_2[..., 0])
        for k in (0, 1):
            a4_4[..., k] = a4_4_std(k)
    if iv == -1:
        a4_2[..., 0] = np.where(a4_2[..., 0] * a4_1[..., 0] <= 0.0, 0.0, a4_2[..., 0])
        for k in (0, 1):
            a4_4[..., k] = a4_4_std(k)
    if iv == 2:
        a4_2[..., 0] = a4_1[..., 0]
        a4_3[..., 0] = a4_1[..., 0]
        a4_4[..., 0] = 0.0
        a4_4[..., 1] = a4_4_std(1)
    if iv < -1 or iv == 1 or iv > 2:
        for k in (0, 1):
            a4_4[..., k] = a4_4_std(k)
    if iv != 2:
        c = posdef_constraint_iv1(a4_1[..., 0], a4_2[..., 0], a4_3[..., 0], a4_4[..., 0])
        a4_2[..., 0], a4_3[..., 0], a4_4[..., 0] = c
    c = remap_constraint(a4_1[..., 1], a4_2[..., 1], a4_3[..., 1], a4_4[..., 1], extm[..., 1])
    a4_2[..., 1], a4_3[..., 1], a4_4[..., 1] = c

    # interior interval(2, -2): reads of a4_2/a4_3/a4_4 at other levels see
    # the values from apply_constraints (parallel semantics)
    b2 = a4_2.copy()
    b3 = a4_3.copy()
    b4 = a4_4.copy()
    for k in range(2, km - 2):
        A1, A2, A3, A4 = a4_1[..., k], a4_2[..., k].copy(), a4_3[..., k].copy(), a4_4[..., k].copy()
        gk, gp1, gp2, gm1 = g[..., k], g[..., k + 1], g[..., k + 2], g[..., k - 1]
        if kord < 9:
            pmp_1 = A1 - gp1
            lac_1 = pmp_1 + 1.5 * gp2
            tmp_min = _pick_min3(A1, pmp_1, lac_1)
            tmp_max0 = np.where(A2 > tmp_min, A2, tmp_min)
            tmp_max = _pick_max3(A1, pmp_1, lac_1)
            A2 = np.where(tmp_max0 < tmp_max, tmp_max0, tmp_max)
            pmp_2 = A1 + 2.0 * gp1
            lac_2 = pmp_2 - 1.5 * gm1
            tmp_min = _pick_min3(A1, pmp_2, lac_2)
            tmp_max0 = np.where(A3 > tmp_min, A3, tmp_min)
            tmp_max = _pick_max3(A1, pmp_2, lac_2)
            A3 = np.where(tmp_max0 < tmp_max, tmp_max0, tmp_max)
            A4 = 3.0 * (2.0 * A1 - (A2 + A3))
        if kord == 9:
            pmp_1 = A1 - 2.0 * gp1
            lac_1 = pmp_1 + 1.5 * gp2
            pmp_2 = A1 + 2.0 * gk
            lac_2 = pmp_2 - 1.5 * gm1
            flat = (extm[..., k] & extm[..., k - 1]) | (extm[..., k] & extm[..., k + 1]) | (
                extm[..., k] & ((qmin > 0.0) & (A1 < qmin))
            )
            A4n = 6.0 * A1 - 3.0 * (A2 + A3)
            big = np.abs(A4n) > np.abs(A2 - A3)
            tmp_min = _pick_min3(A1, pmp_1, lac_1)
            tmp_max0 = np.where(A2 > tmp_min, A2, tmp_min)
            tmp_max = _pick_max3(A1, pmp_1, lac_1)
            A2b = np.where(tmp_max0 < tmp_max, tmp_max0, tmp_max)
            tmp_min = _pick_min3(A1, pmp_2, lac_2)
            tmp_max0 = np.where(A3 > tmp_min, A3, tmp_min)
            tmp_max = _pick_max3(A1, pmp_2, lac_2)
            A3b = np.where(tmp_max0 < tmp_max, tmp_max0, tmp_max)
            A4b = 6.0 * A1 - 3.0 * (A2b + A3b)
            A2 = np.where(flat, A1, np.where(big, A2b, A2))
            A3 = np.where(flat, A1, np.where(big, A3b, A3))
            A4 = np.where(flat, 0.0, np.where(big, A4b, A4n))
        if kord == 10:
            pmp_1 = A1 - 2.0 * gp1
            lac_1 = pmp_1 + 1.5 * gp2
            pmp_2 = A1 + 2.0 * gk
            lac_2 = pmp_2 - 1.5 * gm1
            tmp_min2 = _pick_min3(A1, pmp_1, lac_1)
            tmp_max2 = _pick_max3(A1, pmp_1, lac_1)
            tmp2_ = np.where(A2 > tmp_min2, A2, tmp_min2)
            tmp_min3 = np.where(A1 < pmp_2, A1, pmp_2)
            tmp_min3 = np.where(lac_2 < tmp_min3, lac_2, tmp_min3)
            tmp_max3 = np.where(A1 > pmp_2, A1, pmp_2)
            tmp_max3 = np.where(lac_2 > tmp_max3, lac_2, tmp_max3)
            tmp3_ = np.where(A3 > tmp_min3, A3, tmp_min3)
            e5, e6 = ext5[..., k], ext6[..., k]
            e5n = ext5[..., k - 1] | ext5[..., k + 1]
            e6n = ext6[..., k - 1] | ext6[..., k + 1]
            lim2 = np.where(tmp2_ < tmp_max2, tmp2_, tmp_max2)
            lim3 = np.where(tmp3_ < tmp_max3, tmp3_, tmp_max3)
            flat = e5 & e5n
            lim = (e5 & ~e5n & e6n) | (~e5 & e6 & e5n)
            A2 = np.where(flat, A1, np.where(lim, lim2, A2))
            A3 = np.where(flat, A1, np.where(lim, lim3, A3))
            A4 = 3.0 * (2.0 * A1 - (A2 + A3))
        if iv == 0:
            A2, A3, A4 = posdef_constraint_iv0(A1, A2, A3, A4)
        b2[..., k], b3[..., k], b4[..., k] = A2, A3, A4
    a4_2, a4_3, a4_4 = b2, b3, b4

    k = km - 1
    if iv == 0:
        a4_3[..., k] = np.where(a4_3[..., k] < 0.0, 0.0, a4_3[..., k])
    if iv == -1:
        a4_3[..., k] = np.where(a4_3[..., k] * a4_1[..., k] <= 0.0, 0.0, a4_3[..., k])
    for k in (km - 2, km - 1):
        a4_4[..., k] = a4_4_std(k)
    k = km - 2
    c = remap_constraint(a4_1[..., k], a4_2[..., k], a4_3[..., k], a4_4[..., k], extm[..., k])
    a4_2[..., k], a4_3[..., k], a4_4[..., k] = c
    k = km - 1
    c = posdef_constraint_iv1(a4_1[..., k], a4_2[..., k], a4_3[..., k], a4_4[..., k])
    a4_2[..., k], a4_3[..., k], a4_4[..., k] = c
    return a4_1, a4_2, a4_3, a4_4


def _pick_min3(a, p, l):
    """``a if a < p and a < l else p if p < l else l`` (remap_profile.py:310-316)."""
    return np.where((a < p) & (a < l), a, np.where(p < l, p, l))


def _pick_max3(a, p, l):
    return np.where((a > p) & (a > l), a, np.where(p > l, p, l))


def posdef_constraint_iv0(a4_1, a4_2, a4_3, a4_4):
    """remap_profile.py:43-69."""
    nonpos = a4_1 <= 0.0
    cond = (np.abs(a4_3 - a4_2) < -a4_4)
    with np.errstate(divide="ignore", invalid="ignore"):
        cond = cond & ((a4_1 + 0.25 * (a4_3 - a4_2) ** 2 / a4_4 + a4_4 * (1.0 / 12.0)) < 0.0)
    c1 = cond & (a4_1 < a4_3) & (a4_1 < a4_2)
    c2 = cond & ~c1 & (a4_3 > a4_2)
    c3 = cond & ~c1 & ~c2
    n2 = np.where(c1, a4_1, a4_2)
    n3 = np.where(c1, a4_1, a4_3)
    n4 = np.where(c1, 0.0, a4_4)
    n4 = np.where(c2, 3.0 * (a4_2 - a4_1), n4)
    n3 = np.where(c2, a4_2 - 3.0 * (a4_2 - a4_1), n3)
    n4 = np.where(c3, 3.0 * (a4_3 - a4_1), n4)
    n2 = np.where(c3, a4_3 - 3.0 * (a4_3 - a4_1), n2)
    n2 = np.where(nonpos, a4_1, n2)
    n3 = np.where(nonpos, a4_1, n3)
    n4 = np.where(nonpos, 0.0, n4)
    return n2, n3, n4


def posdef_constraint_iv1(a4_1, a4_2, a4_3, a4_4):
    """remap_profile.py:72-92."""
    da1 = a4_3 - a4_2
    da2 = da1 * da1
    a6da = a4_4 * da1
    c1 = ((a4_1 - a4_2) * (a4_1 - a4_3)) >= 0.0
    c2 = ~c1 & (a6da < -1.0 * da2)
    c3 = ~c1 & ~c2 & (a6da > da2)
    n2 = np.where(c1, a4_1, a4_2)
    n3 = np.where(c1, a4_1, a4_3)
    n4 = np.where(c1, 0.0, a4_4)
    n4 = np.where(c2, 3.0 * (a4_2 - a4_1), n4)
    n3 = np.where(c2, a4_2 - 3.0 * (a4_2 - a4_1), n3)
    n4 = np.where(c3, 3.0 * (a4_3 - a4_1), n4)
    n2 = np.where(c3, a4_3 - 3.0 * (a4_3 - a4_1), n2)
    return n2, n3, n4


def remap_constraint(a4_1, a4_2, a4_3, a4_4, extm):
    """remap_profile.py:95-116."""
    da1 = a4_3 - a4_2
    da2 = da1 * da1
    a6da = a4_4 * da1
    c1 = extm
    c2 = ~c1 & (a6da < -da2)
    c3 = ~c1 & ~c2 & (a6da > da2)
    n2 = np.where(c1, a4_1, a4_2)
    n3 = np.where(c1, a4_1, a4_3)
    n4 = np.where(c1, 0.0, a4_4)
    n4 = np.where(c2, 3.0 * (a4_2 - a4_1), n4)
    n3 = np.where(c2, a4_2 - 3.0 * (a4_2 - a4_1), n3)
    n4 = np.where(c3, 3.0 * (a4_3 - a4_1), n4)
    n2 = np.where(c3, a4_3 - 3.0 * (a4_3 - a4_1), n2)
    return n2, n3, n4


def lagrangian_contributions(pe1, pe2, a4_1, a4_2, a4_3, a4_4, dp1):
    """map_single.py:19-66: conservative integral of the source parabolas over
    each target layer; ``lev`` walks the source layers (scalar loop per column).

    Arrays are ``[ncol, km(+1)]``; returns q ``[ncol, km]``.
    """
    ncol, km = a4_1.shape
    out = np.zeros((ncol, km))
    for c in range(ncol):
        L = 0
        for k in range(km):
            pl = (pe2[c, k] - pe1[c, L]) / dp1[c, L]
            if pe2[c, k + 1] <= pe1[c, L + 1]:
                pr = (pe2[c, k + 1] - pe1[c, L]) / dp1[c, L]
                out[c, k] = (
                    a4_2[c, L]
                    + 0.5 * (a4_4[c, L] + a4_3[c, L] - a4_2[c, L]) * (pr + pl)
                    - _third(a4_4[c, L]) * (pr * (pr + pl) + pl * pl)
                )
            else:
                qsum = (pe1[c, L + 1] - pe2[c, k]) * (
                    a4_2[c, L]
                    + 0.5 * (a4_4[c, L] + a4_3[c, L] - a4_2[c, L]) * (1.0 + pl)
                    - _third(a4_4[c, L]) * (1.0 + pl * (1.0 + pl))
                )
                L += 1
                while pe1[c, L + 1] < pe2[c, k + 1]:
                    qsum += dp1[c, L] * a4_1[c, L]
                    L += 1
                dp = pe2[c, k + 1] - pe1[c, L]
                esl = dp / dp1[c, L]
                qsum += dp * (
                    a4_2[c, L]
                    + 0.5 * esl * (a4_3[c, L] - a4_2[c, L] + a4_4[c, L] * (1.0 - (2.0 / 3.0) * esl))
                )
                out[c, k] = qsum / (pe2[c, k + 1] - pe2[c, k])
    return out


def map_single(q1, pe1, pe2, kord: int, mode: int, i1, i2, j1, j2, qs=None, qmin=0.0):
    """MapSingle.__call__ (map_single.py:130-173): returns the remapped q1.

    ``q1``: ``[I, J, >=km]``, ``pe1``/``pe2``: ``[I, J, >=km+1]``; the box
    ``i1..i2, j1..j2`` is remapped, other entries are returned unchanged.
    ``km`` is ``pe1.shape[2] - 1`` unless q1 has fewer levels.
    """
    out = np.array(q1, dtype=np.float64, copy=True)
    km = pe1.shape[2] - 1
    sl = (slice(i1, i2 + 1), slice(j1, j2 + 1))
    a1 = out[sl][..., :km].copy()
    p1 = np.asarray(pe1, dtype=np.float64)[sl][..., : km + 1]
    p2 = np.asarray(pe2, dtype=np.float64)[sl][..., : km + 1]
    dp1 = p1[..., 1:] - p1[..., :-1]
    qsv = np.zeros(a1.shape[:2]) if qs is None else np.asarray(qs, dtype=np.float64)[sl]
    a4 = remap_profile(qsv, a1, dp1, kord, mode, qmin)
    ni, nj = a1.shape[:2]
    res = lagrangian_contributions(
        p1.reshape(ni * nj, km + 1), p2.reshape(ni * nj, km + 1),
        *(x.reshape(ni * nj, km) for x in a4), dp1.reshape(ni * nj, km),
    )
    out[sl + (slice(0, km),)] = res.reshape(ni, nj, km)
    return out
