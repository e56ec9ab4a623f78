"""CPU oracle for the vertical remapping step: LagrangianToEulerian with its
moist_cv pieces, MapNTracer and FillNegativeTracerValues (fillz).

Test infrastructure only: imported by tests, the smoke check and the CPU
baseline of the benchmark, never by the product path.  Follows
fv3core/stencils/remapping.py:33-628, moist_cv.py:8-228, mapn_tracer.py:10-85
and fillz.py:12-117 statement by statement.  Arrays are one tile
``[I, J, nz + 1]`` (interface storage; layer fields use levels 0..nz-1).  A
``computation`` with several ``interval`` blocks runs them in the order they
are written, each over its levels.
"""
from __future__ import annotations

import numpy as np

from .remap import map_single

RDGAS = 287.05
RVGAS = 461.50
GRAV = 9.80665
CP_AIR = 1004.6
CV_AIR = CP_AIR - RDGAS
RDG = -RDGAS / GRAV
CV_VAP = 3.0 * RVGAS
C_ICE = 1972.0
C_LIQ = 4.1855e3
CONSV_MIN = 0.001
T_MIN = 184.0

TRACERS = ("qvapor", "qliquid", "qrain", "qice", "qsnow", "qgraupel", "qo3mr", "qsgs_tke",
           "qcld")


def moist_cv_nwat6(qvapor, qliquid, qrain, qsnow, qice, qgraupel):
    """moist_cv_nwat6_fn + moist_cvm (moist_cv.py:14-41): (cvm, gz)."""
    ql = qliquid + qrain
    qs = qice + qsnow + qgraupel
    gz = ql + qs
    cvm = (1.0 - (qvapor + gz)) * CV_AIR + qvapor * CV_VAP + ql * C_LIQ + qs * C_ICE
    return cvm, gz


def set_cappa(qvapor, cvm, r_vir):
    """set_cappa (moist_cv.py:8-11)."""
    return RDGAS / (RDGAS + cvm / (1.0 + r_vir * qvapor))


def _nwat6(tr, box, levels):
    return moist_cv_nwat6(*(tr[n][box][..., levels] for n in
                            ("qvapor", "qliquid", "qrain", "qsnow", "qice", "qgraupel")))


def fillz(q, dp):
    """fix_tracer (fillz.py:12-94) on columns ``[..., km]``; returns a new q."""
    q = np.array(q, dtype=np.float64, copy=True)
    dp = np.asarray(dp, dtype=np.float64)
    km = q.shape[-1]
    h = q.shape[:-1]
    zfix = np.zeros(h, dtype=np.int64)
    lower = np.zeros(q.shape)
    upper = np.zeros(q.shape)
    dm = np.zeros(q.shape)
    dm_pos = np.zeros(q.shape)
    # fix_top: interval(0, 1), then interval(1, 2) reading the clamped top
    q[..., 0] = np.where(q[..., 0] < 0, 0.0, q[..., 0])
    dm[..., 0] = q[..., 0] * dp[..., 0]
    q[..., 1] = np.where(q[..., 0] < 0.0, q[..., 1] + q[..., 0] * dp[..., 0] / dp[..., 1],
                         q[..., 1])
    # fix_interior: FORWARD interval(1, -1)
    for k in range(1, km - 1):
        lf = lower[..., k - 1]
        q[..., k] = np.where(lf != 0.0, q[..., k] - lf / dp[..., k], q[..., k])
        neg = q[..., k] < 0.0
        zfix = zfix + neg
        qa = q[..., k - 1]
        cond = neg & (qa > 0.0)
        cand = qa * dp[..., k - 1]
        lim = -(q[..., k] * dp[..., k])
        dq = np.where(cand < lim, cand, lim)
        q[..., k] = np.where(cond, q[..., k] + dq / dp[..., k], q[..., k])
        upper[..., k] = np.where(cond, dq, upper[..., k])
        qb = q[..., k + 1]
        cond2 = neg & (q[..., k] < 0.0) & (qb > 0.0)
        cand = qb * dp[..., k + 1]
        lim = -(q[..., k] * dp[..., k])
        dq = np.where(cand < lim, cand, lim)
        q[..., k] = np.where(cond2, q[..., k] + dq / dp[..., k], q[..., k])
        lower[..., k] = np.where(cond2, dq, lower[..., k])
    # PARALLEL interval(0, -1)
    for k in range(0, km - 1):
        uf = upper[..., k + 1]
        q[..., k] = np.where(uf != 0.0, q[..., k] - uf / dp[..., k], q[..., k])
        dm[..., k] = q[..., k] * dp[..., k]
        dm_pos[..., k] = np.where(dm[..., k] > 0.0, dm[..., k], 0.0)
    # fix_bottom: FORWARD interval(-1, None)
    k = km - 1
    lf = lower[..., k - 1]
    q[..., k] = np.where(lf != 0.0, q[..., k] - lf / dp[..., k], q[..., k])
    qup = q[..., k - 1] * dp[..., k - 1]
    qly = -q[..., k] * dp[..., k]
    dup = np.where(qup < qly, qup, qly)
    cond = (q[..., k] < 0.0) & (q[..., k - 1] > 0.0)
    zfix = zfix + cond
    q[..., k] = np.where(cond, q[..., k] + dup / dp[..., k], q[..., k])
    upper[..., k] = np.where(cond, dup, upper[..., k])
    dm[..., k] = q[..., k] * dp[..., k]
    dm_pos[..., k] = np.where(dm[..., k] > 0.0, dm[..., k], 0.0)
    # PARALLEL interval(-2, -1)
    k = km - 2
    uf = upper[..., k + 1]
    on = uf != 0.0
    q[..., k] = np.where(on, q[..., k] - uf / dp[..., k], q[..., k])
    dm[..., k] = np.where(on, q[..., k] * dp[..., k], dm[..., k])
    dm_pos[..., k] = np.where(on, np.where(dm[..., k] > 0.0, dm[..., k], 0.0), dm_pos[..., k])
    # FORWARD interval(1, None) sums, then final_check
    sum0 = np.zeros(h)
    sum1 = np.zeros(h)
    for k in range(1, km):
        sum0 = sum0 + dm[..., k]
        sum1 = sum1 + dm_pos[..., k]
    with np.errstate(all="ignore"):
        fac = np.where(sum0 > 0.0, sum0 / sum1, 0.0)
        for k in range(1, km):
            val = fac * dm[..., k] / dp[..., k]
            new = np.where(val > 0.0, val, 0.0)
            q[..., k] = np.where((zfix > 0) & (fac > 0.0), new, q[..., k])
    return q


def _map(q, pe1, pe2, kord, mode, qmin=0.0, qs=None):
    """map_single over a whole box ``[ni, nj, nz + 1]``."""
    ni, nj = q.shape[:2]
    return map_single(q, pe1, pe2, kord, mode, 0, ni - 1, 0, nj - 1, qs=qs, qmin=qmin)


def lagrangian_to_eulerian(st, tracers, ak, bk, ptop, akap, zvir, last_step, cfg, nq, t,
                           consv_te=0.0, do_adiabatic_init=False, sat_adjust=None):
    """LagrangianToEulerian.__call__ (remapping.py:416-628) on one tile.

    ``st``: dict of tile arrays [I, J, nz + 1] (pt, delp, delz, peln, u, v, w,
    ua, cappa, q_con, pkz, pk, pe) and IJ arrays (ps, wsd); ``tracers``: dict
    of [I, J, nz + 1] arrays by name.  ``cfg``: kord_tm, kord_tr, kord_wz,
    kord_mt, fill.  ``t``: compute-domain bounds (isc, iec, jsc, jec).
    ``sat_adjust``: None, or the saturation adjustment's inputs (config,
    area, hs, dp1 — its energy scratch, updated in place — mdt, kmp), run
    before the final pt step as remapping.py:581-606 does (tracers must then
    hold qcld).  Returns updated copies (st, tracers).
    """
    st = {k: np.array(v, dtype=np.float64, copy=True) for k, v in st.items()}
    tr = {k: np.array(v, dtype=np.float64, copy=True) for k, v in tracers.items()}
    if cfg["kord_tm"] >= 0:
        raise NotImplementedError("map ppm, untested mode where kord_tm >= 0")
    nz = st["pt"].shape[2] - 1
    K, Ki = slice(0, nz), slice(0, nz + 1)
    i0, i1, j0, j1 = t.isc, t.iec, t.jsc, t.jec
    ni, nj = i1 - i0 + 1, j1 - j0 + 1
    C = (slice(i0, i1 + 1), slice(j0, j1 + 1))
    CJ = (slice(i0, i1 + 1), slice(j0, j1 + 2))   # init_pe / mapu box (extra row)
    CI = (slice(i0, i1 + 2), slice(j0, j1 + 1))   # mapv box (extra column)
    pe, peln = st["pe"], st["peln"]
    # init_pe (remapping.py:33-40) on the box with the extra row; the
    # component's pe2 never receives levels 1..nz-1 of that row
    pe1 = np.array(pe[CJ][..., Ki], copy=True)
    pe2 = np.zeros(pe1.shape)
    pe2[..., 0] = ptop
    pe2[..., nz] = pe1[..., nz]
    pe2c = pe2[:ni, :nj]
    pe1c = pe1[:ni, :nj]
    # moist_cv_pt_pressure (remapping.py:56-125)
    cvm, gz = _nwat6(tr, C, K)
    st["q_con"][C + (K,)] = gz
    cappa = set_cappa(tr["qvapor"][C][..., K], cvm, zvir)
    st["cappa"][C + (K,)] = cappa
    ptc = st["pt"][C][..., K]
    st["pt"][C + (K,)] = ptc * np.exp(cappa / (1.0 - cappa) * np.log(
        RDG * st["delp"][C][..., K] / st["delz"][C][..., K] * ptc))
    st["delz"][C + (K,)] = -st["delz"][C][..., K] / st["delp"][C][..., K]
    ps = pe[C][..., nz].copy()
    st["ps"][C] = ps
    pn2 = np.zeros(pe2c.shape)
    pn2[..., 0] = peln[C][..., 0]
    for k in range(1, nz):
        pe2c[..., k] = ak[k] + bk[k] * ps
    pn2[..., nz] = peln[C][..., nz]
    dp2 = np.zeros(pe2c.shape)
    for k in range(nz - 1, -1, -1):
        dp2[..., k] = pe2c[..., k + 1] - pe2c[..., k]
    st["delp"][C + (K,)] = dp2[..., K]
    # pn2_pk_delp (remapping.py:128-139)
    pn2[..., K] = np.log(pe2c[..., K])
    st["pk"][C + (K,)] = np.exp(akap * pn2[..., K])
    # map_single of pt in log-pressure
    st["pt"][C + (Ki,)] = _map(st["pt"][C][..., Ki], peln[C][..., Ki], pn2,
                               abs(cfg["kord_tm"]), 1, qmin=T_MIN)
    # MapNTracer: kord_tracer[5] is forced to 9 (mapn_tracer.py:41-42)
    for n in range(nq):
        name = TRACERS[n]
        kord = 9 if n == 5 else abs(cfg["kord_tr"])
        tr[name][C + (Ki,)] = _map(tr[name][C][..., Ki], pe1c, pe2c, kord, 0)
    if cfg["fill"]:
        for n in range(nq):
            name = TRACERS[n]
            tr[name][C + (K,)] = fillz(tr[name][C][..., K], dp2[..., K])
    st["w"][C + (Ki,)] = _map(st["w"][C][..., Ki], pe1c, pe2c, cfg["kord_wz"], -2,
                              qs=st["wsd"][C])
    st["delz"][C + (Ki,)] = _map(st["delz"][C][..., Ki], pe1c, pe2c, cfg["kord_wz"], 1)
    # undo_delz_adjust_and_copy_peln (remapping.py:43-53)
    st["delz"][C + (K,)] = -st["delz"][C][..., K] * st["delp"][C][..., K]
    peln[C + (Ki,)] = pn2
    # moist_pkz (moist_cv.py:181-202)
    cvm, gz = _nwat6(tr, C, K)
    st["q_con"][C + (K,)] = gz
    cappa = set_cappa(tr["qvapor"][C][..., K], cvm, zvir)
    st["cappa"][C + (K,)] = cappa
    st["pkz"][C + (K,)] = np.exp(cappa * np.log(RDG * st["delp"][C][..., K]
                                                 / st["delz"][C][..., K] * st["pt"][C][..., K]))
    # pressures_mapu (remapping.py:142-162) + map_single u on [isc, iec] x [jsc, jec+1]
    pe_j = pe[CJ][..., Ki]
    pe_jm = pe[(slice(i0, i1 + 1), slice(j0 - 1, j1 + 1))][..., Ki]
    pe0u = np.zeros(pe_j.shape)
    pe0u[..., 0] = pe_j[..., 0]
    pe0u[..., 1:] = 0.5 * (pe_jm[..., 1:] + pe1[..., 1:])
    bkh = 0.5 * bk[None, None, :nz + 1]
    pe3u = ak[None, None, :nz + 1] + bkh * (pe_jm[..., nz][..., None] + pe_j[..., nz][..., None])
    st["u"][CJ + (Ki,)] = _map(st["u"][CJ][..., Ki], pe0u, pe3u, cfg["kord_mt"], -1)
    # pressures_mapv (remapping.py:165-179) + map_single v on [isc, iec+1] x [jsc, jec]
    pe_i = pe[CI][..., Ki]
    pe_im = pe[(slice(i0 - 1, i1 + 1), slice(j0, j1 + 1))][..., Ki]
    pe0v = np.zeros(pe_i.shape)
    pe3v = np.zeros(pe_i.shape)
    pe0v[..., 0] = pe_i[..., 0]
    pe3v[..., 0] = ak[0]
    pe0v[..., 1:] = 0.5 * (pe_im[..., 1:] + pe_i[..., 1:])
    pe3v[..., 1:] = ak[None, None, 1:nz + 1] + bkh[..., 1:] * (
        pe_im[..., nz][..., None] + pe_i[..., nz][..., None])
    st["v"][CI + (Ki,)] = _map(st["v"][CI][..., Ki], pe0v, pe3v, cfg["kord_mt"], -1)
    # update_ua (remapping.py:182-193) and copy_from_below (196-198)
    ua = st["ua"]
    ua[CJ + (K,)] = pe2[..., 1:nz + 1]
    ua[(slice(i0, i1 + 1), j1 + 1, slice(0, nz - 1))] = pe2[:, nj - 1, 1:nz]
    pe[C + (slice(1, nz),)] = ua[C][..., 0:nz - 1]
    if last_step and not do_adiabatic_init and abs(consv_te) > CONSV_MIN:
        raise NotImplementedError("consv_te beyond CONSV_MIN is not supported")
    if sat_adjust is not None:
        from .satadjust import sat_adjust_3d

        sa = sat_adjust
        out = sat_adjust_3d(sa["config"], sa["area"], sa["dp1"], tr["qvapor"], tr["qliquid"],
                            tr["qice"], tr["qrain"], tr["qsnow"], tr["qgraupel"], tr["qcld"],
                            sa["hs"], st["peln"], st["delp"], st["delz"], st["q_con"], st["pt"],
                            st["pkz"], st["cappa"], zvir, sa["mdt"], False, last_step,
                            sa["kmp"], (i0, i1, j0, j1), nz)
        for name, key in (("qvapor", "qv"), ("qliquid", "ql"), ("qice", "qi"), ("qrain", "qr"),
                          ("qsnow", "qs"), ("qgraupel", "qg"), ("qcld", "qa")):
            tr[name] = out[key]
        for name in ("pt", "q_con", "cappa", "pkz", "delz"):
            st[name] = out[name]
        sa["dp1"][...] = out["te0"]
    if last_step:
        # moist_pt_last_step (moist_cv.py:143-169) on levels 0..nz, dtmp = 0
        g = (tr["qliquid"][C][..., Ki] + tr["qrain"][C][..., Ki] + tr["qice"][C][..., Ki]
             + tr["qsnow"][C][..., Ki] + tr["qgraupel"][C][..., Ki])
        dtmp = 0.0
        st["pt"][C + (Ki,)] = (st["pt"][C][..., Ki] + dtmp * st["pkz"][C][..., Ki]) / (
            (1.0 + zvir * tr["qvapor"][C][..., Ki]) * (1.0 - g))
    else:
        st["pt"][C + (K,)] = st["pt"][C][..., K] / st["pkz"][C][..., K]
    return st, tr
