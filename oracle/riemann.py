"""CPU oracle for the vertical semi-implicit solvers: sim1_solver,
RiemannSolverC and RiemannSolver3.

Test infrastructure only (see oracle/common.py).  Follows
fv3core/stencils/sim1_solver.py:11-101, riem_solver_c.py:13-165 and
riem_solver3.py:24-242 statement by statement.  Column arrays are
``[..., k]`` with ``nk = nz + 1`` levels (interfaces), vectorised over the
leading (horizontal) axes.
"""
from __future__ import annotations

import math

import numpy as np

RDGAS = 287.05
GRAV = 9.80665
RGRAV = 1.0 / GRAV
CP_AIR = 1004.6
KAPPA = RDGAS / CP_AIR


def sim1_solver(w, dm, gm, dz, ptr, pm, pem, wsr, cp3, dt, p_fac):
    """sim1_solver (sim1_solver.py:11-101) on columns of ``nk`` levels.

    Returns new (w, pe, dz).  ``wsr`` is ``[...]`` (no k axis).  Levels are
    indexed as in the reference's domain of nz + 1 levels.
    """
    t1g = 2.0 * dt * dt
    rdt = 1.0 / dt
    w = np.array(w, dtype=np.float64, copy=True)
    dz = np.array(dz, dtype=np.float64, copy=True)
    nk = w.shape[-1]
    nz = nk - 1
    h = w.shape[:-1]
    pe = np.zeros(h + (nk,))
    w1 = np.zeros(h + (nk,))
    pe[..., :nz] = np.exp(gm[..., :nz] * np.log(-dm[..., :nz] / dz[..., :nz] * RDGAS * ptr[..., :nz])) - pm[..., :nz]
    w1[..., :nz] = w[..., :nz]
    g_rat = np.zeros(h + (nk,))
    bb = np.zeros(h + (nk,))
    dd = np.zeros(h + (nk,))
    for k in range(nz - 1):
        g_rat[..., k] = dm[..., k] / dm[..., k + 1]
        bb[..., k] = 2.0 * (1.0 + g_rat[..., k])
        dd[..., k] = 3.0 * (pe[..., k] + g_rat[..., k] * pe[..., k + 1])
    bb[..., nz - 1] = 2.0
    dd[..., nz - 1] = 3.0 * pe[..., nz - 1]
    bet = np.zeros(h + (nk,))
    bet[..., :nz] = bb[..., 0:1]
    pp = np.zeros(h + (nk,))
    pp[..., 0] = 0.0
    pp[..., 1] = dd[..., 0] / bet[..., 1]
    gam = np.zeros(h + (nk,))
    for k in range(1, nz):
        gam[..., k] = g_rat[..., k - 1] / bet[..., k - 1]
        bet[..., k] = bb[..., k] - gam[..., k]
    for k in range(2, nk):
        pp[..., k] = (dd[..., k - 1] - pp[..., k - 1]) / bet[..., k - 1]
    aa = np.zeros(h + (nk,))
    for k in range(nz - 1, 0, -1):
        pp[..., k] = pp[..., k] - gam[..., k] * pp[..., k + 1]
        aa[..., k] = (t1g * 0.5 * (gm[..., k - 1] + gm[..., k]) / (dz[..., k - 1] + dz[..., k])
                      * (pem[..., k] + pp[..., k]))
    bet[..., :] = (dm[..., 0] - aa[..., 1])[..., None]
    w[..., 0] = (dm[..., 0] * w1[..., 0] + dt * pp[..., 1]) / bet[..., 0]
    for k in range(1, nz - 1):
        gam[..., k] = aa[..., k] / bet[..., k - 1]
        bet[..., k] = dm[..., k] - (aa[..., k] + aa[..., k + 1] + aa[..., k] * gam[..., k])
        w[..., k] = (dm[..., k] * w1[..., k] + dt * (pp[..., k + 1] - pp[..., k])
                     - aa[..., k] * w[..., k - 1]) / bet[..., k]
    k = nz - 1
    p1k = t1g * gm[..., k] / dz[..., k] * (pem[..., k + 1] + pp[..., k + 1])
    gam[..., k] = aa[..., k] / bet[..., k - 1]
    bet[..., k] = dm[..., k] - (aa[..., k] + p1k + aa[..., k] * gam[..., k])
    w[..., k] = (dm[..., k] * w1[..., k] + dt * (pp[..., k + 1] - pp[..., k]) - p1k * wsr
                 - aa[..., k] * w[..., k - 1]) / bet[..., k]
    for k in range(nz - 2, -1, -1):
        w[..., k] = w[..., k] - gam[..., k + 1] * w[..., k + 1]
    pe[..., 0] = 0.0
    for k in range(1, nk):
        pe[..., k] = pe[..., k - 1] + dm[..., k - 1] * (w[..., k - 1] - w1[..., k - 1]) * rdt
    p1 = np.zeros(h + (nk,))
    k = nz - 1
    p1[..., k] = (pe[..., k] + 2.0 * pe[..., k + 1]) * 1.0 / 3.0
    for k in range(nz - 2, -1, -1):
        p1[..., k] = ((pe[..., k] + bb[..., k] * pe[..., k + 1] + g_rat[..., k] * pe[..., k + 2])
                      * 1.0 / 3.0 - g_rat[..., k] * p1[..., k + 1])
    s = (slice(None),) * len(h) + (slice(0, nz),)
    maxp = np.where(p_fac * dm[s] > p1[s] + pm[s], p_fac * pm[s], p1[s] + pm[s])
    dz[s] = -dm[s] * RDGAS * ptr[s] * np.exp((cp3[s] - 1.0) * np.log(maxp))
    return w, pe, dz


def riem_solver_c(dt2, cappa, ptop, hs, ws, ptc, q_con, delpc, gz, w3, p_fac):
    """RiemannSolverC.__call__ (riem_solver_c.py:119-165) on columns.

    All 3-D inputs are ``[..., nk]`` (nk = nz + 1), ``hs``/``ws`` ``[...]``.
    Returns (gz, pef); w3 is input only, as in the reference.
    """
    nk = delpc.shape[-1]
    nz = nk - 1
    h = delpc.shape[:-1]
    dm = np.array(delpc, dtype=np.float64, copy=True)
    w = np.array(w3, dtype=np.float64, copy=True)
    pem = np.zeros(h + (nk,))
    peg = np.zeros(h + (nk,))
    pem[..., 0] = ptop
    peg[..., 0] = ptop
    for k in range(1, nk):
        pem[..., k] = pem[..., k - 1] + dm[..., k - 1]
        peg[..., k] = peg[..., k - 1] + dm[..., k - 1] * (1.0 - q_con[..., k - 1])
    dz = np.zeros(h + (nk,))
    dz[..., :nz] = gz[..., 1:] - gz[..., :nz]
    gm = 1.0 / (1.0 - cappa)
    dm = dm / GRAV
    pm = np.zeros(h + (nk,))
    pm[..., :nz] = (peg[..., 1:] - peg[..., :nz]) / np.log(peg[..., 1:] / peg[..., :nz])
    w, pe, dz = sim1_solver(w, dm, gm, dz, ptc, pm, pem, ws, cappa, dt2, p_fac)
    pef = np.zeros(h + (nk,))
    pef[..., 0] = ptop
    pef[..., 1:] = pe[..., 1:] + pem[..., 1:]
    gz_out = np.zeros(h + (nk,))
    gz_out[..., nz] = hs
    for k in range(nz - 1, -1, -1):
        gz_out[..., k] = gz_out[..., k + 1] - dz[..., k] * GRAV
    return gz_out, pef


def riem_solver3(last_call, dt, cappa, ptop, zs, wsd, delz, q_con, delp, pt, zh, pe, ppe, pk3,
                 pk, peln, w, p_fac, use_logp=False, beta=0.0):
    """RiemannSolver3.__call__ (riem_solver3.py:145-242) on columns.

    Returns new (zh, pe, ppe, pk3, pk, peln, w, delz); inputs are not modified.
    """
    nk = delp.shape[-1]
    nz = nk - 1
    h = delp.shape[:-1]
    peln1 = math.log(ptop)
    ptk = math.exp(KAPPA * peln1)
    dm = np.array(delp, dtype=np.float64, copy=True)
    pe_init = np.array(pe, dtype=np.float64, copy=True)
    pem = np.zeros(h + (nk,))
    peln_run = np.zeros(h + (nk,))
    pk3 = np.array(pk3, dtype=np.float64, copy=True)
    peg = np.zeros(h + (nk,))
    pelng = np.zeros(h + (nk,))
    pem[..., 0] = ptop
    peln_run[..., 0] = peln1
    pk3[..., 0] = ptk
    peg[..., 0] = ptop
    pelng[..., 0] = peln1
    for k in range(1, nk):
        pem[..., k] = pem[..., k - 1] + dm[..., k - 1]
        peln_run[..., k] = np.log(pem[..., k])
        peg[..., k] = peg[..., k - 1] + dm[..., k - 1] * (1.0 - q_con[..., k - 1])
        pelng[..., k] = np.log(peg[..., k])
        pk3[..., k] = np.exp(KAPPA * peln_run[..., k])
    gm = 1.0 / (1.0 - cappa)
    dm = dm * RGRAV
    pm = np.zeros(h + (nk,))
    pm[..., :nz] = (peg[..., 1:] - peg[..., :nz]) / (pelng[..., 1:] - pelng[..., :nz])
    delz = np.array(delz, dtype=np.float64, copy=True)
    delz[..., :nz] = zh[..., 1:] - zh[..., :nz]
    w, pe_new, delz = sim1_solver(w, dm, gm, delz, pt, pm, pem, wsd, cappa, dt, p_fac)
    if use_logp:
        pk3 = peln_run.copy()
    ppe = pe_new + pem if beta < -0.1 else pe_new.copy()
    if last_call:
        peln = peln_run.copy()
        pk = pk3.copy()
        pe_out = pem.copy()
    else:
        peln = np.array(peln, copy=True)
        pk = np.array(pk, copy=True)
        pe_out = pe_init
    zh_out = np.zeros(h + (nk,))
    zh_out[..., nz] = zs
    for k in range(nz - 1, -1, -1):
        zh_out[..., k] = zh_out[..., k + 1] - delz[..., k]
    return zh_out, pe_out, ppe, pk3, pk, peln, w, delz
