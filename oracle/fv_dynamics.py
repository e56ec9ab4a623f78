"""CPU oracle for DynamicalCore.step_dynamics (fv3core/stencils/fv_dynamics.py:
193-557), with its wrap-up (neg_adj3, CubedToLatLon; fv_dynamics.py:136-169)
when asked.

Test infrastructure only: imported by tests, the smoke check and the CPU
baseline of the benchmark, never by the product path.  Composes the oracle's
acoustic_dynamics, tracer_advection and lagrangian_to_eulerian in the
reference's order, with fv_setup / pt_adjust (moist_cv.py:197-228,
fv_dynamics.py:32-34) before and set_omega + del2cubed of omga after.
"""
from __future__ import annotations

import numpy as np

from . import remapping as orm
from .dyncore import acoustic_dynamics, del2cubed
from .fvtp2d import Tile
from .tracer import tracer_advection
from .wrapup import c2l_ord2, c2l_ord4, neg_adj3

RDGAS = orm.RDGAS
RVGAS = orm.RVGAS
ZVIR = RVGAS / RDGAS - 1
KAPPA = RDGAS / orm.CP_AIR
NQ = 8


def init_pfull(ak, bk, p_ref=1.0e5):
    """init_pfull (fv_dynamics.py:42-51)."""
    ph1 = ak[:-1] + bk[:-1] * p_ref
    ph2 = ak[1:] + bk[1:] * p_ref
    return (ph2 - ph1) / np.log(ph2 / ph1)


def fv_setup_pt_adjust(S, t, nz):
    """fv_setup (moist_phys, nwat = 6) then pt_adjust on the compute domain, in place."""
    C = (slice(None), slice(t.isc, t.iec + 1), slice(t.jsc, t.jec + 1), slice(0, nz))
    cvm, q_con = orm.moist_cv_nwat6(S["qvapor"][C], S["qliquid"][C], S["qrain"][C],
                                    S["qsnow"][C], S["qice"][C], S["qgraupel"][C])
    dp1 = ZVIR * S["qvapor"][C]
    cappa = RDGAS / (RDGAS + cvm / (1.0 + dp1))
    pkz = np.exp(cappa * np.log(orm.RDG * S["delp"][C] * S["pt"][C] * (1.0 + dp1)
                                * (1.0 - q_con) / S["delz"][C]))
    S["cvm"][C] = cvm
    S["q_con"][C] = q_con
    S["dp1"][C] = dp1
    S["cappa"][C] = cappa
    S["pkz"][C] = pkz
    S["pt"][C] = S["pt"][C] * (1.0 + dp1) * (1.0 - q_con) / pkz


def fv_dynamics(S, cfg, metrics, scalars, ak, bk, halo, tracer_halo, sync, col, nz,
                edges=True, wrapup=False):
    """One step_dynamics on a batch of tiles, in place on ``S``.

    ``edges``: the tile-edge flags (True for the cube tiles, False for a
    doubly periodic plane tile whose halo functions wrap periodically).
    ``S``: dict of [T, I, J, nz + 1] arrays (IJ fields [T, I, J]) holding the
    reference's ArgSpec fields plus the step and acoustic temporaries (cappa,
    dp1, cvm, wsd, te0_2d, ut, vt, gz, zh, pem, pkc, pk3, heat_source, divgd,
    crx, xfx, cry, yfx, ws3, delpc, ptc).  ``cfg``: namelist values plus bdt
    (the step), ptop, ks, k_split, n_split, hord_tr, kord_*, fill, z_tracer,
    nf_omega, p_ref, and optionally do_sat_adj with sat_adjust (a
    SatAdjustConfig).  ``halo(kind, names)`` refreshes state halos in place;
    ``tracer_halo(list)`` returns refreshed copies; ``sync()`` joins the
    D-grid edge winds.  ``wrapup``: end with neg_adj3 and CubedToLatLon
    (order ``cfg["c2l_ord"]``, default 4; the metrics hold a11..a22).
    """
    n = S["delp"].shape[1] - 7
    ntiles = S["delp"].shape[0]
    t = Tile(n, n, 3, edges, edges, edges, edges)
    pfull = init_pfull(ak[:nz + 1], bk[:nz + 1], cfg.get("p_ref", 1.0e5))
    fv_setup_pt_adjust(S, t, nz)
    k_split = cfg["k_split"]
    tr_names = orm.TRACERS[:NQ]
    for n_map in range(k_split):
        last_step = n_map == k_split - 1
        S["dp1"][...] = S["delp"]
        acfg = dict(cfg)
        acfg.update(mdt=cfg["bdt"] / k_split, n_map=n_map + 1)
        acoustic_dynamics(S, acfg, metrics, scalars, ak, bk, pfull, halo, sync, col, nz,
                          edges=edges)
        if cfg.get("z_tracer", True):
            metric6 = {k: np.stack([m[k] for m in metrics]) for k in metrics[0]}
            qs, dp1, mfxd, mfyd, cxd, cyd = tracer_advection(
                [S[k] for k in tr_names], S["dp1"], S["mfxd"], S["mfyd"], S["cxd"], S["cyd"],
                metric6, cfg["hord_tr"], tracer_halo, nz, edges=edges)
            for k, q in zip(tr_names, qs):
                S[k][...] = q
            S["dp1"][...] = dp1
            S["mfxd"][...], S["mfyd"][...], S["cxd"][...], S["cyd"][...] = mfxd, mfyd, cxd, cyd
        if nz > 4:
            rcfg = dict(kord_tm=cfg["kord_tm"], kord_tr=cfg["kord_tr"], kord_wz=cfg["kord_wz"],
                        kord_mt=cfg["kord_mt"], fill=cfg["fill"])
            names = ("pt", "delp", "delz", "peln", "u", "v", "w", "ua", "cappa", "q_con", "pkz",
                     "pk", "pe", "ps", "wsd")
            sat_cfg = cfg.get("sat_adjust") if cfg.get("do_sat_adj") else None
            kmp = nz - 1
            for k in range(nz):
                if pfull[k] > 10.0e2:
                    kmp = k
                    break
            for ti in range(ntiles):
                st = {k: S[k][ti] for k in names}
                trs = {k: S[k][ti] for k in orm.TRACERS}
                sa = None
                if sat_cfg is not None:
                    sa = dict(config=sat_cfg, area=metrics[ti]["area"], hs=S["phis"][ti],
                              dp1=S["dp1"][ti], mdt=cfg["bdt"] / k_split, kmp=kmp)
                o, q = orm.lagrangian_to_eulerian(st, trs, ak, bk, cfg["ptop"], KAPPA, ZVIR,
                                                  last_step, rcfg, NQ, t, sat_adjust=sa)
                for k in names:
                    S[k][ti][...] = o[k]
                for k in orm.TRACERS:
                    S[k][ti][...] = q[k]
            if last_step:
                C = (slice(None), slice(t.isc, t.iec + 1), slice(t.jsc, t.jec + 1),
                     slice(0, nz))
                S["omga"][C] = S["delp"][C] / S["delz"][C] * S["w"][C]
                if cfg.get("nf_omega", 1) > 0:
                    halo("scalar", ["omga"])
                    for ti in range(ntiles):
                        om = S["omga"][ti][..., :nz]
                        del2cubed(om, 0.18 * scalars["da_min"], metrics[ti], t, cfg["nf_omega"])
    if wrapup:
        C = (slice(t.isc, t.iec + 1), slice(t.jsc, t.jec + 1), slice(0, nz))
        names = ("qvapor", "qliquid", "qrain", "qsnow", "qice", "qgraupel", "qcld", "pt")
        for ti in range(ntiles):
            o = neg_adj3(*(S[k][ti][C] for k in names), S["delp"][ti][C])
            for k in names:
                S[k][ti][C] = o[k]
        order = cfg.get("c2l_ord", 4)
        if order == 4:
            halo("vector_d", ["u", "v"])
        c2l = c2l_ord4 if order == 4 else c2l_ord2
        g = 0 if order == 4 else 1
        box = (slice(t.isc - g, t.iec + 1 + g), slice(t.jsc - g, t.jec + 1 + g), slice(0, nz))
        for ti in range(ntiles):
            m = metrics[ti]
            ua, va = c2l(S["u"][ti][..., :nz], S["v"][ti][..., :nz], m["dx"], m["dy"], m["a11"],
                         m["a12"], m["a21"], m["a22"], t.isc, t.iec, t.jsc, t.jec)
            S["ua"][ti][box] = ua[box]
            S["va"][ti][box] = va[box]
