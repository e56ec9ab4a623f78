"""Vertical remapping: LagrangianToEulerian, MapNTracer, fillz.

CPU tests pin the oracle (hand-worked fillz columns, mass conservation,
the dry closed form of moist_cv, and the identity of a remap whose
Lagrangian levels already are the hybrid levels); GPU tests compare the HIP
path with the oracle on deformed columns of the six cube tiles.
"""
import numpy as np
import pytest
import torch

from fv3core_amd._config import DynamicalCoreConfig
from fv3core_amd.constants import CP_AIR, GRAV, KAPPA, RDGAS, ZVIR
from fv3core_amd.workloads import synthetic_hybrid
from oracle import remapping as orm
from tests.dyn_helpers import Case
from tests.helpers import assert_close

NZ = 10
TRACERS = orm.TRACERS


def remap_inputs(case, deform=0.3, plant_negative=True):
    """Deformed Lagrangian columns over hybrid levels, with moist tracers.

    Interior interfaces move by up to ``deform`` / 2 of the thinner adjacent
    hybrid layer; the top stays at ptop and the bottom at ps.  pt is
    T / pm**kappa and delz the hydrostatic thickness, as the dynamics hand
    them over; qliquid changes sign when ``plant_negative`` is set so that
    fillz has work to do.
    """
    nz = case.nz
    ak, bk, ptop = synthetic_hybrid(nz)
    nt = case.ntiles
    ps = [1.0e5 + 800.0 * f[..., 0] for f in case.field(0.0, 1.0)]
    pert = case.field(0.0, 1.0)
    st = {k: [] for k in ("pt", "delp", "delz", "peln", "u", "v", "w", "ua", "cappa", "q_con",
                          "pkz", "pk", "pe", "ps", "wsd")}
    tr = {n: [] for n in TRACERS}
    u = case.field(0.0, 20.0)
    v = case.field(0.0, 20.0)
    w = case.field(0.0, 2.0)
    temp = case.field(250.0, 30.0)
    wsd = case.field(0.0, 0.5)
    qv = case.field(6.0e-3, 5.0e-3)
    qother = {n: case.field(2.0e-5, 1.5e-5) for n in TRACERS[1:]}
    if plant_negative:
        qother["qliquid"] = case.field(0.0, 2.0e-6)
    for t in range(nt):
        pe_e = ak[None, None, :] + bk[None, None, :] * ps[t][..., None]
        dpe = np.diff(pe_e, axis=2)
        shift = np.zeros_like(pe_e)
        shift[..., 1:nz] = 0.5 * deform * pert[t][..., 1:nz] * np.minimum(
            dpe[..., :nz - 1], dpe[..., 1:nz])
        pe = pe_e + shift
        delp = np.zeros_like(pe)
        delp[..., :nz] = np.diff(pe, axis=2)
        delp[..., nz] = delp[..., nz - 1]
        peln = np.log(pe)
        pm = np.ones_like(pe)
        pm[..., :nz] = delp[..., :nz] / (peln[..., 1:] - peln[..., :nz])
        pt = temp[t] * np.exp(-KAPPA * np.log(pm))
        delz = np.full_like(pe, -1.0)
        delz[..., :nz] = -(delp[..., :nz] / GRAV) * RDGAS * pt[..., :nz] * np.exp(
            (KAPPA - 1.0) * np.log(pm[..., :nz]))
        vals = dict(pt=pt, delp=delp, delz=delz, peln=peln, u=u[t], v=v[t], w=w[t],
                    ua=np.zeros_like(pe), cappa=np.full_like(pe, KAPPA),
                    q_con=np.zeros_like(pe), pkz=np.ones_like(pe),
                    pk=np.exp(KAPPA * peln), pe=pe, ps=ps[t].copy(), wsd=wsd[t][..., 0].copy())
        for k, a in vals.items():
            st[k].append(a)
        tr["qvapor"].append(qv[t])
        for n in TRACERS[1:]:
            tr[n].append(qother[n][t])
    return st, tr, ak, bk, ptop


def _cfg(kord=9, fill=True):
    return dict(kord_tm=-kord, kord_tr=kord, kord_wz=kord, kord_mt=kord, fill=fill)


def _oracle(case, st, tr, ak, bk, ptop, last_step, cfg, nq):
    outs, trs = [], []
    for t in range(case.ntiles):
        s_t = {k: v[t] for k, v in st.items()}
        q_t = {k: v[t] for k, v in tr.items()}
        o, q = orm.lagrangian_to_eulerian(s_t, q_t, ak, bk, ptop, KAPPA, ZVIR, last_step, cfg,
                                          nq, case.tile)
        outs.append(o)
        trs.append(q)
    return outs, trs


# ------------------------------------------------------------ oracle pins (CPU)
def test_oracle_fillz_hand_worked_columns():
    dp = np.ones(6)
    # borrow from above: level 1 takes 0.5 from level 0
    q = orm.fillz(np.array([1.0, -0.5, 1.0, 1.0, 1.0, 1.0]), dp)
    np.testing.assert_array_equal(q, [0.5, 0.0, 1.0, 1.0, 1.0, 1.0])
    # nothing above: borrow from below, which pays it back on its own level
    q = orm.fillz(np.array([0.0, -0.5, 1.0, 1.0, 1.0, 1.0]), dp)
    np.testing.assert_array_equal(q, [0.0, 0.0, 0.5, 1.0, 1.0, 1.0])
    # negative bottom layer borrows from the layer above
    q = orm.fillz(np.array([1.0, 1.0, 1.0, 1.0, 1.0, -0.5]), dp)
    np.testing.assert_array_equal(q, [1.0, 1.0, 1.0, 1.0, 0.5, 0.0])
    # neighbours cannot cover the deficit: the column below the top is
    # rescaled by sum(dm) / sum(max(dm, 0)) = 1.2 / 2 and clipped at zero
    q = orm.fillz(np.array([0.0, 0.1, -1.0, 0.1, 1.0, 1.0]), dp)
    np.testing.assert_allclose(q, [0.0, 0.0, 0.0, 0.0, 0.6, 0.6], rtol=1e-14, atol=1e-17)
    # negative top is clipped (mass is added there, as in the reference)
    q = orm.fillz(np.array([-1.0, 1.0, 1.0, 1.0, 1.0, 1.0]), dp)
    np.testing.assert_array_equal(q, [0.0, 1.0, 1.0, 1.0, 1.0, 1.0])
    # uneven thicknesses: level 2 takes min(q1 dp1, -q2 dp2) = 0.2 from level 1
    dp = np.array([1.0, 2.0, 4.0, 1.0, 1.0, 1.0])
    q = orm.fillz(np.array([1.0, 0.1, -0.05, 1.0, 1.0, 1.0]), dp)
    np.testing.assert_allclose(q, [1.0, 0.0, 0.0, 1.0, 1.0, 1.0], atol=1e-17)


def test_oracle_fillz_keeps_mass_and_positivity():
    rng = np.random.default_rng(3)
    q = rng.normal(0.5, 1.0, size=(400, 12))
    dp = rng.uniform(50.0, 500.0, size=(400, 12))
    out = orm.fillz(q, dp)
    q0 = q.copy()
    q0[:, 0] = np.maximum(q0[:, 0], 0.0)
    m0 = (q0 * dp).sum(axis=1)
    m1 = (out * dp).sum(axis=1)
    ok = (q0[:, 1:] * dp[:, 1:]).sum(axis=1) > 0
    np.testing.assert_allclose(m1[ok], m0[ok], rtol=1e-12)
    assert (out[ok] >= 0.0).all()
    pos = (q > 0).all(axis=1)
    np.testing.assert_array_equal(out[pos], q[pos])


def test_oracle_moist_cv_dry_air_is_kappa():
    z = np.zeros(4)
    cvm, gz = orm.moist_cv_nwat6(z, z, z, z, z, z)
    np.testing.assert_allclose(cvm, CP_AIR - RDGAS, rtol=1e-15)
    np.testing.assert_allclose(orm.set_cappa(z, cvm, ZVIR), KAPPA, rtol=1e-15)
    # water loading: vapour replaces dry air at cv_vap, condensate at c_liq / c_ice
    q = np.full(4, 0.01)
    cvm, gz = orm.moist_cv_nwat6(q, q, z, z, 2 * q, z)
    np.testing.assert_allclose(gz, 0.03)
    np.testing.assert_allclose(
        cvm, 0.96 * (CP_AIR - RDGAS) + 0.01 * 3.0 * 461.5 + 0.01 * 4185.5 + 0.02 * 1972.0)


@pytest.mark.parametrize("kord", [9, 10])
def test_oracle_remap_is_identity_on_hybrid_levels(kord):
    """With Lagrangian interfaces equal to ak + bk ps, every remap returns the
    layer means and the moist pt conversion and its inverse cancel."""
    case = Case(6, NZ, cube_tiles=False, seed=11)
    st, tr, ak, bk, ptop = remap_inputs(case, deform=0.0, plant_negative=False)
    outs, trs = _oracle(case, st, tr, ak, bk, ptop, False, _cfg(kord), 8)
    t = case.tile
    c = (slice(t.isc, t.iec + 1), slice(t.jsc, t.jec + 1), slice(0, NZ))
    o = outs[0]
    for k in ("pt", "delz", "delp", "w"):
        np.testing.assert_allclose(o[k][c], st[k][0][c], rtol=1e-11, atol=1e-12, err_msg=k)
    for n in TRACERS[:8]:
        np.testing.assert_allclose(trs[0][n][c], tr[n][0][c], rtol=1e-11, atol=1e-20, err_msg=n)
    cu = (slice(t.isc, t.iec + 1), slice(t.jsc, t.jec + 2), slice(0, NZ))
    cv = (slice(t.isc, t.iec + 2), slice(t.jsc, t.jec + 1), slice(0, NZ))
    np.testing.assert_allclose(o["u"][cu], st["u"][0][cu], rtol=1e-11, atol=1e-10)
    np.testing.assert_allclose(o["v"][cv], st["v"][0][cv], rtol=1e-11, atol=1e-10)
    ci = c[:2] + (slice(0, NZ + 1),)
    np.testing.assert_allclose(o["pe"][ci], st["pe"][0][ci], rtol=1e-14)
    np.testing.assert_allclose(o["peln"][ci], np.log(st["pe"][0][ci]), rtol=1e-14)
    np.testing.assert_allclose(o["pk"][c], np.exp(KAPPA * np.log(st["pe"][0][c])), rtol=1e-14)
    np.testing.assert_allclose(o["ps"][c[:2]], st["pe"][0][c[:2] + (NZ,)], rtol=0)


def test_oracle_remap_conserves_tracer_mass():
    case = Case(6, NZ, cube_tiles=False, seed=12)
    st, tr, ak, bk, ptop = remap_inputs(case, deform=0.4, plant_negative=False)
    outs, trs = _oracle(case, st, tr, ak, bk, ptop, False, _cfg(9), 8)
    t = case.tile
    c = (slice(t.isc, t.iec + 1), slice(t.jsc, t.jec + 1), slice(0, NZ))
    dp_old = st["delp"][0][c]
    dp_new = outs[0]["delp"][c]
    assert np.abs(dp_new - dp_old).max() > 1.0   # the levels did move
    for n in TRACERS[:8]:
        m0 = (tr[n][0][c] * dp_old).sum(axis=2)
        m1 = (trs[0][n][c] * dp_new).sum(axis=2)
        np.testing.assert_allclose(m1, m0, rtol=1e-12, err_msg=n)
    np.testing.assert_allclose(dp_new.sum(axis=2), dp_old.sum(axis=2), rtol=1e-13)


# ------------------------------------------------------------------ GPU parity
def _device_state(case, st, tr):
    from fv3core_amd.utils import gt4py_utils as utils

    dev = {}
    for k, v in st.items():
        if v[0].ndim == 2:
            if case.ntiles == 1:
                t = utils.empty(v[0].shape)
                t.copy_(torch.as_tensor(v[0]))
            else:
                t = utils.empty((6,) + v[0].shape, batch_dims=1)
                t.copy_(torch.as_tensor(np.stack(v)))
            dev[k] = t
        else:
            dev[k] = case.dev(v)
    trd = {k: case.dev(v) for k, v in tr.items()}
    return dev, trd


def _run_gpu(case, st, tr, ak, bk, ptop, last_step, kord, fill, nq):
    from fv3core_amd.stencils.remapping import LagrangianToEulerian

    cfg = DynamicalCoreConfig(npz=case.nz, kord_tm=-kord, kord_tr=kord, kord_wz=kord,
                              kord_mt=kord, fill=fill).remapping
    pfull = 0.5 * (ak[:-1] + ak[1:] + (bk[:-1] + bk[1:]) * 1.0e5)
    l2e = LagrangianToEulerian(case.factory, cfg, None, nq, pfull)
    dev, trd = _device_state(case, st, tr)
    d = dev
    l2e(trd, d["pt"], d["delp"], d["delz"], d["peln"], d["u"], d["v"], d["w"], d["ua"], None,
        d["cappa"], d["q_con"], None, d["pkz"], d["pk"], d["pe"], None, None, d["ps"], d["wsd"],
        None, torch.as_tensor(ak, device="cuda"), torch.as_tensor(bk, device="cuda"), pfull,
        None, ptop, KAPPA, ZVIR, last_step, 0.0, 225.0, 225.0, False, nq)
    torch.cuda.synchronize()
    return dev, trd


@pytest.mark.gpu
@pytest.mark.parametrize("kord,last_step,fill", [(9, False, True), (10, True, True),
                                                 (9, True, False)])
def test_gpu_lagrangian_to_eulerian_matches_oracle(kord, last_step, fill):
    case = Case(12, NZ, cube_tiles=True, seed=20 + kord)
    st, tr, ak, bk, ptop = remap_inputs(case, deform=0.35)
    nq = 8
    outs, trs = _oracle(case, st, tr, ak, bk, ptop, last_step, _cfg(kord, fill), nq)
    dev, trd = _run_gpu(case, st, tr, ak, bk, ptop, last_step, kord, fill, nq)
    for k in st:
        got = dev[k].cpu().numpy()
        want = np.stack([o[k] for o in outs])
        assert_close(got, want, 1e-10, what=k, scale_floor=1e-9)
    for n in TRACERS:
        got = trd[n].cpu().numpy()
        want = np.stack([q[n] for q in trs])
        assert_close(got, want, 1e-10, what=n, scale_floor=1e-9)


@pytest.mark.gpu
def test_gpu_fillz_and_mapn_tracer_match_oracle():
    from fv3core_amd.stencils.mapn_tracer import MapNTracer

    case = Case(12, NZ, cube_tiles=True, seed=31)
    st, tr, ak, bk, ptop = remap_inputs(case, deform=0.4)
    t = case.tile
    nq = 8
    # pe2 = hybrid levels over the Lagrangian surface pressure
    pe1 = st["pe"]
    pe2 = [ak[None, None, :] + bk[None, None, :] * p[..., NZ:NZ + 1] for p in pe1]
    dp2 = [np.concatenate([np.diff(p, axis=2), np.ones(p.shape[:2] + (1,))], axis=2)
           for p in pe2]
    want = {n: [] for n in TRACERS[:nq]}
    c = (slice(t.isc, t.iec + 1), slice(t.jsc, t.jec + 1))
    for ti in range(6):
        for idx, n in enumerate(TRACERS[:nq]):
            q = tr[n][ti].copy()
            kord = 9 if idx == 5 else 10
            q[c] = orm._map(q[c], pe1[ti][c], pe2[ti][c], kord, 0)
            q[c + (slice(0, NZ),)] = orm.fillz(q[c + (slice(0, NZ),)], dp2[ti][c + (slice(0, NZ),)])
            want[n].append(q)
    mapn = MapNTracer(case.factory, 10, nq, t.isc, t.iec, t.jsc, t.jec, fill=True)
    trd = {k: case.dev(v) for k, v in tr.items()}
    mapn(case.dev(pe1), case.dev(pe2), case.dev(dp2), trd, 0.0)
    torch.cuda.synchronize()
    for n in TRACERS[:nq]:
        assert_close(trd[n].cpu().numpy(), np.stack(want[n]), 1e-11, what=n, scale_floor=1e-9)
    np.testing.assert_array_equal(trd["qcld"].cpu().numpy(), np.stack(tr["qcld"]))
