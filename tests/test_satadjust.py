"""SatAdjust3d (fv3core/stencils/saturation_adjustment.py): oracle pins on
the CPU, the HIP kernel against the oracle on the GPU, and the saturation
adjustment inside LagrangianToEulerian."""
import numpy as np
import pytest
import torch

from fv3core_amd._config import SatAdjustConfig
from oracle import satadjust as osa
from tests.helpers import assert_close

WATER = ("qv", "ql", "qi", "qr", "qs", "qg")


def _cells(rng, shape, t_lo=185.0, t_hi=310.0, sat=(0.7, 1.3), neg=0.0, den_range=(0.3, 1.3)):
    """Random moist cells: temperature in [t_lo, t_hi], air density in
    den_range (kg m-3), vapour a random fraction of saturation over water
    (at most 0.03), condensates up to 1e-3 (a share ``neg`` of them slightly
    negative)."""
    t = rng.uniform(t_lo, t_hi, shape)
    dp = rng.uniform(300.0, 1500.0, shape)
    den = rng.uniform(*den_range, shape)  # air density, kg m-3
    delz = -dp / (osa.GRAV * den)
    qsat, _ = osa.wqs2(t, den, True)
    f = {"qv": np.minimum(qsat * rng.uniform(*sat, shape), 0.03)}
    for k in ("ql", "qi", "qr", "qs", "qg"):
        q = rng.uniform(0.0, 1e-3, shape) * (rng.uniform(size=shape) < 0.6)
        if neg:
            q = np.where(rng.uniform(size=shape) < neg, -rng.uniform(0.0, 2e-6, shape), q)
        f[k] = q
    qpz = sum(f[k] for k in WATER)
    q_liq = f["ql"] + f["qr"]
    q_sol = f["qi"] + f["qs"] + f["qg"]
    f["pt"] = t * (1.0 + osa.ZVIR * f["qv"]) * (1.0 - (q_liq + q_sol))
    f.update(dp=dp, delz=delz, te0=np.zeros(shape), q_con=q_liq + q_sol,
             cappa=np.full(shape, 0.28), pkz=np.ones(shape), qa=np.zeros(shape),
             peln=np.zeros(shape[:-1] + (shape[-1] + 1,)))
    del qpz
    return f


def _run(f, cfg, mdt=225.0, last_step=True, consv_te=False, area=1e10, hs=0.0):
    sc = osa.scalars(cfg, mdt)
    return osa.satadjust(f, np.asarray(area), np.asarray(hs), sdt=sc["sdt"], zvir=osa.ZVIR,
                         fac_i2s=sc["fac_i2s"], do_qa=True, consv_te=consv_te,
                         c_air=sc["c_air"], c_vap=sc["c_vap"], mdt=mdt, fac_r2g=sc["fac_r2g"],
                         fac_smlt=sc["fac_smlt"], fac_l2r=sc["fac_l2r"],
                         fac_imlt=sc["fac_imlt"], d0_vap=sc["d0_vap"], lv00=sc["lv00"],
                         fac_v2l=sc["fac_v2l"], fac_l2v=sc["fac_l2v"], last_step=last_step,
                         ext=osa.externals(cfg))


# ------------------------------------------------------------ oracle pins
def test_oracle_tables_match_textbook_vapour_pressures():
    """The closed forms behind the tables against tabulated saturation
    vapour pressures (these pin E00, the latent-heat and heat-capacity
    constants): over water 611.21 Pa at 0 C exactly (the table's anchor),
    2339 Pa at 20 C, 7385 Pa at 40 C, 101.4 kPa at 100 C, 286.5 Pa
    (supercooled) at -10 C; over ice 259.9 Pa at -10 C, 103.2 Pa at -20 C
    and 12.84 Pa at -40 C.  The closed form integrates Clausius-Clapeyron
    with constant heat capacities, exact at 0 C and drifting away from the
    tables by about 0.1 % per 10 K (2 % at 100 C); a wrong constant moves
    the values by tens of percent.  Table index i is the temperature
    TICE - 160 + 0.1 i."""
    idx = lambda c: round((c + 160.0) / 0.1)  # noqa: E731
    assert osa.qs_tablew_fn(1600) == osa.E00
    for celsius, pa, tol in ((20.0, 2339.0, 2e-3), (40.0, 7385.0, 5e-3),
                             (100.0, 101418.0, 3e-2), (-10.0, 286.5, 2e-3)):
        got = float(osa.qs_tablew_fn(idx(celsius)))
        assert abs(got / pa - 1.0) < tol, (celsius, got)
    for celsius, pa, tol in ((-10.0, 259.9, 2e-3), (-20.0, 103.2, 4e-3), (-40.0, 12.84, 5e-3)):
        got = float(osa.qs_table2_fn(idx(celsius)))
        assert abs(got / pa - 1.0) < tol, (celsius, got)
    # the ice curve lies below the supercooled-water curve, meeting it at 0 C
    i = np.arange(1000, 1600)
    assert (osa.qs_table2_fn(i) < osa.qs_tablew_fn(i)).all()
    assert abs(osa.table_ice_oneline(osa.TICE) / osa.E00 - 1.0) < 1e-15


def test_oracle_tables_monotone_and_smooth_at_the_phase_switch():
    """table2 switches from ice to water at index 1600; the two entries next
    to the switch are smoothed, so the table stays increasing, with no step
    larger than its neighbours' there, and the differences are never
    negative."""
    i = np.arange(1590, 1611)
    t2 = osa.qs_table2_fn(i)
    d = np.diff(t2)
    assert (d > 0).all()
    assert d.max() < 1.5 * d.min()
    assert (osa.des2_table(np.arange(-1, 2621)) >= 0).all()
    assert (osa.desw_table(np.arange(-1, 2621)) >= 0).all()
    # the last entry repeats the one before (des_end)
    assert osa.des2_table(2620) == max(0.0, osa.qs_table2_fn(2620) - osa.qs_table2_fn(2619))


@pytest.mark.parametrize("last_step", [True, False])
def test_oracle_total_water_is_conserved(last_step):
    """The adjustment moves water between the six species only."""
    rng = np.random.default_rng(1)
    f = _cells(rng, (40, 30), neg=0.1)
    o = _run(f, SatAdjustConfig(), last_step=last_step)
    before = sum(f[k] for k in WATER)
    after = sum(o[k] for k in WATER)
    np.testing.assert_allclose(after, before, rtol=1e-13, atol=1e-17)
    assert (o["qi"] >= -1e-20).all() and (o["qs"] >= -1e-20).all()


def test_oracle_last_step_brings_warm_vapour_to_saturation():
    """Warm cells without ice: the last-step adjustment is one Newton step of
    qv = qsat(T) (the latent heating counted), so the remaining departure
    from saturation is second order in the temperature change, far below
    the initial 10-30 % supersaturation."""
    rng = np.random.default_rng(2)
    shape = (500,)
    f = _cells(rng, shape, t_lo=280.0, t_hi=295.0, sat=(1.1, 1.3), den_range=(0.9, 1.3))
    assert (f["qv"] < 0.03).all()  # none capped
    for k in ("qi", "qs", "qg"):
        f[k] = np.zeros(shape)
    o = _run(f, SatAdjustConfig(), last_step=True)
    q_liq = o["ql"] + o["qr"]
    t = o["pt"] / ((1.0 + osa.ZVIR * o["qv"]) * (1.0 - q_liq))
    den = -f["dp"] / (osa.GRAV * f["delz"])
    qsat, _ = osa.wqs2(t, den, True)
    dep = np.abs(o["qv"] / qsat - 1.0)
    assert dep.max() < 5e-3, dep.max()
    assert (t > f["pt"] / ((1.0 + osa.ZVIR * f["qv"]) * (1.0 - f["ql"] - f["qr"]))).all()


def test_oracle_hand_worked_cells():
    """Single cells where one branch acts alone: a cell at TICE + 0.05 K (no
    melting, no sublimation) and sub-saturated with no cloud water (no
    condensation, nothing to evaporate) only runs the negative-species fixes:
    negative ice taken from snow, negative snow from graupel, negative cloud
    water from rain; below -48 C all cloud water freezes."""
    cfg = SatAdjustConfig()

    def cell(t, **q):
        f = {k: np.array([q.get(k, 0.0)]) for k in WATER}
        f["qv"] = np.array([q.get("qv", 1e-3)])
        ql_, qs_ = f["ql"] + f["qr"], f["qi"] + f["qs"] + f["qg"]
        f["pt"] = np.array([t]) * (1.0 + osa.ZVIR * f["qv"]) * (1.0 - (ql_ + qs_))
        f.update(dp=np.array([1000.0]), delz=np.array([-300.0]), te0=np.zeros(1),
                 q_con=ql_ + qs_, cappa=np.full(1, 0.28), pkz=np.ones(1), qa=np.zeros(1),
                 peln=np.zeros(2))
        return f

    t0 = osa.TICE + 0.05
    o = _run(cell(t0, qi=-1e-6, qs=5e-6), cfg, last_step=False)
    assert o["qi"][0] == 0.0 and abs(o["qs"][0] - 4e-6) < 1e-21
    o = _run(cell(t0, qs=-1e-6, qg=3e-6), cfg, last_step=False)
    assert o["qs"][0] == 0.0 and abs(o["qg"][0] - 2e-6) < 1e-21
    o = _run(cell(t0, ql=-2e-6, qr=5e-6), cfg, last_step=False)
    assert o["ql"][0] == 0.0 and abs(o["qr"][0] - 3e-6) < 1e-21
    o = _run(cell(200.0, ql=1e-5, qv=1e-7), cfg, last_step=False)
    assert o["ql"][0] == 0.0
    t_after = o["pt"][0] / ((1.0 + osa.ZVIR * o["qv"][0]) * (1.0 - (o["qi"][0] + o["qr"][0])))
    assert t_after > 200.0  # latent heat of fusion


# ------------------------------------------------------------------ GPU
def _device(a):
    from fv3core_amd.utils import gt4py_utils as utils
    t = utils.empty(a.shape, batch_dims=1 if a.ndim == 3 else None)
    t.copy_(torch.as_tensor(np.array(a)))
    return t


@pytest.mark.gpu
@pytest.mark.parametrize("last_step,consv_te,hydrostatic,icloud_f",
                         [(True, False, False, 0), (False, False, False, 0),
                          (True, True, False, 1), (True, False, True, 2)])
def test_gpu_sat_adjust_matches_oracle(last_step, consv_te, hydrostatic, icloud_f):
    """fv3_sat_adjust on six 12 x 12 tiles x 20 levels from kmp = 3, against
    the oracle's SatAdjust3d, every output field at 1e-11 (near-zero floor
    1e-9 of the field's scale): cells from 185 K to 310 K, sub- and
    super-saturated, with negative species."""
    from fv3core_amd.stencils.saturation_adjustment import SatAdjust3d
    from fv3core_amd.utils.stencil import GridIndexing, StencilConfig, StencilFactory

    n, nz, kmp = 12, 20, 3
    rng = np.random.default_rng(10 + icloud_f)
    shape = (6, n + 7, n + 7, nz + 1)
    f = _cells(rng, shape, neg=0.1)
    pe = np.cumsum(np.concatenate([np.full(shape[:-1] + (1,), 300.0), f["dp"][..., :nz]],
                                  axis=-1), axis=-1)
    peln = np.zeros(shape)
    peln[..., :nz + 1] = np.log(pe)
    area = rng.uniform(5e9, 2e10, shape[:3])
    hs = rng.uniform(0.0, 3000.0, shape[:3]) * osa.GRAV
    cfg = SatAdjustConfig(hydrostatic=hydrostatic, icloud_f=icloud_f)
    gi = GridIndexing((n, n, nz), 3, True, True, True, True)
    comp = SatAdjust3d(StencilFactory(StencilConfig(), gi), cfg, _device(area), kmp)
    names = ("te0", "qv", "ql", "qi", "qr", "qs", "qg", "qa")
    dev = {k: _device(f[k]) for k in names + ("dp", "delz", "q_con", "pt", "pkz", "cappa")}
    dpeln = _device(peln)
    comp(dev["te0"], dev["qv"], dev["ql"], dev["qi"], dev["qr"], dev["qs"], dev["qg"], dev["qa"],
         _device(hs), dpeln, dev["dp"], dev["delz"], dev["q_con"], dev["pt"], dev["pkz"],
         dev["cappa"], osa.ZVIR, 225.0, consv_te, last_step, 0.286, kmp)
    torch.cuda.synchronize()
    want = {}
    for t in range(6):
        o = osa.sat_adjust_3d(cfg, area[t], f["te0"][t], f["qv"][t], f["ql"][t], f["qi"][t],
                              f["qr"][t], f["qs"][t], f["qg"][t], f["qa"][t], hs[t], peln[t],
                              f["dp"][t], f["delz"][t], f["q_con"][t], f["pt"][t], f["pkz"][t],
                              f["cappa"][t], osa.ZVIR, 225.0, consv_te, last_step, kmp,
                              (3, n + 2, 3, n + 2), nz)
        for k, v in o.items():
            want.setdefault(k, []).append(v)
    c = (slice(None), slice(3, n + 3), slice(3, n + 3), slice(0, nz))
    key = dict(te0="te0", qv="qv", ql="ql", qi="qi", qr="qr", qs="qs", qg="qg", qa="qa",
               dp="dp", delz="delz", q_con="q_con", pt="pt", pkz="pkz", cappa="cappa")
    for k, dk in key.items():
        got = dev[dk].cpu().numpy()[c]
        ref = np.stack(want[k])[c]
        assert_close(got, ref, 1e-11, near_zero=1e-30, what=k, scale_floor=1e-9)
    # levels above kmp are untouched
    top = (slice(None), slice(3, n + 3), slice(3, n + 3), slice(0, kmp))
    np.testing.assert_array_equal(dev["qv"].cpu().numpy()[top], f["qv"][top])
