"""D-grid step (fxadv, xtp_u/ytp_v, a2b_ord4, divergence damping, d_sw):
oracle pins on the CPU and GPU parity."""
import numpy as np
import pytest
import torch

from fv3core_amd.workloads import plane_metric
from oracle import dsw as ods
from oracle.fvtp2d import Tile
from tests.dyn_helpers import Case
from tests.helpers import assert_close

N, NZ = 12, 4
DX = 1.0e5


def _plane(edges, f=0.0):
    return Tile(N, N, 3, edges, edges, edges, edges), plane_metric(N, f=f)


def _z():
    return np.zeros((N + 7, N + 7, NZ))


def _ij():
    ii = np.arange(N + 7)[:, None, None] * np.ones((1, N + 7, NZ))
    jj = np.arange(N + 7)[None, :, None] * np.ones((N + 7, 1, NZ))
    return ii, jj


# ------------------------------------------------------------ oracle pins (CPU)
@pytest.mark.parametrize("edges", [False, True])
def test_oracle_fxadv_uniform_flow(edges):
    """Constant C-grid winds on the plane: Courant numbers dt U / dx and
    area fluxes dt U dy, on every interface the stencil defines."""
    t, g = _plane(edges)
    U, V, dt = 7.0, -4.0, 30.0
    uc, vc = np.full(_z().shape, U), np.full(_z().shape, V)
    crx, cry, xfx, yfx, ucc, vcc = (_z() for _ in range(6))
    ods.fxadv(uc, vc, crx, cry, xfx, yfx, ucc, vcc, dt, g, t)
    c = (slice(t.isc, t.iec + 2), slice(t.jsc, t.jec + 1))
    np.testing.assert_allclose(crx[c], dt * U / DX, rtol=1e-14)
    np.testing.assert_allclose(xfx[c], dt * U * DX, rtol=1e-14)
    c = (slice(t.isc, t.iec + 1), slice(t.jsc, t.jec + 2))
    np.testing.assert_allclose(cry[c], dt * V / DX, rtol=1e-14)
    np.testing.assert_allclose(yfx[c], dt * V * DX, rtol=1e-14)


@pytest.mark.parametrize("iord", [5, 6, 8])
@pytest.mark.parametrize("c", [0.3, -0.45])
def test_oracle_xtp_u_linear_profile_is_exact(iord, c):
    """PPM advection of a linear u with a constant Courant number c returns
    the mean of u over the swept interval, u at the face shifted by c/2."""
    t, g = _plane(False)
    ii, _ = _ij()
    u = 2.0 + 0.01 * ii
    ub = np.full(u.shape, c * DX / 10.0)  # ub * dt / dx = c with dt = 10
    out = ods.advect_u_along_x(u, ub, g["rdx"], g["dx"], g["dxa"], 10.0, iord, t)
    want = 2.0 + 0.01 * (ii - 0.5 - c / 2.0)
    c_ = (slice(t.isc, t.iec + 2), slice(t.jsc, t.jec + 2))
    if iord == 8:
        # the monotone limiter flattens nothing in a linear profile away from
        # its own end points
        pass
    np.testing.assert_allclose(out[c_], want[c_], rtol=1e-13)


def test_oracle_ytp_v_is_transposed_xtp_u():
    t, g = _plane(True)
    rng = np.random.default_rng(0)
    v = rng.normal(size=_z().shape)
    vb = rng.normal(size=_z().shape) * 1000.0
    a = ods.advect_v_along_y(v, vb, g["rdy"], g["dy"], g["dya"], 20.0, 6, t)
    b = ods.advect_u_along_x(v.transpose(1, 0, 2), vb.transpose(1, 0, 2), g["rdx"].T, g["dx"].T,
                             g["dxa"].T, 20.0, 6, t.transposed()).transpose(1, 0, 2)
    np.testing.assert_array_equal(np.nan_to_num(a, nan=7.0), np.nan_to_num(b, nan=7.0))


def _latlon_plane(g, t):
    """Small-angle lon/lat of centres and corners (1 km cells near the equator)."""
    h = 1.0e3 / 6.3712e6
    ii = np.arange(N + 7)[:, None] * np.ones((1, N + 7))
    jj = np.arange(N + 7)[None, :] * np.ones((N + 7, 1))
    g = dict(g)
    g["agrid1"], g["agrid2"] = (ii + 0.5) * h, (jj + 0.5) * h
    g["bgrid1"], g["bgrid2"] = ii * h, jj * h
    for k in ("edge_w", "edge_e", "edge_s", "edge_n"):
        g[k] = np.full((N + 7, N + 7), 0.5)
    return g


@pytest.mark.parametrize("edges", [False, True])
def test_oracle_a2b_ord4_reproduces_linear_fields(edges):
    """A-grid to B-grid interpolation is exact for linear fields on a
    uniform grid: interior (4th order), tile-edge formulas and the
    great-circle corner extrapolation."""
    t, g = _plane(edges)
    g = _latlon_plane(g, t)
    ii, jj = _ij()
    q = 3.0 + 0.2 * ii - 0.1 * jj
    qout = _z()
    ods.a2b_ord4(q, qout, g, t)
    want = 3.0 + 0.2 * (ii - 0.5) - 0.1 * (jj - 0.5)
    c = (slice(t.isc, t.iec + 2), slice(t.jsc, t.jec + 2))
    np.testing.assert_allclose(qout[c], want[c], rtol=1e-6)
    inner = (slice(t.isc + 1, t.iec + 1), slice(t.jsc + 1, t.jec + 1))
    np.testing.assert_allclose(qout[inner], want[inner], rtol=1e-13)


def test_oracle_divergence_damping_laplacian_of_parabola():
    """With one high-order pass, divg_d = b i^2 becomes the D-grid Laplacian
    -2 b / dx^2 (divg_u = divg_v = 1 on the plane)."""
    t, g = _plane(False)
    ii, jj = _ij()
    beta = 3.0
    divg = beta * ii ** 2
    z = _z
    uc, vc, delpc, ke, vort, ptc = z(), z(), z(), z(), z(), z()
    nord_col = np.ones(NZ + 1)
    ods.divergence_damping(z(), z(), z(), ptc, vort, z(), divg, vc, uc, delpc, ke, z(), 10.0, g, t,
                           nord_col, np.full(NZ + 1, 0.0), 0.0, 0.15, DX * DX, 1)
    c = (slice(t.isc, t.iec + 2), slice(t.jsc, t.jec + 2))
    np.testing.assert_allclose(divg[c], -2.0 * beta / DX ** 2, rtol=1e-12)
    # vc holds the gradient b (2i + 1), uc is zero
    np.testing.assert_allclose(vc[c], beta * (2 * ii[c] + 1), rtol=1e-14)
    np.testing.assert_allclose(uc[c], 0.0, atol=0)
    # ke gets dd8 * divg with dd8 = (da_min_c d4_bg)^2 (damp * delpc is 0)
    np.testing.assert_allclose(ke[c], (DX * DX * 0.15) ** 2 * (-2.0 * beta / DX ** 2), rtol=1e-12)


def test_oracle_column_namelist_c12():
    cfg = dict(ke_bg=0.0, d_con=1.0, nord=3, d2_bg=0.0, vtdm4=0.06, do_vort_damp=True,
               n_sponge=48, d2_bg_k1=0.2, d2_bg_k2=0.1)
    col = ods.column_namelist(cfg, 79)
    np.testing.assert_array_equal(col["nord"][:5], [0, 0, 0, 3, 3])
    np.testing.assert_array_equal(col["nord_v"][:4], [0, 0, 2, 2])
    np.testing.assert_allclose(col["d2_divg"][:4], [0.2, 0.1, 0.02, 0.0])
    np.testing.assert_allclose(col["damp_vt"][:4], [0.1, 0.05, 0.06, 0.06])
    np.testing.assert_allclose(col["damp_w"][:4], [0.2, 0.1, 0.02, 0.06])
    np.testing.assert_array_equal(col["d_con"][:4], [0, 0, 0, 1])


CFG = dict(hord_dp=6, hord_tm=6, hord_mt=6, hord_vt=6, dddmp=0.5, d4_bg=0.15, nord=3,
           d_con=1.0, do_f3d=False, hydrostatic=False, do_skeb=False)
NAMELIST = dict(ke_bg=0.0, d_con=1.0, nord=3, d2_bg=0.0, vtdm4=0.06, do_vort_damp=True,
                n_sponge=48, d2_bg_k1=0.2, d2_bg_k2=0.1)


def dsw_state(case, seed_shift=0):
    f = case.field
    st = dict(delp=f(1000.0, 60.0), pt=f(300.0, 10.0), u=f(12.0, 15.0), v=f(-3.0, 15.0),
              w=f(0.0, 0.5), uc=f(10.0, 15.0), vc=f(-2.0, 15.0), ua=f(11.0, 15.0),
              va=f(-3.0, 15.0), divgd=f(0.0, 1e-5), mfx=f(0.0, 1e6), mfy=f(0.0, 1e6),
              cx=f(0.0, 0.1), cy=f(0.0, 0.1), crx=f(0.0, 0.1), cry=f(0.0, 0.1),
              xfx=f(0.0, 1e8), yfx=f(0.0, 1e8), q_con=f(1e-3, 5e-4), zh=f(5000.0, 100.0),
              heat_source=f(0.0, 1.0), diss_est=f(0.0, 1.0), delpc=f(1000.0, 60.0),
              ptc=f(300.0, 10.0))
    return st


def test_oracle_d_sw_conserves_mass_heat_and_condensate():
    """On a doubly periodic tile the flux-form updates conserve the domain
    sums of delp, pt*delp and q_con*delp (damping fluxes included)."""
    nz = 6
    case = Case(16, nz, cube_tiles=False, seed=4)
    g = case.metric[0]
    st = {k: v[0][..., :nz] for k, v in dsw_state(case).items()}
    col = ods.column_namelist(NAMELIST, nz)
    damp = dict(del6_u=g["del6_u"], del6_v=g["del6_v"], da_min=case.scalars["da_min"],
                da_min_c=case.scalars["da_min_c"])
    t = case.tile
    out = ods.d_sw(st, 30.0, g, t, CFG, col, damp)
    c = (slice(t.isc, t.iec + 1), slice(t.jsc, t.jec + 1))
    area = g["area"][c][:, :, None]
    for name, a0, a1 in (("delp", st["delp"], out["delp"]),
                         ("pt*delp", st["pt"] * st["delp"], out["pt"] * out["delp"]),
                         ("q_con*delp", st["q_con"] * st["delp"], out["q_con"] * out["delp"])):
        s0 = (a0[c] * area).sum(axis=(0, 1))
        s1 = (a1[c] * area).sum(axis=(0, 1))
        np.testing.assert_allclose(s1, s0, rtol=1e-12, err_msg=name)
    assert np.abs(out["delp"][c] - st["delp"][c]).max() > 1e-3
    assert np.isfinite(out["u"][c]).all() and np.isfinite(out["v"][c]).all()


# ------------------------------------------------------------- GPU parity
@pytest.mark.gpu
@pytest.mark.parametrize("cube_tiles", [False, True])
def test_gpu_a2b_ord4_matches_oracle(cube_tiles):
    from fv3core_amd.stencils.a2b_ord4 import AGrid2BGridFourthOrder

    case = Case(16, 5, cube_tiles, seed=31)
    q = case.field(1.0e-5, 3.0e-5)
    out0 = case.field(0.0, 1.0)
    a2b = AGrid2BGridFourthOrder(case.factory, case.grid_data(), 0)
    tq, to = case.dev(q), case.dev(out0)
    a2b(tq, to)
    torch.cuda.synchronize()
    t = case.tile
    nz = case.nz
    for ti in range(case.ntiles):
        ref = out0[ti][..., :nz].copy()
        ods.a2b_ord4(q[ti][..., :nz].copy(), ref, case.metric[ti], t)
        got = case.host(to)[ti][..., :nz]
        full = (slice(0, case.n + 6), slice(0, case.n + 6))
        assert_close(got[full], ref[full], 1e-12, near_zero=1e-300, what=f"tile {ti} qout")


def _dsw_config(nord=3, hord=6, dddmp=0.5):
    from fv3core_amd._config import DynamicalCoreConfig

    cfg = DynamicalCoreConfig(nord=nord, hord_dp=hord, hord_tm=hord, hord_mt=hord, hord_vt=hord,
                              dddmp=dddmp)
    return cfg.d_grid_shallow_water


@pytest.mark.gpu
@pytest.mark.parametrize("cube_tiles,nord,hord,dddmp", [(False, 3, 6, 0.5), (True, 3, 6, 0.5),
                                                        (True, 2, 8, 0.5), (True, 1, 5, 0.0)])
def test_gpu_d_sw_matches_oracle(cube_tiles, nord, hord, dddmp):
    from fv3core_amd.stencils.d_sw import DGridShallowWaterLagrangianDynamics, get_column_namelist
    from fv3core_amd.utils.grid import DampingCoefficients

    nz = 6
    case = Case(16, nz, cube_tiles, seed=41 + nord)
    st = dsw_state(case)
    grid = case.grid_data()
    damp = DampingCoefficients(grid.del6_u, grid.del6_v, case.scalars["da_min"],
                               case.scalars["da_min_c"])
    config = _dsw_config(nord, hord, dddmp)
    col = get_column_namelist(config, nz)
    dsw = DGridShallowWaterLagrangianDynamics(case.factory, grid, damp, col, False, False, config)
    dev = {k: case.dev(v) for k, v in st.items()}
    dt = 37.5
    dsw(dev["delpc"], dev["delp"], dev["ptc"], dev["pt"], dev["u"], dev["v"], dev["w"], dev["uc"],
        dev["vc"], dev["ua"], dev["va"], dev["divgd"], dev["mfx"], dev["mfy"], dev["cx"],
        dev["cy"], dev["crx"], dev["cry"], dev["xfx"], dev["yfx"], dev["q_con"], dev["zh"],
        dev["heat_source"], dev["diss_est"], dt)
    torch.cuda.synchronize()
    t = case.tile
    cfg = dict(hord_dp=hord, hord_tm=hord, hord_mt=hord, hord_vt=hord, dddmp=dddmp, d4_bg=0.15,
               nord=nord, d_con=1.0, do_f3d=False, hydrostatic=False, do_skeb=False)
    ocol = ods.column_namelist(dict(ke_bg=0.0, d_con=1.0, nord=nord, d2_bg=0.0, vtdm4=0.06,
                                    do_vort_damp=True, n_sponge=48, d2_bg_k1=0.2, d2_bg_k2=0.1),
                               nz)
    for k in ocol:
        np.testing.assert_array_equal(ocol[k], col[k], err_msg=k)
    full = (slice(0, case.n + 6), slice(0, case.n + 6))
    for ti in range(case.ntiles):
        g = case.metric[ti]
        damp_d = dict(del6_u=g["del6_u"], del6_v=g["del6_v"], da_min=case.scalars["da_min"],
                      da_min_c=case.scalars["da_min_c"])
        stt = {k: v[ti][..., :nz] for k, v in st.items()}
        ref = ods.d_sw(stt, dt, g, t, cfg, ocol, damp_d)
        for name in _STATE_OUT:
            got = case.host(dev[name])[ti][..., :nz]
            assert_close(got[full], ref[name][full], 1e-11, near_zero=1e-300,
                         what=f"tile {ti} {name}")


_STATE_OUT = ("delp", "pt", "u", "v", "w", "q_con", "uc", "vc", "divgd", "delpc", "ptc", "mfx",
              "mfy", "cx", "cy", "crx", "cry", "xfx", "yfx", "heat_source", "diss_est")
