"""Tangent-linear model and adjoint of fv_tp_2d.

CPU: the torch restatement used as the autodiff reference (tests/torch_fvtp2d.py)
is pinned against the numpy oracle on the periodic workload.

GPU (the HIP kernels, through the C-ABI):
* homogeneity: PPM transport is positively homogeneous of degree one in q
  (every limiter is a min/max/abs/sign of differences of q, every
  smoothness test is scale-free), so at any state the TLM applied to q
  itself returns the forward fluxes: TLM(q; q) == fv_tp_2d(q).  This pins
  the TLM (every partial, every edge branch) against the forward kernel;
* dot-product test <TLM dq, g> == <dq, ADJ g> (transposition, 1e-12);
* central finite differences of the forward kernel along random
  directions against the TLM (1e-6, kink crossings are rare at eps = 1e-7);
* torch autodiff of the torch restatement against the adjoint (1e-10);
* an 8-step 4D-Var cost J = 0.5 ||q_T - q_obs||^2 whose gradient by the
  reverse sweep (each step re-linearised from its saved input) matches
  central finite differences of J.
"""
import numpy as np
import pytest
import torch

from fv3core_amd.utils.stencil import GridIndexing
from fv3core_amd.workloads import transport_inputs
from oracle import fvtp2d as ofv
from tests import torch_fvtp2d as tf

N_HALO = 3


def _periodic(n, nz, workload=1):
    d = transport_inputs(n, nz, workload=workload)
    return d


@pytest.mark.parametrize("hord", [5, 6, 8])
def test_torch_restatement_matches_oracle(hord):
    n, nz = 16, 3
    d = _periodic(n, nz)
    tile = ofv.Tile(n, n, 3, False, False, False, False)
    rfx, rfy = ofv.fvtp2d(d["q"], d["q"], d["crx"], d["cry"], d["x_area_flux"],
                          d["y_area_flux"], d["area"], d["dxa"], d["dya"], hord, tile)
    q = tf.interior_kji(d["q"], n)
    fx, fy = tf.fvtp2d(q, tf.interior_kji(d["crx"], n), tf.interior_kji(d["cry"], n),
                       tf.interior_kji(d["x_area_flux"], n), tf.interior_kji(d["y_area_flux"], n),
                       tf.interior_kji(d["area"], n), hord)
    ofx = tf.interior_kji(rfx, n).numpy()
    ofy = tf.interior_kji(rfy, n).numpy()
    np.testing.assert_allclose(fx.numpy(), ofx, rtol=1e-13, atol=1e-13 * np.abs(ofx).max())
    np.testing.assert_allclose(fy.numpy(), ofy, rtol=1e-13, atol=1e-13 * np.abs(ofy).max())


def test_torch_restatement_is_homogeneous():
    """F(2q) == 2 F(q) exactly for the restatement (the property the GPU test uses)."""
    n, nz = 16, 3
    d = _periodic(n, nz)
    args = [tf.interior_kji(d[k], n) for k in ("crx", "cry", "x_area_flux", "y_area_flux",
                                               "area")]
    q = tf.interior_kji(d["q"], n)
    for hord in (5, 6, 8):
        fx1, fy1 = tf.fvtp2d(q, *args, hord)
        fx2, fy2 = tf.fvtp2d(2.0 * q, *args, hord)
        assert torch.equal(fx2, 2.0 * fx1) and torch.equal(fy2, 2.0 * fy1)


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------


def _dev(a):
    from fv3core_amd.utils.grid import to_storage
    return to_storage(a)


def _lin(n, nz, hord, edges, workload=1):
    from fv3core_amd.adjoint import TransportLinearization
    d = _periodic(n, nz, workload)
    gi = GridIndexing((n, n, nz), N_HALO, *edges)
    dev = {k: _dev(v) for k, v in d.items()}
    nk = d["q"].shape[2]
    lin = TransportLinearization(gi, dev["area"], dev["dxa"], dev["dya"], hord, nk)
    return d, dev, lin


def _flux_masks(n, like):
    h = N_HALO
    mx = torch.zeros_like(like)
    my = torch.zeros_like(like)
    mx[h:h + n + 1, h:h + n, :] = 1.0
    my[h:h + n, h:h + n + 1, :] = 1.0
    return mx, my


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def _ijk(t):
    from fv3core_amd.adjoint import as_ijk_storage
    return as_ijk_storage(t)


@pytest.mark.gpu
@pytest.mark.parametrize("hord", [5, 6, 7, 8])
@pytest.mark.parametrize("edges", [(False,) * 4, (True,) * 4])
def test_gpu_tlm_homogeneity(hord, edges):
    n, nz = 24, 4
    d, dev, lin = _lin(n, nz, hord, edges)
    q = _ijk(dev["q"])
    qy = _ijk(dev["q"] * 1.0)
    aux = (dev["crx"], dev["cry"], dev["x_area_flux"], dev["y_area_flux"])
    fx, fy = lin.forward(q, qy, *aux)
    tfx, tfy = lin.tlm(q, qy, *aux, q, qy)
    torch.cuda.synchronize()
    mx, my = _flux_masks(n, fx)
    for a, b, m, nm in ((fx, tfx, mx, "fx"), (fy, tfy, my, "fy")):
        a, b = (a * m).cpu().numpy(), (b * m).cpu().numpy()
        scale = np.abs(a).max()
        err = np.abs(a - b).max() / scale
        print(f"hord={hord} edges={edges[0]} {nm}: max |TLM(q;q) - F(q)| / max|F| = {err:.2e}")
        assert err < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("hord", [5, 6, 7, 8])
@pytest.mark.parametrize("edges", [(False,) * 4, (True,) * 4])
@pytest.mark.parametrize("mass_flux", [False, True])
def test_gpu_dot_product(hord, edges, mass_flux):
    n, nz = 24, 4
    d, dev, lin = _lin(n, nz, hord, edges)
    g = torch.Generator(device="cpu").manual_seed(100 + hord)
    q = _ijk(dev["q"])
    qy = _ijk(dev["q"] + 0.01 * torch.rand(q.shape, generator=g, dtype=q.dtype).cuda())
    aux = (dev["crx"], dev["cry"], dev["x_area_flux"], dev["y_area_flux"])
    mfx = _ijk(dev["x_area_flux"] * 1.3) if mass_flux else None
    mfy = _ijk(dev["y_area_flux"] * 0.7) if mass_flux else None
    dqx = _ijk(torch.randn(q.shape, generator=g, dtype=q.dtype).cuda())
    dqy = _ijk(torch.randn(q.shape, generator=g, dtype=q.dtype).cuda())
    mx, my = _flux_masks(n, q)
    gfx = _ijk(torch.randn(q.shape, generator=g, dtype=q.dtype).cuda() * mx)
    gfy = _ijk(torch.randn(q.shape, generator=g, dtype=q.dtype).cuda() * my)
    tfx, tfy = lin.tlm(q, qy, *aux, dqx, dqy, mfx, mfy)
    gqx, gqy = lin.adjoint(q, qy, *aux, gfx, gfy, mfx, mfy)
    lhs = float((tfx * gfx).sum() + (tfy * gfy).sum())
    rhs = float((dqx * gqx).sum() + (dqy * gqy).sum())
    print(f"hord={hord} edges={edges[0]} mass={mass_flux}: <L dq, g> = {lhs:.15e}, "
          f"<dq, L^T g> = {rhs:.15e}, rel {_rel(lhs, rhs):.1e}")
    assert _rel(lhs, rhs) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("hord", [6, 8])
@pytest.mark.parametrize("edges", [(False,) * 4, (True,) * 4])
def test_gpu_tlm_finite_differences(hord, edges):
    n, nz = 24, 4
    d, dev, lin = _lin(n, nz, hord, edges)
    g = torch.Generator(device="cpu").manual_seed(7)
    q = _ijk(dev["q"])
    aux = (dev["crx"], dev["cry"], dev["x_area_flux"], dev["y_area_flux"])
    dq = _ijk(torch.randn(q.shape, generator=g, dtype=q.dtype).cuda())
    mx, my = _flux_masks(n, q)
    gfx = torch.randn(q.shape, generator=g, dtype=q.dtype).cuda() * mx
    gfy = torch.randn(q.shape, generator=g, dtype=q.dtype).cuda() * my
    eps = 1e-7
    qp, qm = _ijk(q + eps * dq), _ijk(q - eps * dq)
    fxp, fyp = lin.forward(qp, qp, *aux)
    fxm, fym = lin.forward(qm, qm, *aux)
    fd = float(((fxp - fxm) * gfx).sum() + ((fyp - fym) * gfy).sum()) / (2 * eps)
    tfx, tfy = lin.tlm(q, q, *aux, dq, dq)
    tl = float((tfx * gfx).sum() + (tfy * gfy).sum())
    print(f"hord={hord} edges={edges[0]}: FD {fd:.12e} TLM {tl:.12e} rel {_rel(fd, tl):.1e}")
    assert _rel(fd, tl) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("hord", [5, 6, 8])
def test_gpu_adjoint_matches_torch_autodiff(hord):
    """Gradient of <g, F(q)> w.r.t. the periodic interior q: HIP adjoint through
    FvTp2dFn (with the periodic halo fill) vs torch autodiff of the restatement."""
    from fv3core_amd.adjoint import PeriodicTransport
    n, nz = 24, 4
    d = _periodic(n, nz)
    gi = GridIndexing((n, n, nz), N_HALO, False, False, False, False)
    dev = {k: _dev(v) for k, v in d.items()}
    model = PeriodicTransport(gi, dev, hord)
    g = torch.Generator(device="cpu").manual_seed(11)
    nk = d["q"].shape[2]
    gx = torch.randn((nk, n, n), generator=g, dtype=torch.float64)
    gy = torch.randn((nk, n, n), generator=g, dtype=torch.float64)
    # HIP path
    q_kji = torch.as_tensor(d["q"]).permute(2, 1, 0).contiguous().cuda().requires_grad_(True)
    from fv3core_amd.adjoint import FvTp2dFn, ijk_view
    qf = model.fill(q_kji)
    fx, fy = FvTp2dFn.apply(ijk_view(qf), ijk_view(qf), model.lin, model.aux)
    h = N_HALO
    fxi = fx.permute(2, 1, 0)[:, h:h + n, h:h + n]
    fyi = fy.permute(2, 1, 0)[:, h:h + n, h:h + n]
    obj = (fxi * gx.cuda()).sum() + (fyi * gy.cuda()).sum()
    (grad_hip,) = torch.autograd.grad(obj, q_kji)
    grad_hip = grad_hip[:, h:h + n, h:h + n].cpu()
    # torch reference
    qt = tf.interior_kji(d["q"], n).clone().requires_grad_(True)
    args = [tf.interior_kji(d[k], n) for k in ("crx", "cry", "x_area_flux", "y_area_flux",
                                               "area")]
    tfx, tfy = tf.fvtp2d(qt, *args, hord)
    (grad_ref,) = torch.autograd.grad((tfx * gx).sum() + (tfy * gy).sum(), qt)
    err = (grad_hip - grad_ref).abs().max() / grad_ref.abs().max()
    print(f"hord={hord}: max |adjoint - autodiff| / max |autodiff| = {float(err):.2e}")
    assert float(err) < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("hord", [6, 8])
def test_gpu_multistep_cost_gradient(hord):
    """8-step J = 0.5 ||q_T - q_obs||^2: reverse-sweep gradient vs central FD of J."""
    from fv3core_amd.adjoint import PeriodicTransport, cost_and_gradient
    n, nz, nsteps = 24, 3, 8
    d = _periodic(n, nz)
    gi = GridIndexing((n, n, nz), N_HALO, False, False, False, False)
    dev = {k: _dev(v) for k, v in d.items()}
    model = PeriodicTransport(gi, dev, hord)
    q0 = torch.as_tensor(d["q"]).permute(2, 1, 0).contiguous().cuda()
    g = torch.Generator(device="cpu").manual_seed(3)
    pert = 1e-2 * torch.randn(q0.shape, generator=g, dtype=q0.dtype).cuda()
    with torch.no_grad():
        q_obs = model.run(q0 + pert, nsteps)
    cost, grad = cost_and_gradient(model, q0, q_obs, nsteps)
    for seed in (21, 22):
        gd = torch.Generator(device="cpu").manual_seed(seed)
        dq = torch.randn(q0.shape, generator=gd, dtype=q0.dtype).cuda()
        h = N_HALO
        dq_int = torch.zeros_like(dq)
        dq_int[:, h:h + n, h:h + n] = dq[:, h:h + n, h:h + n]
        eps = 1e-6
        with torch.no_grad():
            jp = 0.5 * float(((model.interior(model.run(q0 + eps * dq_int, nsteps)) -
                               model.interior(q_obs)) ** 2).sum())
            jm = 0.5 * float(((model.interior(model.run(q0 - eps * dq_int, nsteps)) -
                               model.interior(q_obs)) ** 2).sum())
        fd = (jp - jm) / (2 * eps)
        ad = float((grad * dq_int).sum())
        print(f"hord={hord} dir {seed}: J={cost:.6e} FD {fd:.10e} adjoint {ad:.10e} "
              f"rel {_rel(fd, ad):.1e}")
        assert _rel(fd, ad) < 1e-6


# ---------------------------------------------------------------------------
# vertical remap (map_scalar, kord 9)
# ---------------------------------------------------------------------------


def _remap_setup(n, nz, mode, workload=4):
    from fv3core_amd.adjoint import RemapLinearization
    from fv3core_amd.workloads import remap_inputs
    d = remap_inputs(n, nz, workload=workload)
    gi = GridIndexing((n, n, nz), N_HALO, True, True, True, True)
    dev = {k: _dev(d[k]) for k in ("pe1", "pe2", "pt", "q", "w", "qs")}
    return d, dev, RemapLinearization(gi, mode=mode, qmin=0.0)


def _remap_field(dev, mode):
    # signed field for the iv modes that allow negative values, a positive one
    # otherwise; both noisy, so no two neighbouring means tie (a tie makes the
    # kord-9 extremum flag gc(k) gc(k+1) < 0 flip under any perturbation,
    # which finite differences would see as a jump)
    return dev["w"] if mode in (-2, 1, 2) else dev["pt"]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [-2, -1, 0, 1, 2])
def test_gpu_remap_tlm_homogeneity(mode):
    """kord 9 with qmin = 0 is positively homogeneous in q: TLM(q; q) = remap(q)."""
    n, nz = 16, 79
    d, dev, lin = _remap_setup(n, nz, mode)
    q = _ijk(_remap_field(dev, mode))
    qs = dev["qs"] * 0.0 if mode == -2 else None  # qs enters affinely; zero keeps homogeneity
    out = lin.forward(q, dev["pe1"], dev["pe2"], qs)
    out2, dout = lin.tlm(q, q, dev["pe1"], dev["pe2"], qs)
    torch.cuda.synchronize()
    h = N_HALO
    a = out[h:h + n, h:h + n, :nz].cpu().numpy()
    b = dout[h:h + n, h:h + n, :nz].cpu().numpy()
    assert np.array_equal(out.cpu().numpy(), out2.cpu().numpy())
    err = np.abs(a - b).max() / np.abs(a).max()
    print(f"mode={mode}: max |TLM(q;q) - remap(q)| / max|remap| = {err:.2e}")
    assert err < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [-2, -1, 0, 1, 2])
def test_gpu_remap_dot_product(mode):
    """The hand-written reverse sweep against the Dual<1> TLM of the forward kernel."""
    n, nz = 16, 79
    d, dev, lin = _remap_setup(n, nz, mode)
    q = _ijk(_remap_field(dev, mode))
    g = torch.Generator(device="cpu").manual_seed(40 + mode)
    dq = _ijk(torch.randn(q.shape, generator=g, dtype=q.dtype).cuda())
    gout = _ijk(torch.randn(q.shape, generator=g, dtype=q.dtype).cuda())
    qs = dev["qs"] if mode == -2 else None
    _, dout = lin.tlm(q, dq, dev["pe1"], dev["pe2"], qs)
    gin = lin.adjoint(q, gout, dev["pe1"], dev["pe2"], qs)
    lhs = float((dout * gout).sum())
    rhs = float((dq * gin).sum())
    print(f"mode={mode}: <L dq, g> = {lhs:.15e}, <dq, L^T g> = {rhs:.15e}, "
          f"rel {_rel(lhs, rhs):.1e}")
    assert _rel(lhs, rhs) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_gpu_remap_tlm_finite_differences(mode):
    n, nz = 16, 79
    d, dev, lin = _remap_setup(n, nz, mode)
    q = _ijk(_remap_field(dev, mode))
    g = torch.Generator(device="cpu").manual_seed(8)
    dq = _ijk(torch.randn(q.shape, generator=g, dtype=q.dtype).cuda())
    gout = torch.randn(q.shape, generator=g, dtype=q.dtype).cuda()
    # a small step, so that (almost) no column crosses a limiter kink: the
    # kord-9 selections switch where two O(1)-spread quantities tie
    eps = 1e-9 * float(q.abs().max())
    fp = lin.forward(_ijk(q + eps * dq), dev["pe1"], dev["pe2"])
    fm = lin.forward(_ijk(q - eps * dq), dev["pe1"], dev["pe2"])
    fd = float(((fp - fm) * gout).sum()) / (2 * eps)
    _, dout = lin.tlm(q, dq, dev["pe1"], dev["pe2"])
    tl = float((dout * gout).sum())
    print(f"mode={mode}: FD {fd:.12e} TLM {tl:.12e} rel {_rel(fd, tl):.1e}")
    assert _rel(fd, tl) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("hord", [8])
@pytest.mark.parametrize("n", [24, 96])
def test_gpu_multistep_transport_remap_gradient(hord, n):
    """8 steps of fv_tp_2d + map_scalar (kord 9): reverse sweep vs central FD
    of J, on a 24 x 24 tile and on the 96 x 96 x 79 tile of the adjoint
    workload."""
    from fv3core_amd.adjoint import PeriodicTransport, cost_and_gradient
    from fv3core_amd.workloads import remap_inputs
    nz, nsteps = 79, 8
    d = _periodic(n, nz)
    r = remap_inputs(n, nz, workload=5)
    gi = GridIndexing((n, n, nz), N_HALO, False, False, False, False)
    dev = {k: _dev(v) for k, v in d.items()}
    remap = dict(pe1=_dev(r["pe1"]), pe2=_dev(r["pe2"]), mode=1, qmin=0.0)
    model = PeriodicTransport(gi, dev, hord, remap=remap)
    q0 = torch.as_tensor(d["q"]).permute(2, 1, 0).contiguous().cuda()
    g = torch.Generator(device="cpu").manual_seed(13)
    with torch.no_grad():
        q_obs = model.run(q0 + 1e-2 * torch.randn(q0.shape, generator=g, dtype=q0.dtype).cuda(),
                          nsteps)
    cost, grad = cost_and_gradient(model, q0, q_obs, nsteps)
    h = N_HALO
    for seed in (31, 32):
        gd = torch.Generator(device="cpu").manual_seed(seed)
        dq = torch.randn(q0.shape, generator=gd, dtype=q0.dtype).cuda()
        dq_int = torch.zeros_like(dq)
        dq_int[:, h:h + n, h:h + n] = dq[:, h:h + n, h:h + n]
        # kord 9 flattens a layer where two adjacent means tie, a jump in the
        # map: the step is kept small enough that no level pair of the 8
        # steps' ~45k crosses a tie (1e-6 crosses about one)
        eps = 1e-9
        with torch.no_grad():
            jp = 0.5 * float(((model.interior(model.run(q0 + eps * dq_int, nsteps)) -
                               model.interior(q_obs)) ** 2).sum())
            jm = 0.5 * float(((model.interior(model.run(q0 - eps * dq_int, nsteps)) -
                               model.interior(q_obs)) ** 2).sum())
        fd = (jp - jm) / (2 * eps)
        ad = float((grad * dq_int).sum())
        print(f"hord={hord} dir {seed}: J={cost:.6e} FD {fd:.10e} adjoint {ad:.10e} "
              f"rel {_rel(fd, ad):.1e}")
        assert _rel(fd, ad) < 1e-5


# ---------------------------------------------------------------------------
# torch-autodiff reference for the map_scalar adjoint (tests/torch_remap.py)
# ---------------------------------------------------------------------------


def _remap_columns(n, nz, mode, workload=4):
    """Columns [ncol, km] of the remap workload's compute box, as numpy."""
    from fv3core_amd.workloads import remap_inputs
    d = remap_inputs(n, nz, workload=workload)
    h = N_HALO
    box = (slice(h, h + n), slice(h, h + n))
    field = d["w"] if mode in (-2, 1, 2) else d["pt"]
    q = field[box][..., :nz].reshape(n * n, nz)
    pe1 = d["pe1"][box].reshape(n * n, nz + 1)
    pe2 = d["pe2"][box].reshape(n * n, nz + 1)
    qs = d["qs"][box].reshape(n * n)
    return d, q, pe1, pe2, qs


@pytest.mark.parametrize("mode", [-2, -1, 0, 1, 2])
def test_torch_remap_restatement_matches_oracle(mode):
    """The torch restatement of map_scalar (kord 9) equals the numpy oracle's
    MapSingle on the remap workload's columns for every iv mode."""
    from oracle import remap as orp
    from tests import torch_remap as tr
    n, nz = 6, 79
    d, q, pe1, pe2, qs = _remap_columns(n, nz, mode)
    got = tr.map_scalar(torch.as_tensor(q), pe1, pe2, mode,
                        qs=qs if mode == -2 else None).numpy()
    h = N_HALO
    field = d["w"] if mode in (-2, 1, 2) else d["pt"]
    want = orp.map_single(field, d["pe1"], d["pe2"], 9, mode, h, h + n - 1, h, h + n - 1,
                          qs=d["qs"] if mode == -2 else None)
    want = want[h:h + n, h:h + n, :nz].reshape(n * n, nz)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [-2, 0, 1])
def test_gpu_remap_adjoint_matches_torch_autodiff(mode):
    """The hand-written reverse sweep of map_scalar (kord 9) against torch
    autograd of the torch restatement, on the workload's tie-free columns:
    g_q = (d remap / d q)^T g_out to 1e-10."""
    from tests import torch_remap as tr
    n, nz = 16, 79
    d, dev, lin = _remap_setup(n, nz, mode)
    q = _ijk(_remap_field(dev, mode))
    g = torch.Generator(device="cpu").manual_seed(70 + mode)
    gout = _ijk(torch.randn(q.shape, generator=g, dtype=q.dtype).cuda())
    qs = dev["qs"] if mode == -2 else None
    gin = lin.adjoint(q, gout, dev["pe1"], dev["pe2"], qs)
    torch.cuda.synchronize()
    h = N_HALO
    box = (slice(h, h + n), slice(h, h + n))
    qc = torch.tensor(q.cpu().numpy()[box][..., :nz].reshape(n * n, nz), requires_grad=True)
    pe1 = d["pe1"][box].reshape(n * n, nz + 1)
    pe2 = d["pe2"][box].reshape(n * n, nz + 1)
    out = tr.map_scalar(qc, pe1, pe2, mode,
                        qs=d["qs"][box].reshape(n * n) if mode == -2 else None)
    go = torch.as_tensor(gout.cpu().numpy()[box][..., :nz].reshape(n * n, nz))
    (out * go).sum().backward()
    want = qc.grad.numpy()
    got = gin.cpu().numpy()[box][..., :nz].reshape(n * n, nz)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"mode={mode}: max |adjoint - autodiff| / max|autodiff| = {err:.2e}")
    assert err < 1e-10
