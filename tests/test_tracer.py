"""TracerAdvection on the six-tile cube: oracle pin (CPU) and GPU parity."""
import numpy as np
import pytest
import torch

from fv3core_amd import cube
from fv3core_amd.halo import CENTER, apply_map_numpy, build_map
from oracle import tracer as otr

N, NZ = 12, 5


def _inputs(seed=0, const_q=None, ntr=2):
    rng = np.random.default_rng(seed)
    m = cube.tile_metrics(N)
    shape = (6, N + 7, N + 7, NZ + 1)
    ii = np.arange(N + 7)[None, :, None, None]
    jj = np.arange(N + 7)[None, None, :, None]
    kk = np.arange(NZ + 1)[None, None, None, :]
    cx = 0.2 * np.sin(0.4 * ii + 0.3 * jj + 0.5 * kk) * np.ones(shape)
    cy = 0.2 * np.cos(0.3 * ii - 0.5 * jj + 0.2 * kk) * np.ones(shape)
    dp = 1000.0 + 50.0 * rng.uniform(size=shape)
    mfx = cx * m["dy"][..., None] * 500.0
    mfy = cy * m["dx"][..., None] * 500.0
    if const_q is None:
        qs = [rng.uniform(0.0, 1e-2, size=shape) for _ in range(ntr)]
    else:
        qs = [np.full(shape, const_q) for _ in range(ntr)]
    metric = {k: m[k] for k in ("dxa", "dya", "dx", "dy", "sin_sg1", "sin_sg2", "sin_sg3",
                                "sin_sg4", "area", "rarea")}
    return qs, dp, mfx, mfy, cx, cy, metric


def _halo_fn():
    mp = build_map(N, (CENTER,), 3, 3)
    return lambda fields: [apply_map_numpy([f], mp)[0] for f in fields]


def test_oracle_constant_tracer_stays_constant():
    qs, dp, mfx, mfy, cx, cy, metric = _inputs(1, const_q=0.0042)
    halo = _halo_fn()
    out = otr.tracer_advection(qs, dp, mfx, mfy, cx, cy, metric, 8, halo, NZ)
    for q in out[0]:
        np.testing.assert_allclose(q[:, 3:N + 3, 3:N + 3, :NZ], 0.0042, rtol=1e-12)
    # the accumulated fluxes were divided by the sub-cycle count
    np.testing.assert_allclose(out[4][..., :NZ], cx[..., :NZ] / 3.0, rtol=1e-15)


@pytest.mark.gpu
def test_gpu_tracer_advection_matches_oracle():
    from fv3core_amd.halo import CubedSphereCommunicator
    from fv3core_amd.stencils.fvtp2d import FiniteVolumeTransport
    from fv3core_amd.stencils.tracer_2d_1l import TracerAdvection
    from fv3core_amd.utils import gt4py_utils as utils
    from fv3core_amd.utils.grid import GridData
    from fv3core_amd.utils.stencil import GridIndexing, StencilConfig, StencilFactory
    from tests.helpers import assert_close

    qs, dp, mfx, mfy, cx, cy, metric = _inputs(2)
    gi = GridIndexing((N, N, NZ), 3, True, True, True, True)
    factory = StencilFactory(StencilConfig(), gi)

    def dev4(a):
        t = utils.empty(a.shape)
        t.copy_(torch.as_tensor(np.array(a)))
        return t

    def dev3(a):
        t = utils.empty(a.shape, batch_dims=1)
        t.copy_(torch.as_tensor(np.array(a)))
        return t

    grid = GridData({k: dev3(v) for k, v in metric.items()})
    comm = CubedSphereCommunicator(gi)
    transport = FiniteVolumeTransport(factory, grid, None, 0, 8)
    ta = TracerAdvection(factory, transport, comm, len(qs))
    tq = [dev4(q) for q in qs]
    tdp, tmfx, tmfy, tcx, tcy = (dev4(a) for a in (dp, mfx, mfy, cx, cy))
    ta({f"q{i}": q for i, q in enumerate(tq)}, tdp, tmfx, tmfy, tcx, tcy, 225.0)
    torch.cuda.synchronize()
    ref = otr.tracer_advection(qs, dp, mfx, mfy, cx, cy, metric, 8, _halo_fn(), NZ)
    c = (slice(None), slice(3, N + 3), slice(3, N + 3), slice(0, NZ))
    for g, r in zip(tq, ref[0]):
        assert_close(g.cpu().numpy()[c], r[c], 1e-12, near_zero=1e-20, what="tracer")
    assert_close(tdp.cpu().numpy()[c], ref[1][c], 1e-12, what="dp1")
    for g, r, name in zip((tmfx, tmfy, tcx, tcy), ref[2:], ("mfxd", "mfyd", "cxd", "cyd")):
        assert_close(g.cpu().numpy()[..., :NZ], r[..., :NZ], 1e-15, what=name)


@pytest.mark.gpu
@pytest.mark.parametrize("ntr", [1, 8])
def test_gpu_batched_tracer_transport_equals_one_by_one(ntr):
    """The batched sub-cycle transport (one x corner fill, one fv_tp_2d and
    one q update launch for every tracer) returns the bits of the
    per-tracer loop (corner copies, fv_tp_2d, q update for each tracer)."""
    from fv3core_amd.halo import CubedSphereCommunicator
    from fv3core_amd.stencils.fvtp2d import FiniteVolumeTransport
    from fv3core_amd.stencils.tracer_2d_1l import TracerAdvection
    from fv3core_amd.utils import gt4py_utils as utils
    from fv3core_amd.utils.grid import GridData
    from fv3core_amd.utils.stencil import GridIndexing, StencilConfig, StencilFactory

    qs, dp, mfx, mfy, cx, cy, metric = _inputs(3, ntr=ntr)
    gi = GridIndexing((N, N, NZ), 3, True, True, True, True)
    factory = StencilFactory(StencilConfig(), gi)

    def dev(a, bd=None):
        t = utils.empty(a.shape, batch_dims=bd)
        t.copy_(torch.as_tensor(np.array(a)))
        return t

    grid = GridData({k: dev(v, 1) for k, v in metric.items()})
    comm = CubedSphereCommunicator(gi)
    runs = []
    for batched in (True, False):
        transport = FiniteVolumeTransport(factory, grid, None, 0, 8)
        ta = TracerAdvection(factory, transport, comm, ntr)
        assert ta._batched
        ta._batched = batched
        tq = [dev(q) for q in qs]
        args = [dev(a) for a in (dp, mfx, mfy, cx, cy)]
        ta({f"q{i}": q for i, q in enumerate(tq)}, *args, 225.0)
        torch.cuda.synchronize()
        runs.append(tq + args)
    for a, b in zip(*runs):
        assert torch.equal(a, b)
