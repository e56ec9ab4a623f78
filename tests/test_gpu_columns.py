"""GPU parity of the column kernels (remap, Riemann solvers) against the oracle."""
import numpy as np
import pytest
import torch

from fv3core_amd._config import RiemannConfig
from fv3core_amd.stencils.map_single import MapSingle
from fv3core_amd.stencils.riem_solver3 import RiemannSolver3
from fv3core_amd.stencils.riem_solver_c import RiemannSolverC
from fv3core_amd.utils import gt4py_utils as utils
from fv3core_amd.utils.grid import to_storage
from fv3core_amd.utils.stencil import GridIndexing, StencilConfig, StencilFactory
from fv3core_amd.workloads import remap_inputs, riemann_inputs
from oracle import remap as orm
from oracle import riemann as ori
from tests.helpers import assert_close, to_np

pytestmark = pytest.mark.gpu

NZ = 79


def _factory(n, nz=NZ):
    gi = GridIndexing((n, n, nz), 3, True, True, True, True)
    return StencilFactory(StencilConfig(), gi), gi


@pytest.mark.parametrize("kord,mode", [(9, 1), (10, 1), (10, 0), (9, -1), (10, -2), (9, 2),
                                       (5, 1), (-10, 1)])
def test_map_single_parity(kord, mode):
    n = 12
    d = remap_inputs(n, NZ, workload=51)
    factory, gi = _factory(n)
    i1, i2, j1, j2 = gi.isc, gi.iec, gi.jsc, gi.jec
    src = d["q"] if mode == 0 else (d["w"] if mode in (-2, -1) else d["pt"])
    ms = MapSingle(factory, kord, mode, i1, i2, j1, j2)
    q1 = to_storage(src)
    qs = to_storage(d["qs"]) if mode == -2 else None
    ms(q1, to_storage(d["pe1"]), to_storage(d["pe2"]), qs=qs, qmin=0.0)
    torch.cuda.synchronize()
    ref = orm.map_single(src, d["pe1"], d["pe2"], kord, mode, i1, i2, j1, j2,
                         qs=d["qs"] if mode == -2 else None)
    assert_close(to_np(q1)[..., :NZ], ref[..., :NZ], 1e-12, near_zero=1e-18, what="q1")


def test_map_single_batched_and_qmin():
    n, nt = 10, 3
    factory, gi = _factory(n)
    ins = [remap_inputs(n, NZ, workload=60 + t) for t in range(nt)]
    shape = gi.max_shape

    def batch(name):
        b = utils.empty((nt,) + shape)
        for t in range(nt):
            b[t].copy_(torch.as_tensor(ins[t][name]))
        return b

    q1, pe1, pe2 = batch("pt"), batch("pe1"), batch("pe2")
    ms = MapSingle(factory, 9, 1, gi.isc, gi.iec, gi.jsc, gi.jec)
    ms(q1, pe1, pe2, qmin=221.0)
    torch.cuda.synchronize()
    for t in range(nt):
        ref = orm.map_single(ins[t]["pt"], ins[t]["pe1"], ins[t]["pe2"], 9, 1, gi.isc, gi.iec,
                             gi.jsc, gi.jec, qmin=221.0)
        assert_close(to_np(q1[t])[..., :NZ], ref[..., :NZ], 1e-12, what=f"tile {t}")


@pytest.mark.parametrize("seed", [0, 1])
def test_riem_solver_c_parity(seed):
    n = 12
    r = riemann_inputs(n, NZ, workload=70 + seed)
    factory, gi = _factory(n)
    rs = RiemannSolverC(factory, 0.05)
    dev = {k: to_storage(v) for k, v in r.items() if isinstance(v, np.ndarray) and v.ndim >= 2}
    gz = dev["gz"].clone()
    pef = utils.zeros(gi.max_shape)
    rs(r["dt"] / 2, dev["cappa"], r["ptop"], dev["hs"], dev["ws"], dev["pt"], dev["q_con"],
       dev["delp"], gz, pef, dev["w"])
    torch.cuda.synchronize()
    sl = (slice(gi.isc - 1, gi.iec + 2), slice(gi.jsc - 1, gi.jec + 2))
    rgz, rpef = ori.riem_solver_c(r["dt"] / 2, r["cappa"][sl], r["ptop"], r["hs"][sl],
                                  r["ws"][sl], r["pt"][sl], r["q_con"][sl], r["delp"][sl],
                                  r["gz"][sl], r["w"][sl], 0.05)
    assert_close(to_np(gz)[sl], rgz, 1e-12, what="gz")
    assert_close(to_np(pef)[sl], rpef, 1e-12, what="pef")


@pytest.mark.parametrize("last_call,use_logp,beta", [(True, False, 0.0), (False, False, 0.0),
                                                     (True, True, -0.2)])
def test_riem_solver3_parity(last_call, use_logp, beta):
    n = 12
    r = riemann_inputs(n, NZ, workload=80)
    factory, gi = _factory(n)
    cfg = RiemannConfig(p_fac=0.05, a_imp=1.0, use_logp=use_logp, beta=beta)
    rs = RiemannSolver3(factory, cfg)
    rng = np.random.default_rng(3)
    pk3_0 = rng.uniform(size=r["pe"].shape)
    pk_0 = rng.uniform(size=r["pe"].shape)
    peln_0 = rng.uniform(size=r["pe"].shape)
    ppe_0 = np.zeros(r["pe"].shape)
    t = {k: to_storage(v) for k, v in r.items() if isinstance(v, np.ndarray) and v.ndim >= 2}
    delz, zh, pe, w = t["delz"].clone(), t["zh"].clone(), t["pe"].clone(), t["w"].clone()
    ppe, pk3, pk, peln = (to_storage(a) for a in (ppe_0, pk3_0, pk_0, peln_0))
    wsd = to_storage(np.full(r["zs"].shape, 0.01))
    rs(last_call, r["dt"], t["cappa"], r["ptop"], t["zs"], wsd, delz, t["q_con"], t["delp"],
       t["pt"], zh, pe, ppe, pk3, pk, peln, w)
    torch.cuda.synchronize()
    sl = (slice(gi.isc, gi.iec + 1), slice(gi.jsc, gi.jec + 1))
    ref = ori.riem_solver3(last_call, r["dt"], r["cappa"][sl], r["ptop"], r["zs"][sl],
                           np.full(r["zs"].shape, 0.01)[sl], r["delz"][sl], r["q_con"][sl],
                           r["delp"][sl], r["pt"][sl], r["zh"][sl], r["pe"][sl], ppe_0[sl],
                           pk3_0[sl], pk_0[sl], peln_0[sl], r["w"][sl], 0.05, use_logp=use_logp,
                           beta=beta)
    names = ["zh", "pe", "ppe", "pk3", "pk", "peln", "w", "delz"]
    got = [zh, pe, ppe, pk3, pk, peln, w, delz]
    # ppe and w come out of small differences (w - w1 summed into ppe): their
    # error is measured against the field's largest magnitude, so that
    # last-bit exp/log differences amplified by the cancellation do not count
    floors = {"ppe": 0.5, "w": 0.5}
    for name, g, rr in zip(names, got, ref):
        assert_close(to_np(g)[sl], rr, 1e-12, near_zero=1e-12, what=name,
                     scale_floor=floors.get(name, 0.0))
