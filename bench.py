"""Benchmark of the FV3 hot path on one MI355X.

Prints one JSON line.  A step is one pass of the hot path over one batch of
synthetic input with every input already resident in HBM.

Workload "dyncore" (default): AcousticDynamics (dyn_core: c_sw, d_sw,
riem_solver_c, riem_solver3, updatedzc/updatedzd, nh_p_grad, Rayleigh
damping, the on-device six-tile halo updates) on six C96 tiles 96x96x79 in
FP64, n_split = 6 acoustic substeps per call, dt = 225 s, from the
hydrostatic jet state of ``workloads.cube_state`` (hybrid levels, ps = 1e5,
T = 288 K lapsing to 216 K, u = 35 sin^2(2 lat), seeded 1 K pt bump, w = 0).
One step is one call; cell-level updates per step = 6 * 96 * 96 * 79 * 6
(every cell advanced through every acoustic substep).

Workload "fv_dynamics": DynamicalCore.step_dynamics on six C128 tiles
128x128x79 (the same initial state at C128 plus 8 smooth tracers in
[0, 1e-2]): 6 acoustic substeps, tracer advection (hord 8) and the remap
with kord 10; cell-level updates per step = 6 * 128 * 128 * 79 (the acoustic
count, 6x that, is given beside it).

Workload "ensemble" (the default): 32 perturbed members of six C48 tiles in
two batches of 96 tiles plus four uneven members (C24, C64, C96, C128),
AcousticDynamics on every batch per step, FP64 and mixed, every batch of a
call in one HIP graph with all but the first batch as parallel branches.

Workload "adjoint": 8 steps of flux-form fv_tp_2d (hord 8) transport plus
map_scalar remap (kord 9) on one doubly periodic 96x96x79 tile and the
gradient of J = 0.5 ||q_T - q_obs||^2 by the hand-written adjoints; the line adds "adjoint_over_forward", the time
of the reverse sweep (re-linearisation included) over the forward run.

Workload "fvtp2d": FiniteVolumeTransport (fv_tp_2d, hord=8, area fluxes) on
6 tiles of 96x96x79, FP64, one fused kernel per step; cell-level updates per
step = 6 * 96 * 96 * 79.

The roofline object prices one kernel of the step (the fused fv_tp_2d
kernel unless --roofline-kernel names another) with its algorithmic bytes
over its average launch time, measured with HIP events recorded on the
kernel's own stream by the library during the timed region.  The CPU
baseline times the numpy oracle on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
import types

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
MIXED_ERROR_FIELDS = ("delp", "pt", "delz", "w", "u", "v", "zh", "pe", "pk", "peln", "pkz", "ua",
                      "va", "uc", "vc", "omga")
MIXED_FORM = ("FP32 storages and arithmetic for the bulk fields (u, v, w, delp, pt, q_con, "
              "cappa, C-grid winds, fluxes and accumulators) in c_sw, d_sw (fv_tp_2d, del-n "
              "and divergence damping, a2b_ord4), Rayleigh damping, del2cubed and their halo "
              "exchange; FP64 heights, pressures (pe, peln, pk, pk3, pkc, pkz, zh, gz, delz), "
              "Riemann solvers, height updates and pressure gradients")


def _parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None)
    p.add_argument("--warmup", type=int, default=None)
    p.add_argument("--no-adjoint", action="store_true",
                   help="ensemble: skip the adjoint/forward ratio run")
    p.add_argument("--workload", default="ensemble",
                   choices=["dyncore", "fv_dynamics", "fvtp2d", "adjoint", "ensemble"])
    p.add_argument("--roofline-kernel", default="fvtp2d_kernel",
                   choices=["fvtp2d_kernel", "riem3_kernel", "riem_c_kernel",
                            "map_single_kernel", "delnflux_kernel", "a2b_interp_kernel"])
    p.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    p.add_argument("--graph", action="store_true",
                   help="dyncore: also time the step replayed from a captured HIP graph")
    p.add_argument("--no-graph", action="store_true",
                   help="ensemble: time only the eager calls (default: also the one-graph "
                        "form, which gives value)")
    p.add_argument("--mixed", action="store_true",
                   help="dyncore / fv_dynamics: the mixed-precision form (FP32 bulk fields); "
                        "reports its error against the FP64 form")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    p.add_argument("--copy-calibration", action="store_true",
                   help="also run a plain device copy of known size (for PMC scaling)")
    return p.parse_args()


def cpu_baseline(seconds: float):
    """Oracle fv_tp_2d (numpy, FP64, one core) on the 48x48x79 periodic tile."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from fv3core_amd.workloads import transport_inputs
    from oracle import fvtp2d as ofv

    n, nz = 48, 79
    d = transport_inputs(n, nz, workload=1)
    t = ofv.Tile(n, n, 3, False, False, False, False)
    args = (d["q"], d["q"], d["crx"], d["cry"], d["x_area_flux"], d["y_area_flux"], d["area"],
            d["dxa"], d["dya"], 8, t)
    ofv.fvtp2d(*args, check_domains=False)
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        ofv.fvtp2d(*args, check_domains=False)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    cells = n * n * nz
    return {
        "value": cells / med,
        "unit": "cell-level-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"numpy oracle fv_tp_2d hord=8, one 48x48x79 periodic tile, median of "
        f"{len(times)} calls ({med * 1e3:.1f} ms/call), host {os.cpu_count()} cpus visible",
    }


def _oracle_acoustic_sample(seconds: float):
    """(median seconds per call, calls, cell-level updates per call) of the
    oracle AcousticDynamics on six C24 tiles x 79 levels, n_split = 2."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    import warnings

    from fv3core_amd._config import DynamicalCoreConfig
    from fv3core_amd.stencils.d_sw import get_column_namelist
    from fv3core_amd.workloads import cube_state
    from oracle import dyncore as odc
    from oracle.fv_dynamics import init_pfull
    from oracle.host_halo import HostHalo, acoustic_state

    n, nz, n_split = 24, 79, 2
    init = cube_state(n, nz, seed=0)
    st0, m = init["state"], init["metric"]
    ak, bk, ptop = init["ak"], init["bk"], init["ptop"]
    cfg = DynamicalCoreConfig(npz=nz, n_split=n_split, k_split=1, nord=3, dt_atmos=225)
    metrics = [{k: v[t] for k, v in m.items() if isinstance(v, np.ndarray) and v.ndim == 3}
               for t in range(6)]
    col = get_column_namelist(cfg.d_grid_shallow_water, nz)
    ocfg = dict(mdt=float(cfg.dt_atmos), n_split=n_split, k_split=1, n_map=1, ptop=ptop,
                hord_dp=cfg.hord_dp, hord_tm=cfg.hord_tm, hord_mt=cfg.hord_mt,
                hord_vt=cfg.hord_vt, dddmp=cfg.dddmp, d4_bg=cfg.d4_bg, nord=cfg.nord,
                d_con=cfg.d_con, convert_ke=cfg.convert_ke, vtdm4=cfg.vtdm4, p_fac=cfg.p_fac,
                rf_cutoff=cfg.rf_cutoff, tau=cfg.tau, delt_max=cfg.delt_max)
    pf = init_pfull(ak, bk)
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or not times:
        S = acoustic_state(st0, init)
        halo = HostHalo(S, n)
        t0 = time.perf_counter()
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            odc.acoustic_dynamics(S, ocfg, metrics, dict(da_min=m["da_min"],
                                                         da_min_c=m["da_min_c"]),
                                  ak, bk, pf, halo, halo.sync, col, nz)
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), len(times), 6 * n * n * nz * n_split


def _oracle_worker(seconds):
    return _oracle_acoustic_sample(seconds)


def cpu_baseline_dyncore(seconds: float, all_cores: bool = True):
    """Oracle AcousticDynamics (numpy, FP64) on six C24 tiles x 79 levels,
    n_split = 2 (a bounded sample of the acoustic workloads): one core, and
    with ``all_cores`` one independent member per worker process on up to 16
    cores at once (members need no exchange, as in the batched ensemble).
    Runs before the process touches the GPU (the pool forks)."""
    import multiprocessing

    med, calls, cells = _oracle_acoustic_sample(seconds)
    out = {
        "value": cells / med,
        "unit": "cell-level-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"numpy oracle AcousticDynamics on 6 x C24 x 79 levels, n_split=2, "
        f"same initial state and namelist, median of {calls} calls ({med:.2f} s/call), "
        f"host {os.cpu_count()} cpus visible",
    }
    if all_cores:
        workers = max(1, min(16, os.cpu_count() or 1))
        t0 = time.perf_counter()
        with multiprocessing.get_context("fork").Pool(workers) as pool:
            res = pool.map(_oracle_worker, [seconds] * workers)
        elapsed = time.perf_counter() - t0
        rate = sum(c / m for m, _, c in res)
        out["all_cores"] = {
            "value": rate, "unit": "cell-level-updates/s", "cores": workers,
            "sample": f"{workers} worker processes, each timing its own member of the same "
            f"sample concurrently (median per worker), {elapsed:.1f} s wall",
        }
    return out


def _copy_calibration():
    """A plain device copy of known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters of a profiled run (256 MiB read + 256 MiB written)."""
    import torch

    a = torch.ones(256 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    return a.numel() * 8


def _profiled_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` in `workload` from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summaries (copy-calibrated), or None."""
    import glob
    import re

    pat = re.compile(rf"^traffic_bytes_per_launch {re.escape(kernel)} {re.escape(workload)}:"
                     r"\s*([0-9.eE+]+)\s*$")
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = pat.match(line.strip())
                if m:
                    return float(m.group(1))
    return None


def _timing(lib, name):
    import ctypes

    cnt, ms, byt = ctypes.c_int64(0), ctypes.c_double(0.0), ctypes.c_double(0.0)
    rc = lib.fv3_timing_read(ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(byt))
    if rc != 0:
        raise RuntimeError(lib.fv3_last_error().decode())
    return cnt.value, ms.value, byt.value


def _mixed_errors(make, state_of, n, nz):
    """Max 2|a-b|/(|a|+|b|) (denominators floored at 1e-6 of the field's max)
    of every output field after one AcousticDynamics call of the mixed form
    against the FP64 form from the same initial state; ps is pe at the
    surface."""
    import torch

    from fv3core_amd.mixed import mixed_state

    plain, mixed = make(False), make(True)
    sp, sm = state_of(), mixed_state(state_of())
    plain(sp)
    mixed(sm)
    torch.cuda.synchronize()
    c = (slice(None), slice(3, n + 3), slice(3, n + 3))

    def rel(a, b):
        a, b = a.double(), b.double()
        scale = float(torch.maximum(a.abs().max(), b.abs().max()))
        den = torch.clamp(a.abs() + b.abs(), min=1e-6 * scale + 1e-300)
        return float((2.0 * (a - b).abs() / den).max())

    errs = {k: rel(getattr(sm, k)[c + (slice(0, nz),)], getattr(sp, k)[c + (slice(0, nz),)])
            for k in MIXED_ERROR_FIELDS}
    errs["ps"] = rel(sm.pe[c + (nz,)], sp.pe[c + (nz,)])
    errs["scaled"] = _scaled_errors(sm, sp, n, nz)
    return errs


def _scaled_errors(sm, sp, n, nz):
    """max|a - b| / max|b| per field: the error against the field's scale,
    which stays meaningful for fields that pass through zero (w, winds)."""
    import torch

    c = (slice(None), slice(3, n + 3), slice(3, n + 3))
    out = {}
    for k in MIXED_ERROR_FIELDS:
        a = getattr(sm, k)[c + (slice(0, nz),)].double()
        b = getattr(sp, k)[c + (slice(0, nz),)].double()
        out[k] = float((a - b).abs().max() / torch.clamp(b.abs().max(), min=1e-300))
    return out


def bench_dyncore(steps: int, warmup: int, roofline_kernel: str, graph: bool = False,
                  copy_calibration: bool = False, mixed: bool = False):
    import ctypes

    import torch

    from fv3core_amd import _lib
    from fv3core_amd._config import DynamicalCoreConfig
    from fv3core_amd.halo import CubedSphereCommunicator
    from fv3core_amd.stencils.dyn_core import AcousticDynamics
    from fv3core_amd.stencils.fv_dynamics import init_pfull
    from fv3core_amd.utils import gt4py_utils as utils
    from fv3core_amd.utils.grid import DampingCoefficients, GridData
    from fv3core_amd.utils.stencil import GridIndexing, StencilConfig, StencilFactory
    from fv3core_amd.workloads import cube_state

    n, nz, n_split = 96, 79, 6
    init = cube_state(n, nz, seed=0)
    st0, m = init["state"], init["metric"]
    ak, bk, ptop = init["ak"], init["bk"], init["ptop"]
    cfg = DynamicalCoreConfig(npz=nz, n_split=n_split, k_split=1, nord=3, dt_atmos=225)
    gi = GridIndexing((n, n, nz), 3, True, True, True, True)
    factory = StencilFactory(StencilConfig(), gi)

    def dev(a, bd):
        t = utils.empty(a.shape, batch_dims=bd)
        t.copy_(torch.as_tensor(a))
        return t

    grid = GridData({k: dev(v, 1) for k, v in m.items()
                     if isinstance(v, np.ndarray) and v.ndim == 3})
    damp = DampingCoefficients(grid.del6_u, grid.del6_v, m["da_min"], m["da_min_c"])
    comm = CubedSphereCommunicator(gi)
    phis = dev(init["phis"], 1)
    def make(mixed_form):
        return AcousticDynamics(comm, factory, grid, damp, 0, False, False,
                                cfg.acoustic_dynamics, ak, bk, init_pfull(ak, bk, 1.0e5), phis,
                                mixed_precision=mixed_form)

    def state_of():
        st = types.SimpleNamespace(**{k: dev(v, None) for k, v in st0.items()})
        st.phis = phis
        st.wsd = dev(init["wsd"], 1)
        st.ptop, st.ks, st.mdt, st.n_map = ptop, 0, float(cfg.dt_atmos), 1
        return st

    mixed_errs = _mixed_errors(make, state_of, n, nz) if mixed else None
    acoustic = make(mixed)
    state = state_of()
    if mixed:
        from fv3core_amd.mixed import mixed_state

        state = mixed_state(state)
    lib = _lib.lib()
    for _ in range(warmup):
        acoustic(state)
    torch.cuda.synchronize()
    lib.fv3_timing_arm(roofline_kernel.encode())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        acoustic(state)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cnt, kern_ms, kern_bytes = _timing(lib, roofline_kernel)
    lib.fv3_timing_arm(None)
    finite = bool(torch.isfinite(state.pt[:, 3:n + 3, 3:n + 3, :nz]).all().item())
    graph_wall = None
    if graph:
        acoustic._use_graph = True
        for _ in range(max(warmup, 0) + 2):
            acoustic(state)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            acoustic(state)
        torch.cuda.synchronize()
        graph_wall = time.perf_counter() - t0
    copy_bytes = _copy_calibration() if copy_calibration else None
    cells = 6 * n * n * nz * n_split
    avg_s = kern_ms / max(cnt, 1) * 1e-3
    achieved = kern_bytes / max(kern_ms * 1e-3, 1e-30) / 1e9
    return {
        "value": steps * cells / (graph_wall if graph_wall else wall),
        "ms_per_step": (graph_wall if graph_wall else wall) / steps * 1e3,
        "eager_ms_per_step": wall / steps * 1e3,
        "graph_ms_per_step": graph_wall / steps * 1e3 if graph_wall else None,
        "config": {
            "workload": f"AcousticDynamics (dyn_core) on 6 cube tiles x {n}x{n}x{nz}, "
            f"n_split={n_split}, dt=225 s, nonhydrostatic, jet + 1 K pt bump initial state" +
            (", mixed precision: " + MIXED_FORM if mixed else ""),
            "tiles": 6, "n": n, "nz": nz, "n_split": n_split,
            "cells_per_step": cells, "state_finite_after_timed_steps": finite,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": _profiled_traffic(roofline_kernel, "dyncore"),
            "kernel": roofline_kernel,
            "kernel_launches_timed": cnt,
            "kernel_avg_us": avg_s * 1e6,
            "algorithmic_bytes_per_launch": kern_bytes / max(cnt, 1),
        },
        "dtype": "f64+f32" if mixed else "f64",
        "copy_calibration_bytes": copy_bytes,
        "mixed_vs_fp64": mixed_errs,
        "mixed_gate_ok": (max(mixed_errs[k] for k in ("ps", "pt", "delp")) <= 1e-5
                          if mixed else None),
    }


def bench_fv_dynamics(steps: int, warmup: int, roofline_kernel: str, mixed: bool = False):
    """DynamicalCore.step_dynamics on six C128 tiles x 79 levels: n_split = 6
    acoustic substeps, 8 tracers advected (hord_tr = 8), remap with kord 10.
    With ``mixed`` both forms are timed, ``value`` is the mixed form's rate
    (acoustic substeps in the mixed form, tracers and remap in FP64), and
    the line adds the FP64 rate, their ratio and the mixed form's errors in
    ps, pt and delp against the FP64 form after one step from the same
    state."""
    import gc

    import torch

    from fv3core_amd import _lib
    from fv3core_amd._config import DynamicalCoreConfig
    from fv3core_amd.halo import CubedSphereCommunicator
    from fv3core_amd.stencils.fv_dynamics import DynamicalCore
    from fv3core_amd.utils import gt4py_utils as utils
    from fv3core_amd.utils.grid import DampingCoefficients, GridData
    from fv3core_amd.utils.stencil import GridIndexing, StencilConfig, StencilFactory
    from fv3core_amd.workloads import cube_state, cube_tracers

    n, nz, n_split = 128, 79, 6
    init = cube_state(n, nz, seed=0)
    st0, m = init["state"], init["metric"]
    ak, bk, ptop = init["ak"], init["bk"], init["ptop"]
    cfg = DynamicalCoreConfig(npz=nz, n_split=n_split, k_split=1, nord=3, dt_atmos=225,
                              kord_tm=-10, kord_tr=10, kord_wz=10, kord_mt=10, hord_tr=8)
    gi = GridIndexing((n, n, nz), 3, True, True, True, True)
    factory = StencilFactory(StencilConfig(), gi)

    def dev(a, bd):
        t = utils.empty(a.shape, batch_dims=bd)
        t.copy_(torch.as_tensor(a))
        return t

    grid = GridData({k: dev(v, 1) for k, v in m.items()
                     if isinstance(v, np.ndarray) and v.ndim == 3})
    damp = DampingCoefficients(grid.del6_u, grid.del6_v, m["da_min"], m["da_min_c"])
    comm = CubedSphereCommunicator(gi)
    phis = dev(init["phis"], 1)
    tracers0 = cube_tracers(st0["delp"].shape)

    def fresh():
        state = {k: dev(v, None) for k, v in st0.items() if k != "cappa"}
        state.update({k: dev(v, None) for k, v in tracers0.items()})
        state["phis"] = phis
        state["ps"] = dev(st0["pe"][..., nz].copy(), 1)
        return state

    bdt = float(cfg.dt_atmos)
    lib = _lib.lib()

    def run(form_mixed):
        core = DynamicalCore(comm, grid, factory, damp, cfg, ak, bk, phis, skip_wrapup=True,
                             mixed_precision=form_mixed)
        state = fresh()
        core.step_dynamics(state, 0.0, False, bdt, ptop, n_split, 0)
        torch.cuda.synchronize()
        one_step = {k: state[k].clone() for k in ("pt", "delp", "pe")}
        for _ in range(warmup):
            core.step_dynamics(state, 0.0, False, bdt, ptop, n_split, 0)
        torch.cuda.synchronize()
        lib.fv3_timing_arm(roofline_kernel.encode())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            core.step_dynamics(state, 0.0, False, bdt, ptop, n_split, 0)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        timing = _timing(lib, roofline_kernel)
        lib.fv3_timing_arm(None)
        finite = bool(torch.isfinite(state["pt"][:, 3:n + 3, 3:n + 3, :nz]).all().item())
        del core, state
        gc.collect()
        torch.cuda.empty_cache()
        return wall, timing, finite, one_step

    results = {form: run(form == "mixed") for form in (("fp64", "mixed") if mixed else ("fp64",))}
    wall, (cnt, kern_ms, kern_bytes), finite, one = results["mixed" if mixed else "fp64"]
    cells = 6 * n * n * nz
    achieved = kern_bytes / max(kern_ms * 1e-3, 1e-30) / 1e9
    out = {
        "value": steps * cells / wall,
        "ms_per_step": wall / steps * 1e3,
        "config": {
            "workload": f"DynamicalCore.step_dynamics on 6 cube tiles x {n}x{n}x{nz}: "
            f"n_split={n_split} acoustic substeps, 8 tracers (hord_tr=8), remap kord=10 "
            "with moist_cv, dt=225 s" + ("; mixed form: acoustic substeps with " + MIXED_FORM +
                                         ", tracer transport and remap FP64" if mixed else ""),
            "tiles": 6, "n": n, "nz": nz, "n_split": n_split, "tracers": 8,
            "cells_per_step": cells, "acoustic_cells_per_step": cells * n_split,
            "state_finite_after_timed_steps": finite,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": roofline_kernel,
            "kernel_launches_timed": cnt,
            "kernel_avg_us": kern_ms / max(cnt, 1) * 1e3,
            "algorithmic_bytes_per_launch": kern_bytes / max(cnt, 1),
        },
        "dtype": "f64+f32" if mixed else "f64",
    }
    if mixed:
        wall64, _, finite64, one64 = results["fp64"]
        c = (slice(None), slice(3, n + 3), slice(3, n + 3))

        def rel(a, b):
            a, b = a.double(), b.double()
            den = torch.clamp(a.abs() + b.abs(), min=1e-6 * float(b.abs().max()) + 1e-300)
            return float((2.0 * (a - b).abs() / den).max())

        def share(a, b, tol=1e-5):
            a, b = a.double(), b.double()
            den = torch.clamp(a.abs() + b.abs(), min=1e-6 * float(b.abs().max()) + 1e-300)
            return float((2.0 * (a - b).abs() / den > tol).double().mean())

        pairs = {k: (one[k][c + (slice(0, nz),)], one64[k][c + (slice(0, nz),)])
                 for k in ("pt", "delp")}
        pairs["ps"] = (one["pe"][c + (nz,)], one64["pe"][c + (nz,)])
        errs = {k: rel(a, b) for k, (a, b) in pairs.items()}
        # the kord-10 remap switches its reconstruction discontinuously (one
        # ulp on its inputs moves single levels by up to 1e-4), so the
        # pointwise maximum is set by a few switched points: the share of
        # points above the gate is given beside it
        errs["share_above_1e-5"] = {k: share(a, b) for k, (a, b) in pairs.items()}
        out.update({
            "fp64_value": steps * cells / wall64,
            "fp64_ms_per_step": wall64 / steps * 1e3,
            "mixed_over_fp64": wall64 / wall,
            "mixed_vs_fp64": errs,
            "mixed_gate_ok": max(errs[k] for k in ("pt", "delp", "ps")) <= 1e-5,
        })
        out["config"]["state_finite_after_timed_steps"] = finite and finite64
    return out


def bench_fvtp2d(steps: int, warmup: int, dtype_name: str, copy_calibration: bool):
    import torch

    from fv3core_amd.stencils.fvtp2d import CopiedCorners, FiniteVolumeTransport
    from fv3core_amd.utils import gt4py_utils as utils
    from fv3core_amd.utils.grid import grid_from_numpy
    from fv3core_amd.utils.stencil import GridIndexing, StencilConfig, StencilFactory
    from fv3core_amd.workloads import transport_inputs

    dtype = torch.float64 if dtype_name == "f64" else torch.float32
    n, nz, ntiles, hord = 96, 79, 6, 8
    gi = GridIndexing((n, n, nz), 3, False, False, False, False)
    factory = StencilFactory(StencilConfig(), gi)
    shape = gi.max_shape
    fields = {}
    for name in ("q", "crx", "cry", "x_area_flux", "y_area_flux"):
        fields[name] = utils.empty((ntiles,) + shape, dtype=dtype)
    metric = None
    for t in range(ntiles):
        d = transport_inputs(n, nz, workload=2 + 10 * t)
        if metric is None:
            metric = d
        for name in fields:
            fields[name][t].copy_(torch.as_tensor(d[name], dtype=dtype))
    grid, damp = grid_from_numpy(metric, dtype=dtype)
    fx = utils.zeros((ntiles,) + shape, dtype=dtype)
    fy = utils.zeros((ntiles,) + shape, dtype=dtype)
    fvt = FiniteVolumeTransport(factory, grid, damp, 0, hord)
    q = CopiedCorners(fields["q"], fields["q"], fields["q"])
    call = lambda: fvt(q, fields["crx"], fields["cry"], fields["x_area_flux"],  # noqa: E731
                       fields["y_area_flux"], fx, fy)
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        ev[s][0].record(stream)
        call()
        ev[s][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [a.elapsed_time(b) for a, b in ev]
    cells = ntiles * n * n * nz
    bytes_per_cell = 7 * (8 if dtype == torch.float64 else 4)
    avg_kernel_s = float(np.mean(kern_ms)) * 1e-3
    achieved = cells * bytes_per_cell / avg_kernel_s / 1e9
    out = {
        "value": steps * cells / wall,
        "ms_per_step": wall / steps * 1e3,
        "config": {
            "workload": f"fv_tp_2d hord={hord} (area fluxes, no damping), {ntiles} tiles x "
            f"{n}x{n}x{nz}, doubly periodic synthetic tiles, one fused kernel per step",
            "tiles": ntiles,
            "n": n,
            "nz": nz,
            "hord": hord,
            "cells_per_step": cells,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": "fvtp2d_kernel",
            "kernel_avg_us": avg_kernel_s * 1e6,
            "algorithmic_bytes_per_launch": cells * bytes_per_cell,
        },
        "dtype": dtype_name,
    }
    if copy_calibration:
        a = torch.empty(256 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        out["copy_calibration_bytes"] = a.numel() * 8
    return out


# algorithmic HBM bytes per cell and acoustic substep of the whole step
# (every state field read and written once, the accumulators read and
# written, the C-grid winds written): 40 FP64 transfers
STEP_BYTES_PER_CELL_FP64 = 320.0


def _ensemble_parts(cfg, mixed):
    from fv3core_amd.ensemble import ensemble_acoustic
    from fv3core_amd.workloads import cube_state

    # 32 perturbed C48 members in two batches of 16 (run side by side, one
    # batch's kernel tails overlap the other's kernels), then the uneven
    # members built the same way without perturbation
    plan = [(48, 16, 0.5, 1.0, 48000), (48, 16, 0.5, 1.0, 48001)] + [
        (n, 1, 0.0, 0.0, 1000 * n) for n in (24, 64, 96, 128)]
    parts, cells = [], 0
    for n, members, pt_sd, uv_sd, seed in plan:
        init = cube_state(n, cfg.npz, seed=0)
        parts.append(ensemble_acoustic(init, members, cfg, seed=seed, pt_sd=pt_sd,
                                       uv_sd=uv_sd, mixed_precision=mixed))
        cells += members * 6 * n * n * cfg.npz * cfg.n_split
        del init
    return parts, cells


def _time_ensemble(parts, steps, warmup, roofline_kernel, graph):
    import torch

    from fv3core_amd import _lib
    from fv3core_amd.ensemble import UnevenEnsemble

    ens = UnevenEnsemble(parts)
    lib = _lib.lib()
    for _ in range(warmup):
        ens()
    torch.cuda.synchronize()
    lib.fv3_timing_arm(roofline_kernel.encode())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        # the roofline kernel is timed on the first part (16 C48 members):
        # around the short launches of the small members an event pair also
        # holds the host's submission gap
        for p, (acoustic, state) in enumerate(ens.parts):
            lib.fv3_timing_hold(int(p != 0))
            acoustic(state)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cnt, kern_ms, kern_bytes = _timing(lib, roofline_kernel)
    lib.fv3_timing_arm(None)
    graph_wall = graph_seq_wall = None
    if graph:
        # every batch in one HIP graph, one launch per call (eager, capture,
        # replays): the small batches as parallel branches beside the large
        # one (the value), and in sequence
        walls = []
        for concurrent in (True, False):
            gens = UnevenEnsemble(parts, use_graph=True, concurrent=concurrent)
            for _ in range(3):
                gens()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                gens()
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t0)
            del gens
        graph_wall, graph_seq_wall = walls
    nz = parts[0][1].pt.shape[-1] - 1
    finite = all(bool(torch.isfinite(st.pt[:, 3:-4, 3:-4, :nz]).all().item()) for _, st in parts)
    return wall, (graph_wall, graph_seq_wall), cnt, kern_ms, kern_bytes, finite


def _ensemble_mixed_errors(cfg, n=96, pt_sd=0.0, uv_sd=0.0, seed=96000):
    """Errors of the mixed form against the FP64 form after one call on one
    member at C``n`` (the same initial state and perturbation for both)."""
    import torch

    from fv3core_amd.ensemble import ensemble_acoustic
    from fv3core_amd.workloads import cube_state

    init = cube_state(n, cfg.npz, seed=0)
    built = {}

    def make(mixed):
        acoustic, state = ensemble_acoustic(init, 1, cfg, seed=seed, pt_sd=pt_sd, uv_sd=uv_sd,
                                            mixed_precision=mixed)
        built[mixed] = state
        return acoustic

    plain, mixed = make(False), make(True)
    sp, sm = built[False], built[True]  # one initial state, FP64 and mixed storages
    plain(sp)
    mixed(sm)
    torch.cuda.synchronize()
    nz = cfg.npz
    c = (slice(None), slice(3, n + 3), slice(3, n + 3))

    def rel(a, b):
        a, b = a.double(), b.double()
        scale = float(torch.maximum(a.abs().max(), b.abs().max()))
        den = torch.clamp(a.abs() + b.abs(), min=1e-6 * scale + 1e-300)
        return float((2.0 * (a - b).abs() / den).max())

    errs = {k: rel(getattr(sm, k)[c + (slice(0, nz),)], getattr(sp, k)[c + (slice(0, nz),)])
            for k in MIXED_ERROR_FIELDS}
    errs["ps"] = rel(sm.pe[c + (nz,)], sp.pe[c + (nz,)])
    errs["scaled"] = _scaled_errors(sm, sp, n, nz)
    return errs


def bench_ensemble(steps: int, warmup: int, graph: bool, roofline_kernel: str,
                   adjoint: bool = True):
    """Batched ensemble: 32 members of six C48 tiles x 79 levels (the C96
    workload's initial state at C48, each member with seeded N(0, 0.5 K) pt
    and N(0, 1 m/s) u, v perturbations) in two batches of 96 tiles, plus
    four uneven members at C24, C64, C96 and C128, each its own batch.  A step is
    one AcousticDynamics call (n_split = 6, dt = 225 s) of every batch;
    cell-level updates per step = sum over members of 6 n^2 79 6.

    Both forms are timed: plain FP64 and the mixed form (MIXED_FORM).
    ``value`` is the mixed form's rate, with the FP64 rate, their ratio, the
    mixed form's per-field error against the FP64 form (one call on the C96
    member) and, unless ``adjoint`` is off, the adjoint/forward time ratio of
    the transport + remap adjoint workload beside it."""
    import gc

    import torch

    from fv3core_amd._config import DynamicalCoreConfig

    nz, n_split = 79, 6
    cfg = DynamicalCoreConfig(npz=nz, n_split=n_split, k_split=1, nord=3, dt_atmos=225)
    # the gate is read on the unperturbed C96 member; a perturbed C48 member
    # (grid-scale noise, the worst case for the limiters) is reported beside it
    errs = _ensemble_mixed_errors(cfg)
    errs_pert = _ensemble_mixed_errors(cfg, n=48, pt_sd=0.5, uv_sd=1.0, seed=48000)
    results = {}
    cells = 0
    for form in ("fp64", "mixed"):
        parts, cells = _ensemble_parts(cfg, form == "mixed")
        results[form] = _time_ensemble(parts, steps, warmup, roofline_kernel, graph)
        del parts
        gc.collect()
        torch.cuda.empty_cache()
    wall64, (graph64, graph64_seq), _, _, _, finite64 = results["fp64"]
    wall, (graph_wall, graph_seq), cnt, kern_ms, kern_bytes, finite = results["mixed"]
    used = graph_wall if graph_wall else wall
    used64 = graph64 if graph64 else wall64
    avg_s = kern_ms / max(cnt, 1) * 1e-3
    achieved = kern_bytes / max(kern_ms * 1e-3, 1e-30) / 1e9
    step_bytes = cells * STEP_BYTES_PER_CELL_FP64
    out = {
        "value": steps * cells / used,
        "ms_per_step": used / steps * 1e3,
        "eager_ms_per_step": wall / steps * 1e3,
        "graph_ms_per_step": graph_wall / steps * 1e3 if graph_wall else None,
        "graph_sequential_ms_per_step": graph_seq / steps * 1e3 if graph_seq else None,
        "fp64_graph_sequential_ms_per_step": graph64_seq / steps * 1e3 if graph64_seq else None,
        "fp64_value": steps * cells / used64,
        "fp64_ms_per_step": used64 / steps * 1e3,
        "fp64_eager_ms_per_step": wall64 / steps * 1e3,
        "mixed_over_fp64": used64 / used,
        "mixed_vs_fp64": errs,
        "mixed_gate_ok": max(errs[k] for k in ("ps", "pt", "delp")) <= 1e-5,
        "mixed_vs_fp64_perturbed_c48": errs_pert,
        "mixed_gate_ok_perturbed_c48": max(errs_pert[k] for k in ("ps", "pt", "delp")) <= 1e-5,
        "step_roofline": {
            "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBPS,
            "bytes_per_cell_substep": STEP_BYTES_PER_CELL_FP64,
            "achieved": step_bytes / (used / steps) / 1e9,
            "frac": step_bytes / (used / steps) / 1e9 / HBM_PEAK_GBPS,
            "fp64_achieved": step_bytes / (used64 / steps) / 1e9,
        },
        "config": {
            "workload": "batched ensemble: 32 members x 6 x C48 x 79 in two batches of 16 + uneven "
            "members C24/C64/C96/C128 (one cube each), AcousticDynamics n_split=6, dt=225 s, "
            "perturbed jet states; mixed form: " + MIXED_FORM +
            (", every batch of a call in one HIP graph (one launch per call), the five "
             "other batches as parallel branches beside the first" if graph
             else ""),
            "members": 36, "tiles": 6 * 36, "nz": nz, "n_split": n_split,
            "cells_per_step": cells,
            "state_finite_after_timed_steps": bool(finite and finite64),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": _profiled_traffic(roofline_kernel, "ensemble"),
            "kernel": roofline_kernel,
            "kernel_launches_timed": cnt,
            "kernel_launches_of": "the first 16-member C48 batch (mixed form, eager timed steps)",
            "kernel_avg_us": avg_s * 1e6,
            "algorithmic_bytes_per_launch": kern_bytes / max(cnt, 1),
        },
        "dtype": "f64+f32",
    }
    if adjoint:
        adj = bench_adjoint(3, 1)
        out["adjoint_over_forward"] = adj["adjoint_over_forward"]
        out["adjoint_workload"] = adj["config"]["workload"]
    return out


def bench_adjoint(steps: int, warmup: int):
    """Adjoint of PPM transport + vertical remap: one doubly periodic C96 tile
    96x96x79; each of 8 steps is the flux-form fv_tp_2d (hord 8) transport of
    q (dt fixed by the Courant numbers of the periodic workload) followed by
    map_scalar (kord 9, iv 1) from perturbed Lagrangian to hybrid interfaces;
    J = 0.5 ||q_T - q_obs||^2 with q_obs the run from a seeded 1e-2
    perturbation of q0.  A step is one 8-step forward run; the
    reverse sweep (per step: re-linearisation from the saved input, then the
    adjoint kernels) is timed separately and reported as its ratio to the
    forward run."""
    import torch

    from fv3core_amd import _lib
    from fv3core_amd.adjoint import PeriodicTransport
    from fv3core_amd.utils.grid import to_storage
    from fv3core_amd.utils.stencil import GridIndexing
    from fv3core_amd.workloads import remap_inputs, transport_inputs

    n, nz, hord, nsteps = 96, 79, 8, 8
    d = transport_inputs(n, nz, workload=4)
    r = remap_inputs(n, nz, workload=4)
    gi = GridIndexing((n, n, nz), 3, False, False, False, False)
    dev = {k: to_storage(v) for k, v in d.items()}
    remap = dict(pe1=to_storage(r["pe1"]), pe2=to_storage(r["pe2"]), mode=1, qmin=0.0)
    model = PeriodicTransport(gi, dev, hord, remap=remap)
    q0 = torch.as_tensor(d["q"]).permute(2, 1, 0).contiguous().cuda()
    g = torch.Generator(device="cpu").manual_seed(4)
    with torch.no_grad():
        q_obs = model.run(q0 + 1e-2 * torch.randn(q0.shape, generator=g,
                                                  dtype=q0.dtype).cuda(), nsteps)

    def forward_run():
        with torch.no_grad():
            return model.run(q0, nsteps)

    def recorded():
        x = q0.detach().clone().requires_grad_(True)
        r = model.interior(model.run(x, nsteps)) - model.interior(q_obs)
        return x, 0.5 * (r * r).sum()

    for _ in range(warmup):
        forward_run()
        x, cost = recorded()
        torch.autograd.grad(cost, x)
    torch.cuda.synchronize()
    lib = _lib.lib()
    lib.fv3_timing_arm(b"fvtp2d_kernel")
    t0 = time.perf_counter()
    for _ in range(steps):
        forward_run()
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / steps
    cnt, kern_ms, kern_bytes = _timing(lib, "fvtp2d_kernel")
    lib.fv3_timing_arm(None)
    rev = []
    for _ in range(steps):
        x, cost = recorded()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        torch.autograd.grad(cost, x)
        torch.cuda.synchronize()
        rev.append(time.perf_counter() - t1)
    rev_s = float(np.median(rev))
    cells = n * n * nz * nsteps
    avg_s = kern_ms / max(cnt, 1) * 1e-3
    achieved = kern_bytes / max(kern_ms * 1e-3, 1e-30) / 1e9
    return {
        "value": cells / fwd,
        "ms_per_step": fwd * 1e3,
        "adjoint_over_forward": rev_s / fwd,
        "reverse_ms_per_step": rev_s * 1e3,
        "config": {
            "workload": f"adjoint of fv_tp_2d (hord={hord}) + map_scalar remap (kord 9): one "
            f"doubly periodic tile {n}x{n}x{nz}, {nsteps} steps of flux-form transport then "
            "remap from perturbed Lagrangian to hybrid interfaces, J = 0.5||q_T - q_obs||^2, "
            "reverse sweep re-linearising each step from its saved input",
            "tiles": 1, "n": n, "nz": nz, "hord": hord, "transport_steps": nsteps,
            "cells_per_step": cells,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None,
            "kernel": "fvtp2d_kernel",
            "kernel_launches_timed": cnt,
            "kernel_avg_us": avg_s * 1e6,
            "algorithmic_bytes_per_launch": kern_bytes / max(cnt, 1),
        },
        "dtype": "f64",
    }


def main():
    args = _parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (all six tiles and ensemble "
                                   "members share one MI355X); --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "dyncore":
        steps = 20 if args.steps is None else args.steps
        warmup = 3 if args.warmup is None else args.warmup
        cpu = None if args.no_cpu_baseline else cpu_baseline_dyncore(args.cpu_seconds)
        res = bench_dyncore(steps, warmup, args.roofline_kernel, args.graph,
                            args.copy_calibration, args.mixed)
    elif args.workload == "fv_dynamics":
        steps = 5 if args.steps is None else args.steps
        warmup = 1 if args.warmup is None else args.warmup
        res = bench_fv_dynamics(steps, warmup, args.roofline_kernel, args.mixed)
        cpu = None
    elif args.workload == "ensemble":
        steps = 3 if args.steps is None else args.steps
        warmup = 1 if args.warmup is None else args.warmup
        cpu = None if args.no_cpu_baseline else cpu_baseline_dyncore(args.cpu_seconds)
        res = bench_ensemble(steps, warmup, not args.no_graph, args.roofline_kernel,
                             adjoint=not args.no_adjoint)
        if args.copy_calibration:
            res["copy_calibration_bytes"] = _copy_calibration()
    elif args.workload == "adjoint":
        steps = 10 if args.steps is None else args.steps
        warmup = 2 if args.warmup is None else args.warmup
        res = bench_adjoint(steps, warmup)
        cpu = None
    else:
        steps = 200 if args.steps is None else args.steps
        warmup = 20 if args.warmup is None else args.warmup
        res = bench_fvtp2d(steps, warmup, args.dtype, args.copy_calibration)
        cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_seconds)
    args.steps, args.warmup = steps, warmup
    line = {
        "metric": "cell-level-updates/s",
        "value": res["value"],
        "unit": "cell-level-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": res["dtype"],
        "data": "synthetic",
        "config": res["config"],
        "roofline": res["roofline"],
        "cpu_baseline": cpu,
    }
    for extra in ("eager_ms_per_step", "graph_ms_per_step", "fp64_value", "fp64_ms_per_step",
                  "mixed_over_fp64", "step_roofline", "adjoint_over_forward",
                  "adjoint_workload", "reverse_ms_per_step", "mixed_vs_fp64", "mixed_gate_ok",
                  "mixed_vs_fp64_perturbed_c48", "mixed_gate_ok_perturbed_c48",
                  "fp64_eager_ms_per_step", "graph_sequential_ms_per_step",
                  "fp64_graph_sequential_ms_per_step"):
        if res.get(extra) is not None:
            line[extra] = res[extra]
    if res.get("copy_calibration_bytes"):
        line["copy_calibration_bytes"] = res["copy_calibration_bytes"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
