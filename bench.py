"""Benchmark of the FV3 hot path on one MI355X.

Prints one JSON line.  A step is one pass of the hot path over one batch of
synthetic input with every input already resident in HBM.

Workload "fvtp2d": FiniteVolumeTransport (fv_tp_2d, hord=8, area fluxes) on
6 tiles of 96x96x79 (the C96 cube shape; each tile doubly periodic, inputs
built like the single-tile transport workload), FP64.  One step is one
call = one fused kernel launch.  Cell-level updates per step =
6 * 96 * 96 * 79.

The roofline object prices the fused kernel with its algorithmic bytes
(q, crx, cry, x_area_flux, y_area_flux read, q_x_flux, q_y_flux written:
7 fields x 8 B = 56 B per cell) over its average launch time measured with
HIP events on the launch stream.  The CPU baseline times the numpy oracle on
one 48x48x79 tile (single core).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def _parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--workload", default="fvtp2d", choices=["fvtp2d"])
    p.add_argument("--dtype", default="f64", choices=["f64", "f32"])
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


def main():
    args = _parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (all six tiles and ensemble "
                                   "members share one MI355X); --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    res = bench_fvtp2d(args.steps, args.warmup, args.dtype, args.copy_calibration)
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_seconds)
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
    if "copy_calibration_bytes" in res:
        line["copy_calibration_bytes"] = res["copy_calibration_bytes"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
