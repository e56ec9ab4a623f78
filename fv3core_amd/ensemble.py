"""Ensembles of cubed-sphere states in one batch.

A member is a whole six-tile cube.  Members are stacked along the storages'
tile-batch axis (member m owns tiles 6m .. 6m+5), and every kernel of the
path treats tiles independently while the halo update applies its six-tile
map to each member, so one call of a component (AcousticDynamics, d_sw, the
Riemann solvers, ...) advances every member with one launch per kernel.
Members of different resolutions cannot share a storage; ``UnevenEnsemble``
keeps one AcousticDynamics per resolution and runs them back to back,
optionally all captured in one HIP graph launched once per call.

The reference runs one cube per MPI job; ensembles are outside it.
"""
from __future__ import annotations

import types
from typing import Dict, List, Sequence

import numpy as np
import torch

from ._config import DynamicalCoreConfig
from .utils import gt4py_utils as utils

_STATE_FIELDS_2D = ("phis", "wsd")


def _dev(a: np.ndarray, batch_dims, device=None) -> torch.Tensor:
    t = utils.empty(a.shape, batch_dims=batch_dims, device=device)
    t.copy_(torch.as_tensor(a))
    return t


def replicate(t: torch.Tensor, members: int) -> torch.Tensor:
    """A (6, ...) device storage repeated ``members`` times along the batch
    axis, keeping the i-fastest layout (built with the storage allocator)."""
    shape = (t.shape[0] * members,) + tuple(t.shape[1:])
    out = utils.empty(shape, dtype=t.dtype, device=t.device,
                      batch_dims=1 if t.dim() == 3 else None)
    for m in range(members):
        out[t.shape[0] * m:t.shape[0] * (m + 1)].copy_(t)
    return out


def ensemble_acoustic(init: Dict, members: int, cfg: DynamicalCoreConfig, seed: int = 0,
                      pt_sd: float = 0.5, uv_sd: float = 1.0, use_graph: bool = False,
                      mixed_precision: bool = False):
    """(AcousticDynamics, state) for ``members`` perturbed copies of a cube state
    (with ``mixed_precision``, the mixed form and its state: bulk fields FP32).

    ``init`` is what :func:`fv3core_amd.workloads.cube_state` returns.  Member
    m gets pt multiplied by (1 + N(0, pt_sd) / 288) (about pt_sd kelvin near
    the surface) and N(0, uv_sd) m/s added to u and v, from a generator
    seeded with ``seed + m``; halos are then refreshed by the on-device
    halo update.
    """
    from .halo import CubedSphereCommunicator
    from .stencils.dyn_core import AcousticDynamics
    from .stencils.fv_dynamics import init_pfull
    from .utils.grid import DampingCoefficients, GridData
    from .utils.stencil import GridIndexing, StencilConfig, StencilFactory

    st0, m = init["state"], init["metric"]
    n = st0["delp"].shape[1] - 7
    nz = cfg.npz
    gi = GridIndexing((n, n, nz), 3, True, True, True, True)
    factory = StencilFactory(StencilConfig(), gi)
    grid = GridData({k: replicate(_dev(v, 1), members) for k, v in m.items()
                     if isinstance(v, np.ndarray) and v.ndim == 3})
    damp = DampingCoefficients(grid.del6_u, grid.del6_v, m["da_min"], m["da_min_c"])
    comm = CubedSphereCommunicator(gi)
    phis = replicate(_dev(init["phis"], 1), members)
    state = types.SimpleNamespace(**{k: replicate(_dev(v, None), members)
                                     for k, v in st0.items()})
    state.phis = phis
    state.wsd = replicate(_dev(init["wsd"], 1), members)
    state.ptop, state.ks, state.mdt, state.n_map = init["ptop"], 0, float(cfg.dt_atmos), 1
    for mm in range(members):
        g = torch.Generator(device="cpu").manual_seed(seed + mm)
        sl = slice(6 * mm, 6 * mm + 6)
        shp = tuple(state.pt[sl].shape)
        dpt = torch.randn(shp, generator=g, dtype=torch.float64).to(state.pt.device)
        du = torch.randn(shp, generator=g, dtype=torch.float64).to(state.pt.device)
        dv = torch.randn(shp, generator=g, dtype=torch.float64).to(state.pt.device)
        state.pt[sl].mul_(1.0 + pt_sd * dpt / 288.0)
        state.u[sl].add_(uv_sd * du)
        state.v[sl].add_(uv_sd * dv)
    comm.get_scalar_halo_updater().update([state.pt])
    comm.get_vector_halo_updater().update([state.u], [state.v])
    acoustic = AcousticDynamics(comm, factory, grid, damp, 0, False, False, cfg.acoustic_dynamics,
                                init["ak"], init["bk"], init_pfull(init["ak"], init["bk"], 1.0e5),
                                phis, use_graph=use_graph, mixed_precision=mixed_precision)
    if mixed_precision:
        from .mixed import mixed_state

        state = mixed_state(state)
    return acoustic, state


def member_slice(state, m: int):
    """A six-tile state holding copies of member m's fields."""
    out = types.SimpleNamespace()
    for k, v in vars(state).items():
        if isinstance(v, torch.Tensor):
            s = v[6 * m:6 * m + 6]
            t = utils.empty(tuple(s.shape), dtype=s.dtype, device=s.device,
                            batch_dims=1 if s.dim() == 3 else None)
            t.copy_(s)
            setattr(out, k, t)
        else:
            setattr(out, k, v)
    return out


class UnevenEnsemble:
    """Members of several resolutions: one batched AcousticDynamics per
    resolution, called by :meth:`__call__`.

    With ``use_graph=True`` every resolution's call is captured into one HIP
    graph, so that a call of the whole ensemble (all its acoustic substeps,
    halo exchanges and components, for every batch) is one graph launch: the
    first call runs eagerly (it fills the component caches), the second
    captures and replays, later calls replay.  The parts' own graph option
    must be off; the storages must stay the same between calls.

    With ``concurrent=True`` the first part runs on the caller's stream and
    every other part on a stream of its own (default priority), forked from
    and joined back into the caller's stream, so that the small batches,
    whose launches cannot fill the GPU, run beside the large one; in a
    captured graph they are parallel branches.  The parts share no storage,
    so the results are those of the sequential call bit for bit."""

    def __init__(self, parts: Sequence, use_graph: bool = False, concurrent: bool = False):
        self.parts: List = list(parts)  # (acoustic, state) pairs
        self._use_graph = bool(use_graph)
        self._concurrent = bool(concurrent)
        self._streams = None
        self._calls = 0
        self.graph = None

    def _run(self):
        if not self._concurrent or len(self.parts) < 2:
            for acoustic, state in self.parts:
                acoustic(state)
            return
        main = torch.cuda.current_stream()
        if self._streams is None:
            self._streams = [torch.cuda.Stream(device=main.device) for _ in self.parts[1:]]
        for s in self._streams:
            s.wait_stream(main)
        acoustic, state = self.parts[0]
        acoustic(state)
        for s, (acoustic, state) in zip(self._streams, self.parts[1:]):
            with torch.cuda.stream(s):
                acoustic(state)
        for s in self._streams:
            main.wait_stream(s)

    def __call__(self):
        if not self._use_graph:
            self._run()
            return
        if any(getattr(a, "_use_graph", False) for a, _ in self.parts):
            raise ValueError("UnevenEnsemble(use_graph=True) captures the parts itself: "
                             "build them with use_graph=False")
        self._calls += 1
        if self.graph is not None:
            self.graph.replay()
            return
        if self._calls == 1:
            self._run()
            return
        graph = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(graph):
            self._run()
        torch.cuda.synchronize()
        self.graph = graph
        graph.replay()
