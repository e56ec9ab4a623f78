"""DynamicalCore (fv3core/stencils/fv_dynamics.py:193-557) on the GPU.

One call of ``step_dynamics`` runs the reference's sequence on the batched
six-tile storages: the preamble (fv_setup + pt_adjust, one fused kernel),
then k_split times the copy delp -> dp1, AcousticDynamics, TracerAdvection
and LagrangianToEulerian, and after the last remap set_omega, the omega
halo update and HyperdiffusionDamping of omga.

The step ends with the reference's wrap-up (fv_dynamics.py:136-169):
AdjustNegativeTracerMixingRatio (neg_adj3) corrects negative qvapor,
qliquid, qrain, qsnow, qice, qgraupel and qcld, moving water between species
and adjusting pt, and CubedToLatLon recomputes ua and va from the D-grid
winds (after a vector halo update of u and v for the 4th-order form).
``skip_wrapup=True`` (no counterpart in the reference) leaves those nine
fields as the last remap leaves them.
"""
from __future__ import annotations

import ctypes
import types
from typing import Dict, Mapping

import numpy as np
import torch

from .. import _lib
from .. import mixed as _mixed
from ..constants import KAPPA, ZVIR
from ..utils import gt4py_utils as utils
from ..utils.grid import DampingCoefficients, GridData
from ..utils.stencil import StencilFactory
from .c2l_ord import CubedToLatLon
from .del2cubed import HyperdiffusionDamping
from .dyn_core import AcousticDynamics
from .fvtp2d import FiniteVolumeTransport
from .neg_adj3 import AdjustNegativeTracerMixingRatio
from .remapping import LagrangianToEulerian
from .tracer_2d_1l import TracerAdvection

# tracers handled by the step (fv_dynamics.py:25-29)
NQ = 8

ARG_NAMES = (
    "qvapor", "qliquid", "qrain", "qsnow", "qice", "qgraupel", "qo3mr", "qsgs_tke", "qcld",
    "pt", "delp", "delz", "peln", "u", "v", "w", "ua", "va", "uc", "vc", "q_con", "pe", "phis",
    "pk", "pkz", "ps", "omga", "ak", "bk", "mfxd", "mfyd", "cxd", "cyd", "diss_estd",
)

_OP_COPY = 3


class NullTimer:
    """Stand-in for fv3gfs.util.NullTimer: ``clock(name)`` is a no-op context."""

    class _Clock:
        def __enter__(self):
            return self

        def __exit__(self, *exc):
            return False

    def clock(self, name):
        return self._Clock()


def init_pfull(ak, bk, p_ref: float):
    """init_pfull (fv_dynamics.py:42-51): log-mean pressure of each layer."""
    ph1 = ak[:-1] + bk[:-1] * p_ref
    ph2 = ak[1:] + bk[1:] * p_ref
    return (ph2 - ph1) / np.log(ph2 / ph1)


def _storage(q):
    return q if isinstance(q, torch.Tensor) else q.storage


def _host(a):
    a = _storage(a) if not isinstance(a, np.ndarray) else a
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float64)


class DynamicalCore:
    """Fortran fv_dynamics."""

    def __init__(self, comm, grid_data: GridData, stencil_factory: StencilFactory,
                 damping_coefficients: DampingCoefficients, config, ak, bk, phis,
                 use_graph: bool = False, skip_wrapup: bool = False,
                 mixed_precision: bool = False):
        """``mixed_precision`` (no counterpart in the reference, which is FP64
        only): run the acoustic substeps in the mixed form (FP32 bulk
        storages, FP64 solvers, pressures and heights); the step's state stays
        FP64: before each k_split iteration's acoustic call its bulk fields
        are rounded into FP32 copies the component owns, and after it widened
        back, so the tracer transport and the remap run in FP64 on the
        acoustic step's results.
        ``use_graph`` (no counterpart in the reference): capture a whole
        step (every k_split iteration of the acoustic substeps, tracer
        transport and remap, and the wrap-up) in one HIP graph and replay it
        on later steps with the same storages and scalars; the first step for
        a new set runs eagerly, the second captures and replays.
        ``skip_wrapup`` (no counterpart in the reference): end the step after
        the last remap, without neg_adj3 and CubedToLatLon."""
        gi = stencil_factory.grid_indexing
        self._skip_wrapup = bool(skip_wrapup)
        self._use_graph = bool(use_graph)
        self._mixed = bool(mixed_precision)
        self._bulk32: Dict[str, torch.Tensor] = {}
        self._graphs = {}
        self._seen_keys = set()
        if not config.moist_phys:
            raise AssertionError("fvsetup is only implemented for moist_phys=true")
        if config.nwat != 6:
            raise AssertionError("Only nwat=6 has been implemented and tested")
        if config.hydrostatic:
            raise NotImplementedError("Hydrostatic is not implemented")
        if config.inline_q or NQ == 0:
            raise NotImplementedError("tracer_2d not implemented, turn on z_tracer")
        self.comm = comm
        self.grid_data = grid_data
        self.grid_indexing = gi
        self._da_min = float(damping_coefficients.da_min)
        self.config = config
        self._ak_host = _host(ak)
        self._bk_host = _host(bk)
        npz = gi.domain[2]
        self._pfull = init_pfull(self._ak_host[:npz + 1], self._bk_host[:npz + 1], config.p_ref)
        self._phis = _storage(phis)
        dev = self._phis.device
        self._ak = torch.as_tensor(self._ak_host, device=dev)
        self._bk = torch.as_tensor(self._bk_host, device=dev)
        tracer_transport = FiniteVolumeTransport(
            stencil_factory=stencil_factory, grid_data=grid_data,
            damping_coefficients=damping_coefficients, grid_type=config.grid_type,
            hord=config.hord_tr)
        self.tracer_advection = TracerAdvection(stencil_factory, tracer_transport, comm, NQ)
        self.acoustic_dynamics = AcousticDynamics(
            comm, stencil_factory, grid_data, damping_coefficients, config.grid_type, False,
            False, config.acoustic_dynamics, self._ak_host, self._bk_host, self._pfull,
            self._phis, mixed_precision=self._mixed)
        self._hyperdiffusion = HyperdiffusionDamping(stencil_factory, damping_coefficients,
                                                     grid_data.rarea, config.nf_omega)
        area_64 = grid_data.area_64 if grid_data.has("area_64") else grid_data.area
        self._lagrangian_to_eulerian_obj = LagrangianToEulerian(
            stencil_factory, config.remapping, area_64, NQ, self._pfull)
        self._omega_halo_updater = comm.get_scalar_halo_updater()
        if not self._skip_wrapup:
            self._cubed_to_latlon = CubedToLatLon(stencil_factory, grid_data,
                                                  order=config.c2l_ord)
            self._adjust_tracer_mixing_ratio = AdjustNegativeTracerMixingRatio(
                stencil_factory, config.check_negative, config.hydrostatic)
        self._temporaries = None

    # ------------------------------------------------------------ helpers
    def _acoustic(self, state):
        """AcousticDynamics on the FP64 state, or on its mixed-precision form."""
        if not self._mixed:
            self.acoustic_dynamics(state)
            return
        ms = types.SimpleNamespace()
        for k, v in vars(state).items():
            if k in _mixed.BULK_FIELDS and isinstance(v, torch.Tensor):
                t = self._bulk32.get(k)
                if t is None or t.shape != v.shape or t.stride() != v.stride():
                    t = torch.empty_like(v, dtype=torch.float32)
                    self._bulk32[k] = t
                _mixed.cast_into(v, t)
                setattr(ms, k, t)
            else:
                setattr(ms, k, v)
        self.acoustic_dynamics(ms)
        for k, t in self._bulk32.items():
            if isinstance(state.__dict__.get(k), torch.Tensor):
                _mixed.cast_into(t, state.__dict__[k])

    def _make_temporaries(self, like: torch.Tensor):
        """fvdyn_temporaries (fv_dynamics.py:172-190)."""
        shape3 = tuple(like.shape)
        bd = 1 if like.dim() == 4 else 0
        shape2 = shape3[:-1]
        tmps = {}
        for name in ("te_2d", "te0_2d", "wsd"):
            tmps[name] = utils.zeros(shape2, dtype=like.dtype, device=like.device, batch_dims=bd)
        for name in ("cappa", "dp1", "cvm"):
            tmps[name] = utils.zeros(shape3, dtype=like.dtype, device=like.device, batch_dims=bd)
        tmps["gz"] = torch.zeros(shape3[-1], dtype=like.dtype, device=like.device)
        return tmps

    def _fv_setup(self, state):
        gi = self.grid_indexing
        nz = gi.domain[2]
        q6 = [state.qvapor, state.qliquid, state.qrain, state.qice, state.qsnow, state.qgraupel]
        fields = [state.delp, state.delz, state.pt, state.q_con, state.cvm, state.pkz,
                  state.cappa, state.dp1]
        d = _lib.make_domain(gi, state.pt, nz, others3=q6 + fields)
        ptrs = (ctypes.c_void_p * 6)(*[q.data_ptr() for q in q6])
        _lib.check(_lib.lib().fv3_fv_setup_f64(
            ctypes.byref(d), ptrs, _lib.ptr(state.delp), _lib.ptr(state.delz), _lib.ptr(state.pt),
            _lib.ptr(state.q_con), _lib.ptr(state.cvm), _lib.ptr(state.pkz),
            _lib.ptr(state.cappa), _lib.ptr(state.dp1), ctypes.c_void_p(_lib.stream_handle())))

    def _copy(self, src, dst):
        nk = min(src.shape[-1], dst.shape[-1])
        d = _lib.make_domain(self.grid_indexing, dst, nk, others3=(src,))
        _lib.check(_lib.lib().fv3_dyncore_column_f64(
            ctypes.byref(d), ctypes.c_int32(_OP_COPY), ctypes.c_double(0.0), _lib.ptr(src), None,
            _lib.ptr(dst), ctypes.c_void_p(_lib.stream_handle())))

    def _set_omega(self, state):
        nz = self.grid_indexing.domain[2]
        d = _lib.make_domain(self.grid_indexing, state.omga, nz,
                             others3=(state.delp, state.delz, state.w))
        _lib.check(_lib.lib().fv3_set_omega_f64(
            ctypes.byref(d), _lib.ptr(state.delp), _lib.ptr(state.delz), _lib.ptr(state.w),
            _lib.ptr(state.omga), ctypes.c_void_p(_lib.stream_handle())))

    # ---------------------------------------------------------------- step
    def step_dynamics(self, state: Mapping, conserve_total_energy: bool,
                      do_adiabatic_init: bool, timestep: float, ptop, n_split: int, ks: int,
                      timer=None):
        """Step the model state forward by one timestep (all fields in place).

        ``state``: mapping (or namespace) of the reference's ArgSpec fields
        as device storages or objects with a ``.storage``.
        """
        timer = timer if timer is not None else NullTimer()
        get = state.__getitem__ if isinstance(state, Mapping) else \
            (lambda n: getattr(state, n))
        ns = types.SimpleNamespace()
        for name in ARG_NAMES:
            if name in ("ak", "bk"):
                continue
            ns.__dict__[name] = _storage(get(name))
        _lib.require_device(*[ns.__dict__[n] for n in ARG_NAMES if n not in ("ak", "bk")])
        ns.__dict__.update({
            "consv_te": conserve_total_energy, "bdt": timestep,
            "mdt": timestep / self.config.k_split, "do_adiabatic_init": do_adiabatic_init,
            "ptop": ptop, "n_split": n_split, "k_split": self.config.k_split, "ks": ks,
        })
        if not self._use_graph:
            self._compute(ns, timer)
            return
        key = (tuple((n, ns.__dict__[n].data_ptr()) for n in ARG_NAMES if n not in ("ak", "bk")),
               float(timestep), float(ptop), int(n_split), int(ks), bool(do_adiabatic_init),
               float(conserve_total_energy))
        graph = self._graphs.get(key)
        if graph is not None:
            graph.replay()
            return
        if key not in self._seen_keys:
            # first step for these storages: eager, fills every cache
            self._seen_keys.add(key)
            self._compute(ns, timer)
            return
        graph = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(graph):
            self._compute(ns, NullTimer())
        torch.cuda.synchronize()
        self._graphs[key] = graph
        graph.replay()

    def _compute(self, state, timer):
        if self._temporaries is None:
            self._temporaries = self._make_temporaries(state.pt)
        state.__dict__.update(self._temporaries)
        tracers = {name: state.__dict__[name] for name in utils.tracer_variables[0:NQ]}
        state.ak = self._ak
        state.bk = self._bk
        config = self.config
        # compute_preamble (fv_dynamics.py:54-106)
        if state.consv_te > 0 and not state.do_adiabatic_init:
            raise NotImplementedError(
                "compute total energy is not implemented, it needs an allReduce")
        if (not config.rf_fast) and config.tau != 0:
            raise NotImplementedError("Rayleigh_Super, called when rf_fast=False and tau !=0")
        if config.adiabatic and config.kord_tm > 0:
            raise NotImplementedError(
                "unimplemented namelist options adiabatic with positive kord_tm")
        self._fv_setup(state)
        for n_map in range(state.k_split):
            state.n_map = n_map + 1
            last_step = n_map == state.k_split - 1
            self._copy(state.delp, state.dp1)
            with timer.clock("DynCore"):
                self._acoustic(state)
            if config.z_tracer:
                with timer.clock("TracerAdvection"):
                    self.tracer_advection(tracers, state.dp1, state.mfxd, state.mfyd, state.cxd,
                                          state.cyd, state.mdt)
            if self.grid_indexing.domain[2] > 4:
                with timer.clock("Remapping"):
                    self._lagrangian_to_eulerian_obj(
                        tracers, state.pt, state.delp, state.delz, state.peln, state.u, state.v,
                        state.w, state.ua, state.va, state.cappa, state.q_con, state.qcld,
                        state.pkz, state.pk, state.pe, state.phis, state.te0_2d, state.ps,
                        state.wsd, state.omga, self._ak, self._bk, self._pfull, state.dp1,
                        state.ptop, KAPPA, ZVIR, last_step, state.consv_te,
                        state.bdt / state.k_split, state.bdt, state.do_adiabatic_init, NQ)
                if last_step:
                    # post_remap (fv_dynamics.py:109-133)
                    self._set_omega(state)
                    if config.nf_omega > 0:
                        self._omega_halo_updater.update([state.omga])
                        self._hyperdiffusion(state.omga, 0.18 * self._da_min)
        if not self._skip_wrapup:
            # wrapup (fv_dynamics.py:136-169)
            self._adjust_tracer_mixing_ratio(
                state.qvapor, state.qliquid, state.qrain, state.qsnow, state.qice,
                state.qgraupel, state.qcld, state.pt, state.delp, state.delz, state.peln)
            self._cubed_to_latlon(state.u, state.v, state.ua, state.va, self.comm)
