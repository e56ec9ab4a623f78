"""AcousticDynamics (fv3core/stencils/dyn_core.py:210-761) on the GPU.

The host loop over the n_split acoustic substeps follows the reference's
order of components and halo exchanges exactly; every component is a
gfx950 kernel sequence of this package, the halo exchanges are on-device
gathers (fv3core_amd.halo), and nothing synchronises the host with the GPU
inside a call.  The state is a namespace of storages as in the reference;
the component's temporaries are added to it on the first call, as
``state.__dict__.update(self._temporaries)`` does there.

With ``mixed_precision=True`` the step runs the mixed form of
fv3core_amd.mixed: the state's bulk fields (``mixed.BULK_FIELDS``) are FP32
storages advanced by the FP32 instantiations of c_sw, d_sw, Rayleigh
damping, del2cubed and the halo exchange on FP32 metric terms, while the
Riemann solvers, the height updates, the pressure gradients and the pk / pe
halos run in FP64 on FP64 heights and pressures (reading the FP32 fields
directly or through FP64 copies).

With ``use_graph=True`` the whole call (every substep, halo exchange and
component) is captured once into a HIP graph and replayed on later calls
that pass the same storages and scalars; the first call for a new set runs
eagerly (it fills the component caches), the second captures and replays.
"""
from __future__ import annotations

import ctypes
from typing import Dict

import numpy as np
import torch

from .. import _lib
from .. import mixed as _mixed
from .._config import AcousticDynamicsConfig
from ..constants import GRAV, KAPPA
from ..utils import gt4py_utils as utils
from ..utils.grid import DampingCoefficients, GridData
from ..utils.stencil import StencilFactory
from .c_sw import CGridShallowWaterDynamics
from .d_sw import DGridShallowWaterLagrangianDynamics, get_column_namelist
from .del2cubed import HyperdiffusionDamping
from .nh_p_grad import NonHydrostaticPressureGradient
from .pk3_halo import PK3Halo, edge_pe
from .ray_fast import RayleighDamping
from .riem_solver3 import RiemannSolver3
from .riem_solver_c import RiemannSolverC
from .temperature_adjust import compute_pkz_tempadjust
from .updatedzc import UpdateGeopotentialHeightOnCGrid
from .updatedzd import UpdateHeightOnDGrid

HUGE_R = 1.0e40
CNST_0P20 = 0.2

_OP_SET_GZ, _OP_SET_PEM, _OP_GEOPOTENTIAL, _OP_COPY, _OP_FILL = range(5)


def get_nk_heat_dissipation(config, npz: int) -> int:
    """dyn_core.py:148-163."""
    if config.convert_ke or config.vtdm4 > 1.0e-4:
        return npz
    if config.d2_bg_k1 < 1.0e-3:
        return 0
    if config.d2_bg_k2 < 1.0e-3:
        return 1
    return 2


def k_bounds():
    """d_sw.k_bounds (d_sw.py:39-40) for the damping bands."""
    return [[0, 1], [1, 1], [2, 1], [3, None]]


class AcousticDynamics:
    """Fortran dyn_core: n_split Lagrangian acoustic substeps (Lin 2004)."""

    def __init__(self, comm, stencil_factory: StencilFactory, grid_data: GridData,
                 damping_coefficients: DampingCoefficients, grid_type, nested, stretched_grid,
                 config: AcousticDynamicsConfig, ak, bk, pfull, phis, use_graph: bool = False,
                 mixed_precision: bool = False):
        gi = stencil_factory.grid_indexing
        self._use_graph = bool(use_graph)
        self._graph = None  # the graph recorded last
        self._graphs = {}
        self._seen_keys = set()
        if config.d_ext != 0:
            raise NotImplementedError("d_ext != 0 is not implemented")
        if config.beta != 0:
            raise NotImplementedError("beta != 0 is not implemented")
        if config.use_logp:
            raise NotImplementedError("use_logp=True is not implemented")
        self.comm = comm
        self.config = config
        self._gi = gi
        self._da_min = damping_coefficients.da_min
        self.grid_data = grid_data
        npz = gi.domain[2]
        self._npz = npz
        self._pfull = np.asarray(pfull.detach().cpu().numpy() if isinstance(pfull, torch.Tensor)
                                 else pfull, dtype=np.float64)
        dsw_cfg = config.d_grid_shallow_water
        self._nk_heat_dissipation = get_nk_heat_dissipation(dsw_cfg, npz)
        self.nonhydrostatic_pressure_gradient = NonHydrostaticPressureGradient(
            stencil_factory, grid_data, config.grid_type)
        lead = tuple(grid_data.dx.shape[:-2])
        dev, dt_ = grid_data.dx.device, grid_data.dx.dtype
        shape3 = lead + tuple(gi.max_shape)
        shape2 = lead + tuple(gi.max_shape[:2])
        self._mixed = bool(mixed_precision)
        # storage type of the bulk fields (FP32 in the mixed form)
        bulk = torch.float32 if self._mixed else dt_
        self._temporaries: Dict[str, torch.Tensor] = {}
        for name in ("ut", "vt", "heat_source", "divgd", "crx", "xfx", "cry", "yfx"):
            self._temporaries[name] = utils.zeros(shape3, dtype=bulk, device=dev)
        for name in ("gz", "zh", "pem", "pkc", "pk3"):
            self._temporaries[name] = utils.zeros(shape3, dtype=dt_, device=dev)
        self._temporaries["ws3"] = utils.zeros(shape2, dtype=dt_, device=dev, batch_dims=len(lead))
        # FP64 copy of delp for the FP64 kernels that read it through a copy
        # (the pressure halos and nh_p_grad); the height updates read the FP32
        # winds, Courant numbers and area fluxes directly
        self._f64 = {}
        if self._mixed:
            self._f64["delp"] = utils.zeros(shape3, dtype=dt_, device=dev)
        grid_bulk = _mixed.grid_fp32(grid_data) if self._mixed else grid_data
        damp_bulk = (DampingCoefficients(grid_bulk.del6_u, grid_bulk.del6_v,
                                         damping_coefficients.da_min,
                                         damping_coefficients.da_min_c)
                     if self._mixed else damping_coefficients)
        self._temporaries["gz"].fill_(HUGE_R)
        if not config.hydrostatic:
            self._temporaries["pk3"].fill_(HUGE_R)
        column_namelist = get_column_namelist(dsw_cfg, npz)
        ak_h = np.asarray(ak.detach().cpu().numpy() if isinstance(ak, torch.Tensor) else ak,
                          dtype=np.float64)
        bk_h = np.asarray(bk.detach().cpu().numpy() if isinstance(bk, torch.Tensor) else bk,
                          dtype=np.float64)
        # dp_ref_compute (dyn_core.py:66-77): npz layer thicknesses (the
        # (npz+1)-th entry stays zero as in the reference's storage)
        dp_ref = np.zeros(npz + 1)
        dp_ref[:npz] = ak_h[1:npz + 1] - ak_h[:npz] + (bk_h[1:npz + 1] - bk_h[:npz]) * 1.0e5
        self._dp_ref_host = dp_ref
        self._dp_ref = torch.as_tensor(dp_ref, device=dev)
        self._zs = (phis * (1.0 / GRAV)).contiguous() if isinstance(phis, torch.Tensor) else None
        if self._zs is None:
            raise TypeError("phis must be a device storage")
        self._zs = utils.empty(phis.shape, dtype=dt_, device=dev, batch_dims=len(lead))
        self._zs.copy_(phis * (1.0 / GRAV))
        if not config.hydrostatic:
            self.update_height_on_d_grid = UpdateHeightOnDGrid(
                stencil_factory, damping_coefficients, grid_data, grid_type, config.hord_tm,
                dp_ref, column_namelist, k_bounds())
            self.riem_solver3 = RiemannSolver3(stencil_factory, config.riemann)
            self.riem_solver_c = RiemannSolverC(stencil_factory, p_fac=config.p_fac)
        self.dgrid_shallow_water_lagrangian_dynamics = DGridShallowWaterLagrangianDynamics(
            stencil_factory, grid_bulk, damp_bulk, column_namelist, nested, stretched_grid,
            dsw_cfg)
        self.cgrid_shallow_water_lagrangian_dynamics = CGridShallowWaterDynamics(
            stencil_factory, grid_bulk, nested, config.grid_type, config.nord)
        self.update_geopotential_height_on_c_grid = UpdateGeopotentialHeightOnCGrid(
            stencil_factory, grid_data.area)
        self._do_del2cubed = self._nk_heat_dissipation != 0 and config.d_con > 1.0e-5
        if self._do_del2cubed:
            self._hyperdiffusion = HyperdiffusionDamping(
                stencil_factory, damp_bulk, grid_bulk.rarea, nmax=min(3, config.nord + 1))
        if config.rf_fast:
            self._rayleigh_damping = RayleighDamping(stencil_factory, rf_cutoff=config.rf_cutoff,
                                                     tau=config.tau,
                                                     hydrostatic=config.hydrostatic)
        self._pk3_halo = PK3Halo(stencil_factory)
        # halo updaters (dyn_core.py:216-286)
        self._h_scalar = comm.get_scalar_halo_updater()
        self._h_vector_d = comm.get_vector_halo_updater(kind="vector_d")
        self._h_vector_c = comm.get_vector_halo_updater(kind="vector_c")
        self._h_corner = comm.get_vector_halo_updater(kind="corner")
        self._h_pkc = comm.get_scalar_halo_updater(
            n_halo=2 if gi.domain[0] == gi.domain[1] else gi.n_halo)

    # ------------------------------------------------------------ helpers
    def _column(self, op, nk, out, a=None, b=None, scalar=0.0):
        d = _lib.make_domain(self._gi, out, nk)
        _lib.check(_lib.fn("fv3_dyncore_column", out.dtype)(
            ctypes.byref(d), ctypes.c_int32(op), ctypes.c_double(scalar), _lib.ptr(a),
            _lib.ptr(b), _lib.ptr(out), ctypes.c_void_p(_lib.stream_handle())))

    def _zero_data(self, state, first_timestep: bool):
        """zero_data (dyn_core.py:47-62)."""
        for f in (state.mfxd, state.mfyd, state.cxd, state.cyd):
            self._column(_OP_FILL, self._npz, f, scalar=0.0)
        if first_timestep:
            c = self._gi
            sl = (Ellipsis, slice(c.isc, c.iec + 1), slice(c.jsc, c.jec + 1), slice(0, self._npz))
            state.diss_estd[sl] = 0.0

    def _p_grad_c(self, uc, vc, delpc, pkc, gz, dt2):
        g = self.grid_data
        d = _lib.make_domain(self._gi, uc, self._npz, g.rdxc, others3=(vc, delpc, pkc, gz),
                             others2=(g.rdyc,))
        name = "fv3_p_grad_c_mixed" if uc.dtype == torch.float32 else "fv3_p_grad_c_f64"
        _lib.check(getattr(_lib.lib(), name)(
            ctypes.byref(d), _lib.ptr(g.rdxc), _lib.ptr(g.rdyc),
            ctypes.c_int32(int(bool(self.config.hydrostatic))), _lib.ptr(uc), _lib.ptr(vc),
            _lib.ptr(delpc), _lib.ptr(pkc), _lib.ptr(gz), ctypes.c_double(dt2),
            ctypes.c_void_p(_lib.stream_handle())))

    # ---------------------------------------------------------------- step
    def _key(self, state):
        """Identity of a call for graph replay: storage addresses and scalars."""
        ptrs = tuple(sorted((k, v.data_ptr()) for k, v in state.__dict__.items()
                            if isinstance(v, torch.Tensor)))
        return ptrs, float(state.mdt), int(state.n_map), float(state.ptop), int(state.ks)

    def __call__(self, state):
        if not self._use_graph:
            return self._step(state)
        state.__dict__.update(self._temporaries)
        key = self._key(state)
        graph = self._graphs.get(key)
        if graph is not None:
            graph.replay()
            return None
        if key not in self._seen_keys:
            # first call for these storages: eager, fills every cache
            self._seen_keys.add(key)
            return self._step(state)
        # one graph per (storages, n_map, dt ...) combination: with k_split > 1
        # the calls of a step alternate n_map and each keeps its own graph
        graph = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(graph):
            self._step(state)
        torch.cuda.synchronize()
        self._graphs[key] = graph
        self._graph = graph
        graph.replay()
        return None

    def _step(self, state):
        config = self.config
        npz = self._npz
        end_step = state.n_map == config.k_split
        akap = KAPPA
        dt = state.mdt / config.n_split
        dt2 = 0.5 * dt
        n_split = config.n_split
        hydro = config.hydrostatic
        mixed = self._mixed
        if mixed:
            _mixed.check_mixed_state(state)
        f64 = self._f64

        def as64(name, t):
            """t itself (FP64 form) or its FP64 copy (mixed form)."""
            return _mixed.cast_into(t, f64[name]) if mixed else t

        self._h_scalar.update([state.q_con, state.cappa])
        self._h_scalar.update([state.delp, state.pt])
        self._h_vector_d.update([state.u], [state.v])
        state.__dict__.update(self._temporaries)
        self._zero_data(state, state.n_map == 1)
        for it in range(n_split):
            remap_step = config.breed_vortex_inline or (it == n_split - 1)
            if not hydro:
                self._h_scalar.update([state.w])
                if it == 0:
                    self._column(_OP_SET_GZ, npz + 1, state.gz, a=self._zs, b=state.delz)
                    self._h_scalar.update([state.gz])
            if it == n_split - 1 and end_step and config.use_old_omega:
                self._column(_OP_SET_PEM, npz, state.pem, a=as64("delp", state.delp),
                             scalar=state.ptop)
            state.delpc, state.ptc = self.cgrid_shallow_water_lagrangian_dynamics(
                state.delp, state.pt, state.u, state.v, state.w, state.uc, state.vc, state.ua,
                state.va, state.ut, state.vt, state.divgd, state.omga, dt2)
            if config.nord > 0:
                self._h_corner.update([state.divgd])
            if not hydro:
                if it == 0:
                    self._column(_OP_COPY, npz + 1, state.zh, a=state.gz)
                else:
                    self._column(_OP_COPY, npz + 1, state.gz, a=state.zh)
                self.update_geopotential_height_on_c_grid(
                    self._dp_ref, self._zs, state.ut, state.vt, state.gz, state.ws3, dt2)
                self.riem_solver_c(dt2, state.cappa, state.ptop, state.phis, state.ws3,
                                   state.ptc, state.q_con, state.delpc, state.gz, state.pkc,
                                   state.omga)
            self._p_grad_c(state.uc, state.vc, state.delpc, state.pkc, state.gz, dt2)
            self._h_vector_c.update([state.uc], [state.vc])
            self.dgrid_shallow_water_lagrangian_dynamics(
                state.vt, state.delp, state.ptc, state.pt, state.u, state.v, state.w, state.uc,
                state.vc, state.ua, state.va, state.divgd, state.mfxd, state.mfyd, state.cxd,
                state.cyd, state.crx, state.cry, state.xfx, state.yfx, state.q_con, state.zh,
                state.heat_source, state.diss_estd, dt)
            self._h_scalar.update([state.delp, state.pt, state.q_con])
            if not hydro:
                self.update_height_on_d_grid(self._zs, state.zh, state.crx, state.cry,
                                             state.xfx, state.yfx, state.wsd, dt)
                self.riem_solver3(remap_step, dt, state.cappa, state.ptop, self._zs, state.wsd,
                                  state.delz, state.q_con, state.delp, state.pt, state.zh,
                                  state.pe, state.pkc, state.pk3, state.pk, state.peln,
                                  state.w)
                self._h_scalar.update([state.zh])
                self._h_pkc.update([state.pkc])
                delp64 = as64("delp", state.delp)
                if remap_step:
                    edge_pe(self._gi, state.pe, delp64, state.ptop)
                self._pk3_halo(state.pk3, delp64, state.ptop, akap)
                self._column(_OP_GEOPOTENTIAL, npz + 1, state.gz, a=state.zh)
                self.nonhydrostatic_pressure_gradient(state.u, state.v, state.pkc, state.gz,
                                                      state.pk3, delp64, dt, state.ptop, akap)
            if config.rf_fast:
                self._rayleigh_damping(state.u, state.v, state.w, self._dp_ref_host[:npz],
                                       self._pfull, dt, state.ptop, state.ks)
            if it != n_split - 1:
                self._h_vector_d.update([state.u], [state.v])
            elif config.grid_type < 4:
                self.comm.synchronize_vector_interfaces(state.u, state.v)
        if self._do_del2cubed:
            self._h_scalar.update([state.heat_source])
            cd = CNST_0P20 * self._da_min
            self._hyperdiffusion(state.heat_source, cd)
            if not hydro:
                delt_time_factor = abs(dt * config.delt_max)
                compute_pkz_tempadjust(self._gi, self._nk_heat_dissipation, state.delp,
                                       state.delz, state.cappa, state.heat_source, state.pt,
                                       state.pkz, delt_time_factor)
