"""TracerAdvection (fv3core/stencils/tracer_2d_1l.py:120-300) on the GPU.

The sub-cycle count is fixed at 3 as in the reference (cmax = 2.0,
tracer_2d_1l.py:222-249).  Per call: one fused flux-preparation kernel, then
per sub-cycle one dp kernel and one batched transport of all tracers
(``fv3_tracer_transport``: their corner copies, one fv_tp_2d launch over
every tracer and tile, one q update launch), and the on-device halo gather
of all tracers between sub-cycles.  Where the domain's corner layout rules
the batch out, each tracer gets its own corner copy, fv_tp_2d and update.
"""
from __future__ import annotations

import ctypes
import math

import torch

from .. import _lib
from ..utils import gt4py_utils as utils
from ..utils.stencil import X_DIM, Y_DIM, Z_DIM, StencilFactory
from .fvtp2d import FiniteVolumeTransport, PreAllocatedCopiedCornersFactory


def _storage(q):
    """A Quantity's storage, or the tensor itself."""
    return q if isinstance(q, torch.Tensor) else q.storage


class TracerAdvection:
    """Fortran tracer_2D_1L."""

    def __init__(self, stencil_factory: StencilFactory, transport: FiniteVolumeTransport, comm,
                 tracer_count, grid_data=None):
        gi = stencil_factory.grid_indexing
        self._gi = gi
        self._tracer_count = tracer_count
        self.comm = comm
        # the reference reads the module-level grid (spec.grid); here the
        # metric terms come from the transport the caller hands in
        if grid_data is None:
            grid_data = getattr(transport, "grid_data", None)
        if grid_data is None:
            raise ValueError("transport carries no grid_data (GridData with dxa, dya, dx, dy, "
                             "sin_sg1..4, rarea)")
        self._grid = grid_data
        self._stencil_factory = stencil_factory
        self.finite_volume_transport = transport
        self._tmp = None
        self._halo = comm.get_scalar_halo_updater() if comm is not None else None
        self._copy_corners = None
        # one batched transport of every tracer per sub-cycle when the
        # transport carries no damping of its own
        self._batched = getattr(transport, "delnflux", None) is None

    def _temporaries(self, like):
        key = (tuple(like.shape), like.dtype)
        if self._tmp is None or self._tmp[0] != key:
            shape = tuple(like.shape)
            bd = 1 if like.dim() == 4 else 0
            mk = lambda: utils.zeros(shape, dtype=like.dtype, device=like.device,  # noqa: E731
                                     batch_dims=bd)
            tmps = dict(xfx=mk(), yfx=mk(), fx=mk(), fy=mk(), dp=mk(), ytmp=mk())
            tmps["fxs"] = [mk() for _ in range(self._tracer_count)]
            tmps["fys"] = [mk() for _ in range(self._tracer_count)]
            self._copy_corners = PreAllocatedCopiedCornersFactory(
                self._stencil_factory, dims=[X_DIM, Y_DIM, Z_DIM], y_temporary=tmps["ytmp"])
            self._tmp = (key, tmps)
        return self._tmp[1]

    def _transport_batch(self, qs, tmp, cxd, cyd, mfxd, mfyd, dp1, dp2, nz, stream):
        """fv_tp_2d and q_adjust of every tracer in one batched call; turns
        the batch off for good when the library rules it out."""
        fvt = self.finite_volume_transport
        g = self._grid
        n = len(qs)
        d = _lib.make_domain(self._gi, dp1, nz, g.area,
                             others3=list(qs) + tmp["fxs"][:n] + tmp["fys"][:n] +
                             [cxd, cyd, tmp["xfx"], tmp["yfx"], mfxd, mfyd, dp2],
                             others2=(g.dxa, g.dya, g.rarea))
        table = ctypes.c_void_p * n
        rc = _lib.lib().fv3_tracer_transport_f64(
            ctypes.byref(d), int(fvt._hord), n, table(*[q.data_ptr() for q in qs]),
            table(*[f.data_ptr() for f in tmp["fxs"][:n]]),
            table(*[f.data_ptr() for f in tmp["fys"][:n]]), _lib.ptr(cxd), _lib.ptr(cyd),
            _lib.ptr(tmp["xfx"]), _lib.ptr(tmp["yfx"]), _lib.ptr(mfxd), _lib.ptr(mfyd),
            _lib.ptr(g.area), _lib.ptr(g.dxa), _lib.ptr(g.dya), _lib.ptr(g.rarea), _lib.ptr(dp1),
            _lib.ptr(dp2), stream)
        if rc == _lib.E_UNSUPPORTED:
            self._batched = False
            return
        _lib.check(rc)

    def __call__(self, tracers, dp1, mfxd, mfyd, cxd, cyd, mdt):
        if len(tracers) != self._tracer_count:
            raise ValueError(
                f"incorrect number of tracers, {self._tracer_count} was "
                f"specified on init but {len(tracers)} were passed"
            )
        qs = [_storage(q) for q in (tracers.values() if hasattr(tracers, "values") else tracers)]
        _lib.require_device(dp1, mfxd, mfyd, cxd, cyd, *qs)
        tmp = self._temporaries(dp1)
        g = self._grid
        nz = self._gi.domain[2]
        d = _lib.make_domain(self._gi, cxd, nz, g.dxa,
                             others3=(cyd, mfxd, mfyd, tmp["xfx"], tmp["yfx"]))
        cmax = 2.0
        n_split = int(math.floor(1.0 + cmax))
        _lib.check(_lib.lib().fv3_tracer_flux_prep_f64(
            ctypes.byref(d), n_split, _lib.ptr(cxd), _lib.ptr(cyd), _lib.ptr(mfxd),
            _lib.ptr(mfyd), _lib.ptr(tmp["xfx"]), _lib.ptr(tmp["yfx"]), _lib.ptr(g.dxa),
            _lib.ptr(g.dya), _lib.ptr(g.dx), _lib.ptr(g.dy), _lib.ptr(g.sin_sg1),
            _lib.ptr(g.sin_sg2), _lib.ptr(g.sin_sg3), _lib.ptr(g.sin_sg4),
            ctypes.c_void_p(_lib.stream_handle())))
        if self._halo is not None:
            self._halo.update(qs)
        dp2 = tmp["dp"]
        s = ctypes.c_void_p(_lib.stream_handle())
        for it in range(n_split):
            last_call = it == n_split - 1
            da = _lib.make_domain(self._gi, dp1, nz, g.rarea, others3=(dp2, mfxd, mfyd))
            _lib.check(_lib.lib().fv3_tracer_adjust_f64(
                ctypes.byref(da), 0, _lib.ptr(dp2), _lib.ptr(dp1), _lib.ptr(mfxd),
                _lib.ptr(mfyd), _lib.ptr(g.rarea), None, s))
            if self._batched:
                self._transport_batch(qs, tmp, cxd, cyd, mfxd, mfyd, dp1, dp2, nz, s)
            if not self._batched:
                for q in qs:
                    self.finite_volume_transport(
                        self._copy_corners(q), cxd, cyd, tmp["xfx"], tmp["yfx"], tmp["fx"],
                        tmp["fy"], x_mass_flux=mfxd, y_mass_flux=mfyd)
                    _lib.check(_lib.lib().fv3_tracer_adjust_f64(
                        ctypes.byref(da), 1, _lib.ptr(q), _lib.ptr(dp1), _lib.ptr(tmp["fx"]),
                        _lib.ptr(tmp["fy"]), _lib.ptr(g.rarea), _lib.ptr(dp2), s))
            if not last_call:
                if self._halo is not None:
                    self._halo.update(qs)
                dp1, dp2 = dp2, dp1
