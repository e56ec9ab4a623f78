"""Adjoint of PPM transport (fv_tp_2d) behind ``torch.autograd``.

``FvTp2dFn`` is a :class:`torch.autograd.Function` whose forward runs the
fused transport kernel ``fv3_fvtp2d`` and whose backward runs the
hand-written adjoint ``fv3_fvtp2d_adjoint``.  The adjoint re-linearises the
forward at the saved input (every hord limiter branch, smoothness mask and
upwind choice is taken as the forward took it), so each call keeps only its
input state: the call is its own checkpoint.  ``TransportLinearization``
also exposes the tangent-linear model (``fv3_fvtp2d_tlm``) for dot-product
tests.

``RemapLinearization`` / ``MapScalarFn`` do the same for the vertical remap
map_scalar (MapSingle with kord 9) through ``fv3_map_single_adjoint``.

``PeriodicTransport`` steps a scalar on a doubly periodic tile with the
flux-form update ``q += (fx(i) - fx(i+1) + fy(j) - fy(j+1)) / area`` (the
``q_adjust`` form of fv3core/stencils/tracer_2d_1l.py:108-117 with unit
layer mass), optionally followed by a remap from fixed Lagrangian to fixed
Eulerian interfaces, and :func:`cost_and_gradient` evaluates the 4D-Var-style cost
``J = 0.5 ||q_T - q_obs||^2`` over ``n`` steps and ``dJ/dq0`` by the reverse
sweep.  Courant numbers, area fluxes and metric terms are held fixed.

Tensors handed to the kernels are i-fastest ``(I, J, K)`` storages.  The
autograd-side state is kept as row-major ``(K, J, I)`` tensors, whose
``permute(2, 1, 0)`` view is exactly that layout.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import torch
import torch.nn.functional as F

from . import _lib


def ijk_view(kji: torch.Tensor) -> torch.Tensor:
    """(I, J, K) i-fastest view of a row-major (K, J, I) tensor."""
    return kji.permute(2, 1, 0)


def as_ijk_storage(t: torch.Tensor) -> torch.Tensor:
    """``t`` (I, J, K) with i-fastest strides (copied if it has other strides)."""
    if t.dim() == 3 and t.stride(0) == 1 and t.stride(1) == t.shape[0] and \
            t.stride(2) == t.shape[0] * t.shape[1]:
        return t
    return ijk_view(t.permute(2, 1, 0).contiguous())


def new_ijk(like: torch.Tensor, zero: bool = True) -> torch.Tensor:
    """(I, J, K) i-fastest storage shaped like ``like`` (zeroed unless told not to)."""
    ni, nj, nk = like.shape
    make = torch.zeros if zero else torch.empty
    return ijk_view(make((nk, nj, ni), dtype=like.dtype, device=like.device))


class TransportLinearization:
    """fv_tp_2d forward, tangent-linear and adjoint on one (I, J, K) storage.

    ``grid_indexing`` gives the compute domain and tile-edge anchors (as for
    FiniteVolumeTransport); ``area``, ``dxa``, ``dya`` are IJ metric fields;
    ``hord`` is 5, 6, 7 or 8; ``nk`` the number of levels transported.
    """

    def __init__(self, grid_indexing, area, dxa, dya, hord: int, nk: int):
        if hord not in (5, 6, 7, 8):
            raise NotImplementedError(f"hord={hord} is not supported (5, 6, 7, 8)")
        _lib.require_device(area, dxa, dya)
        if area.dtype != torch.float64:
            raise TypeError("the transport adjoint computes in float64")
        self.gi = grid_indexing
        self.area, self.dxa, self.dya = area, dxa, dya
        self.hord = int(hord)
        self.nk = int(nk)
        self._work = None
        self._slab = 0

    def _domain(self, qx, *others):
        return _lib.make_domain(self.gi, qx, self.nk, self.area, others3=others,
                                others2=(self.dxa, self.dya))

    def _workspace(self, ref):
        ntiles, ni, nj, _, sj, sk, st = _lib.layout3(ref)
        need = (ntiles - 1) * st + self.nk * sk
        need = (need + 63) // 64 * 64
        nslabs = _lib.lib().fv3_fvtp2d_workspace_slabs()
        if self._work is None or self._slab < need or self._work.device != ref.device:
            self._work = torch.empty(nslabs * need, dtype=torch.float64, device=ref.device)
            self._slab = need
        return self._work, self._slab

    def forward(self, qx, qy, crx, cry, xfx, yfx, xmf=None, ymf=None):
        """(fx, fy) = fv_tp_2d(q) on fresh zeroed storages."""
        _lib.require_device(qx, qy, crx, cry, xfx, yfx, xmf, ymf)
        fx, fy = new_ijk(qx), new_ijk(qx)
        d = self._domain(qx, qy, crx, cry, xfx, yfx, xmf, ymf, fx, fy)
        _lib.check(_lib.lib().fv3_fvtp2d_f64(
            ctypes.byref(d), self.hord, _lib.ptr(qx), _lib.ptr(qy), _lib.ptr(crx),
            _lib.ptr(cry), _lib.ptr(xfx), _lib.ptr(yfx), _lib.ptr(xmf), _lib.ptr(ymf),
            _lib.ptr(self.area), _lib.ptr(self.dxa), _lib.ptr(self.dya), _lib.ptr(fx),
            _lib.ptr(fy), ctypes.c_void_p(_lib.stream_handle())))
        return fx, fy

    def tlm(self, qx, qy, crx, cry, xfx, yfx, dqx, dqy, xmf=None, ymf=None):
        """(dfx, dfy): the transport linearised at (qx, qy) applied to (dqx, dqy)."""
        _lib.require_device(qx, qy, crx, cry, xfx, yfx, xmf, ymf, dqx, dqy)
        dfx, dfy = new_ijk(qx), new_ijk(qx)
        d = self._domain(qx, qy, crx, cry, xfx, yfx, xmf, ymf, dqx, dqy, dfx, dfy)
        work, slab = self._workspace(qx)
        _lib.check(_lib.lib().fv3_fvtp2d_tlm_f64(
            ctypes.byref(d), self.hord, _lib.ptr(qx), _lib.ptr(qy), _lib.ptr(crx),
            _lib.ptr(cry), _lib.ptr(xfx), _lib.ptr(yfx), _lib.ptr(xmf), _lib.ptr(ymf),
            _lib.ptr(self.area), _lib.ptr(self.dxa), _lib.ptr(self.dya), _lib.ptr(dqx),
            _lib.ptr(dqy), _lib.ptr(dfx), _lib.ptr(dfy), _lib.ptr(work), ctypes.c_int64(slab),
            ctypes.c_void_p(_lib.stream_handle())))
        return dfx, dfy

    def adjoint(self, qx, qy, crx, cry, xfx, yfx, gfx, gfy, xmf=None, ymf=None):
        """(g_qx, g_qy): the transposed linearisation applied to (gfx, gfy)."""
        _lib.require_device(qx, qy, crx, cry, xfx, yfx, xmf, ymf, gfx, gfy)
        gqx, gqy = new_ijk(qx, zero=False), new_ijk(qx, zero=False)  # fully overwritten
        d = self._domain(qx, qy, crx, cry, xfx, yfx, xmf, ymf, gfx, gfy, gqx, gqy)
        work, slab = self._workspace(qx)
        _lib.check(_lib.lib().fv3_fvtp2d_adjoint_f64(
            ctypes.byref(d), self.hord, _lib.ptr(qx), _lib.ptr(qy), _lib.ptr(crx),
            _lib.ptr(cry), _lib.ptr(xfx), _lib.ptr(yfx), _lib.ptr(xmf), _lib.ptr(ymf),
            _lib.ptr(self.area), _lib.ptr(self.dxa), _lib.ptr(self.dya), _lib.ptr(gfx),
            _lib.ptr(gfy), _lib.ptr(gqx), _lib.ptr(gqy), _lib.ptr(work), ctypes.c_int64(slab),
            ctypes.c_void_p(_lib.stream_handle())))
        return gqx, gqy


class RemapLinearization:
    """map_scalar (MapSingle with kord 9, mode ``iv``) forward, tangent-linear
    and adjoint with respect to the remapped field, on the columns of the
    compute domain of ``grid_indexing`` (km = its vertical extent).  pe1, pe2
    (interfaces), qs (IJ, used by iv = -2) and qmin are held fixed."""

    def __init__(self, grid_indexing, mode: int = 1, qmin: float = 0.0, kord: int = 9):
        if abs(int(kord)) != 9:
            raise NotImplementedError("the remap derivative exists for kord 9")
        self.gi = grid_indexing
        self.kord = int(kord)
        self.mode = int(mode)
        self.qmin = float(qmin)
        self.km = grid_indexing.domain[2]
        self.box = (grid_indexing.isc, grid_indexing.iec, grid_indexing.jsc, grid_indexing.jec)
        self._work = {}

    def _domain(self, q, pe1, pe2, qs, *others):
        if pe1.shape[-1] < self.km + 1 or pe2.shape[-1] < self.km + 1:
            raise ValueError("pe1 and pe2 need km + 1 interface levels")
        return _lib.make_domain(self.gi, q, self.km, qs, others3=(pe1, pe2) + others)

    def _workspace(self, ref, nslabs: int):
        ntiles, ni, nj, _, sj, sk, st = _lib.layout3(ref)
        need = (ntiles - 1) * st + (self.km + 1) * sk
        need = (need + 63) // 64 * 64
        w = self._work.get(nslabs)
        if w is None or w[1] < need or w[0].device != ref.device:
            w = (torch.empty(nslabs * need, dtype=torch.float64, device=ref.device), need)
            self._work[nslabs] = w
        return w

    def forward(self, q, pe1, pe2, qs=None):
        """A remapped copy of q (columns outside the box and level km unchanged)."""
        _lib.require_device(q, pe1, pe2, qs)
        out = new_ijk(q, zero=False)
        out.copy_(q)
        d = self._domain(out, pe1, pe2, qs)
        work, slab = self._workspace(out, 7)
        i1, i2, j1, j2 = self.box
        _lib.check(_lib.lib().fv3_map_single_f64(
            ctypes.byref(d), self.kord, self.mode, i1, i2, j1, j2, _lib.ptr(out), _lib.ptr(pe1),
            _lib.ptr(pe2), _lib.ptr(qs), ctypes.c_double(self.qmin), _lib.ptr(work),
            ctypes.c_int64(slab), ctypes.c_void_p(_lib.stream_handle())))
        return out

    def tlm(self, q, dq, pe1, pe2, qs=None):
        """(remapped q, its perturbation for the perturbation dq of q)."""
        _lib.require_device(q, dq, pe1, pe2, qs)
        out, dout = new_ijk(q, zero=False), new_ijk(q, zero=False)
        out.copy_(q)
        dout.copy_(dq)
        d = self._domain(out, pe1, pe2, qs, dout)
        work, slab = self._workspace(out, _lib.lib().fv3_map_single_tlm_workspace_slabs())
        i1, i2, j1, j2 = self.box
        _lib.check(_lib.lib().fv3_map_single_tlm_f64(
            ctypes.byref(d), self.kord, self.mode, i1, i2, j1, j2, _lib.ptr(out), _lib.ptr(dout),
            _lib.ptr(pe1), _lib.ptr(pe2), _lib.ptr(qs), ctypes.c_double(self.qmin),
            _lib.ptr(work), ctypes.c_int64(slab), ctypes.c_void_p(_lib.stream_handle())))
        return out, dout

    def adjoint(self, q, g_out, pe1, pe2, qs=None):
        """Cotangent of q for the cotangent g_out of the remapped field."""
        _lib.require_device(q, g_out, pe1, pe2, qs)
        g_in = new_ijk(q, zero=False)
        g_in.copy_(g_out)  # identity outside the remapped columns and levels
        d = self._domain(q, pe1, pe2, qs, g_out, g_in)
        work, slab = self._workspace(q, _lib.lib().fv3_map_single_adjoint_workspace_slabs())
        i1, i2, j1, j2 = self.box
        _lib.check(_lib.lib().fv3_map_single_adjoint_f64(
            ctypes.byref(d), self.kord, self.mode, i1, i2, j1, j2, _lib.ptr(q), _lib.ptr(pe1),
            _lib.ptr(pe2), _lib.ptr(qs), ctypes.c_double(self.qmin), _lib.ptr(g_out),
            _lib.ptr(g_in), _lib.ptr(work), ctypes.c_int64(slab),
            ctypes.c_void_p(_lib.stream_handle())))
        return g_in


class MapScalarFn(torch.autograd.Function):
    """q_out = map_scalar(q) with the hand-written adjoint as backward."""

    @staticmethod
    def forward(ctx, q, lin: RemapLinearization, pe1, pe2, qs):
        q_s = as_ijk_storage(q)
        out = lin.forward(q_s, pe1, pe2, qs)
        ctx.save_for_backward(q_s)
        ctx.lin = lin
        ctx.aux = (pe1, pe2, qs)
        return out

    @staticmethod
    def backward(ctx, g):
        (q,) = ctx.saved_tensors
        return ctx.lin.adjoint(q, as_ijk_storage(g), *ctx.aux), None, None, None, None


class FvTp2dFn(torch.autograd.Function):
    """(fx, fy) = fv_tp_2d(q_x, q_y) with the hand-written adjoint as backward.

    ``aux`` = (crx, cry, x_area_flux, y_area_flux, x_mass_flux, y_mass_flux)
    is held fixed (no gradient flows to it).
    """

    @staticmethod
    def forward(ctx, qx, qy, lin: TransportLinearization, aux):
        qx_s, qy_s = as_ijk_storage(qx), as_ijk_storage(qy)
        fx, fy = lin.forward(qx_s, qy_s, *aux[:4], aux[4], aux[5])
        ctx.save_for_backward(qx_s, qy_s)
        ctx.lin = lin
        ctx.aux = aux
        return fx, fy

    @staticmethod
    def backward(ctx, gfx, gfy):
        qx, qy = ctx.saved_tensors
        aux = ctx.aux
        gfx = torch.zeros_like(qx) if gfx is None else gfx
        gfy = torch.zeros_like(qx) if gfy is None else gfy
        gqx, gqy = ctx.lin.adjoint(qx, qy, *aux[:4], as_ijk_storage(gfx),
                                   as_ijk_storage(gfy), aux[4], aux[5])
        return gqx, gqy, None, None


class PeriodicTransport:
    """Flux-form PPM transport of a scalar on one doubly periodic tile.

    ``inputs`` holds device (I, J, K) storages ``crx``, ``cry``,
    ``x_area_flux``, ``y_area_flux`` and IJ fields ``area``, ``dxa``, ``dya``
    (as :func:`fv3core_amd.workloads.transport_inputs` builds them).  The
    state is a row-major (K, J, I) tensor of the storage shape; only its
    compute interior is meaningful, the halo is refilled by wrapping before
    every step.
    """

    def __init__(self, grid_indexing, inputs: dict, hord: int, n_halo: int = 3,
                 remap: Optional[dict] = None):
        gi = grid_indexing
        self.gi = gi
        self.h = n_halo
        self.n = gi.domain[0]
        if gi.domain[0] != gi.domain[1]:
            raise ValueError("a square tile is expected")
        crx = as_ijk_storage(inputs["crx"])
        ni, nj, nk = crx.shape
        self.shape = (ni, nj, nk)
        self.lin = TransportLinearization(gi, inputs["area"], inputs["dxa"], inputs["dya"],
                                          hord, nk)
        self.aux = (crx, as_ijk_storage(inputs["cry"]), as_ijk_storage(inputs["x_area_flux"]),
                    as_ijk_storage(inputs["y_area_flux"]), None, None)
        h, n = self.h, self.n
        if n < max(h, ni - h - n, nj - h - n):
            raise ValueError("the tile is narrower than its halo")
        # optional vertical remap after each transport step: dict(pe1, pe2,
        # mode=1, qmin=0.0, qs=None) with (I, J, K) interface storages
        self.remap = None
        if remap is not None:
            self.remap = (RemapLinearization(gi, remap.get("mode", 1), remap.get("qmin", 0.0)),
                          as_ijk_storage(remap["pe1"]), as_ijk_storage(remap["pe2"]),
                          remap.get("qs"))
        rarea = 1.0 / inputs["area"]                       # (I, J)
        self._rarea = rarea.t()[None, h:h + n, h:h + n]   # (1, J, I) interior

    def fill(self, q_kji: torch.Tensor) -> torch.Tensor:
        """Halo of a (K, J, I) state refilled by wrapping (differentiable).

        Built from slices and concatenations, whose backward is slicing and
        addition (no index sort as for a general gather).
        """
        h, n = self.h, self.n
        ni, nj, _ = self.shape
        c = q_kji[:, h:h + n, h:h + n]
        c = torch.cat([c[:, n - h:n], c, c[:, 0:nj - h - n]], dim=1)
        return torch.cat([c[:, :, n - h:n], c, c[:, :, 0:ni - h - n]], dim=2)

    def step(self, q_kji: torch.Tensor) -> torch.Tensor:
        h, n = self.h, self.n
        qf = self.fill(q_kji)
        q = ijk_view(qf)
        fx, fy = FvTp2dFn.apply(q, q, self.lin, self.aux)
        fx_kji = fx.permute(2, 1, 0)
        fy_kji = fy.permute(2, 1, 0)
        c = slice(h, h + n)
        c1 = slice(h + 1, h + n + 1)
        dq = (fx_kji[:, c, c] - fx_kji[:, c, c1] + fy_kji[:, c, c] - fy_kji[:, c1, c]) * \
            self._rarea
        ni, nj, _ = self.shape
        dq = F.pad(dq, (h, ni - h - n, h, nj - h - n))
        q_new = qf + dq
        if self.remap is not None:
            lin, pe1, pe2, qs = self.remap
            q_new = MapScalarFn.apply(ijk_view(q_new), lin, pe1, pe2, qs).permute(2, 1, 0)
        return q_new

    def run(self, q0_kji: torch.Tensor, nsteps: int) -> torch.Tensor:
        q = q0_kji
        for _ in range(nsteps):
            q = self.step(q)
        return q

    def interior(self, q_kji: torch.Tensor) -> torch.Tensor:
        h, n = self.h, self.n
        return q_kji[:, h:h + n, h:h + n]


def cost_and_gradient(model: PeriodicTransport, q0_kji: torch.Tensor, q_obs_kji: torch.Tensor,
                      nsteps: int) -> Tuple[float, torch.Tensor]:
    """J = 0.5 ||q_T - q_obs||^2 over the compute interior, and dJ/dq0."""
    q0 = q0_kji.detach().clone().requires_grad_(True)
    qt = model.run(q0, nsteps)
    r = model.interior(qt) - model.interior(q_obs_kji)
    cost = 0.5 * (r * r).sum()
    (grad,) = torch.autograd.grad(cost, q0)
    return float(cost.detach()), grad
