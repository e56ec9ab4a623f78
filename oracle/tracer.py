"""CPU oracle for TracerAdvection (fv3core/stencils/tracer_2d_1l.py:120-300).

Test infrastructure only (see oracle/common.py).  Six tiles are arrays
``[6, I, J, K]``; halos are refreshed by the host gather of the halo map
(``fv3core_amd.halo.apply_map_numpy``), which the cube tests pin against the
geometry separately.
"""
from __future__ import annotations

import math

import numpy as np

from .common import box, finite_on, sh
from .fvtp2d import Tile, copy_corners_xy, fvtp2d


def flux_compute(cx, cy, g, t: Tile):
    """flux_compute (tracer_2d_1l.py:19-57) on one tile; returns (xfx, yfx)."""
    dxa, dya, dx, dy = g["dxa"][:, :, None], g["dya"][:, :, None], g["dx"][:, :, None], g["dy"][:, :, None]
    s1, s2, s3, s4 = (g[f"sin_sg{k}"][:, :, None] for k in (1, 2, 3, 4))
    xfx = np.full(cx.shape, np.nan)
    yfx = np.full(cy.shape, np.nan)
    xv = np.where(cx > 0, cx * sh(dxa, -1, 0) * dy * sh(s3, -1, 0), cx * dxa * dy * s1)
    yv = np.where(cy > 0, cy * sh(dya, 0, -1) * dx * sh(s4, 0, -1), cy * dya * dx * s2)
    xfx[t.isc:t.iec + 2, t.jsc - 3:t.jec + 4] = xv[t.isc:t.iec + 2, t.jsc - 3:t.jec + 4]
    yfx[t.isc - 3:t.iec + 4, t.jsc:t.jec + 2] = yv[t.isc - 3:t.iec + 4, t.jsc:t.jec + 2]
    return xfx, yfx


def tracer_advection(tracers, dp1, mfxd, mfyd, cxd, cyd, metric, hord, halo, nz, edges=True):
    """Returns (tracers, dp1_storage, mfxd, mfyd, cxd, cyd) after the call.

    ``metric``: dict of [6, I, J] arrays; ``halo(list_of_fields)`` returns the
    fields with refreshed halos.  Arrays are copied, not modified.
    """
    ntile = cxd.shape[0]
    n = cxd.shape[1] - 7
    tile = Tile(n, n, 3, edges, edges, edges, edges)
    qs = [np.array(q, copy=True) for q in tracers]
    dp1 = np.array(dp1, copy=True)
    mfxd, mfyd, cxd, cyd = (np.array(a, copy=True) for a in (mfxd, mfyd, cxd, cyd))
    kz = slice(0, nz)
    xfx = np.zeros(cxd.shape)
    yfx = np.zeros(cyd.shape)
    for t in range(ntile):
        g = {k: v[t] for k, v in metric.items()}
        a, b = flux_compute(cxd[t][..., kz], cyd[t][..., kz], g, tile)
        xfx[t][..., kz] = np.nan_to_num(a)
        yfx[t][..., kz] = np.nan_to_num(b)
    n_split = math.floor(1.0 + 2.0)
    frac = 1.0 / n_split
    for arr in (cxd, xfx, mfxd, cyd, yfx, mfyd):
        arr[..., kz] = arr[..., kz] * frac
    qs = halo(qs)
    dp2 = np.zeros(dp1.shape)
    a_dp1, a_dp2 = dp1, dp2  # storage identities, as in the reference
    ci, cj = slice(tile.isc, tile.iec + 1), slice(tile.jsc, tile.jec + 1)
    cbox = box(cxd.shape[1:3], tile.isc, tile.iec, tile.jsc, tile.jec)
    for it in range(n_split):
        last = it == n_split - 1
        for t in range(ntile):
            ra = metric["rarea"][t][:, :, None]
            div = (mfxd[t] - sh(mfxd[t], 1, 0) + mfyd[t] - sh(mfyd[t], 0, 1)) * ra
            a_dp2[t][ci, cj, kz] = (a_dp1[t] + div)[ci, cj, kz]
        for q in qs:
            for t in range(ntile):
                g = {k: v[t] for k, v in metric.items()}
                qx, qy = copy_corners_xy(q[t], tile)
                q[t] = qx  # the x-differentiable copy shares memory with q
                fx, fy = fvtp2d(qx, qy, cxd[t], cyd[t], xfx[t], yfx[t], g["area"], g["dxa"],
                                g["dya"], hord, tile, x_mass_flux=mfxd[t], y_mass_flux=mfyd[t])
                ra = g["rarea"][:, :, None]
                # dp2 is only formed on the compute box (zero in the halo)
                with np.errstate(divide="ignore", invalid="ignore"):
                    new = (q[t] * a_dp1[t] + (fx - sh(fx, 1, 0) + fy - sh(fy, 0, 1)) * ra) \
                        / a_dp2[t]
                finite_on(new[..., kz], cbox, "q_adjust")
                q[t][ci, cj, kz] = new[ci, cj, kz]
        if not last:
            qs = halo(qs)
            a_dp1, a_dp2 = a_dp2, a_dp1
    return qs, dp1, mfxd, mfyd, cxd, cyd
