"""Seeded synthetic inputs for the transport and column workloads.

All generators return numpy FP64 arrays in the reference's storage index
space ``[i, j, k]`` (compute origin ``(n_halo, n_halo, 0)``, storage
``(N + 2*n_halo + 1, N + 2*n_halo + 1, nz + 1)``).  The seed of field ``f``
of workload ``c`` is ``1000*c + f``.
"""
from __future__ import annotations

import numpy as np


def storage_shape(n: int, nz: int, n_halo: int = 3):
    return (n + 2 * n_halo + 1, n + 2 * n_halo + 1, nz + 1)


def fill_periodic(a: np.ndarray, n: int, n_halo: int = 3) -> np.ndarray:
    """Fill every non-compute storage entry by wrapping with period ``n`` in i and j.

    Interface fields (extent ``n + 1``) get ``a[iec+1] == a[isc]``, which is
    what a doubly periodic tile implies.
    """
    ni, nj = a.shape[0], a.shape[1]
    ii = n_halo + (np.arange(ni) - n_halo) % n
    jj = n_halo + (np.arange(nj) - n_halo) % n
    return a[ii][:, jj].copy()


def _smooth_sinusoids(rng, x, y, nz, modes, amp_range, phase=True):
    f = np.zeros(x.shape[:2] + (nz,))
    amps = rng.uniform(amp_range[0], amp_range[1], size=(len(modes), nz))
    ph = rng.uniform(0.0, 2.0 * np.pi, size=(len(modes), 2, nz))
    for mi, m in enumerate(modes):
        f += (
            amps[mi][None, None, :]
            * np.sin(2.0 * np.pi * m * x[:, :, None] + ph[mi, 0][None, None, :])
            * np.sin(2.0 * np.pi * m * y[:, :, None] + ph[mi, 1][None, None, :])
        )
    return f


def transport_inputs(n: int = 48, nz: int = 79, workload: int = 1, n_halo: int = 3):
    """Doubly periodic single-tile fv_tp_2d inputs on a uniform 100 km grid.

    q: sum of four random-phase sinusoids (wavenumbers 1..4) scaled to [0, 1]
    plus N(0, 1e-3) noise.  crx, cry: 0.4 x smooth seeded fields normalised to
    max |s| = 1.5 (so |c| <= 0.6).  x_area_flux = crx*dy, y_area_flux = cry*dx.
    Every storage level (nz + 1 of them) is filled, since the reference's
    transport stencils run over nz + 1 levels.
    """
    shape = storage_shape(n, nz, n_halo)
    nk = shape[2]
    ni, nj = shape[0], shape[1]
    i = np.arange(ni)
    j = np.arange(nj)
    x = ((i - n_halo + 0.5) / n)[:, None] * np.ones((1, nj))
    y = np.ones((ni, 1)) * ((j - n_halo + 0.5) / n)[None, :]
    base = 1000 * workload
    f = _smooth_sinusoids(np.random.default_rng(base + 1), x, y, nk, (1, 2, 3, 4), (0.5, 1.0))
    fmin = f.min(axis=(0, 1), keepdims=True)
    fmax = f.max(axis=(0, 1), keepdims=True)
    q = (f - fmin) / (fmax - fmin)
    q = q + np.random.default_rng(base + 2).normal(0.0, 1e-3, size=q.shape)

    def courant(seed, face_axis):
        rng = np.random.default_rng(seed)
        xf = ((i - n_halo) / n)[:, None] * np.ones((1, nj)) if face_axis == 0 else x
        yf = np.ones((ni, 1)) * ((j - n_halo) / n)[None, :] if face_axis == 1 else y
        s = np.zeros(shape)
        b = rng.uniform(0.5, 1.0, size=(2, nk))
        th = rng.uniform(0.0, 2.0 * np.pi, size=(2, 2, nk))
        for mi, m in enumerate((1, 2)):
            s += (
                b[mi][None, None, :]
                * np.cos(2 * np.pi * m * xf[:, :, None] + th[mi, 0][None, None, :])
                * np.cos(2 * np.pi * m * yf[:, :, None] + th[mi, 1][None, None, :])
            )
        s = s / np.abs(s).max(axis=(0, 1), keepdims=True) * 1.5
        return 0.4 * s

    crx = courant(base + 3, 0)
    cry = courant(base + 4, 1)
    q = fill_periodic(q, n, n_halo)
    crx = fill_periodic(crx, n, n_halo)
    cry = fill_periodic(cry, n, n_halo)
    dx = dy = 1.0e5
    metric = uniform_metric(n, n_halo, dx)
    return dict(
        q=q,
        crx=crx,
        cry=cry,
        x_area_flux=crx * dy,
        y_area_flux=cry * dx,
        **metric,
    )


def uniform_metric(n: int, n_halo: int = 3, dx: float = 1.0e5):
    """IJ metric terms of a uniform Cartesian tile (sin_sg = 1, cos_sg = 0)."""
    shp = (n + 2 * n_halo + 1, n + 2 * n_halo + 1)
    ones = np.ones(shp)
    area = dx * dx * ones
    return dict(
        area=area,
        rarea=1.0 / area,
        dxa=dx * ones,
        dya=dx * ones,
        dx=dx * ones,
        dy=dx * ones,
        dxc=dx * ones,
        dyc=dx * ones,
        # del6_u = sina_v * dx / dyc, del6_v = sina_u * dy / dxc
        # (fv3core/grid/geometry.py:527-530) with sina = 1
        del6_u=ones * (dx / dx),
        del6_v=ones * (dx / dx),
        da_min=float(dx * dx),
    )
