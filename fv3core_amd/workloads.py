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


def plane_metric(n: int, n_halo: int = 3, dx: float = 1.0e5, f: float = 0.0):
    """Every IJ metric term of an orthogonal uniform tile (a plane).

    sin_sg = 1, cos_sg = 0, cosa = 0, sina = 1 and all spacings ``dx``; the
    Coriolis parameter fC is the constant ``f``.
    """
    m = uniform_metric(n, n_halo, dx)
    ones = np.ones_like(m["area"])
    for k in (1, 2, 3, 4):
        m[f"sin_sg{k}"] = ones.copy()
        m[f"cos_sg{k}"] = 0.0 * ones
    for name in ("cosa", "cosa_u", "cosa_v", "cosa_s"):
        m[name] = 0.0 * ones
    for name in ("sina", "sina_u", "sina_v", "rsina", "rsin_u", "rsin_v", "rsin2"):
        m[name] = ones.copy()
    m["area_c"] = dx * dx * ones
    m["rarea_c"] = 1.0 / m["area_c"]
    for r, b in (("rdx", "dx"), ("rdy", "dy"), ("rdxc", "dxc"), ("rdyc", "dyc"),
                 ("rdxa", "dxa"), ("rdya", "dya")):
        m[r] = 1.0 / m[b]
    m["divg_u"] = ones.copy()
    m["divg_v"] = ones.copy()
    m["fC"] = f * ones
    m["f0"] = f * ones
    m["da_min_c"] = float(dx * dx)
    # small-angle longitude/latitude of corners and centres near the equator
    h = dx / 6.3712e6
    ii = np.arange(ones.shape[0])[:, None] * ones
    jj = np.arange(ones.shape[1])[None, :] * ones
    m["bgrid1"], m["bgrid2"] = ii * h, jj * h
    m["agrid1"], m["agrid2"] = (ii + 0.5) * h, (jj + 0.5) * h
    for name in ("edge_w", "edge_e", "edge_s", "edge_n"):
        m[name] = 0.5 * ones
    return m


def hybrid_coefficients(km: int = 79):
    """(ak, bk, ptop) of the reference's 79-layer hybrid grid (fv3core/grid/eta.py:38-217)."""
    if km != 79:
        raise NotImplementedError("Only a 79 vertical level grid is tabulated")
    import pathlib

    d = np.load(pathlib.Path(__file__).with_name("_eta79.npz"))
    ak, bk = d["ak"].copy(), d["bk"].copy()
    return ak, bk, float(ak[0])


def _smooth_ij(rng, shape2, amp):
    ni, nj = shape2
    x = np.arange(ni)[:, None] / max(ni - 1, 1)
    y = np.arange(nj)[None, :] / max(nj - 1, 1)
    f = np.zeros(shape2)
    for m in (1, 2):
        a, p1, p2 = rng.uniform(0.3, 1.0), rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)
        f += a * np.sin(2 * np.pi * m * x + p1) * np.cos(2 * np.pi * m * y + p2)
    return amp * f / np.abs(f).max()


def remap_inputs(n: int = 48, nz: int = 79, workload: int = 4, n_halo: int = 3):
    """Eulerian and perturbed Lagrangian interfaces plus fields to remap.

    pe2 = ak + bk*ps with ps = 1e5 Pa + a smooth +-500 Pa field; pe1 = pe2
    plus a smooth interior displacement (zero at top and bottom, at most 30%
    of the local layer thickness), so both grids span the same column.
    Returns dict(pe1, pe2, pt (K-like scalar), q (tracer in [0, 1e-2] with
    zeros), w (signed), qs (IJ bottom value), ptop).
    """
    ak, bk, ptop = hybrid_coefficients(nz)
    shape = storage_shape(n, nz, n_halo)
    ni, nj, nk = shape
    base = 1000 * workload
    rng = np.random.default_rng(base + 5)
    ps = 1.0e5 + _smooth_ij(rng, (ni, nj), 500.0)
    pe2 = ak[None, None, :] + bk[None, None, :] * ps[:, :, None]
    dp2 = np.diff(pe2, axis=2)
    rng2 = np.random.default_rng(base + 6)
    disp = rng2.uniform(-0.3, 0.3, size=(ni, nj, nk))
    disp[:, :, 0] = 0.0
    disp[:, :, -1] = 0.0
    lim = np.minimum(np.concatenate([dp2[:, :, :1], dp2], axis=2),
                     np.concatenate([dp2, dp2[:, :, -1:]], axis=2))
    pe1 = pe2 + disp * lim
    rng3 = np.random.default_rng(base + 7)
    kk = np.arange(nz)[None, None, :]
    pt = 220.0 + 80.0 * (kk / nz) ** 2 + rng3.normal(0, 1.0, size=(ni, nj, nz))
    q = np.clip(1e-2 * (kk / nz) ** 3 + rng3.normal(0, 1e-3, size=(ni, nj, nz)), 0.0, None)
    w = rng3.normal(0.0, 1.0, size=(ni, nj, nz))
    pad = lambda a: np.concatenate([a, a[:, :, -1:]], axis=2)  # noqa: E731
    return dict(pe1=pe1, pe2=pe2, pt=pad(pt), q=pad(q), w=pad(w),
                qs=rng3.normal(0.0, 0.5, size=(ni, nj)), ptop=ptop, ak=ak, bk=bk)


def riemann_inputs(n: int = 24, nz: int = 79, workload: int = 2, n_halo: int = 3,
                   dt: float = 37.5):
    """A hydrostatically consistent column state plus perturbations.

    delp from the hybrid grid with ps ~ 1e5 Pa; cappa = Rd/cp with a small
    moisture-like variation; T ~ 210-290 K and pt = T / p**cappa; delz from the discrete hydrostatic
    relation the solver uses (so the unperturbed state is at rest), then pt
    perturbed by 1 % and w by N(0, 0.5 m/s).
    """
    RD, CP, G = 287.05, 1004.6, 9.80665
    ak, bk, ptop = hybrid_coefficients(nz)
    shape = storage_shape(n, nz, n_halo)
    ni, nj, nk = shape
    base = 1000 * workload
    rng = np.random.default_rng(base + 11)
    ps = 1.0e5 + _smooth_ij(rng, (ni, nj), 800.0)
    pe = ak[None, None, :] + bk[None, None, :] * ps[:, :, None]
    delp = np.zeros(shape)
    delp[:, :, :nz] = np.diff(pe, axis=2)
    delp[:, :, nz] = delp[:, :, nz - 1]
    q_con = np.zeros(shape)
    q_con[:, :, :nz] = rng.uniform(0.0, 2e-3, size=(ni, nj, nz))
    cappa = RD / CP * (1.0 + rng.uniform(-0.01, 0.01, size=shape))
    kk = np.arange(nk)[None, None, :]
    temp = 210.0 + 80.0 * (kk / nz) + rng.normal(0, 1.0, size=shape)
    # hydrostatic dz for the solver's equation of state
    peg = np.zeros(shape)
    peg[:, :, 0] = ptop
    for k in range(1, nk):
        peg[:, :, k] = peg[:, :, k - 1] + delp[:, :, k - 1] * (1.0 - q_con[:, :, k - 1])
    pm = np.zeros(shape)
    pm[:, :, :nz] = (peg[:, :, 1:] - peg[:, :, :nz]) / np.log(peg[:, :, 1:] / peg[:, :, :nz])
    pm[:, :, nz] = pm[:, :, nz - 1]
    # the acoustic solver's pt is T / p**cappa (p in Pa)
    pt = temp * np.exp(-cappa * np.log(pm))
    dm = delp / G
    delz = -dm * RD * pt * np.exp((cappa - 1.0) * np.log(pm))
    delz[:, :, nz] = 0.0
    zs = np.zeros((ni, nj))
    zh = np.zeros(shape)
    zh[:, :, nz] = zs
    for k in range(nz - 1, -1, -1):
        zh[:, :, k] = zh[:, :, k + 1] - delz[:, :, k]
    pt = pt * (1.0 + rng.uniform(-0.01, 0.01, size=shape))
    w = rng.normal(0.0, 0.5, size=shape)
    # the C-grid solver reads gz as heights (its dz is gz[k+1] - gz[k]) and
    # returns geopotential (riem_solver_c.py:36, 57-64)
    gz = zh.copy()
    return dict(ak=ak, bk=bk, ptop=ptop, delp=delp, q_con=q_con, cappa=cappa, pt=pt,
                delz=delz, zh=zh, zs=zs, hs=zs * G, w=w, ws=np.zeros((ni, nj)), gz=gz,
                pe=pe.copy() if pe.shape[2] == nk else pe, dt=dt)


def synthetic_hybrid(nz: int, ptop: float = 300.0, p0: float = 1.0e5):
    """(ak, bk) of a smooth hybrid grid for any nz: pure pressure at the top,
    terrain-following at the bottom (bk = eta^3, ak = p_ref - bk p0 >= 0)."""
    eta = np.arange(nz + 1) / nz
    p_ref = ptop + (p0 - ptop) * eta ** 1.5
    bk = eta ** 3
    ak = p_ref - bk * p0
    return ak, bk, ptop


def cube_state(n: int = 96, nz: int = 79, seed: int = 0, ps: float = 1.0e5):
    """The DynCore initial state on the six cube tiles (arrays [6, I, J, K]).

    Hybrid ak/bk levels (the 79-level grid for nz = 79), ps = 1e5 Pa,
    T = 288 K with a 6.5 K/km lapse rate down to a 216 K tropopause
    (expressed through pressure), a zonal jet u = 35 sin^2(2 lat) m/s
    projected on the D-grid edges, a seeded 1 K Gaussian pt bump, w = 0,
    dry air (cappa = Rd/cp, q_con = 0), flat surface.  pt is T / p**cappa,
    delz the hydrostatic thickness, pe/peln/pk/pkz consistent with delp.
    """
    from . import cube
    from .constants import CP_AIR, GRAV, KAPPA, RDGAS

    rng = np.random.default_rng(seed)
    if nz == 79:
        ak, bk, ptop = hybrid_coefficients(79)
    else:
        ak, bk, ptop = synthetic_hybrid(nz)
    m = cube.tile_metrics(n)
    h = 3
    ni = n + 2 * h + 1
    nk = nz + 1
    shp = (6, ni, ni, nk)
    pe = np.zeros(shp)
    pe[...] = ak[None, None, None, :] + bk[None, None, None, :] * ps
    delp = np.zeros(shp)
    delp[..., :nz] = np.diff(pe, axis=3)
    delp[..., nz] = delp[..., nz - 1]
    peln = np.log(pe)
    pk = np.exp(KAPPA * peln)
    pm = np.ones(shp)
    pm[..., :nz] = delp[..., :nz] / (peln[..., 1:] - peln[..., :nz])
    gamma = 0.0065 * RDGAS / GRAV
    temp = np.maximum(216.0, 288.0 * (pm / ps) ** gamma)
    cappa = np.full(shp, RDGAS / CP_AIR)
    pt = temp * np.exp(-cappa * np.log(pm))
    # seeded Gaussian bump of 1 K centred at a random point of tile 0..5
    P = m["agrid3"]
    c = rng.normal(size=3)
    c /= np.linalg.norm(c)
    dist = np.arccos(np.clip(np.einsum("tijx,x->tij", P, c), -1, 1))
    bump = np.exp(-(dist / 0.2) ** 2)[..., None]
    pt = pt * (1.0 + bump / temp)
    delz = np.zeros(shp)
    delz[..., :nz] = -(delp[..., :nz] / GRAV) * RDGAS * pt[..., :nz] * \
        np.exp((cappa[..., :nz] - 1.0) * np.log(pm[..., :nz]))
    zh = np.zeros(shp)
    for k in range(nz - 1, -1, -1):
        zh[..., k] = zh[..., k + 1] - delz[..., k]
    pkz = np.ones(shp)
    pkz[..., :nz] = (pk[..., 1:] - pk[..., :nz]) / (KAPPA * (peln[..., 1:] - peln[..., :nz]))
    # zonal jet on the D-grid: u on y-faces (edge P(i,j) -> P(i+1,j)), v on x-faces
    G = m["grid3"]

    def edge_wind(p1, p2):
        mid = p1 + p2
        mid = mid / np.linalg.norm(mid, axis=-1, keepdims=True)
        tang = p2 - p1
        tang = tang / np.linalg.norm(tang, axis=-1, keepdims=True)
        lat = np.arcsin(np.clip(mid[..., 2], -1, 1))
        east = np.stack([-mid[..., 1], mid[..., 0], np.zeros_like(mid[..., 0])], axis=-1)
        east = east / np.maximum(np.linalg.norm(east, axis=-1, keepdims=True), 1e-12)
        U = 35.0 * np.sin(2.0 * lat) ** 2
        return U * np.einsum("...x,...x->...", east, tang)

    u = np.zeros(shp)
    v = np.zeros(shp)
    u[:, :ni - 1, :, :] = edge_wind(G[:, :-1, :], G[:, 1:, :])[..., None]
    v[:, :, :ni - 1, :] = edge_wind(G[:, :, :-1], G[:, :, 1:])[..., None]
    z = np.zeros(shp)
    state = dict(delp=delp, pt=pt, u=u, v=v, w=z.copy(), delz=delz, q_con=z.copy(),
                 cappa=cappa, uc=z.copy(), vc=z.copy(), ua=z.copy(), va=z.copy(),
                 omga=z.copy(), mfxd=z.copy(), mfyd=z.copy(), cxd=z.copy(), cyd=z.copy(),
                 diss_estd=z.copy(), pe=pe, pk=pk, peln=peln, pkz=pkz, zh=zh)
    return dict(state=state, ak=ak, bk=bk, ptop=ptop, metric=m,
                phis=np.zeros((6, ni, ni)), wsd=np.zeros((6, ni, ni)))


def cube_tracers(shape, seed: int = 5000, names=None):
    """Nine tracer fields of a batched storage shape (6, I, J, K): smooth seeded
    fields in [0, 1e-2] kg/kg (tracer t from rng(seed + 1 + t)), the last
    (cloud fraction) in [0, 1]."""
    from .constants import TRACER_VARIABLES

    names = list(names or TRACER_VARIABLES)
    nt, ni, nj, nk = shape
    ii = np.arange(ni)[None, :, None, None] / ni
    jj = np.arange(nj)[None, None, :, None] / nj
    kk = np.arange(nk)[None, None, None, :] / nk
    tt = np.arange(nt)[:, None, None, None]
    out = {}
    for n, name in enumerate(names):
        rng = np.random.default_rng(seed + 1 + n)
        p = rng.uniform(0.0, 2.0 * np.pi, size=4)
        f = (np.sin(2 * np.pi * ii + p[0] + tt) * np.cos(2 * np.pi * jj + p[1])
             * np.cos(np.pi * kk + p[2]) + 0.5 * np.sin(2 * np.pi * (ii + jj) + p[3]))
        f = (f - f.min()) / (f.max() - f.min())
        scale = 1.0 if name == "qcld" else 1.0e-2
        out[name] = np.ascontiguousarray(scale * f)
    return out
