"""The six-tile halo map against the cube geometry (CPU), and the GPU gather
against the host restatement of the same map (bit-exact)."""
import numpy as np
import pytest
import torch

from fv3core_amd import cube
from fv3core_amd.halo import (
    CENTER,
    CORNER,
    X_FACE,
    Y_FACE,
    HaloUpdater,
    apply_map_numpy,
    build_map,
)

N = 8
H = 3


@pytest.fixture(scope="module")
def corners():
    return cube.tile_corners(N, H)


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_connectivity_is_symmetric_and_matches_geometry(corners):
    P = corners
    lo, hi = H, H + N
    for t in range(6):
        for edge in ("east", "west", "north", "south"):
            tb, eb, rot = cube.neighbour(t, edge)
            back = cube.neighbour(tb, eb)
            assert back[0] == t and back[2] == rot
            # the corner points along the shared edge coincide
            if edge == "east":
                a = P[t, hi, lo:hi + 1]
            elif edge == "west":
                a = P[t, lo, lo:hi + 1]
            elif edge == "north":
                a = P[t, lo:hi + 1, hi]
            else:
                a = P[t, lo:hi + 1, lo]
            if eb == "east":
                b = P[tb, hi, lo:hi + 1]
            elif eb == "west":
                b = P[tb, lo, lo:hi + 1]
            elif eb == "north":
                b = P[tb, lo:hi + 1, hi]
            else:
                b = P[tb, lo:hi + 1, lo]
            if rot:
                b = b[::-1]
            np.testing.assert_allclose(a, b, atol=1e-14)


def test_scalar_halo_cells_are_the_adjacent_neighbour_cells(corners):
    P = corners
    mp = build_map(N, (CENTER,), 3, H)
    assert len(mp) == 6 * 4 * 3 * N
    # every halo cell shares exactly one edge (two corners) with the cell one
    # step closer to the tile, walking outward from the tile's own edge cell
    def cell_corners(t, i, j):
        return [P[t, i, j], P[t, i + 1, j], P[t, i + 1, j + 1], P[t, i, j + 1]]

    src = {(r[0], r[1], r[2]): (r[4], r[5], r[6]) for r in mp}

    def corners_of(t, i, j):
        if (t, i, j) in src:
            return cell_corners(*src[(t, i, j)])
        return cell_corners(t, i, j)

    for (t, i, j) in src:
        # inward neighbour in A's index space
        if i < H:
            inner = (t, i + 1, j)
        elif i >= H + N:
            inner = (t, i - 1, j)
        elif j < H:
            inner = (t, i, j + 1)
        else:
            inner = (t, i, j - 1)
        a = np.array(corners_of(*(src.get((t, i, j), (t, i, j)) and (t, i, j))))
        b = np.array(corners_of(*inner))
        shared = sum(np.min(np.linalg.norm(b - p, axis=1)) < 1e-12 for p in a)
        assert shared == 2, ((t, i, j), inner)
    # and the halo cells along an edge are each other's neighbours in A's order
    for (t, i, j) in src:
        nxt = (t, i, j + 1) if (i < H or i >= H + N) else (t, i + 1, j)
        if nxt not in src:
            continue
        a = np.array(corners_of(t, i, j))
        b = np.array(corners_of(*nxt))
        shared = sum(np.min(np.linalg.norm(b - p, axis=1)) < 1e-12 for p in a)
        assert shared == 2, ((t, i, j), nxt)


def _velocity(p):
    # solid-body rotation about a tilted axis: smooth and nonzero on every tile
    axis = _unit(np.array([0.3, -0.5, 0.8]))
    return np.cross(axis, p)


@pytest.mark.parametrize("kind", ["vector_d", "vector_c"])
def test_vector_halo_components_follow_the_rotation(kind):
    # Near the middle of an edge the grid lines of the two tiles continue
    # each other without a kink, so A's axis directions extended into the halo
    # match the neighbour's axes there: the mapped neighbour component (with
    # its sign) must equal the velocity component along A's extended axis.
    # Away from the edge middle the lines kink (up to 60 degrees at cube
    # corners), so only the first halo row at the two middle points of every
    # edge is compared, on a fine tile (n = 64) where those points sit within
    # a degree of the kink-free middle.  A swapped or negated component is off
    # by O(|V|).
    n = 64
    P = cube.tile_corners(n, H)
    stg = (Y_FACE, X_FACE) if kind == "vector_d" else (X_FACE, Y_FACE)
    mp = build_map(n, stg, 3, H)

    def axis_dir(t, i, j, s, ax):
        # tangent of the face-coordinate line `ax` through the point
        if s == Y_FACE:
            mid = _unit(P[t, i, j] + P[t, i + 1, j])
        else:
            mid = _unit(P[t, i, j] + P[t, i, j + 1])
        if ax == 0:
            e = (P[t, i + 1, j] - P[t, i, j]) if s == Y_FACE else (
                P[t, i + 1, j] + P[t, i + 1, j + 1] - P[t, i - 1, j] - P[t, i - 1, j + 1])
        else:
            e = (P[t, i, j + 1] - P[t, i, j]) if s == X_FACE else (
                P[t, i, j + 1] + P[t, i + 1, j + 1] - P[t, i, j - 1] - P[t, i + 1, j - 1])
        e = e - mid * (e @ mid)
        return mid, _unit(e)

    checked = 0
    for r in mp:
        t, i, j, c, tb, ib, jb, sc = (int(v) for v in r)
        s_ = stg[c]
        along = j if (i < H or i > H + n) else i
        across_i = i < H or i > H + n - (1 if s_[0] == 0.5 else 0)
        depth = (H - i if i < H else i - (H + n) + (1 if s_[0] == 0.5 else 0)) if across_i \
            else (H - j if j < H else j - (H + n) + (1 if s_[1] == 0.5 else 0))
        if depth != 1 or abs(along - (H + n / 2 - 0.5)) > 1.0:
            continue
        src_comp, neg = sc // 2, sc % 2
        mid_b, e_b = axis_dir(tb, ib, jb, stg[src_comp], src_comp)
        mid_a, e_a = axis_dir(t, i, j, s_, c)
        np.testing.assert_allclose(mid_a, mid_b, atol=2e-2)
        got = float(_velocity(mid_b) @ e_b) * (-1 if neg else 1)
        want = float(_velocity(mid_a) @ e_a)
        assert abs(got - want) < 0.05, (t, i, j, c, got, want)
        checked += 1
    assert checked >= 6 * 4 * 2


def test_corner_map_keeps_corners(corners):
    # a corner point one grid line beyond the edge comes from the neighbour's
    # first interior grid line, one grid spacing from the shared edge point
    P = corners
    mp = build_map(N, (CORNER,), 3, H)
    lo, hi = H, H + N
    n_checked = 0
    for r in mp:
        t, i, j, _, tb, ib, jb, _ = (int(v) for v in r)
        if i == lo - 1:
            e, step = P[t, lo, j], np.linalg.norm(P[t, lo + 1, j] - P[t, lo, j])
        elif i == hi + 1:
            e, step = P[t, hi, j], np.linalg.norm(P[t, hi, j] - P[t, hi - 1, j])
        elif j == lo - 1:
            e, step = P[t, i, lo], np.linalg.norm(P[t, i, lo + 1] - P[t, i, lo])
        elif j == hi + 1:
            e, step = P[t, i, hi], np.linalg.norm(P[t, i, hi] - P[t, i, hi - 1])
        else:
            continue
        d = np.linalg.norm(P[tb, ib, jb] - e)
        assert 0.7 * step < d < 1.4 * step, (t, i, j, d / step)
        n_checked += 1
    assert n_checked == 6 * 4 * (N + 1)


def test_sources_are_never_halo_points():
    for stg in ((CENTER,), (Y_FACE, X_FACE), (X_FACE, Y_FACE), (CORNER,)):
        mp = build_map(N, stg, 3, H)
        dst = {(r[0], r[1], r[2], r[3]) for r in mp}
        for r in mp:
            assert (r[4], r[5], r[6], r[7] // 2) not in dst
            assert H <= r[5] <= H + N and H <= r[6] <= H + N


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["scalar", "vector_d", "vector_c", "corner"])
def test_gpu_halo_update_matches_host_gather(kind):
    from fv3core_amd.utils import gt4py_utils as utils
    from fv3core_amd.utils.stencil import GridIndexing

    gi = GridIndexing((N, N, 5), 3, True, True, True, True)
    ncomp = 1 if kind in ("scalar", "corner") else 2
    rng = np.random.default_rng(11)
    shape = (6,) + gi.max_shape
    host = [rng.normal(size=shape) for _ in range(2 * ncomp)]
    dev = []
    for h in host:
        t = utils.empty(shape)
        t.copy_(torch.as_tensor(h))
        dev.append(t)
    upd = HaloUpdater(gi, kind)
    upd.update(*dev)
    torch.cuda.synchronize()
    stg = HaloUpdater._KINDS[kind]
    mp = build_map(N, stg, 3, 3)
    for g in range(2):
        want = apply_map_numpy(host[g * ncomp:(g + 1) * ncomp], mp)
        for c in range(ncomp):
            np.testing.assert_array_equal(dev[g * ncomp + c].cpu().numpy(), want[c])
