"""Inputs shared by the acoustic-step tests: smooth seeded fields on one
periodic plane tile or on the six tiles of the cube, and device copies."""
import numpy as np
import torch

from fv3core_amd import cube
from fv3core_amd.utils import gt4py_utils as utils
from fv3core_amd.utils.grid import GridData
from fv3core_amd.utils.stencil import GridIndexing, StencilConfig, StencilFactory
from fv3core_amd.workloads import fill_periodic, plane_metric
from oracle.fvtp2d import Tile


def smooth(rng, shape, mean, amp, modes=(1, 2, 3)):
    """Sum of random-phase sinusoids over the (i, j, k) index space."""
    ni, nj, nk = shape[-3:]
    x = np.arange(ni)[:, None, None] / ni
    y = np.arange(nj)[None, :, None] / nj
    z = np.arange(nk)[None, None, :] / max(nk, 1)
    f = np.zeros((ni, nj, nk))
    for m in modes:
        a = rng.uniform(0.3, 1.0) / m
        p = rng.uniform(0, 2 * np.pi, size=3)
        f += a * np.sin(2 * np.pi * m * x + p[0]) * np.cos(2 * np.pi * m * y + p[1]) * \
            np.cos(np.pi * z + p[2])
    return mean + amp * f / np.abs(f).max()


class Case:
    """A single periodic plane tile (edges off) or the six cube tiles (edges on)."""

    def __init__(self, n, nz, cube_tiles, seed=0, f=1.0e-4):
        self.n, self.nz, self.cube = n, nz, cube_tiles
        self.ntiles = 6 if cube_tiles else 1
        self.rng = np.random.default_rng(seed)
        edges = bool(cube_tiles)
        self.tile = Tile(n, n, 3, edges, edges, edges, edges)
        self.gi = GridIndexing((n, n, nz), 3, edges, edges, edges, edges)
        self.factory = StencilFactory(StencilConfig(), self.gi)
        if cube_tiles:
            m = cube.tile_metrics(n)
            self.metric = [{k: v[t] for k, v in m.items() if isinstance(v, np.ndarray)}
                           for t in range(6)]
        else:
            m = plane_metric(n, f=f)
            self.metric = [{k: v for k, v in m.items() if isinstance(v, np.ndarray)}]
        # scalar terms (da_min, da_min_c ...), the same for every tile
        self.scalars = {k: float(v) for k, v in m.items() if not isinstance(v, np.ndarray)}
        self.shape = (n + 7, n + 7, nz + 1)

    def field(self, mean, amp):
        out = []
        for _ in range(self.ntiles):
            a = smooth(self.rng, self.shape, mean, amp)
            if not self.cube:
                a = fill_periodic(a, self.n)
            out.append(a)
        return out

    def dev(self, arrs):
        """Device storage of one field: (I, J, K) for one tile, (6, I, J, K) for six."""
        if self.ntiles == 1:
            t = utils.empty(arrs[0].shape)
            t.copy_(torch.as_tensor(arrs[0]))
            return t
        t = utils.empty((6,) + arrs[0].shape)
        t.copy_(torch.as_tensor(np.stack(arrs)))
        return t

    def host(self, t):
        a = t.detach().cpu().numpy()
        return [a] if self.ntiles == 1 else [a[i] for i in range(6)]

    def grid_data(self, names=None):
        fields = {}
        keys = names or self.metric[0].keys()
        for k in keys:
            if self.ntiles == 1:
                t = utils.empty(self.metric[0][k].shape)
                t.copy_(torch.as_tensor(self.metric[0][k]))
            else:
                t = utils.empty((6,) + self.metric[0][k].shape, batch_dims=1)
                t.copy_(torch.as_tensor(np.stack([m[k] for m in self.metric])))
            fields[k] = t
        return GridData(fields)
