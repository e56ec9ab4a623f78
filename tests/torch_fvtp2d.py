"""Torch restatement of fv_tp_2d on a doubly periodic tile (autodiff reference).

Test infrastructure only: an independent torch form of the interior PPM
transport (no tile-edge regions), differentiated by torch autograd to give a
reference gradient for the hand-written HIP adjoint.  It follows
fv3core/stencils/xppm.py:31-44 (fx1_fn), 84-102 (dm_iord8plus, al and
blbr for hord 8), 147-181 (compute_al for hord < 8), 269-280
(compute_x_flux) and fvtp2d.py:33-92 (q_i, q_j, final_fluxes); yppm is
xppm with i and j exchanged (convert_xppm_yppm.sh:9-57).

Arrays are periodic interiors laid out (K, J, I); ``sh(a, di, dj)`` is a at
(i + di, j + dj) with wrap-around.  x-face i lies between cells i-1 and i.
"""
from __future__ import annotations

import torch

P1 = 7.0 / 12.0
P2 = -1.0 / 12.0


def sh(a, di=0, dj=0):
    return torch.roll(a, shifts=(-dj, -di), dims=(-2, -1))


def sign(a, b):
    """|a| if b > 0 else -|a| (basic_operations.py:33-40)."""
    return torch.where(b > 0, a.abs(), -a.abs())


def xppm(q, c, hord: int):
    if hord == 8:
        xt = 0.25 * (sh(q, 1) - sh(q, -1))
        dqr = torch.maximum(torch.maximum(q, sh(q, -1)), sh(q, 1)) - q
        dql = q - torch.minimum(torch.minimum(q, sh(q, -1)), sh(q, 1))
        dm = sign(torch.minimum(torch.minimum(xt.abs(), dqr), dql), xt)
        al = 0.5 * (sh(q, -1) + q) + 1.0 / 3.0 * (sh(dm, -1) - dm)
        x2 = 2.0 * dm
        bl = -1.0 * sign(torch.minimum(x2.abs(), (al - q).abs()), x2)
        br = sign(torch.minimum(x2.abs(), (sh(al, 1) - q).abs()), x2)
        b0 = bl + br
        pos = (1.0 - c) * (sh(br, -1) - c * sh(b0, -1))
        neg = (1.0 + c) * (bl + c * b0)
        return torch.where(c > 0.0, sh(q, -1) + pos, q + neg)
    al = P1 * (sh(q, -1) + q) + P2 * (sh(q, -2) + sh(q, 1))
    bl = al - q
    br = sh(al, 1) - q
    b0 = bl + br
    smt = (bl * br < 0) if hord == 5 else ((3.0 * b0.abs()) < (bl - br).abs())
    mask = torch.where(sh(smt, -1) | smt, 1.0, 0.0).to(q.dtype)
    pos = (1.0 - c) * (sh(br, -1) - c * sh(b0, -1))
    neg = (1.0 + c) * (bl + c * b0)
    return torch.where(c > 0.0, sh(q, -1) + pos * mask, q + neg * mask)


def _t(a):
    return a.transpose(-1, -2)


def yppm(q, c, hord: int):
    return _t(xppm(_t(q), _t(c), hord))


def fvtp2d(q, crx, cry, xfx, yfx, area, hord: int):
    """(fx, fy) on the periodic interior; ``area`` is (J, I)."""
    ymean = yppm(q, cry, hord)
    fyy = yfx * ymean
    q_i = (q * area + fyy - sh(fyy, 0, 1)) / (area + yfx - sh(yfx, 0, 1))
    xouter = xppm(q_i, crx, hord)
    xmean = xppm(q, crx, hord)
    fx1 = xfx * xmean
    q_j = (q * area + fx1 - sh(fx1, 1, 0)) / (area + xfx - sh(xfx, 1, 0))
    youter = yppm(q_j, cry, hord)
    return 0.5 * (xouter + xmean) * xfx, 0.5 * (youter + ymean) * yfx


def interior_kji(a, n: int, h: int = 3):
    """(K, J, I) periodic interior of an (I, J, K) numpy/torch storage."""
    t = torch.as_tensor(a)
    if t.dim() == 2:
        return t[h:h + n, h:h + n].transpose(0, 1).contiguous()
    return t[h:h + n, h:h + n, :].permute(2, 1, 0).contiguous()
