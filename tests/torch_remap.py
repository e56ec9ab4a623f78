"""Torch restatement of map_scalar with kord 9: the autodiff reference for the
hand-written remap adjoint.

Written from the reference's text, RemapProfile (remap_profile.py:43-116
constraints, 119-207 set_initial_vals, 210-280 apply_constraints, 283-495
set_interpolation_coefficients with the kord-9 inner limiter, 376-430) and
lagrangian_contributions (map_single.py:19-66), with every branch as a
``torch.where`` so that autograd differentiates the branch the forward took.
The pressures are constants: the overlap walk of lagrangian_contributions
depends only on them, so it is formed once as a linear operator on the four
parabola coefficients of every source layer (numpy) and applied with torch.

Columns are rows of ``[ncol, km]`` tensors (``[ncol, km + 1]`` for edges).
Test infrastructure only, like ``tests/torch_fvtp2d.py``.
"""
from __future__ import annotations

import numpy as np
import torch


def _std_a4(a1, a2, a3):
    return 3.0 * (2.0 * a1 - (a2 + a3))


def _posdef_iv1(a1, a2, a3, a4):
    """posdef_constraint_iv1 (remap_profile.py:72-92)."""
    da1 = a3 - a2
    da2 = da1 * da1
    a6da = a4 * da1
    c1 = ((a1 - a2) * (a1 - a3)) >= 0.0
    c2 = ~c1 & (a6da < -1.0 * da2)
    c3 = ~c1 & ~c2 & (a6da > da2)
    b4 = torch.where(c2, 3.0 * (a2 - a1), torch.where(c3, 3.0 * (a3 - a1), a4))
    b3 = torch.where(c2, a2 - b4, a3)
    b2 = torch.where(c3, a3 - b4, a2)
    return (torch.where(c1, a1, b2), torch.where(c1, a1, b3),
            torch.where(c1, torch.zeros_like(a4), b4))


def _remap_constraint(a1, a2, a3, a4, extm):
    """remap_constraint (remap_profile.py:95-116)."""
    da1 = a3 - a2
    da2 = da1 * da1
    a6da = a4 * da1
    c2 = ~extm & (a6da < -da2)
    c3 = ~extm & ~c2 & (a6da > da2)
    b4 = torch.where(c2, 3.0 * (a2 - a1), torch.where(c3, 3.0 * (a3 - a1), a4))
    b3 = torch.where(c2, a2 - b4, a3)
    b2 = torch.where(c3, a3 - b4, a2)
    return (torch.where(extm, a1, b2), torch.where(extm, a1, b3),
            torch.where(extm, torch.zeros_like(a4), b4))


def _posdef_iv0(a1, a2, a3, a4):
    """posdef_constraint_iv0 (remap_profile.py:43-69)."""
    nonpos = a1 <= 0.0
    with torch.no_grad():
        ratio = torch.where(a4 != 0.0, 0.25 * (a3 - a2) ** 2 / torch.where(a4 != 0.0, a4, 1.0),
                            torch.full_like(a4, float("inf")))
        cond = (torch.abs(a3 - a2) < -a4) & ((a1 + ratio + a4 * (1.0 / 12.0)) < 0.0)
    c1 = cond & (a1 < a3) & (a1 < a2)
    c2 = cond & ~c1 & (a3 > a2)
    c3 = cond & ~c1 & ~c2
    b4 = torch.where(c1, torch.zeros_like(a4),
                     torch.where(c2, 3.0 * (a2 - a1), torch.where(c3, 3.0 * (a3 - a1), a4)))
    b3 = torch.where(c1, a1, torch.where(c2, a2 - b4, a3))
    b2 = torch.where(c1, a1, torch.where(c3, a3 - b4, a2))
    return (torch.where(nonpos, a1, b2), torch.where(nonpos, a1, b3),
            torch.where(nonpos, torch.zeros_like(a4), b4))


def _min3(a, p, l):
    return torch.where((a < p) & (a < l), a, torch.where(p < l, p, l))


def _max3(a, p, l):
    return torch.where((a > p) & (a > l), a, torch.where(p > l, p, l))


def profile_kord9(qs, a1, dp, iv, qmin=0.0):
    """cs_profile with kord 9: (a4_1, a4_2, a4_3, a4_4), each [ncol, km]."""
    ncol, km = a1.shape
    A = [a1[:, k] for k in range(km)]
    D = [dp[:, k] for k in range(km)]
    q = [None] * (km + 1)
    gam = [None] * (km + 1)
    # set_initial_vals: edge values from the tridiagonal system
    if iv == -2:
        half = torch.full_like(A[0], 0.5)
        q[0] = 1.5 * A[0]
        gam[1] = half
        gr1 = D[0] / D[1]
        q[1] = (3.0 * (A[0] + A[1]) - q[0]) / (2.0 + gr1 + gr1 - gam[1])
        grid = [None] * (km + 1)
        for k in range(2, km):
            old = D[k - 2] / D[k - 1]
            gam[k] = old / (2.0 + old + old - gam[k - 1])
            grid[k] = D[k - 1] / D[k]
        for k in range(2, km - 1):
            q[k] = (3.0 * (A[k - 1] + A[k]) - q[k - 1]) / (2.0 + grid[k] + grid[k] - gam[k])
        k = km - 1
        q[k] = (3.0 * (A[k - 1] + A[k]) - grid[k] * qs - q[k - 1]) / \
            (2.0 + grid[k] + grid[k] - gam[k])
        q[km] = qs
        for k in range(km - 2, -1, -1):
            q[k] = q[k] - gam[k + 1] * q[k + 1]
        gam[0] = torch.zeros_like(A[0])
    else:
        gr = D[1] / D[0]
        bet = gr * (gr + 0.5)
        q[0] = ((gr + gr) * (gr + 1.0) * A[0] + A[1]) / bet
        gam[0] = (1.0 + gr * (gr + 1.5)) / bet
        for k in range(1, km):
            d4 = D[k - 1] / D[k]
            bet = 2.0 + d4 + d4 - gam[k - 1]
            q[k] = (3.0 * (A[k - 1] + d4 * A[k]) - q[k - 1]) / bet
            gam[k] = d4 / bet
        d4 = D[km - 2] / D[km - 1]
        a_bot = 1.0 + d4 * (d4 + 1.5)
        q[km] = (2.0 * d4 * (d4 + 1.0) * A[km - 1] + A[km - 2] - a_bot * q[km - 1]) / \
            (d4 * (d4 + 0.5) - a_bot * gam[km - 1])
        for k in range(km - 1, -1, -1):
            q[k] = q[k] - gam[k] * q[k + 1]
    # apply_constraints: clamp the edge values between neighbouring means
    hi, lo, g = [None] * km, [None] * km, list(gam)
    for k in range(1, km):
        hi[k] = torch.where(A[k - 1] > A[k], A[k - 1], A[k])
        lo[k] = torch.where(A[k - 1] < A[k], A[k - 1], A[k])
        g[k] = A[k] - A[k - 1]

    def cap_hi(x, k):
        return torch.where(x >= hi[k], hi[k], x)

    def cap_lo(x, k):
        return torch.where(x <= lo[k], lo[k], x)

    q[1] = cap_lo(cap_hi(q[1], 1), 1)
    for k in range(2, km - 1):
        both = (g[k - 1] * g[k + 1]) > 0
        upmax = g[k - 1] > 0
        low_branch = cap_hi(q[k], k)
        if iv == 0:
            low_branch = torch.where(low_branch < 0.0, torch.zeros_like(low_branch), low_branch)
        q[k] = torch.where(both, cap_lo(cap_hi(q[k], k), k),
                           torch.where(upmax, cap_lo(q[k], k), low_branch))
    q[km - 1] = cap_lo(cap_hi(q[km - 1], km - 1), km - 1)
    a2 = [q[k] for k in range(km)]
    a3 = [q[k + 1] for k in range(km)]
    a4 = [torch.zeros_like(A[0]) for _ in range(km)]
    extm = [None] * km
    extm[0] = (a2[0] - A[0]) * (a3[0] - A[0]) > 0.0
    for k in range(1, km - 1):
        extm[k] = g[k] * g[k + 1] < 0.0
    extm[km - 1] = (a2[km - 1] - A[km - 1]) * (a3[km - 1] - A[km - 1]) > 0.0
    # set_interpolation_coefficients: top two levels by iv
    zero = torch.zeros_like(A[0])
    if iv == 0:
        a2[0] = torch.where(a2[0] < 0.0, zero, a2[0])
    if iv == -1:
        a2[0] = torch.where(a2[0] * A[0] <= 0.0, zero, a2[0])
    if iv == 2:
        a2[0], a3[0], a4[0] = A[0], A[0], zero
        a4[1] = _std_a4(A[1], a2[1], a3[1])
    else:
        a4[0] = _std_a4(A[0], a2[0], a3[0])
        a4[1] = _std_a4(A[1], a2[1], a3[1])
        a2[0], a3[0], a4[0] = _posdef_iv1(A[0], a2[0], a3[0], a4[0])
    a2[1], a3[1], a4[1] = _remap_constraint(A[1], a2[1], a3[1], a4[1], extm[1])
    # interior: the kord-9 limiter
    for k in range(2, km - 2):
        a1k, a2k, a3k = A[k], a2[k], a3[k]
        pmp_1 = a1k - 2.0 * g[k + 1]
        lac_1 = pmp_1 + 1.5 * g[k + 2]
        pmp_2 = a1k + 2.0 * g[k]
        lac_2 = pmp_2 - 1.5 * g[k - 1]
        flat = (extm[k] & extm[k - 1]) | (extm[k] & extm[k + 1])
        if qmin > 0.0:
            flat = flat | (extm[k] & (a1k < qmin))
        a4n = 6.0 * a1k - 3.0 * (a2k + a3k)
        steep = torch.abs(a4n) > torch.abs(a2k - a3k)
        lim2 = torch.where(torch.where(a2k > _min3(a1k, pmp_1, lac_1), a2k, _min3(a1k, pmp_1, lac_1))
                           < _max3(a1k, pmp_1, lac_1),
                           torch.where(a2k > _min3(a1k, pmp_1, lac_1), a2k,
                                       _min3(a1k, pmp_1, lac_1)),
                           _max3(a1k, pmp_1, lac_1))
        lim3 = torch.where(torch.where(a3k > _min3(a1k, pmp_2, lac_2), a3k, _min3(a1k, pmp_2, lac_2))
                           < _max3(a1k, pmp_2, lac_2),
                           torch.where(a3k > _min3(a1k, pmp_2, lac_2), a3k,
                                       _min3(a1k, pmp_2, lac_2)),
                           _max3(a1k, pmp_2, lac_2))
        n2 = torch.where(steep, lim2, a2k)
        n3 = torch.where(steep, lim3, a3k)
        n4 = torch.where(steep, 6.0 * a1k - 3.0 * (lim2 + lim3), a4n)
        a2[k] = torch.where(flat, a1k, n2)
        a3[k] = torch.where(flat, a1k, n3)
        a4[k] = torch.where(flat, zero, n4)
        if iv == 0:
            a2[k], a3[k], a4[k] = _posdef_iv0(a1k, a2[k], a3[k], a4[k])
    # bottom two levels
    b = km - 1
    if iv == 0:
        a3[b] = torch.where(a3[b] < 0.0, zero, a3[b])
    if iv == -1:
        a3[b] = torch.where(a3[b] * A[b] <= 0.0, zero, a3[b])
    for k in (km - 2, km - 1):
        a4[k] = _std_a4(A[k], a2[k], a3[k])
    a2[b - 1], a3[b - 1], a4[b - 1] = _remap_constraint(A[b - 1], a2[b - 1], a3[b - 1],
                                                        a4[b - 1], extm[b - 1])
    a2[b], a3[b], a4[b] = _posdef_iv1(A[b], a2[b], a3[b], a4[b])
    return (a1, torch.stack(a2, dim=1), torch.stack(a3, dim=1), torch.stack(a4, dim=1))


def overlap_operator(pe1, pe2):
    """Weights W [ncol, km, km, 4] with q_out[c, k] = sum_L sum_m
    W[c, k, L, m] a4_{m+1}[c, L]: lagrangian_contributions (map_single.py:
    19-66) written out as the integral of each source parabola
    a2 + x (a3 - a2 + a4 (1 - x)) ... over the part of the source layer that
    the target layer covers (x the position within the source layer)."""
    pe1 = np.asarray(pe1, dtype=np.float64)
    pe2 = np.asarray(pe2, dtype=np.float64)
    ncol, kp1 = pe1.shape
    km = kp1 - 1
    dp1 = pe1[:, 1:] - pe1[:, :-1]
    W = np.zeros((ncol, km, km, 4))
    for c in range(ncol):
        L = 0
        for k in range(km):
            pl = (pe2[c, k] - pe1[c, L]) / dp1[c, L]
            if pe2[c, k + 1] <= pe1[c, L + 1]:
                pr = (pe2[c, k + 1] - pe1[c, L]) / dp1[c, L]
                s = pr + pl
                W[c, k, L] += [0.0, 1.0 - 0.5 * s, 0.5 * s,
                               0.5 * s - (pr * s + pl * pl) / 3.0]
                continue
            span = pe2[c, k + 1] - pe2[c, k]
            w = (pe1[c, L + 1] - pe2[c, k]) / span
            s = 1.0 + pl
            W[c, k, L] += w * np.array([0.0, 1.0 - 0.5 * s, 0.5 * s,
                                        0.5 * s - (1.0 + pl * s) / 3.0])
            L += 1
            while pe1[c, L + 1] < pe2[c, k + 1]:
                W[c, k, L, 0] += dp1[c, L] / span
                L += 1
            d = pe2[c, k + 1] - pe1[c, L]
            esl = d / dp1[c, L]
            W[c, k, L] += (d / span) * np.array([0.0, 1.0 - 0.5 * esl, 0.5 * esl,
                                                 0.5 * esl * (1.0 - (2.0 / 3.0) * esl)])
    return W


def map_scalar(q, pe1, pe2, iv, qs=None, W=None):
    """MapSingle with kord 9 on columns: q [ncol, km] (torch) from the edges
    pe1 to pe2 ([ncol, km + 1], numpy); returns [ncol, km]."""
    dp = torch.as_tensor(np.diff(np.asarray(pe1), axis=1), dtype=q.dtype, device=q.device)
    qs_t = torch.zeros_like(q[:, 0]) if qs is None else torch.as_tensor(qs, dtype=q.dtype,
                                                                       device=q.device)
    a = profile_kord9(qs_t, q, dp, iv)
    if W is None:
        W = overlap_operator(pe1, pe2)
    Wt = torch.as_tensor(W, dtype=q.dtype, device=q.device)
    coeffs = torch.stack(a, dim=2)  # [ncol, km, 4]
    return torch.einsum("cklm,clm->ck", Wt, coeffs)
