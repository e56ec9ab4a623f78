"""Shared helpers of the CPU oracle (test infrastructure only).

The oracle is a numpy FP64 restatement of the reference's stencils.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker; the product path never
calls it.

GT4Py semantics restated here (gt4py v36, the reference's pinned version,
GT4PY_VERSION.txt:1):
* inside ``computation(PARALLEL)`` each statement finishes over its whole
  (extended) domain before the next starts: the oracle evaluates every
  statement on the full array;
* ``horizontal(region[...])`` writes only at absolute storage indices derived
  from the tile-edge anchors (fv3core/utils/stencil.py:457-495): the oracle
  uses boolean masks on the absolute index;
* offset reads that fall outside the array read NaN here (``sh``), so a valid
  output that depends on data nobody computed shows up as NaN;
* ``FORWARD``/``BACKWARD`` k loops are plain Python loops over k.
"""
from __future__ import annotations

import numpy as np


def sh(a: np.ndarray, di: int = 0, dj: int = 0) -> np.ndarray:
    """Return ``b`` with ``b[i, j, ...] = a[i + di, j + dj, ...]``, NaN outside."""
    out = np.full(a.shape, np.nan, dtype=np.float64)
    ni, nj = a.shape[0], a.shape[1]
    i0, i1 = max(0, -di), min(ni, ni - di)
    j0, j1 = max(0, -dj), min(nj, nj - dj)
    if i1 > i0 and j1 > j0:
        out[i0:i1, j0:j1] = a[i0 + di : i1 + di, j0 + dj : j1 + dj]
    return out


def shk(a: np.ndarray, dk: int) -> np.ndarray:
    """Return ``b`` with ``b[..., k] = a[..., k + dk]``, NaN outside."""
    out = np.full(a.shape, np.nan, dtype=np.float64)
    nk = a.shape[-1]
    k0, k1 = max(0, -dk), min(nk, nk - dk)
    if k1 > k0:
        out[..., k0:k1] = a[..., k0 + dk : k1 + dk]
    return out


def imask(shape, i) -> np.ndarray:
    """Mask of the cells whose first index is ``i`` (False everywhere if None)."""
    m = np.zeros(shape, dtype=bool)
    if i is not None and 0 <= i < shape[0]:
        m[i] = True
    return m


def jmask(shape, j) -> np.ndarray:
    m = np.zeros(shape, dtype=bool)
    if j is not None and 0 <= j < shape[1]:
        m[:, j] = True
    return m


def box(shape, i0, i1, j0, j1) -> np.ndarray:
    """Mask of the absolute box [i0, i1] x [j0, j1] (inclusive)."""
    m = np.zeros(shape, dtype=bool)
    i0c, j0c = max(i0, 0), max(j0, 0)
    if i1 >= i0c and j1 >= j0c:
        m[i0c : i1 + 1, j0c : j1 + 1] = True
    return m


def restrict(a: np.ndarray, i0, i1, j0, j1) -> np.ndarray:
    """NaN everywhere outside the absolute box [i0, i1] x [j0, j1].

    Used to model the fact that a stencil only writes its own domain, so
    any later read outside it would see stale memory in the reference.
    """
    out = np.full(a.shape, np.nan, dtype=np.float64)
    i0c, j0c = max(i0, 0), max(j0, 0)
    out[i0c : i1 + 1, j0c : j1 + 1] = a[i0c : i1 + 1, j0c : j1 + 1]
    return out


def sign(a, b):
    """``|a|`` if ``b > 0`` else ``-|a|`` (fv3core/stencils/basic_operations.py:33-40)."""
    return np.where(b > 0, np.abs(a), -np.abs(a))


def ij3(a: np.ndarray) -> np.ndarray:
    """Broadcast an IJ field against IJK fields."""
    return a[:, :, None] if a.ndim == 2 else a


def transpose_ij(a):
    if a is None:
        return None
    if a.ndim == 2:
        return a.T
    return np.swapaxes(a, 0, 1)


def finite_on(a: np.ndarray, mask2: np.ndarray, what: str) -> np.ndarray:
    """Return ``a`` after checking that it is finite wherever ``mask2`` (an IJ
    mask) is set.  Expressions evaluated over the whole array divide by
    zero-valued halo points outside the region a stencil writes; they are
    computed under ``np.errstate`` and this check makes sure no inf or NaN
    reaches the written region."""
    bad = ~np.isfinite(a[mask2])
    if bad.any():
        raise FloatingPointError(f"{what}: {int(bad.sum())} non-finite values in the written region")
    return a
