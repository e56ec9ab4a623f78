"""Comparison helpers shared by the tests.

The relative error is the reference's translate-test metric
2|c - r| / (|c| + |r|), zero where both are zero
(tests/savepoint/test_translate.py:26-33 of the reference).
"""
import numpy as np
import torch


def rel_err(c, r):
    c = np.asarray(c, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    denom = np.abs(c) + np.abs(r)
    with np.errstate(invalid="ignore", divide="ignore"):
        e = np.where(denom > 0, 2.0 * np.abs(c - r) / denom, 0.0)
    both_nan = np.isnan(c) & np.isnan(r)
    return np.where(both_nan, 0.0, e)


def assert_close(c, r, max_error, near_zero=0.0, what="", scale_floor=0.0):
    """Pass where 2|c-r|/(|c|+|r|) <= max_error.

    ``scale_floor`` (a fraction) floors the denominator at
    ``scale_floor * 2 max|r|``: used only for fields that are small
    differences of large terms, where the last-bit differences of exp/log
    between libraries are amplified by cancellation.
    """
    c = np.asarray(c, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    assert c.shape == r.shape, f"{what}: shape {c.shape} vs {r.shape}"
    assert not np.isnan(r).any(), f"{what}: reference has NaN"
    if scale_floor > 0:
        floor = scale_floor * 2.0 * np.abs(r).max()
        denom = np.maximum(np.abs(c) + np.abs(r), floor)
        e = np.where(denom > 0, 2.0 * np.abs(c - r) / np.where(denom > 0, denom, 1.0), 0.0)
    else:
        e = rel_err(c, r)
    if near_zero > 0:
        e = np.where((np.abs(c) < near_zero) & (np.abs(r) < near_zero), 0.0, e)
    bad = e > max_error
    if bad.any():
        idx = np.argwhere(bad)[:5]
        raise AssertionError(
            f"{what}: {bad.sum()} of {bad.size} points exceed {max_error} "
            f"(max {e.max():.3e}); first {idx.tolist()}: "
            f"got {[c[tuple(i)] for i in idx]}, want {[r[tuple(i)] for i in idx]}"
        )
    return float(e.max()) if e.size else 0.0


def to_np(t):
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().numpy()
    return np.asarray(t)
