"""Storage helpers with the reference's names and semantics, on torch tensors.

Mirrors fv3core/utils/gt4py_utils.py:22-316: ``halo = 3``, default origin
``(3, 3, 0)``, ``make_storage_from_shape`` (optionally memoised by
``cache_key``) and ``make_storage_data``.  A storage is a torch tensor of the
requested shape whose memory is i-fastest: a 3-D ``(I, J, K)`` storage has
strides ``(1, I, I*J)``.  That is the layout the HIP kernels coalesce on.
Storages live on the current GPU when one is present (the kernels need it);
on a machine without a GPU they are CPU tensors, usable for host logic only.
"""
from __future__ import annotations

from typing import Hashable, Optional, Tuple

import numpy as np
import torch

from ..constants import TRACER_VARIABLES

halo = 3
origin = (halo, halo, 0)
tracer_variables = list(TRACER_VARIABLES)

_DTYPES = {
    np.float64: torch.float64,
    np.float32: torch.float32,
    np.int32: torch.int32,
    np.int64: torch.int64,
    np.bool_: torch.bool,
    float: torch.float64,
    int: torch.int64,
    bool: torch.bool,
}


def default_device() -> torch.device:
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


def to_torch_dtype(dtype) -> torch.dtype:
    if isinstance(dtype, torch.dtype):
        return dtype
    if dtype in _DTYPES:
        return _DTYPES[dtype]
    return _DTYPES[np.dtype(dtype).type]


def empty(shape, dtype=np.float64, device=None, batch_dims=None) -> torch.Tensor:
    """Uninitialised storage with i-fastest memory order.

    The last ``len(shape) - batch_dims`` axes are the field axes (I, J[, K]),
    stored with I fastest; leading batch axes (tiles, members) are slowest.
    By default a 4-D shape has one batch axis and 1- to 3-D shapes none.
    """
    device = default_device() if device is None else device
    tdt = to_torch_dtype(dtype)
    shape = tuple(int(s) for s in shape)
    if batch_dims is None:
        batch_dims = max(0, len(shape) - 3)
    field = shape[batch_dims:]
    if len(field) <= 1:
        return torch.empty(shape, dtype=tdt, device=device)
    # allocate the field axes reversed and permute them back: I becomes fastest
    rev = torch.empty(shape[:batch_dims] + field[::-1], dtype=tdt, device=device)
    perm = list(range(batch_dims)) + [batch_dims + len(field) - 1 - a for a in range(len(field))]
    return rev.permute(*perm)


def zeros(shape, dtype=np.float64, device=None, batch_dims=None) -> torch.Tensor:
    t = empty(shape, dtype, device, batch_dims)
    t.zero_()
    return t


def make_storage_from_shape_uncached(
    shape: Tuple[int, ...],
    origin: Tuple[int, ...] = origin,
    *,
    dtype=np.float64,
    init: bool = False,
    mask=None,
    device=None,
) -> torch.Tensor:
    """New storage of ``shape`` (gt4py_utils.py:221-262); zeroed if ``init``."""
    return zeros(shape, dtype, device) if init else empty(shape, dtype, device)


_storage_cache = {}


def make_storage_from_shape(
    shape: Tuple[int, ...],
    origin: Tuple[int, ...] = origin,
    *,
    dtype=np.float64,
    init: bool = False,
    mask=None,
    cache_key: Optional[Hashable] = None,
    device=None,
) -> torch.Tensor:
    """Storage of ``shape``, memoised by ``cache_key`` (gt4py_utils.py:268-316)."""
    if cache_key is None:
        return make_storage_from_shape_uncached(
            shape, origin, dtype=dtype, init=init, mask=mask, device=device
        )
    key = (tuple(shape), tuple(origin), cache_key, str(dtype), init, mask)
    if key not in _storage_cache:
        _storage_cache[key] = make_storage_from_shape_uncached(
            shape, origin, dtype=dtype, init=init, mask=mask, device=device
        )
    out = _storage_cache[key]
    if init:
        out.zero_()
    return out


def make_storage_data(
    data,
    shape: Optional[Tuple[int, ...]] = None,
    origin: Tuple[int, ...] = origin,
    *,
    dtype=np.float64,
    mask=None,
    start: Tuple[int, ...] = (0, 0, 0),
    device=None,
) -> torch.Tensor:
    """Storage holding ``data`` at ``start`` inside ``shape`` (zero elsewhere)."""
    arr = data.detach().cpu().numpy() if isinstance(data, torch.Tensor) else np.asarray(data)
    if shape is None:
        shape = arr.shape
    out = zeros(shape, dtype, device)
    sl = tuple(slice(s, s + n) for s, n in zip(start, arr.shape))
    out[sl] = torch.as_tensor(arr, dtype=out.dtype).to(out.device)
    return out


def asarray(t) -> np.ndarray:
    """numpy copy of a storage."""
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().numpy()
    return np.asarray(t)


def device_sync() -> None:
    if torch.cuda.is_available():
        torch.cuda.synchronize()
