"""ctypes binding of ``libfv3core_amd.so`` and tensor -> C-ABI marshalling.

The library is built in-tree by :mod:`fv3core_amd.build`.  Loading it never
falls back to anything else: a missing library or a CPU tensor raises.
"""
from __future__ import annotations

import ctypes
import pathlib
from typing import Optional, Sequence

import torch

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libfv3core_amd.so"

FV3_NO_EDGE_LO = -(1 << 28)
FV3_NO_EDGE_HI = 1 << 28

DELN_ADD = 0
DELN_MASS = 1
DELN_STORE = 2
DELN_STORE_MASS = 3


class Domain(ctypes.Structure):
    """Mirror of ``fv3_domain`` in include/fv3core_amd.h."""

    _fields_ = [
        ("isc", ctypes.c_int32),
        ("iec", ctypes.c_int32),
        ("jsc", ctypes.c_int32),
        ("jec", ctypes.c_int32),
        ("i_start", ctypes.c_int32),
        ("i_end", ctypes.c_int32),
        ("j_start", ctypes.c_int32),
        ("j_end", ctypes.c_int32),
        ("ni", ctypes.c_int32),
        ("nj", ctypes.c_int32),
        ("nk", ctypes.c_int32),
        ("ntiles", ctypes.c_int32),
        ("sj", ctypes.c_int64),
        ("sk", ctypes.c_int64),
        ("st", ctypes.c_int64),
        ("sj2", ctypes.c_int64),
        ("st2", ctypes.c_int64),
    ]


_lib: Optional[ctypes.CDLL] = None

_P = ctypes.c_void_p
_I = ctypes.c_int32

# name -> argtypes (after the domain pointer where one is taken)
_SIGNATURES = {
    "fv3_version": ([], ctypes.c_char_p),
    "fv3_last_error": ([], ctypes.c_char_p),
    "fv3_device_count": ([], _I),
    "fv3_fvtp2d_workspace_slabs": ([], _I),
    "fv3_map_single_tlm_workspace_slabs": ([], _I),
    "fv3_map_single_adjoint_workspace_slabs": ([], _I),
}


def _declare(lib: ctypes.CDLL) -> None:
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def exported_symbols() -> Sequence[str]:
    """Names of every function include/fv3core_amd.h declares."""
    header = _HERE.parent / "include" / "fv3core_amd.h"
    import re

    text = header.read_text()
    return sorted(set(re.findall(r"\b(fv3_[a-z0-9_]+)\s*\(", text)))


def lib() -> ctypes.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m fv3core_amd.build`"
            )
        handle = ctypes.CDLL(str(LIB_PATH))
        _declare(handle)
        _lib = handle
    return _lib


# return codes of the C ABI (include/fv3core_amd.h)
E_ARG, E_UNSUPPORTED, E_HIP = -1, -2, -3


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().fv3_last_error().decode()
        if rc == -2:
            raise NotImplementedError(msg)
        raise RuntimeError(f"fv3core_amd error {rc}: {msg}")


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def require_device(*tensors) -> None:
    for t in tensors:
        if t is None:
            continue
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"expected a torch.Tensor, got {type(t).__name__}")
        if not t.is_cuda:
            raise RuntimeError(
                "fv3core_amd kernels run on the GPU only; got a tensor on "
                f"{t.device} (there is no CPU path)"
            )


def ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def layout3(t: torch.Tensor):
    """(ntiles, ni, nj, nk, sj, sk, st) of a 3-D (I,J,K) or 4-D (T,I,J,K) field."""
    if t.dim() == 3:
        ni, nj, nk = t.shape
        si, sj, sk = t.stride()
        ntiles, st = 1, 0
    elif t.dim() == 4:
        ntiles, ni, nj, nk = t.shape
        st, si, sj, sk = t.stride()
    else:
        raise ValueError(f"expected a 3-D or 4-D field, got shape {tuple(t.shape)}")
    if si != 1:
        raise ValueError("fields must have unit stride in i (i-fastest layout)")
    return ntiles, ni, nj, nk, sj, sk, st


def layout2(t: torch.Tensor):
    """(ntiles, ni, nj, sj2, st2) of an IJ field (I,J) or (T,I,J)."""
    if t.dim() == 2:
        ni, nj = t.shape
        si, sj = t.stride()
        ntiles, st = 1, 0
    elif t.dim() == 3:
        ntiles, ni, nj = t.shape
        st, si, sj = t.stride()
    else:
        raise ValueError(f"expected an IJ field, got shape {tuple(t.shape)}")
    if si != 1:
        raise ValueError("IJ fields must have unit stride in i")
    return ntiles, ni, nj, sj, st


def make_domain(grid_indexing, ref3: torch.Tensor, nk: int, ref2: Optional[torch.Tensor] = None,
                others3: Sequence[Optional[torch.Tensor]] = (),
                others2: Sequence[Optional[torch.Tensor]] = ()) -> Domain:
    """Domain descriptor from a GridIndexing and the fields of one call.

    Every 3-D field of the call must share ``ref3``'s strides and every IJ
    field ``ref2``'s, since the kernels address them with one set of strides.
    """
    ntiles, ni, nj, nk_store, sj, sk, st = layout3(ref3)
    if nk > nk_store:
        raise ValueError(f"{nk} levels requested but the storage has {nk_store}")
    for o in others3:
        if o is None:
            continue
        lo = layout3(o)
        if lo[0] != ntiles or lo[1:3] != (ni, nj) or lo[4:] != (sj, sk, st) or lo[3] < nk:
            raise ValueError(
                f"field layout {tuple(o.shape)}/{o.stride()} differs from "
                f"{tuple(ref3.shape)}/{ref3.stride()}"
            )
    sj2, st2 = sj, 0
    if ref2 is not None:
        nt2, ni2, nj2, sj2, st2 = layout2(ref2)
        if (ni2, nj2) != (ni, nj) or (nt2 != ntiles and st2 != 0):
            raise ValueError("IJ field does not match the 3-D fields' horizontal shape")
        if nt2 == 1 and ntiles > 1:
            st2 = 0
        for o in others2:
            if o is None:
                continue
            lo = layout2(o)
            if lo[1:3] != (ni, nj) or lo[3] != sj2 or (lo[4] != st2 and lo[0] != 1):
                raise ValueError("IJ fields of one call must share a layout")
    g = grid_indexing
    edges = g.edge_anchors()
    d = Domain()
    d.isc, d.iec, d.jsc, d.jec = g.isc, g.iec, g.jsc, g.jec
    d.i_start = edges["i_start"] if edges["i_start"] is not None else FV3_NO_EDGE_LO
    d.i_end = edges["i_end"] if edges["i_end"] is not None else FV3_NO_EDGE_HI
    d.j_start = edges["j_start"] if edges["j_start"] is not None else FV3_NO_EDGE_LO
    d.j_end = edges["j_end"] if edges["j_end"] is not None else FV3_NO_EDGE_HI
    d.ni, d.nj, d.nk, d.ntiles = ni, nj, nk, ntiles
    d.sj, d.sk, d.st, d.sj2, d.st2 = sj, sk, st, sj2, st2
    return d


def fn(name: str, dtype: torch.dtype):
    """The ``_f64``/``_f32`` variant of an entry point for a dtype."""
    suffix = {torch.float64: "_f64", torch.float32: "_f32"}.get(dtype)
    if suffix is None:
        raise TypeError(f"unsupported dtype {dtype}")
    return getattr(lib(), name + suffix)


METRIC_NAMES = (
    "dx", "dy", "dxa", "dya", "dxc", "dyc",
    "rdx", "rdy", "rdxa", "rdya", "rdxc", "rdyc",
    "area", "rarea", "area_c", "rarea_c",
    "sin_sg1", "sin_sg2", "sin_sg3", "sin_sg4",
    "cos_sg1", "cos_sg2", "cos_sg3", "cos_sg4",
    "cosa", "sina", "cosa_u", "cosa_v", "cosa_s", "sina_u", "sina_v",
    "rsin_u", "rsin_v", "rsina", "rsin2",
    "fC", "f0", "del6_u", "del6_v", "divg_u", "divg_v",
    "agrid1", "agrid2", "bgrid1", "bgrid2",
    "edge_w", "edge_e", "edge_s", "edge_n",
)


class Metrics(ctypes.Structure):
    """Mirror of ``fv3_metrics`` in include/fv3core_amd.h."""

    _fields_ = [(n, ctypes.c_void_p) for n in METRIC_NAMES]


def metrics_struct(grid_data, dtype: torch.dtype, names: Sequence[str]):
    """(Metrics, ref2) for the named GridData fields, checked for one layout.

    Every named field must exist, be a device tensor of ``dtype`` and share
    the first field's IJ layout; the others stay NULL.
    """
    m = Metrics()
    ref2 = None
    for n in names:
        if not grid_data.has(n):
            raise ValueError(f"GridData lacks the metric term {n!r}")
        t = getattr(grid_data, n)
        require_device(t)
        if t.dtype != dtype:
            raise TypeError(f"metric {n} is {t.dtype}, the call computes in {dtype}")
        if ref2 is None:
            ref2 = t
        elif layout2(t)[1:] != layout2(ref2)[1:] or t.shape != ref2.shape:
            raise ValueError(f"metric {n} does not share the IJ layout of {names[0]}")
        setattr(m, n, t.data_ptr())
    return m, ref2
