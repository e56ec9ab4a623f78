"""The remap entry point that allocates nothing, against the one that takes a workspace.

fv3_map_single_noalloc_f64 runs the same column kernel as fv3_map_single_f64
with its slabs in the library's file-scope scratch array, so the two must
agree bit for bit.  Two calls in a row, the second on a smaller input, check
that the shared scratch carries nothing over from one call to the next.
"""
import ctypes

import numpy as np
import pytest
import torch

from fv3core_amd import _lib
from fv3core_amd.utils import gt4py_utils as utils
from fv3core_amd.utils.stencil import GridIndexing
from fv3core_amd.workloads import remap_inputs
from oracle import remap as orm
from tests.helpers import assert_close, to_np

pytestmark = pytest.mark.gpu

NZ = 79
_P = ctypes.c_void_p
_I = ctypes.c_int32
SCRATCH_BYTES = 96 << 20


def _noalloc():
    fn = _lib.lib().fv3_map_single_noalloc_f64
    fn.argtypes = [_P, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, ctypes.c_double, _P]
    fn.restype = _I
    return fn


def _with_work():
    fn = _lib.lib().fv3_map_single_f64
    fn.argtypes = [_P, _I, _I, _I, _I, _I, _I, _P, _P, _P, _P, ctypes.c_double, _P,
                   ctypes.c_int64, _P]
    fn.restype = _I
    return fn


def _batch(n, ntiles, workload):
    gi = GridIndexing((n, n, NZ), 3, True, True, True, True)
    ins = [remap_inputs(n, NZ, workload=workload + t) for t in range(ntiles)]

    def stack(name):
        b = utils.empty((ntiles,) + gi.max_shape)
        for t in range(ntiles):
            b[t].copy_(torch.as_tensor(ins[t][name]))
        return b

    qs = utils.empty((ntiles,) + gi.max_shape[:2], batch_dims=1)
    for t in range(ntiles):
        qs[t].copy_(torch.as_tensor(ins[t]["qs"]))
    return gi, ins, stack("pt"), stack("w"), stack("pe1"), stack("pe2"), qs


def _run_both(gi, q_src, pe1, pe2, qs, kord, mode, qmin):
    """(no-alloc result, workspace result) of one remap of q_src."""
    d = _lib.make_domain(gi, q_src, NZ, qs, others3=(pe1, pe2))
    box = (gi.isc, gi.iec, gi.jsc, gi.jec)
    stream = ctypes.c_void_p(_lib.stream_handle())
    q_a = q_src.clone()
    rc = _noalloc()(ctypes.byref(d), kord, mode, *box, _lib.ptr(q_a), _lib.ptr(pe1),
                    _lib.ptr(pe2), _lib.ptr(qs), ctypes.c_double(qmin), stream)
    _lib.check(rc)
    _, _, _, _, _, sk, st = _lib.layout3(q_src)
    slab = (q_src.shape[0] - 1) * st + (NZ + 1) * sk
    work = torch.empty(6 * slab, dtype=torch.float64, device=q_src.device)
    q_b = q_src.clone()
    rc = _with_work()(ctypes.byref(d), kord, mode, *box, _lib.ptr(q_b), _lib.ptr(pe1),
                      _lib.ptr(pe2), _lib.ptr(qs), ctypes.c_double(qmin), _lib.ptr(work),
                      ctypes.c_int64(slab), stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    return q_a, q_b


def test_noalloc_remap_equals_workspace_form_on_two_calls():
    # call 1: three 40 x 40 tiles, kord 10, iv 1 (fills most of the scratch
    # rows the second call reuses)
    gi, _, pt, _, pe1, pe2, qs = _batch(40, 3, workload=300)
    a, b = _run_both(gi, pt, pe1, pe2, qs, 10, 1, 0.0)
    assert torch.equal(a, b), "call 1: no-alloc remap differs from the workspace form"
    # call 2: one smaller tile, another limiter and the iv = -2 bottom value
    gi2, ins2, _, w2, pe1b, pe2b, qs2 = _batch(12, 1, workload=310)
    a2, b2 = _run_both(gi2, w2, pe1b, pe2b, qs2, 9, -2, 0.0)
    assert torch.equal(a2, b2), "call 2: no-alloc remap differs from the workspace form"
    # and both are the oracle's remap
    ref = orm.map_single(ins2[0]["w"], ins2[0]["pe1"], ins2[0]["pe2"], 9, -2, gi2.isc, gi2.iec,
                         gi2.jsc, gi2.jec, qs=ins2[0]["qs"])
    assert_close(to_np(a2[0])[..., :NZ], ref[..., :NZ], 1e-12, near_zero=1e-18, what="w")


def test_noalloc_remap_rejects_inputs_beyond_the_scratch():
    # six 128 x 128 tiles need 6 slabs of (5*st + 80*sk) doubles: far above 96 MiB
    n, nt = 128, 6
    gi = GridIndexing((n, n, NZ), 3, True, True, True, True)
    q = utils.zeros((nt,) + gi.max_shape)
    pe = utils.zeros((nt,) + gi.max_shape)
    d = _lib.make_domain(gi, q, NZ, None, others3=(pe, pe))
    _, _, _, _, _, sk, st = _lib.layout3(q)
    assert 6 * 8 * ((nt - 1) * st + (NZ + 1) * sk) > SCRATCH_BYTES
    before = q.clone()
    rc = _noalloc()(ctypes.byref(d), 9, 1, gi.isc, gi.iec, gi.jsc, gi.jec, _lib.ptr(q),
                    _lib.ptr(pe), _lib.ptr(pe), None, ctypes.c_double(0.0),
                    ctypes.c_void_p(_lib.stream_handle()))
    assert rc == _lib.E_ARG
    assert "maximum" in _lib.lib().fv3_last_error().decode()
    torch.cuda.synchronize()
    assert torch.equal(q, before)
    assert np.isfinite(to_np(q)).all()
