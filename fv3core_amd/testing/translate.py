"""Savepoint (translate) comparison harness with the reference's semantics.

The reference checks each component against Fortran savepoints: it reads a
component's inputs at an "-In" savepoint, runs the Python component, and
compares every output field with the "-Out" savepoint by the metric
``2 |c - r| / (|c| + |r|)`` against a per-component ``max_error``, with
near-zero values optionally ignored and thresholds overridable per backend
and platform from a YAML file (tests/savepoint/test_translate.py:26-225,
fv3core/testing/translate.py:26-248, translate/overrides/*.yaml).  This
module restates that machinery for storages held as torch tensors:

* the metric and the pass / failure-report functions;
* the threshold-override YAML (safe loader);
* ``TranslateFortranData2Py``, the base of the per-component translate
  classes: serialized names, start indices, k-axis moves, storage
  placement and output slicing;
* ``TranslateGrid``, the index facts and GridData a translate class reads.

Readers of the savepoint data are in :mod:`fv3core_amd.testing.savepoint`.
"""
from __future__ import annotations

import os
from typing import Dict, List, Optional, Union

import numpy as np
import torch

from ..utils import gt4py_utils as utils

GPU_MAX_ERR = 1e-10
GPU_NEAR_ZERO = 1e-15


# ------------------------------------------------------------ the metric
def compare_arr(computed_data, ref_data):
    """2 |c - r| / (|c| + |r|), zero where both are zero."""
    computed_data = np.asarray(computed_data, dtype=np.float64)
    ref_data = np.asarray(ref_data, dtype=np.float64)
    denom = np.abs(ref_data) + np.abs(computed_data)
    with np.errstate(divide="ignore", invalid="ignore"):
        compare = 2.0 * np.abs(computed_data - ref_data) / denom
    compare[denom == 0] = 0.0
    return compare


def compare_scalar(computed_data, ref_data):
    return compare_arr(np.atleast_1d(computed_data), np.atleast_1d(ref_data))[0]


def success_array(computed_data, ref_data, eps, ignore_near_zero_errors: Union[dict, bool],
                  near_zero: float):
    """Points that pass: both NaN, metric below eps, or (when asked) both
    below the near-zero threshold in magnitude."""
    computed_data = np.asarray(computed_data, dtype=np.float64)
    ref_data = np.asarray(ref_data, dtype=np.float64)
    ok = np.logical_or(np.logical_and(np.isnan(computed_data), np.isnan(ref_data)),
                       compare_arr(computed_data, ref_data) < eps)
    if isinstance(ignore_near_zero_errors, dict):
        if ignore_near_zero_errors.keys():
            near_zero = ignore_near_zero_errors["near_zero"]
            ok = ok | ((np.abs(computed_data) < near_zero) & (np.abs(ref_data) < near_zero))
    elif ignore_near_zero_errors:
        ok = ok | ((np.abs(computed_data) < near_zero) & (np.abs(ref_data) < near_zero))
    return ok


def success(computed_data, ref_data, eps, ignore_near_zero_errors, near_zero=0.0) -> bool:
    return bool(np.all(success_array(computed_data, ref_data, eps, ignore_near_zero_errors,
                                     near_zero)))


def sample_wherefail(computed_data, ref_data, eps, print_failures, failure_stride, test_name,
                     ignore_near_zero_errors, near_zero, xy_indices=False) -> str:
    """Failure report: (every failure_stride-th) failing point, the count,
    the worst point by the metric, and optionally the failing (i, j)."""
    computed_data = np.asarray(computed_data, dtype=np.float64)
    ref_data = np.asarray(ref_data, dtype=np.float64)
    bad = ~success_array(computed_data, ref_data, eps, ignore_near_zero_errors, near_zero)
    found = np.where(bad)
    lines: List[str] = []
    worst = (-1.0, None, 0.0, 0.0, 0.0)
    for b in range(len(found[0])):
        idx = tuple(int(f[b]) for f in found)
        c, r = computed_data[idx], ref_data[idx]
        metric = compare_scalar(c, r)
        absd = abs(c - r)
        if print_failures and b % failure_stride == 0:
            lines.append(f"index: {list(idx)}, computed {c}, reference {r}, absolute diff "
                         f"{absd:.3e}, metric diff: {metric:.3e}")
        if metric > worst[0]:
            worst = (metric, idx, absd, c, r)
    full = ref_data.size
    lines.append(f"Failed count: {len(found[0])}/{full} "
                 f"({round(100.0 * len(found[0]) / max(full, 1), 2)}%),\n"
                 f"Worst failed index {list(worst[1]) if worst[1] is not None else None}\n"
                 f"\tcomputed:{worst[3]}\n\treference: {worst[4]}\n"
                 f"\tabsolute diff: {worst[2]:.3e}\n\tmetric diff: {worst[0]:.3e}\n")
    if xy_indices and computed_data.ndim >= 3:
        axes = tuple(range(2, computed_data.ndim))
        xy = np.where(np.any(bad, axis=axes))
        lines.append("failed horizontal indices:" + str(list(zip(*[a.tolist() for a in xy]))))
    return "\n".join(lines)


# ------------------------------------------------------- threshold overrides
def platform() -> str:
    return "docker" if os.environ.get("IN_DOCKER", False) else "metal"


def load_threshold_overrides(path) -> Optional[dict]:
    """The override YAML (translate/overrides/*.yaml), read with the safe loader."""
    if path is None:
        return None
    import yaml

    with open(path) as f:
        return yaml.safe_load(f)


def process_override(threshold_overrides, testobj, test_name: str, backend: str,
                     platform_name: Optional[str] = None) -> None:
    """Apply the one override entry matching (backend, platform) of
    ``test_name`` to ``testobj`` (max_error, near_zero,
    ignore_near_zero_errors as a list of fields or a dict of per-field
    near-zero thresholds with an optional all_other_near_zero)."""
    override = (threshold_overrides or {}).get(test_name)
    if override is None:
        return
    plat = platform_name or platform()
    for spec in override:
        spec.setdefault("platform", plat)
        spec.setdefault("backend", backend)
    matches = [s for s in override if s["backend"] == backend and s["platform"] == plat]
    if len(matches) > 1:
        raise ValueError("misconfigured threshold overrides file, more than 1 specification "
                         f"for {test_name} with backend={backend}, platform={plat}")
    if not matches:
        return
    match = matches[0]
    if "max_error" in match:
        testobj.max_error = float(match["max_error"])
    if "near_zero" in match:
        testobj.near_zero = float(match["near_zero"])
    if "ignore_near_zero_errors" in match:
        parsed = match["ignore_near_zero_errors"]
        if isinstance(parsed, list):
            testobj.ignore_near_zero_errors = {field: True for field in parsed}
        elif isinstance(parsed, dict):
            testobj.ignore_near_zero_errors = {
                field: {"near_zero": float(v)} for field, v in parsed.items()}
            if "all_other_near_zero" in match:
                for key in testobj.out_vars:
                    if key not in testobj.ignore_near_zero_errors:
                        testobj.ignore_near_zero_errors[key] = {
                            "near_zero": float(match["all_other_near_zero"])}
        else:
            raise TypeError("ignore_near_zero_errors is either a list or a dict")


# ------------------------------------------------------------------ the grid
class TranslateGrid:
    """Index facts of one tile (the reference Grid's is_/ie/js/je/isd/...
    and its index dictionaries, fv3core/utils/grid.py:230-378) plus the
    objects a component is built from."""

    def __init__(self, stencil_factory, grid_data=None, damping_coefficients=None,
                 grid_type: int = 0):
        gi = stencil_factory.grid_indexing
        self.stencil_factory = stencil_factory
        self.grid_indexing = gi
        self.grid_data = grid_data
        self.damping_coefficients = damping_coefficients
        self.grid_type = grid_type
        h = gi.n_halo
        self.halo = h
        self.nic, self.njc, self.npz = gi.domain
        self.is_, self.js = h, h
        self.ie, self.je = h + self.nic - 1, h + self.njc - 1
        self.isd, self.jsd = 0, 0
        self.ied, self.jed = self.ie + h, self.je + h
        self.nid, self.njd = self.nic + 2 * h, self.njc + 2 * h

    def full_origin(self):
        return (self.isd, self.jsd, 0)

    def domain_shape_full(self, *, add=(0, 0, 0)):
        return (self.nid + add[0], self.njd + add[1], self.npz + add[2])

    def domain_shape_compute(self, *, add=(0, 0, 0)):
        return (self.nic + add[0], self.njc + add[1], self.npz + add[2])

    def default_domain_dict(self):
        return {"istart": self.isd, "iend": self.ied, "jstart": self.jsd, "jend": self.jed,
                "kstart": 0, "kend": self.npz - 1}

    def default_dict_buffer_2d(self):
        d = self.default_domain_dict()
        d["iend"] += 1
        d["jend"] += 1
        return d

    def compute_dict(self):
        return {"istart": self.is_, "iend": self.ie, "jstart": self.js, "jend": self.je,
                "kstart": 0, "kend": self.npz - 1}

    def compute_dict_buffer_2d(self):
        d = self.compute_dict()
        d["iend"] += 1
        d["jend"] += 1
        return d

    def default_buffer_k_dict(self):
        d = self.default_domain_dict()
        d["kend"] = self.npz
        return d

    def compute_buffer_k_dict(self):
        d = self.compute_dict()
        d["kend"] = self.npz
        return d

    def x3d_compute_dict(self):
        return {**self.default_domain_dict(), "istart": self.is_, "iend": self.ie + 1,
                "jstart": self.js, "jend": self.je}

    def y3d_compute_dict(self):
        return {**self.default_domain_dict(), "istart": self.is_, "iend": self.ie,
                "jstart": self.js, "jend": self.je + 1}

    def x3d_compute_domain_y_dict(self):
        return {**self.default_domain_dict(), "istart": self.is_, "iend": self.ie + 1,
                "jstart": self.jsd, "jend": self.jed}

    def y3d_compute_domain_x_dict(self):
        return {**self.default_domain_dict(), "istart": self.isd, "iend": self.ied,
                "jstart": self.js, "jend": self.je + 1}

    def horizontal_starts_from_shape(self, shape):
        """Where serialized data of this horizontal shape starts in a storage:
        the compute origin for compute-sized (or +1 staggered) data, one
        point earlier for compute + 2, else the storage corner."""
        if tuple(shape[0:2]) in [self.domain_shape_compute(add=a)[0:2]
                                 for a in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0))]:
            return self.is_, self.js
        if tuple(shape[0:2]) == (self.nic + 2, self.njc + 2):
            return self.is_ - 1, self.js - 1
        return 0, 0

    @staticmethod
    def slice_dict(d, ndim: int = 3):
        sl = (slice(d["istart"], d["iend"] + 1), slice(d["jstart"], d["jend"] + 1),
              slice(d["kstart"], d["kend"] + 1))
        return sl[:ndim]


# ----------------------------------------------------------- translate base
def read_serialized_data(serializer, savepoint, variable):
    data = serializer.read(variable, savepoint)
    if np.asarray(data).size == 1:
        return np.asarray(data).reshape(-1)[0]
    return data


class TranslateFortranData2Py:
    """Base of a component's translate class: declares the serialized input
    and output fields (``in_vars``/``out_vars``), turns the savepoint's
    arrays into storages at the right start indices, runs the component and
    slices its outputs back to the serialized extents."""

    max_error = 1e-14
    near_zero = 1e-18

    def __init__(self, grid: TranslateGrid, origin=utils.origin):
        self.origin = origin
        self.in_vars: Dict = {"data_vars": {}, "parameters": []}
        self.out_vars: Dict = {}
        self.write_vars: List[str] = []
        self.grid = grid
        self.maxshape = grid.domain_shape_full(add=(1, 1, 1))
        self.ignore_near_zero_errors: Dict = {}

    # -- inputs
    def serialnames(self, d: Dict) -> List[str]:
        return [info.get("serialname", name) for name, info in d.items()]

    def collect_input_data(self, serializer, savepoint) -> Dict:
        out = {}
        for name in self.serialnames(self.in_vars["data_vars"]) + self.in_vars["parameters"]:
            out[name] = read_serialized_data(serializer, savepoint, name)
        return out

    def make_storage_data(self, array, istart=0, jstart=0, kstart=0, dummy_axes=None,
                          full_shape=False) -> torch.Tensor:
        """A storage of the tile's max shape (IJ arrays: (I, J); K arrays:
        (K,)) holding ``array`` from (istart, jstart, kstart) on."""
        a = np.asarray(array, dtype=np.float64)
        shape = list(self.maxshape)
        if a.ndim == 3:
            return utils.make_storage_data(a, tuple(shape), start=(istart, jstart, kstart))
        if a.ndim == 2:
            return utils.make_storage_data(a, tuple(shape[:2]), start=(istart, jstart))
        if a.ndim == 1:
            return utils.make_storage_data(a, (shape[2],), start=(kstart,))
        return torch.as_tensor(a)

    def _update_info(self, info, inputs):
        for k, v in list(info.items()):
            if k == "serialname" or isinstance(v, list):
                continue
            if isinstance(v, str) and v in inputs:
                info[k] = inputs[v]

    def make_storage_data_input_vars(self, inputs: Dict) -> None:
        for p in self.in_vars["parameters"]:
            if isinstance(inputs[p], (np.integer,)):
                inputs[p] = int(inputs[p])
            elif isinstance(inputs[p], np.floating):
                inputs[p] = float(inputs[p])
        for name, info in self.in_vars["data_vars"].items():
            serialname = info.get("serialname", name)
            self._update_info(info, inputs)
            arr = np.asarray(inputs[serialname])
            if "kaxis" in info:
                arr = np.moveaxis(arr, info["kaxis"], 2)
            istart, jstart = self.grid.horizontal_starts_from_shape(arr.shape)
            istart = info.get("istart", istart)
            jstart = info.get("jstart", jstart)
            kstart = info.get("kstart", 0)
            inputs[name] = self.make_storage_data(np.squeeze(arr), istart, jstart, kstart,
                                                  full_shape="full_shape" in info)
            if name != serialname:
                del inputs[serialname]

    # -- run
    def compute_func(self, **inputs):
        raise NotImplementedError("a translate class defines compute_func")

    def compute_from_storage(self, inputs: Dict) -> Dict:
        outputs = self.compute_func(**inputs)
        if outputs is not None:
            inputs.update(outputs)
        return inputs

    def compute(self, inputs: Dict) -> Dict:
        self.make_storage_data_input_vars(inputs)
        return self.slice_output(self.compute_from_storage(inputs))

    # -- outputs
    def slice_output(self, inputs: Dict) -> Dict:
        utils.device_sync()
        out = {}
        for var, info in self.out_vars.items():
            self._update_info(info, inputs)
            serialname = info.get("serialname", var)
            ds = self.grid.default_domain_dict()
            ds.update({k: v for k, v in info.items() if k in ds})
            data = utils.asarray(inputs[var])
            out[serialname] = np.squeeze(data[self.grid.slice_dict(ds, data.ndim)])
            if "kaxis" in info:
                out[serialname] = np.moveaxis(out[serialname], 2, info["kaxis"])
        return out

    # -- comparison
    def check(self, serializer, savepoint_out, output: Dict) -> Dict[str, Optional[str]]:
        """Per output field: None when it passes, else the failure report."""
        result = {}
        for varname in self.serialnames(self.out_vars):
            ref = np.asarray(serializer.read(varname, savepoint_out))
            got = np.asarray(output[varname])
            ignore = self.ignore_near_zero_errors.get(varname, False)
            if success(got, ref, self.max_error, ignore, self.near_zero):
                result[varname] = None
            else:
                result[varname] = sample_wherefail(got, ref, self.max_error, True, 1, varname,
                                                   ignore, self.near_zero, xy_indices=True)
        return result
