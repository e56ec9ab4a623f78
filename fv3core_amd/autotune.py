"""Offline tile-shape autotuner for the LDS-tiled stencil kernels.

The workgroup shapes of ``delnflux_kernel`` and ``a2b_interp_kernel`` are
macros in ``csrc/tile_config.hpp``.  A shape only changes how the work is
split, never the arithmetic, so every candidate returns the same bits; the
tuner picks the fastest on the GPU it runs on:

1. ``build``: for each candidate, compile the kernel's source with the
   shape's ``-D`` flags and link it with the other objects of the in-tree
   build into ``_tune/lib_<kernel>_<a>_<b>.so`` (CPU only: hipcc
   cross-compiles for gfx950).
2. ``measure``: run a benchmark workload once per library, each in its own
   process, timing the kernel with the library's HIP-event hooks
   (``bench.py --roofline-kernel <kernel>``), and write the mean launch
   times to ``_tune/results.json``.  ``--workload ensemble`` (the default)
   times the launches of the 32-member C48 batch in the mixed form;
   ``--workload dyncore`` those of the FP64 acoustic step on six C96 tiles.
3. ``apply``: rewrite the defaults in ``tile_config.hpp`` with the fastest
   shape of each kernel (rebuild the library afterwards).

``python -m fv3core_amd.autotune build``, then on a GPU
``python -m fv3core_amd.autotune measure``, then
``python -m fv3core_amd.autotune apply``.
"""
from __future__ import annotations

import argparse
import json
import pathlib
import re
import subprocess
import sys
from typing import Dict, List, Tuple

from . import build as _build

HERE = pathlib.Path(__file__).resolve().parent
TUNE_DIR = HERE / "_tune"
CONFIG = _build.CSRC / "tile_config.hpp"
REPO = HERE.parent

# kernel -> (source, timing name, macro names, candidate (rows, waves) pairs)
KERNELS: Dict[str, Tuple[str, str, Tuple[str, str], List[Tuple[int, int]]]] = {
    "delnflux": ("delnflux.hip", "delnflux_kernel", ("FV3_DN_TJ", "FV3_DN_NY"),
                 [(4, 4), (8, 2), (8, 4), (8, 8), (12, 4), (16, 4), (16, 8), (24, 4)]),
    "a2b": ("a2b.hip", "a2b_interp_kernel", ("FV3_AB_TJ", "FV3_AB_NY"),
            [(8, 4), (12, 4), (16, 4), (16, 8), (24, 8), (32, 8), (32, 16)]),
}


def variant_lib(kernel: str, shape: Tuple[int, int]) -> pathlib.Path:
    return TUNE_DIR / f"lib_{kernel}_{shape[0]}_{shape[1]}.so"


def build_variants(kernels=None) -> List[pathlib.Path]:
    """Compile one library per candidate shape (needs the in-tree objects)."""
    lib = _build.build(verbose=False)
    objdir = lib.parent / "_build"
    TUNE_DIR.mkdir(exist_ok=True)
    made = []
    for kernel in kernels or KERNELS:
        src, _, macros, shapes = KERNELS[kernel]
        others = [str(o) for o in sorted(objdir.glob("*.o")) if o.stem != pathlib.Path(src).stem]
        for shape in shapes:
            obj = TUNE_DIR / f"{kernel}_{shape[0]}_{shape[1]}.o"
            defs = [f"-D{macros[0]}={shape[0]}", f"-D{macros[1]}={shape[1]}"]
            cmd = [_build.hipcc()] + _build.COMMON_FLAGS + defs + ["-c", str(_build.CSRC / src),
                                                                 "-o", str(obj)]
            res = subprocess.run(cmd, capture_output=True, text=True)
            if res.returncode != 0:
                raise RuntimeError(f"hipcc failed for {kernel} {shape}:\n{res.stderr}")
            out = variant_lib(kernel, shape)
            res = subprocess.run([_build.hipcc(), "-shared", "-fPIC",
                                  f"--offload-arch={_build.OFFLOAD_ARCH}"] + others +
                                 [str(obj), "-o", str(out)], capture_output=True, text=True)
            if res.returncode != 0:
                raise RuntimeError(f"link failed for {kernel} {shape}:\n{res.stderr}")
            obj.unlink()
            made.append(out)
    return made


_RUNNER = """
import pathlib, runpy, sys
sys.path.insert(0, {repo!r})
from fv3core_amd import _lib
_lib.LIB_PATH = pathlib.Path({lib!r})
sys.argv = ["bench.py", "--workload", {workload!r}, "--steps", "{steps}", "--warmup", "1",
            "--no-cpu-baseline", "--roofline-kernel", {name!r}] + {extra!r}
runpy.run_path({bench!r}, run_name="__main__")
"""


# bench.py options that leave out what the kernel timing does not need
_WORKLOAD_ARGS = {"ensemble": ["--no-graph", "--no-adjoint"], "dyncore": []}


def measure(kernels=None, steps: int = 3, timeout: int = 300,
            workload: str = "ensemble") -> dict:
    """Mean launch time (us) of each variant's kernel inside the workload."""
    if workload not in _WORKLOAD_ARGS:
        raise ValueError(f"unknown workload {workload}")
    results = {}
    for kernel in kernels or KERNELS:
        _, name, _, shapes = KERNELS[kernel]
        for shape in shapes:
            lib = variant_lib(kernel, shape)
            if not lib.exists():
                raise FileNotFoundError(f"{lib} is missing: run the build step first")
            code = _RUNNER.format(repo=str(REPO), lib=str(lib), steps=steps, name=name,
                                  bench=str(REPO / "bench.py"), workload=workload,
                                  extra=_WORKLOAD_ARGS[workload])
            res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                                 timeout=timeout)
            if res.returncode != 0:
                raise RuntimeError(f"timing run failed for {kernel} {shape}:\n{res.stderr[-2000:]}")
            line = [ln for ln in res.stdout.splitlines() if ln.startswith("{")][-1]
            us = float(json.loads(line)["roofline"]["kernel_avg_us"])
            results.setdefault(kernel, {})[f"{shape[0]}x{shape[1]}"] = us
            print(f"{kernel} rows={shape[0]} waves={shape[1]}: {us:.1f} us", flush=True)
    TUNE_DIR.mkdir(exist_ok=True)
    (TUNE_DIR / "results.json").write_text(json.dumps(results, indent=1))
    return results


def best_shapes(results: dict) -> Dict[str, Tuple[int, int]]:
    best = {}
    for kernel, times in results.items():
        key = min(times, key=times.get)
        a, b = key.split("x")
        best[kernel] = (int(a), int(b))
    return best


def rewrite_config(text: str, best: Dict[str, Tuple[int, int]]) -> str:
    """tile_config.hpp text with the defaults of the tuned kernels replaced."""
    for kernel, shape in best.items():
        macros = KERNELS[kernel][2]
        for macro, value in zip(macros, shape):
            pat = re.compile(rf"(#ifndef {macro}\n#define {macro} )(\d+)")
            if not pat.search(text):
                raise ValueError(f"{macro} has no default in tile_config.hpp")
            text = pat.sub(lambda m: m.group(1) + str(value), text)
    return text


def apply(results_path: pathlib.Path = TUNE_DIR / "results.json") -> Dict[str, Tuple[int, int]]:
    best = best_shapes(json.loads(results_path.read_text()))
    CONFIG.write_text(rewrite_config(CONFIG.read_text(), best))
    return best


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("step", choices=["build", "measure", "apply"])
    p.add_argument("--kernel", action="append", choices=sorted(KERNELS))
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--workload", default="ensemble", choices=sorted(_WORKLOAD_ARGS))
    a = p.parse_args(argv)
    if a.step == "build":
        for lib in build_variants(a.kernel):
            print(lib)
    elif a.step == "measure":
        measure(a.kernel, a.steps, workload=a.workload)
    else:
        for kernel, shape in apply().items():
            print(f"{kernel}: rows={shape[0]} waves={shape[1]}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
