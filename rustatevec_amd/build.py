"""Build recipe for libqsvh.so, the engine's HIP library (gfx950).

Every source is compiled by hipcc for ``--offload-arch=gfx950`` and linked into one
shared library next to this file (kept in-tree so it travels with the sources).
Packed-FP32 instruction selection is disabled for all device code.  The exact-precision
kernels and the host IR are compiled without floating-point contraction so that their
complex arithmetic matches the reference operation for operation.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqsvh.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value"]
# file -> extra flags
SOURCES = {
    "ir.cpp": ["-ffp-contract=off"],
    "api.cpp": [],
    "plan.cpp": [],
    "kernels_plain.hip": ["-ffp-contract=off"],
    "kernels_compose.hip": ["-ffp-contract=off"],
    "kernels_util.hip": [],
    "kernels_mfma.hip": [],
    "kernels_tile.hip": ["-ffp-contract=off"],
    "kernels_batch.hip": ["-ffp-contract=off"],
    "batch.cpp": [],
    "kernels_mlp.hip": [],
    "mlp.cpp": [],
    "planner.cpp": ["-ffp-contract=off"],
}


def _compile(name: str, flags: list[str], verbose: bool) -> str:
    src = os.path.join(CSRC, name)
    obj = os.path.join(BUILD, name + ".o")
    lang = ["-x", "hip"] if name.endswith(".hip") else []
    cmd = [HIPCC, *COMMON, *flags, *lang, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {name}:\n{r.stderr}")
    return obj


def build(verbose: bool = False, jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda kv: _compile(kv[0], kv[1], verbose), SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
