"""Build recipe for the native parts of subsweep_amd.

  libsubsweep_amd.so  HIP library (gfx950 device code + C ABI host runtime)
  libssw_grid.so      host grid / input generators (g++, OpenMP)

Both are built in-tree next to this file. Run `python -m subsweep_amd.build`.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SWEEP_LIB = os.path.join(HERE, "libsubsweep_amd.so")
GRID_LIB = os.path.join(HERE, "libssw_grid.so")

# -ffp-contract=off: the reference's Rust arithmetic is never contracted into
# FMAs, and the upwind/downwind sign tests must be evaluated exactly as written.
# -packed-fp32-ops off: no packed FP32 instructions are generated.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_sweep() -> str:
    src = os.path.join(CSRC, "sweep.hip")
    placement = os.path.join(CSRC, "placement.hip")
    placement_obj = os.path.join(CSRC, "placement.o")
    server = os.path.join(CSRC, "server.cpp")
    sources = os.path.join(CSRC, "sources.cpp")
    sweep_obj = os.path.join(CSRC, "sweep.o")
    server_obj = os.path.join(CSRC, "server.o")
    sources_obj = os.path.join(CSRC, "sources.o")
    tmp = SWEEP_LIB + ".tmp"
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    _run([HIPCC, *compile_flags, "-c", src, "-o", sweep_obj])
    _run([HIPCC, *compile_flags, "-c", placement, "-o", placement_obj])
    # the request loop and the star-particle sources are host-only C++
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-I", os.path.join(ROOT, "include"), "-c", server,
          "-o", server_obj])
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"), "-c",
          sources, "-o", sources_obj])
    _run([HIPCC, "-shared", "-fPIC", sweep_obj, placement_obj, server_obj, sources_obj, "-lpthread", "-o", tmp])
    for o in (sweep_obj, placement_obj, server_obj, sources_obj):
        os.remove(o)
    os.replace(tmp, SWEEP_LIB)
    return SWEEP_LIB


def build_grid() -> str:
    src = os.path.join(CSRC, "gridgen.cpp")
    tmp = GRID_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", src, "-o", tmp])
    os.replace(tmp, GRID_LIB)
    return GRID_LIB


def build_all() -> None:
    build_grid()
    build_sweep()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    else:
        if "grid" in what:
            build_grid()
        if "sweep" in what:
            build_sweep()
