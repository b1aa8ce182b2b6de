"""Builds the native libraries of the package in-tree (``dmplex_amd/lib``).

* ``libdmplex_host.so``: C++/OpenMP setup code (mesh, Plex, TRiSK tables), g++.
* ``libtrisk_hip.so``: the gfx950 kernels and the C ABI of ``include/trisk_hip.h``,
  hipcc for gfx950 only.

Run ``python -m dmplex_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "lib"
HOST_SOURCES = ["csrc/host/mesh.cpp", "csrc/host/trisk_grid.cpp", "csrc/host/netcdf.cpp",
                "csrc/host/icon_grid.cpp", "csrc/host/mpas_tables.cpp"]
HIP_SOURCES = ["csrc/hip/trisk_device.hip", "csrc/hip/trisk_ensemble.hip", "csrc/hip/trisk_adjoint.hip",
               "csrc/hip/icon_device.hip", "csrc/hip/mpas_device.hip", "csrc/hip/lloyd_device.hip"]
HIP_HOST_SOURCES = ["csrc/hip/trisk_layout.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed-FP32 instructions are kept out of the device code; FP contraction is off so
# that the FP64 kernels round exactly as the sequential reference formulas do.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def _run(cmd: list[str], cwd: Path) -> None:
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build_host() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libdmplex_host.so"
    tmp = LIB / "libdmplex_host.so.tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-Wall", "-o", str(tmp)] + HOST_SOURCES + ["-ldl"], PKG)
    os.replace(tmp, out)
    return out


def build_device() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libtrisk_hip.so"
    tmp = LIB / "libtrisk_hip.so.tmp"
    cmds = []
    for src in HIP_SOURCES:
        cmds.append([HIPCC, *HIP_FLAGS, "-c", "-o", str(LIB / (Path(src).stem + ".o")), src])
    for src in HIP_HOST_SOURCES:
        cmds.append(["g++", "-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall",
                     "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-c", "-o", str(LIB / (Path(src).stem + ".o")),
                     src])
    objs = [c[c.index("-o") + 1] for c in cmds]
    # one compiler per source, at most as many at once as there are cores (and 8)
    with ThreadPoolExecutor(max_workers=max(1, min(8, os.cpu_count() or 1))) as pool:
        list(pool.map(lambda c: _run(c, PKG), cmds))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-fopenmp", "-o", str(tmp), *objs], PKG)
    for o in objs:
        os.remove(o)
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_host()
    build_device()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"host": build_host, "device": build_device, "all": build_all}[what]()
