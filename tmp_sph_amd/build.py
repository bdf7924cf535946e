"""Build recipe of the gfx950 shared library ``tmp_sph_amd/libsph_mi355x.so``.

Every ``csrc/*.hip`` file is compiled by hipcc for gfx950 only and the host
``csrc/*.cpp`` files by g++; the objects are linked by hipcc into one shared
library in the package directory (in-tree, so it travels with the sources).

Run ``python -m tmp_sph_amd.build`` or call :func:`build`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_obj"
LIB = PKG / "libsph_mi355x.so"
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 instructions stay off (their results were not reliable on these parts)
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]


def _hipcc() -> str:
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(h).exists():
        raise RuntimeError("hipcc not found: the gfx950 library cannot be built")
    return h


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile(src: Path) -> Path:
    out = OBJ / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [_hipcc(), *HIP_FLAGS, "-c", str(src), "-o", str(out)]
    else:
        cmd = ["g++", *CXX_FLAGS, "-c", str(src), "-o", str(out)]
    _run(cmd)
    return out


def build(verbose: bool = False) -> Path:
    """Compile every source and link the shared library; returns its path."""
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = min(8, len(srcs)) or 1
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
