"""Build recipe of the CPU oracle (test infrastructure, not the product).

Compiles ``oracle/sph_oracle.cpp`` with g++ and OpenMP into
``oracle/liboracle_sph.so``.  ``-ffp-contract=off`` keeps every floating-point
operation rounded once, as written.  ``-march=x86-64-v3`` (AVX2) rather than
``-march=native`` so the library also runs on a different host CPU.

The reference itself (CUDA C++ needing nvcc, the CUDA runtime and Thrust) is
not buildable in this image, so there is no ``oracle/_ref`` build.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRCS = [HERE / "sph_oracle.cpp", HERE / "gnn_oracle.cpp"]
LIB = HERE / "liboracle_sph.so"

FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
         "-shared", "-std=c++17", "-Wall"]


def build(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["g++", *FLAGS, *map(str, SRCS), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
