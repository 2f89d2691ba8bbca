"""Build the C oracle (TEST INFRASTRUCTURE) into oracle/liboracle_hdb.so.

    python -m oracle.build

The oracle is compiled with gcc, OpenMP and -ffp-contract=off.  The
reference's own path is CUDA C++ (nvcc, thrust, cuRAND, cuBLAS) and is not
buildable in this image, so there is no oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_hdb.so"


def build(verbose: bool = True) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-std=c11",
           "-Wall", str(HERE / "hdb_oracle.c"), "-o", str(tmp), "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
