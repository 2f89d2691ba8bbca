"""Build libhdbscan_mi355x.so (gfx950) in-tree.

    python -m gpu_hdbscan_amd.build        # or gpu_hdbscan_amd.build.build()

Every .hip file under csrc/ is compiled with hipcc for gfx950 with packed
FP32 instructions disabled; labels.cpp (host only) is compiled with g++
without floating-point contraction; the objects are linked into
gpu_hdbscan_amd/libhdbscan_mi355x.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_build"
LIB = PKG / "libhdbscan_mi355x.so"
ARCH = os.environ.get("HDB_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "is not a recognized feature for this target" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build(verbose: bool = True) -> Path:
    OUT_DIR.mkdir(exist_ok=True)
    jobs = []
    for src in sorted(CSRC.glob("*.hip")):
        obj = OUT_DIR / (src.stem + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for src in sorted(CSRC.glob("*.cpp")):
        obj = OUT_DIR / (src.stem + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    # C++ driver over include/gpu_hdbscan.hpp (mirrors HDBSCAN/_new.cu)
    drv = PKG / "hdbscan_pipeline"
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{PKG.parent / 'include'}", str(PKG / "cpp" / "hdbscan_pipeline.cpp"),
          "-o", str(drv) + ".tmp", f"-L{PKG}", "-lhdbscan_mi355x", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath-link,/opt/rocm/lib"])
    os.replace(str(drv) + ".tmp", drv)
    if verbose:
        print(f"built {LIB} and {drv}")
    return LIB


if __name__ == "__main__":
    build()
