"""Builds libe2edp.so (and the drop-in drivers) in-tree with hipcc for gfx950.

Usage: python -m e2e_gpu_dp_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = PKG / "libe2edp.so"
BIN = PKG / "bin"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("E2E_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no a*b+c fusion, so the lane arithmetic rounds exactly where the
# reference's source does; denormals kept and f32 divide/sqrt correctly rounded, as on
# the host. Packed FP32 instructions are disabled for all device code.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unknown-warning-option",
]

LIB_SOURCES = ["abi.cpp", "devpool.cpp", "build.hip", "scan.hip", "vi.hip", "spmm_vi.hip", "rollout.hip", "envgen.hip",
               "server.cpp"]
DRIVERS = {"build_model": "build_model_main.cpp", "spvi": "spvi_main.cpp"}


def _compile(src: Path) -> Path:
    obj = OUT / (src.name + ".o")
    cmd = [HIPCC, *COMMON, "-I", str(CSRC), "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC, *COMMON, "-x", "hip", "-I", str(CSRC), "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 0, verbose: bool = True) -> Path:
    if not Path(HIPCC).exists() and shutil.which(HIPCC) is None:
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OUT.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in LIB_SOURCES]
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    # rocBLAS: the plain GEMMs of the input generator's Gram matrices and subspace iteration
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_drivers()
    if verbose:
        print(f"built {LIB} and {', '.join(DRIVERS)} in {BIN}")
    return LIB


def build_drivers() -> None:
    """The drop-in executables: plain C++ over the C ABI (no HIP types), linked to
    libe2edp.so next to them."""
    BIN.mkdir(exist_ok=True)
    for exe, src in DRIVERS.items():
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", str(PKG.parent / "include"),
               "-o", str(BIN / exe), str(CSRC / src), "-L", str(PKG), "-le2edp",
               "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    a = ap.parse_args(argv)
    build(a.jobs)


if __name__ == "__main__":
    sys.exit(main())
