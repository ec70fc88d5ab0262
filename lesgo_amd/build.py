"""Build liblesgo_amd.so (gfx950) in-tree with hipcc.

Usage: python -m lesgo_amd.build   (or lesgo_amd.build.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblesgo_amd.so")
SOURCES = ["case.hip", "passes_f64.hip", "passes_f32.hip", "tune.hip", "turbines.hip", "capi.hip",
           "server.hip", "stats.hip", "calibrate.hip", "tridag.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 ops stay off (wrong results seen with some forms on this pool)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


FSRC = os.path.join(HERE, "fortran")
FORTRAN_LIB = ["lesgo_amd_c.f90", "lesgo_state.f90", "lesgo_gpu.f90"]
# executables: the time-loop driver and the drop-in check program
PROGRAMS = {"lesgo_amd_main": "main.f90", "lesgo_amd_dropins": "dropins.f90"}
DRIVER = os.path.join(HERE, "lesgo_amd_main")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


def build_driver(verbose: bool = True) -> str:
    """Fortran programs (amdflang) linked against liblesgo_amd.so."""
    for exe, main in PROGRAMS.items():
        mod = os.path.join(OBJ, "fmod_" + exe)
        os.makedirs(mod, exist_ok=True)
        out = os.path.join(HERE, exe)
        srcs = [os.path.join(FSRC, f) for f in FORTRAN_LIB + [main]]
        cmd = [FC, "-O2", "-module-dir", mod, *srcs, "-L" + HERE, "-llesgo_amd",
               "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-o", out + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=OBJ)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {main}:\n{r.stderr[-6000:]}")
        os.replace(out + ".tmp", out)
        if verbose:
            print(f"built {out}")
    return DRIVER


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_driver(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
