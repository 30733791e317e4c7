"""Builds libforge_hip.so (gfx950) in-tree with hipcc, then the Fortran host
(forge_amd/_lib/forge_main) with amdflang against it.

Usage: python -m forge_amd.build   (or forge_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT, "libforge_hip.so")
SOURCES = ["operators.hip", "tridiag.hip", "twostage.hip", "dc.hip", "backtr.hip", "gemm.hip", "density.hip",
           "c3.hip", "sym.hip", "valley.hip", "mlp.hip", "scf_device.hip", "capi.hip", "noalloc.hip"]
FSRC = os.path.join(HERE, "fortran")
# dependency order: the C-ABI binding, parameters, geometry, shims, driver
FORTRAN = ["forge_c.f90", "forge_setup.f90", "forge_geometry.f90", "forge_hartreefock.f90",
           "forge_main.f90"]
FORTRAN_EXE = os.path.join(OUT, "forge_main")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         # packed FP32 instructions stay off on this hardware
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-variable"]


def _compile(src: str) -> str:
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(LIB + ".tmp", LIB)
    for o in objs:
        os.remove(o)
    if verbose:
        print("built", LIB)
    if os.path.exists(FLANG):
        build_fortran()
        if verbose:
            print("built", FORTRAN_EXE)
    return LIB


def build_fortran() -> str:
    """The Fortran host, linked against libforge_hip.so (rpath $ORIGIN)."""
    mod = os.path.join(OUT, "fmod")
    os.makedirs(mod, exist_ok=True)
    cmd = [FLANG, "-O2", "-ffp-contract=off", "-module-dir", mod,
           *[os.path.join(FSRC, f) for f in FORTRAN], "-o", FORTRAN_EXE + ".tmp",
           f"-L{OUT}", "-lforge_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stdout}\n{r.stderr}")
    os.replace(FORTRAN_EXE + ".tmp", FORTRAN_EXE)
    return FORTRAN_EXE


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
