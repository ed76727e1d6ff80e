"""Builds libsod2d_hip.so in-tree with hipcc for gfx950.

Each .hip translation unit under csrc/ is compiled to an object in
csrc/_build/ and the objects are linked into sod2d_refmap_amd/libsod2d_hip.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libsod2d_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SOD_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-command-line-argument",
    # no packed FP32 instructions (see the design notes on gfx950 code generation)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", f"-I{INCLUDE}", f"-I{CSRC}",
]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4) // 2))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", flush=True)
    return LIB


CXX = os.environ.get("SOD_CXX", "g++")
HDF5_ROOT = os.environ.get("SOD_HDF5_ROOT", "/opt/conda")
IO_LIB = os.path.join(HERE, "libsod2d_io.so")


def build_io(verbose: bool = True) -> str:
    """Compiles the host-only HDF5 mesh/restart library (csrc/io_h5.cpp) into
    sod2d_refmap_amd/libsod2d_io.so against the HDF5 C library."""
    inc, lib = os.path.join(HDF5_ROOT, "include"), os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        raise RuntimeError(f"hdf5.h not found under {inc} (set SOD_HDF5_ROOT)")
    tmp = IO_LIB + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{inc}",
           os.path.join(CSRC, "io_h5.cpp"), f"-L{lib}", "-lhdf5", f"-Wl,-rpath,{lib}", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"io library build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, IO_LIB)
    if verbose:
        print(f"built {IO_LIB}", flush=True)
    return IO_LIB


FC = os.environ.get("SOD_FC", "/opt/rocm/bin/amdflang")
FORTRAN = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FORTRAN, "_build")
DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "shim_driver.f90")
DRIVER = os.path.join(ROOT, "tests", "fortran", "_build", "shim_driver")
IO_DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "io_driver.f90")
IO_DRIVER = os.path.join(ROOT, "tests", "fortran", "_build", "io_driver")


def build_fortran(verbose: bool = True) -> str:
    """Compiles the ISO_C_BINDING module (fortran/mod_sod2d_hip.f90) to an
    object + .mod in fortran/_build/, and links the test driver that calls
    the library through it (tests/fortran/shim_driver.f90)."""
    os.makedirs(FBUILD, exist_ok=True)
    obj = os.path.join(FBUILD, "mod_sod2d_hip.o")
    cmd = [FC, "-O2", "-fPIC", "-c", os.path.join(FORTRAN, "mod_sod2d_hip.f90"), "-module-dir", FBUILD, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if os.path.exists(DRIVER_SRC):
        os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
        rpath = os.path.relpath(HERE, os.path.dirname(DRIVER))
        cmd = [FC, "-O2", DRIVER_SRC, obj, "-I", FBUILD, "-module-dir", os.path.dirname(DRIVER), "-L", HERE,
               "-lsod2d_hip", f"-Wl,-rpath,$ORIGIN/{rpath}", "-o", DRIVER]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    # the HDF5 file module and its test driver (needs libsod2d_io.so)
    io_obj = os.path.join(FBUILD, "mod_sod2d_io.o")
    cmd = [FC, "-O2", "-fPIC", "-c", os.path.join(FORTRAN, "mod_sod2d_io.f90"), "-module-dir", FBUILD, "-o", io_obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if os.path.exists(IO_DRIVER_SRC) and os.path.exists(IO_LIB):
        rpath = os.path.relpath(HERE, os.path.dirname(IO_DRIVER))
        cmd = [FC, "-O2", IO_DRIVER_SRC, io_obj, "-I", FBUILD, "-module-dir", os.path.dirname(IO_DRIVER), "-L", HERE,
               "-lsod2d_io", f"-Wl,-rpath,$ORIGIN/{rpath}", "-o", IO_DRIVER]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {obj}, {io_obj}, {DRIVER} and {IO_DRIVER}", flush=True)
    return obj


if __name__ == "__main__":
    build()
    build_io()
    build_fortran()
    sys.exit(0)
