"""Build step of the Fortran host side, with the flang of the ROCm toolchain.

Products (in-tree, next to the HIP library):
  2pcf_amd/libtpcf_gpu_fortran.so   modules tpcf_gpu_c and tpcf_gpu
                                    (2pcf_amd/fortran/tpcf_gpu.f90), linked
                                    against lib2pcf_amd.so
  2pcf_amd/fortran/include/*.mod    the module files a Fortran driver compiles
                                    against (flang's module format)
  2pcf_amd/fortran/fortran_driver   tests/fortran/fortran_driver.f90 and
  2pcf_amd/fortran/fortran_patches  tests/fortran/fortran_patches.f90, the
                                    Fortran host programs the GPU tests run
Both shared objects find each other through an $ORIGIN run path.  Skipped
(with a message) when flang is absent.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
FDIR = os.path.join(PKG, "fortran")
MODDIR = os.path.join(FDIR, "include")
BUILD = os.path.join(ROOT, "_build", "fortran")
LIB = os.path.join(PKG, "libtpcf_gpu_fortran.so")
DRIVER = os.path.join(FDIR, "fortran_driver")
PROGRAMS = ("fortran_driver", "fortran_patches")
FC = os.environ.get("FC", "/opt/rocm/llvm/bin/flang")
FFLAGS = ["-O2", "-fPIC"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))


def build() -> str | None:
    if not os.path.exists(FC):
        print(f"fortran_build: {FC} not found; Fortran host side skipped")
        return None
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(MODDIR, exist_ok=True)
    obj = os.path.join(BUILD, "tpcf_gpu.o")
    _run([FC, *FFLAGS, "-c", os.path.join(FDIR, "tpcf_gpu.f90"), "-module-dir", MODDIR, "-o", obj])
    tmp = LIB + ".tmp"
    _run([FC, *FFLAGS, "-shared", "-o", tmp, obj, "-L" + PKG, "-l2pcf_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, LIB)
    for prog in PROGRAMS:
        obj_p = os.path.join(BUILD, prog + ".o")
        _run([FC, *FFLAGS, "-c", os.path.join(ROOT, "tests", "fortran", prog + ".f90"), "-I", MODDIR,
              "-module-dir", BUILD, "-o", obj_p])
        exe = os.path.join(FDIR, prog)
        tmp = exe + ".tmp"
        _run([FC, *FFLAGS, "-o", tmp, obj_p, "-L" + PKG, "-ltpcf_gpu_fortran", "-l2pcf_amd",
              "-Wl,-rpath,$ORIGIN/.."])
        os.replace(tmp, exe)
    return LIB


if __name__ == "__main__":
    print(build())
