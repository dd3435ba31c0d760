"""Builds the Fortran host layer with flang, in-tree:

  lib/libmgbf_fortran.so  modules mg_parameter, mg_intstate, mg_entrymod,
                          mg_transfer, mg_filtering, mgbf_capi (ISO_C_BINDING)
  lib/fmod/*.mod          module files for Fortran callers
  lib/mgbf_driver         the RBETA_TEST-style driver program
  lib/mgbf_nml_check      prints the sizes derived from a namelist (no device)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
FSRC = os.path.join(PKG, "fortran")
LIBDIR = os.path.join(PKG, "lib")
MODDIR = os.path.join(LIBDIR, "fmod")
SOURCES = ["mgbf_capi.f90", "mg_modules.f90"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def flang() -> str:
    for cand in (os.environ.get("FC"), os.path.join(ROCM, "lib/llvm/bin/flang"), shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("flang not found")


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError("flang failed")


def build() -> str:
    os.makedirs(MODDIR, exist_ok=True)
    fc = flang()
    rpath = "-Wl,-rpath,$ORIGIN"
    hiplib = os.path.join(ROCM, "lib")
    lib = os.path.join(LIBDIR, "libmgbf_fortran.so")
    _run([fc, "-O2", "-fPIC", "-shared", f"-J{MODDIR}"] +
         [os.path.join(FSRC, s) for s in SOURCES] +
         ["-o", lib, f"-L{LIBDIR}", "-lmgbf", f"-L{hiplib}", "-lamdhip64", rpath,
          f"-Wl,-rpath,{hiplib}"])
    drv = os.path.join(LIBDIR, "mgbf_driver")
    _run([fc, "-O2", f"-I{MODDIR}", os.path.join(FSRC, "mgbf_driver.f90"), "-o", drv,
          f"-L{LIBDIR}", "-lmgbf_fortran", "-lmgbf", f"-L{hiplib}", "-lamdhip64",
          rpath, f"-Wl,-rpath,{hiplib}"])
    chk = os.path.join(LIBDIR, "mgbf_nml_check")
    _run([fc, "-O2", f"-I{MODDIR}", os.path.join(FSRC, "mgbf_nml_check.f90"), "-o", chk,
          f"-L{LIBDIR}", "-lmgbf_fortran", "-lmgbf", f"-L{hiplib}", "-lamdhip64",
          rpath, f"-Wl,-rpath,{hiplib}"])
    return drv


if __name__ == "__main__":
    print(build())
