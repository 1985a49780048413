"""Build recipe of libfesom_gpu.so (gfx950) with hipcc, in-tree.

The library is plain HIP C++ behind a C ABI; no torch extension machinery.
Flags: FP64 kernels keep the written evaluation order (-ffp-contract=off) and
packed FP32 instructions are disabled for the whole build.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfesom_gpu.so")
SOURCES = ["fgpu_runtime.hip", "fgpu_transport.hip", "fgpu_step.hip", "fgpu_ssh.hip", "fgpu_mpnn.hip",
           "fgpu_dyn.hip", "fgpu_diff.hip"]
ARCH = os.environ.get("FGPU_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_SOURCES = ["fesom_gpu_c.f90", "fesom_gpu_host.f90"]
FORTRAN_DRIVER = os.path.join(FORTRAN_DIR, "fgpu_case_driver")
# the drop-in entry points with the reference's argument lists, over stand-ins
# of the reference's type modules, and the program that runs a case through them
DROPIN_SOURCES = ["fesom_types.f90", "fesom_gpu_dropin.f90"]
DROPIN_DRIVER = os.path.join(FORTRAN_DIR, "fgpu_dropin_driver")


def flang():
    for cand in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang", shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def flags():
    return ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
            "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-I", os.path.join(HERE, "..", "include")]


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [cc, *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False):
    """Compile the Fortran host modules and link the case driver against
    libfesom_gpu.so (rpath to the package directory). Returns the driver path,
    or None when no Fortran compiler is installed."""
    fc = flang()
    if fc is None:
        return None
    objdir = os.path.join(HERE, "_build", "fortran")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in FORTRAN_SOURCES + ["fgpu_case_driver.f90"]:
        obj = os.path.join(objdir, src.replace(".f90", ".o"))
        cmd = [fc, "-O2", "-fPIC", "-module-dir", objdir, "-I", objdir, "-c", os.path.join(FORTRAN_DIR, src),
               "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed on {src}:\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    tmp = FORTRAN_DRIVER + ".tmp"
    cmd = [fc, "-O2", "-o", tmp, *objs, "-L", HERE, "-lfesom_gpu", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"flang link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, FORTRAN_DRIVER)
    # drop-in: its own module directory, since its type modules carry the
    # reference's module names
    ddir = os.path.join(HERE, "_build", "fortran_dropin")
    os.makedirs(ddir, exist_ok=True)
    dobjs = [os.path.join(objdir, "fesom_gpu_c.o")]
    for src in DROPIN_SOURCES + ["fgpu_dropin_driver.f90"]:
        obj = os.path.join(ddir, src.replace(".f90", ".o"))
        cmd = [fc, "-O2", "-fPIC", "-module-dir", ddir, "-I", ddir, "-I", objdir, "-c",
               os.path.join(FORTRAN_DIR, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed on {src}:\n{r.stdout}\n{r.stderr}")
        dobjs.append(obj)
    tmp = DROPIN_DRIVER + ".tmp"
    cmd = [fc, "-O2", "-o", tmp, *dobjs, "-L", HERE, "-lfesom_gpu", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"flang link failed (drop-in driver):\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, DROPIN_DRIVER)
    return FORTRAN_DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
