"""Builds libffpm.so (the gfx950 HIP library) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libffpm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "ffpm_runtime.hip",
    "ffpm_stairstep_kernels.hip",
    "ffpm_pcg_kernels.hip",
    "ffpm_stairstep_host.hip",
    "ffpm_laplace.hip",
    "ffpm_batch.hip",
    "ffpm_unet.hip",
    "ffpm_partition.hip",
    "ffpm_mg.hip",
]

# -ffp-contract=off keeps every expression's rounding identical to the
# reference's serial arithmetic (no fused multiply-adds); packed FP32 ops are
# disabled for all device code.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# roctx ranges (the reference's trace module's names) for rocprofv3 timelines
ROCM_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
LIBS = ["-L", ROCM_LIB, "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM_LIB}"]

FC = os.environ.get("FFPM_FC", "/opt/rocm/bin/amdflang")
FORTRAN_DIR = os.path.join(PKG, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN_DIR, "_build")
FORTRAN_MODULES = ["fpm_ffpm_status.f90", "fpm_ffpm_frame.f90", "fpm_trace.f90",
                   "fpm_stairstep_mpi_3d.f90", "fpm_laplace_3d.f90", "fpm_terms_3d.f90",
                   "fpm_partition_2d.f90"]
FORTRAN_TESTS = ["test_stairstep_exact_3d.f90", "test_fpm_3d.f90", "partition_frame_driver.f90",
                 "test_frame_residency.f90"]


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), *srcs, *LIBS, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False) -> str:
    """Compiles the Fortran drop-in modules (same module names as the
    reference) into libfpm_dropin.a plus .mod files, and links the Fortran
    test programs of tests/fortran against it and libffpm.so."""
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    objs = []
    for src in FORTRAN_MODULES:
        obj = os.path.join(FORTRAN_BUILD, src.replace(".f90", ".o"))
        cmd = [FC, "-O2", "-fPIC", "-c", os.path.join(FORTRAN_DIR, src), "-module-dir",
               FORTRAN_BUILD, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    archive = os.path.join(FORTRAN_BUILD, "libfpm_dropin.a")
    if os.path.exists(archive):
        os.remove(archive)
    subprocess.run(["ar", "rcs", archive, *objs], check=True)
    tests_dir = os.path.join(ROOT, "tests", "fortran")
    out_dir = os.path.join(tests_dir, "_build")
    os.makedirs(out_dir, exist_ok=True)
    for src in FORTRAN_TESTS:
        exe = os.path.join(out_dir, src.replace(".f90", ""))
        cmd = [FC, "-O2", os.path.join(tests_dir, src), "-I", FORTRAN_BUILD, "-module-dir",
               out_dir, archive, "-L", LIBDIR, "-lffpm", f"-Wl,-rpath,{LIBDIR}",
               "-Wl,-rpath,$ORIGIN/../../../flowfpmpp_amd/lib", "-o", exe]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return archive


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
