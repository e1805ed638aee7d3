"""Build recipe of the HIP engine library (gfx950) and the Fortran interface module.

hipcc flags: -ffp-contract=off keeps floating-point expressions unfused so the
engine's FP64 results are reproducible against a CPU evaluation in the same
order; packed-FP32 instruction selection is disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "ngpu_engine.hip")
HOST_SRC = [os.path.join(HERE, "csrc", "ngpu_fcidump.cpp")]   # host-only parts of the library
OUT = os.path.join(HERE, "libneci_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + os.path.join(ROOT, "include")]


def build_library(verbose=False):
    cmd = [HIPCC] + FLAGS + ["-o", OUT + ".tmp", SRC] + HOST_SRC
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for " + SRC)
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built", OUT)
    return OUT


def build_fortran(verbose=False):
    """Compile the ISO_C_BINDING module and the Fortran driver with amdflang when present."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    fdir = os.path.join(ROOT, "fortran")
    if fc is None or not os.path.isdir(fdir):
        return None
    srcs = [os.path.join(fdir, f) for f in ("neci_gpu_mod.F90", "neci_gpu_driver.F90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    exe = os.path.join(fdir, "neci_gpu_driver")
    cmd = [fc, "-O2", "-J", fdir, "-o", exe] + srcs + ["-L" + HERE, "-lneci_gpu", "-Wl,-rpath," + HERE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("amdflang failed")
    if verbose:
        print("built", exe)
    return exe


if __name__ == "__main__":
    build_library(verbose=True)
    build_fortran(verbose=True)
