"""Build the gfx950 hot-path library (and its Fortran binding) in-tree.

    python 2dmatching-cf4_amd/build.py

Outputs:
  ipm2d/_lib/libipm_hip.so      HIP kernels + C ABI (include/ipm_hip.h)
  ipm2d/_lib/libipm_probe.so    bandwidth calibration kernels the benchmark runs
  fortran/build/                iso_c_binding module and the OneStep driver
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "ipm2d", "_lib")
LIB = os.path.join(LIBDIR, "libipm_hip.so")
PROBE_LIB = os.path.join(LIBDIR, "libipm_probe.so")
FDIR = os.path.join(PKG, "fortran")
FBUILD = os.path.join(FDIR, "build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # every product and sum rounded as written: no FMA contraction
    "-ffp-contract=off",
    # FP64 atomics in LDS and L2 as single instructions
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_hip():
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, f) for f in ("ipm_hip.hip", "ipm_window.hip", "ipm_recombine.hip")]
    tmp = LIB + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs, "-o", tmp])
    os.replace(tmp, LIB)
    tmp = PROBE_LIB + ".tmp"
    run([HIPCC, *HIP_FLAGS, os.path.join(CSRC, "ipm_probe.hip"), "-o", tmp])
    os.replace(tmp, PROBE_LIB)
    return LIB


def build_fortran():
    """Compile the iso_c_binding module and the OneStep driver against the library."""
    if not os.path.exists(FLANG):
        print("flang not found; skipping the Fortran binding", flush=True)
        return None
    os.makedirs(FBUILD, exist_ok=True)
    mod = os.path.join(FDIR, "ipm_hip_mod.f90")
    drv = os.path.join(FDIR, "ipm_onestep_driver.f90")
    if not os.path.exists(mod):
        return None
    run([FLANG, "-O2", "-c", mod, "-module-dir", FBUILD, "-o", os.path.join(FBUILD, "ipm_hip_mod.o")])
    if os.path.exists(drv):
        exe = os.path.join(FBUILD, "ipm_onestep_driver")
        run([FLANG, "-O2", "-I", FBUILD, drv, os.path.join(FBUILD, "ipm_hip_mod.o"), "-L", LIBDIR, "-lipm_hip",
             "-Wl,-rpath," + LIBDIR, "-o", exe])
        return exe
    return None


def main():
    build_hip()
    build_fortran()


if __name__ == "__main__":
    sys.exit(main())
