"""Builds the gfx950 library (and the Fortran binding) in-tree.

    python simple3.0_amd/build.py            # libpftcc_hip.so + Fortran module/test program

Outputs go to simple3.0_amd/lib/ (git-ignored, shipped with the tree).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FC", "/opt/rocm/bin/amdflang")

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             # no packed FP32 instructions (v_pk_*_f32) anywhere in device code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libpftcc_hip.so")
    tmp = out + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "pftcc.hip")])
    os.replace(tmp, out)
    return out


def build_synth():
    out = os.path.join(LIB, "libpftcc_synth.so")
    tmp = out + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "synth.hip")])
    os.replace(tmp, out)
    return out


def build_fortran():
    """Fortran module simple_polarft_corrcalc (ISO_C_BINDING over the C ABI) + test driver."""
    src_mod = os.path.join(FSRC, "simple_polarft_corrcalc.f90")
    src_drv = os.path.join(FSRC, "pftcc_fortran_driver.f90")
    if not (os.path.exists(src_mod) and os.path.exists(FLANG)):
        return None
    mod_dir = os.path.join(LIB, "fmod")
    os.makedirs(mod_dir, exist_ok=True)
    obj = os.path.join(mod_dir, "simple_polarft_corrcalc.o")
    run([FLANG, "-O2", "-fPIC", "-c", src_mod, "-module-dir", mod_dir, "-o", obj])
    if os.path.exists(src_drv):
        exe = os.path.join(LIB, "pftcc_fortran_driver")
        run([FLANG, "-O2", "-fopenmp", src_drv, obj, "-I" + mod_dir, "-module-dir", mod_dir,
             "-L" + LIB, "-lpftcc_hip", "-Wl,-rpath,$ORIGIN", "-o", exe])
        return exe
    return obj


def main():
    build_hip()
    build_synth()
    build_fortran()


if __name__ == "__main__":
    sys.exit(main())
