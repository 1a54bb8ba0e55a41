"""Builds the native libraries of fgt2d_amd in-tree.

* ``fgt2d_amd/libfgtpw.so`` -- HIP kernels (gfx950) and the C++ host runtime
  behind the C-ABI of ``include/fgtpw.h``;
* ``fgt2d_amd/libfgt2d.so`` -- the Fortran drop-in ``fgt2d``/``fgt3d``
  subroutines (fortran/fgt_dropin.f90) over that C-ABI, when amdflang exists.

Run ``python -m fgt2d_amd.build`` (or ``__graft_entry__.build()``).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfgtpw.so")
FLIB = os.path.join(HERE, "libfgt2d.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FGTPW_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
          "-Wall", "-Wno-unused-result"]
# device code: no packed-FP32 instructions; gfx950 only
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]
HOST = ["-ffp-contract=off"]

SOURCES = [("kernels2d.hip", DEVICE), ("kernels3d.hip", DEVICE), ("api.cpp", DEVICE),
           ("tree.cpp", HOST)]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def _compile(item):
    src, flags = item
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    _run([HIPCC] + COMMON + flags + ["-c", os.path.join(CSRC, src), "-o", out])
    return out


def build_hip(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-pthread", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_fortran(verbose=False):
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        if verbose:
            print("amdflang not found; Fortran drop-in not built")
        return None
    src = os.path.join(HERE, "fortran", "fgt_dropin.f90")
    if not os.path.exists(src):
        return None
    os.makedirs(OBJ, exist_ok=True)
    tmp = FLIB + ".tmp"
    _run([fc, "-O2", "-fPIC", "-shared", "-module-dir", OBJ, "-o", tmp, src, "-L" + HERE,
          "-lfgtpw", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FLIB)
    if verbose:
        print("built", FLIB)
    return FLIB


def build(verbose=False):
    build_hip(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
