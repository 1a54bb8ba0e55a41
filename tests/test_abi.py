"""The C-ABI library loads and exports every function include/fgtpw.h
declares, and the Fortran drop-in exports the reference's symbols
(fgt2d_, fgt2dpart_direct_vec_ and the prini family the reference test
driver links, test_fgt2d.f).  No compute calls: runs without a GPU."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

import fgt2d_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fgtpw.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fgtpw_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared()
    for want in ("fgtpw_fgt2d", "fgtpw_fgt2d_device", "fgtpw_fgt2dpart_direct",
                 "fgtpw_last_error", "fgtpw_init"):
        assert want in names


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(fgt2d_amd.library_path())
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_error_text_without_a_call():
    lib = ctypes.CDLL(fgt2d_amd.library_path())
    lib.fgtpw_last_error.restype = ctypes.c_char_p
    assert lib.fgtpw_last_error() == b""


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "fgt2d_amd", "libfgt2d.so")),
                    reason="Fortran drop-in not built (no amdflang)")
def test_fortran_dropin_exports_reference_symbols():
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", os.path.join(ROOT, "fgt2d_amd", "libfgt2d.so")],
                         capture_output=True, text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for s in ("fgt2d_", "fgt2dpart_direct_vec_", "fgt3d_", "fgt3dpart_direct_vec_", "prini_",
              "prin2_", "prinf_", "prinf_long_"):
        assert s in syms, s


def test_submit_without_a_server_is_an_error():
    import numpy as np
    src = np.random.default_rng(0).random((10, 2))
    with pytest.raises(fgt2d_amd.FGTError, match="not running"):
        fgt2d_amd.Server.submit(type("S", (), {"_lib": fgt2d_amd.load(), "_out": {}})(), 2,
                                1e-3, 1e-6, src, np.ones(10))
