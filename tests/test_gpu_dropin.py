"""The reference's own test program (2d/planewave/test/test_fgt2d.f), built
unchanged against this build's Fortran drop-in libfgt2d.so by
`make -C oracle dropin`, runs on the GPU: 1M sources, delta = 6.78e-6,
eps = 5e-11, potentials at the sources.  Its own check -- the direct sum
at the first 20 sources (test_fgt2d.f:161-182) -- must meet the requested
precision."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "test_fgt2d_dropin")
EXE3 = os.path.join(ROOT, "oracle", "_ref", "test_fgt3d_dropin")


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in test driver not built "
                    "(make -C oracle dropin, where the reference sources exist)")
def test_reference_test_driver_on_the_dropin(tmp_path):
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    m = re.search(r"pot err =\s*\n\s*([0-9.Ee+-]+)", r.stdout)
    assert m, r.stdout[-2000:]
    assert float(m.group(1)) < 1e-9


@pytest.mark.skipif(not os.path.exists(EXE3), reason="3D drop-in test driver not built")
def test_reference_3d_test_driver_on_the_dropin(tmp_path):
    """3d/planewave/test/test_fgt3d.f unchanged: 300k sources, delta 1.2e-3,
    eps 1e-6, potentials at the sources, checked by the program itself
    against the direct sum at 20 sources."""
    r = subprocess.run([EXE3], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    m = re.search(r"pot err =\s*\n\s*([0-9.Ee+-]+)", r.stdout)
    assert m, r.stdout[-2000:]
    assert float(m.group(1)) < 1e-5
