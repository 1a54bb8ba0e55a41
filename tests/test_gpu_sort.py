"""The library's stable LSD radix sort of the Morton keys (tuning key
sort=4, tiles of 2048 keys, and sort=5, tiles of 4096) against rocPRIM's
onesweep sort (sort=0): all are stable, so the permutations, and with them
the trees and every output bit, must be the same.  Each sort runs in a child process (the tuning table is read when the
library loads).  Key widths from 15 to 28 bits: two and three passes.
The script asserts the depth each case was chosen for."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASES = [
    # (dim, kind, ns, nt, delta, eps, ifpgh, ifpghtarg)
    (2, "mixture", 300000, 100000, 1e-5, 1e-6, 1, 1),   # 18 key bits: two 9-bit passes
    (2, "clump", 60000, 20000, 1e-8, 1e-6, 1, 1),       # nlmax 14, 28 bits: three 10-bit passes
    (2, "uniform", 50000, 0, 1e-4, 1e-9, 3, 0),
    (3, "cluster", 60000, 20000, 1e-5, 1e-6, 1, 1),
    (3, "tight", 40000, 10000, 1e-5, 1e-6, 1, 1),       # nlmax 9, 27 bits: three 9-bit passes
]


def run_case(i):
    """Outputs and tree of CASES[i] in this process."""
    sys.path.insert(0, ROOT)
    import fgt2d_amd
    from tests import _pts
    dim, kind, ns, nt, delta, eps, ifs, ift = CASES[i]
    rng = np.random.default_rng(ns + nt + dim)
    if dim == 2 and kind == "uniform":
        pts = lambda n: _pts.uniform(rng, n)
    elif dim == 2 and kind == "mixture":
        pts = lambda n: _pts.mixture(rng, n)
    elif dim == 2:  # a tight clump with the unit square's corners
        pts = lambda n: np.concatenate([np.clip(0.5 + 2e-4 * rng.normal(size=(n - 2, 2)), 0, 1),
                                        [[0.0, 0.0], [1.0, 1.0]]])
    elif kind == "cluster":
        pts = lambda n: np.clip(0.5 + 0.005 * rng.normal(size=(n, 3)), 0, 1)
    else:
        pts = lambda n: np.concatenate([np.clip(0.5 + 1e-3 * rng.normal(size=(n - 2, 3)), 0, 1),
                                        [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]]])
    src = pts(ns)
    tg = pts(nt) if nt else None
    chg = rng.normal(size=(ns, 1))
    f = fgt2d_amd.fgt2d if dim == 2 else fgt2d_amd.fgt3d
    out = f(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift)
    t = (fgt2d_amd.last_tree2d if dim == 2 else fgt2d_amd.last_tree3d)(ns, nt if nt else 1)
    res = {"out%d" % k: np.asarray(v) for k, v in enumerate(out) if v is not None}
    for k in ("laddr", "par", "nch", "isrc", "isrcse", "itarg", "itargse"):
        res["tree_" + k] = np.asarray(t[k])
    res["nlmax"] = np.array([fgt2d_amd.last_stats()["nlmax"]])
    return res


def _child(i, sort, path):
    env = dict(os.environ, FGTPW_TUNE="sort=%d" % sort)
    code = ("import sys, numpy as np; sys.path.insert(0, %r); from tests import test_gpu_sort as T; "
            "np.savez(%r, **T.run_case(%d))" % (ROOT, path, i))
    subprocess.run([sys.executable, "-c", code], env=env, cwd=ROOT, check=True, timeout=240)
    return np.load(path)


@pytest.mark.parametrize("i", range(len(CASES)), ids=[str(i) for i in range(len(CASES))])
def test_lsd_sort_matches_rocprim_sort(i, tmp_path):
    print("case", CASES[i], flush=True)
    a = _child(i, 0, str(tmp_path / "rocprim.npz"))
    nl = int(a["nlmax"][0])
    print("nlmax", nl, flush=True)
    if i == 1:
        assert nl == 14
    if i == 4:
        assert nl == 9
    for shape in (4, 5):
        b = _child(i, shape, str(tmp_path / ("lsd%d.npz" % shape)))
        assert sorted(a.files) == sorted(b.files)
        for k in a.files:
            assert np.array_equal(a[k], b[k]), (shape, k)
