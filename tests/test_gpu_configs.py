"""GPU parity on the benchmark workloads at their full sizes, and on trees
deep enough to need 64-bit Morton keys.

* 2D, 2M mixture sources + 2M mixture targets, delta 1e-5, eps 1e-6,
  potentials at the targets: tree, permutations and potentials against the
  2D oracle (the C restatement, equal to the reference bit-for-bit).
* 3D, 1M uniform sources = targets, delta 1e-3, eps 1e-6, potentials at the
  sources: against the 3D oracle, octree exact.
* 3D, 8M sources on a sphere shell (radius 0.4, radial sd 1e-3) + 8M uniform
  targets, delta 1e-5, eps 1e-9, potentials at the targets: the reference
  cannot run it (its expansion store overflows int32 offsets), and the
  oracle cannot hold it either, so the check is the requested precision
  against the FP64 direct sum without cutoff on 10k targets (the reference
  test's own check, test_fgt3d.f:142-152), the GPU direct sum itself checked
  against the oracle's on 50 of them.
* Trees with nlmax above the 32-bit key width (2D nlmax 16, 3D nlmax 11)
  that still form expansions: the u64 key, sort, leaf-depth, leaf-record and
  leaf-start code against the oracle, trees exact.
"""
import numpy as np
import pytest

import bench
import fgt2d_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _same_tree(g, t, dim):
    assert g["nboxes"] == t["nboxes"] and g["nlevels"] == t["nlevels"]
    for k in ("laddr", "lev", "par", "nch", "ch", "ncol", "isrc", "isrcse", "itarg", "itargse"):
        assert np.array_equal(g[k], t[k]), k
    assert np.array_equal(g["cen"], t["cen"])


def test_mixture_2m_matches_oracle():
    print("2D mixture 2M + 2M, delta 1e-5, eps 1e-6", flush=True)
    src, chg, tg = bench.workload()
    got = fgt2d_amd.fgt2d(1e-5, 1e-6, src, chg, tg, ifpgh=0, ifpghtarg=1)
    st = fgt2d_amd.last_stats()
    assert st["npw"] == 26 and st["nlevels"] == 9, st
    g = fgt2d_amd.last_tree2d(len(src), len(tg))
    want = O.oracle_fgt2d(1e-5, 1e-6, src, chg, tg, ifpgh=0, ifpghtarg=1)
    assert _rel(got[3], want[3]) < 1e-11, _rel(got[3], want[3])
    _same_tree(g, O.oracle_tree2d(1e-5, 1e-6, src, tg), 2)
    # the requested precision against the direct sum on 2000 targets
    sel = np.random.default_rng(1).permutation(len(tg))[:2000]
    ex = fgt2d_amd.fgt2dpart_direct(1e-5, 1e-100, src, chg, tg[sel])[0]
    assert O.derr(ex, got[3][sel]) < 1e-5


def test_uniform3d_1m_matches_oracle():
    print("3D uniform 1M, delta 1e-3, eps 1e-6, pot at sources", flush=True)
    rng = np.random.default_rng(bench.SEED)
    src = rng.random((1_000_000, 3))
    chg = rng.uniform(-1, 1, (1_000_000, 1))
    got = fgt2d_amd.fgt3d(1e-3, 1e-6, src, chg, None, ifpgh=1, ifpghtarg=0)
    st = fgt2d_amd.last_stats()
    assert st["npw"] == 26 and st["npwlevel"] == 4 and st["nlevels"] == 6, st
    g = fgt2d_amd.last_tree3d(len(src), 0)
    want = O.oracle_fgt3d(1e-3, 1e-6, src, chg, None, ifpgh=1, ifpghtarg=0)
    assert _rel(got[0], want[0]) < 1e-11, _rel(got[0], want[0])
    t = O.oracle_tree3d(1e-3, 1e-6, src, np.zeros((0, 3)))
    assert g["nboxes"] == t["nboxes"] and g["nlevels"] == t["nlevels"]
    for k in ("laddr", "lev", "par", "nch", "ch", "ncol", "isrc", "isrcse"):
        assert np.array_equal(g[k], t[k]), k
    assert np.array_equal(g["cen"], t["cen"])


def test_shell3d_8m_against_direct_sum():
    print("3D shell 8M + 8M uniform targets, delta 1e-5, eps 1e-9", flush=True)
    src, chg, tg = bench.shell_workload()
    got = fgt2d_amd.fgt3d(1e-5, 1e-9, src, chg, tg, ifpgh=0, ifpghtarg=1)[3]
    st = fgt2d_amd.last_stats()
    # the workload reaches the expansion levels: npw 38, cutoff level 7,
    # leaves down to level 9
    assert st["npw"] == 38 and st["npwlevel"] == 7 and st["nlevels"] == 9, st
    sel = np.random.default_rng(1).permutation(len(tg))[:10000]
    ex = fgt2d_amd.fgt3dpart_direct(1e-5, 1e30, src, chg, tg[sel])[0]
    # the GPU direct sum against the oracle's on 50 targets
    ex50 = O.direct3d(1e-5, 1e30, src, chg, tg[sel[:50]])[0]
    assert _rel(ex[:50], ex50) < 1e-12
    err = O.derr(ex, got[sel])
    print("rel l2 error vs direct sum", err, flush=True)
    assert err < 3e-8, err


def _cluster2d(rng, n, nt):
    src = np.vstack([rng.normal(0.5, 2e-4, (n, 2)), [[0.0, 0.0], [1.0, 1.0]]])
    tg = np.vstack([rng.normal(0.5, 2e-4, (nt, 2)), [[0.2, 0.9]]])
    return src, rng.normal(size=(len(src), 1)), tg


def test_u64_keys_2d_matches_oracle():
    """nlmax 16 (32 key bits): the 64-bit key path, with expansions at
    levels 14-16."""
    rng = np.random.default_rng(3)
    src, chg, tg = _cluster2d(rng, 30000, 8000)
    got = fgt2d_amd.fgt2d(3e-10, 1e-6, src, chg, tg, ifpgh=1, ifpghtarg=1)
    st = fgt2d_amd.last_stats()
    assert st["nlmax"] == 16 and st["npwlevel"] == 14 and st["nlevels"] == 16, st
    want = O.oracle_fgt2d(3e-10, 1e-6, src, chg, tg, ifpgh=1, ifpghtarg=1)
    assert _rel(got[0], want[0]) < 1e-11 and _rel(got[3], want[3]) < 1e-11
    g = fgt2d_amd.last_tree2d(len(src), len(tg))
    _same_tree(g, O.oracle_tree2d(3e-10, 1e-6, src, tg), 2)


def test_u64_keys_3d_matches_oracle():
    """nlmax 11 (33 key bits): the 64-bit octree key path, with expansions
    at levels 9-11."""
    rng = np.random.default_rng(4)
    src = np.vstack([rng.normal(0.5, 1e-3, (20000, 3)), [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]]])
    tg = np.vstack([rng.normal(0.5, 1e-3, (4000, 3)), [[0.3, 0.6, 0.2]]])
    chg = rng.normal(size=(len(src), 1))
    got = fgt2d_amd.fgt3d(1e-6, 1e-6, src, chg, tg, ifpgh=1, ifpghtarg=1)
    st = fgt2d_amd.last_stats()
    assert st["nlmax"] == 11 and st["npwlevel"] == 9 and st["nlevels"] == 11, st
    want = O.oracle_fgt3d(1e-6, 1e-6, src, chg, tg, ifpgh=1, ifpghtarg=1)
    assert _rel(got[0], want[0]) < 1e-11 and _rel(got[3], want[3]) < 1e-11
    g = fgt2d_amd.last_tree3d(len(src), len(tg))
    t = O.oracle_tree3d(1e-6, 1e-6, src, tg)
    _same_tree(g, t, 3)
    for b in range(t["nboxes"]):
        n = t["ncol"][b]
        assert np.array_equal(g["col"][b, :n], t["col"][b, :n])
