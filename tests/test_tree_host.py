"""Host logic of the quadtree: the runtime's box tree, built from the list of
non-empty adaptive-tree leaves that the GPU stage produces, must equal the
reference's tree (pts_tree_build + pts_tree_fix_lr, pts_tree2d4.f:363-996)
box for box: same numbering, levels, parents, children, colleagues and
bit-identical centres.  The GPU stage is restated here in numpy (tests/_pts.py)."""
import numpy as np
import pytest

import fgt2d_amd
from oracle import oracle as O
from tests import _pts

CASES = [
    ("uniform", 3000, 1e-4, 1e-6),
    ("uniform", 20000, 1e-3, 1e-9),
    ("mixture", 30000, 1e-5, 1e-6),
    ("mixture", 20000, 3e-6, 1e-3),
    ("uniform", 500, 1e-2, 1e-3),
    ("mixture", 40000, 1e-4, 1e-9),
    ("ring", 20000, 1e-5, 1e-6),
]


def _points(kind, rng, n):
    if kind == "uniform":
        return _pts.uniform(rng, n)
    if kind == "mixture":
        return _pts.mixture(rng, n)
    t = rng.random(n) * 2 * np.pi
    r = 0.3 + 0.002 * rng.normal(size=n)
    return np.stack([0.5 + r * np.cos(t), 0.5 + r * np.sin(t)], axis=1)


def _host_tree(delta, eps, src, tg):
    npwlevel, bs0, cen0 = O.boxsize0(delta, eps, src, tg)
    t = O.oracle_tree2d(delta, eps, src, tg)
    nlmax = t["nlmax"]
    keys = np.sort(_pts.keys2d(src, cen0, bs0, nlmax), kind="stable")
    lk, lv = _pts.leaf_runs(keys, nlmax, npwlevel)
    return fgt2d_amd.tree2d_from_leaves(nlmax, npwlevel, bs0, cen0, lk, lv), t


@pytest.mark.parametrize("kind,n,delta,eps", CASES)
def test_host_tree_matches_oracle_tree(kind, n, delta, eps):
    rng = np.random.default_rng(n)
    src = _points(kind, rng, n)
    tg = _points(kind, rng, n // 2)
    h, t = _host_tree(delta, eps, src, tg)
    assert h["nlevels"] == t["nlevels"] and h["nboxes"] == t["nboxes"]
    for k in ("laddr", "lev", "par", "nch", "ch", "ncol"):
        assert np.array_equal(h[k], t[k]), k
    assert np.array_equal(h["cen"], t["cen"])  # bit-identical FP64 centres
    for b in range(t["nboxes"]):
        n_ = t["ncol"][b]
        assert np.array_equal(h["col"][b, :n_], t["col"][b, :n_])


@pytest.mark.skipif(not O.have_ref(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("kind,n,delta,eps", CASES[:4])
def test_host_tree_matches_reference_tree(kind, n, delta, eps):
    rng = np.random.default_rng(n + 1)
    src = _points(kind, rng, n)
    tg = _points(kind, rng, n // 3)
    h, _ = _host_tree(delta, eps, src, tg)
    r = O.ref_tree2d(delta, eps, src, tg)
    assert h["nboxes"] == r["nboxes"] and h["nlevels"] == r["nlevels"]
    for k in ("laddr", "lev", "par", "nch", "ch", "ncol"):
        assert np.array_equal(h[k], r[k]), k
    assert np.array_equal(h["cen"], r["cen"])


def test_level_restriction_happens():
    # a dense clump next to empty space forces the level-restriction pass to
    # add boxes; the host tree must reproduce them
    rng = np.random.default_rng(3)
    src = np.concatenate([0.5 + 0.003 * rng.normal(size=(20000, 2)), rng.random((50, 2))])
    tg = rng.random((10, 2))
    h, t = _host_tree(1e-5, 1e-6, src, tg)
    assert np.array_equal(h["par"], t["par"])
    assert np.array_equal(h["cen"], t["cen"])
    # more boxes than the adaptive tree alone: some leaf got refined by LR
    nonempty_leaves = len(_pts.leaf_runs(np.sort(_pts.keys2d(src, O.boxsize0(1e-5, 1e-6, src, tg)[2],
                                                             O.boxsize0(1e-5, 1e-6, src, tg)[1],
                                                             t["nlmax"])), t["nlmax"],
                                         t["npwlevel"])[0])
    assert h["nboxes"] > nonempty_leaves


def test_host_tree_large_matches_oracle_tree():
    # 73k leaves: large enough that the builder forks its order-independent
    # loops onto worker threads; the tree must still be the reference's
    rng = np.random.default_rng(7)
    src = _pts.uniform(rng, 2_000_000)
    tg = _pts.uniform(rng, 1000)
    h, t = _host_tree(1e-6, 1e-6, src, tg)
    assert len(h["lev"]) >= 100_000
    assert h["nlevels"] == t["nlevels"] and h["nboxes"] == t["nboxes"]
    for k in ("laddr", "lev", "par", "nch", "ch", "ncol"):
        assert np.array_equal(h[k], t[k]), k
    assert np.array_equal(h["cen"], t["cen"])
    lev = t["lev"]
    sel = np.flatnonzero(lev <= t["npwlevel"])
    for b in sel[:: max(1, len(sel) // 2000)]:
        n_ = t["ncol"][b]
        assert np.array_equal(h["col"][b, :n_], t["col"][b, :n_])
