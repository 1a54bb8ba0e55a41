"""GPU parity of the 3D transform (fgt3d) through the C-ABI: against the
3D oracle (oracle/fgt3d_oracle.c, itself equal to the 3D reference
bit-for-bit) on seeded DLARAN inputs, against the reference's own golden
outputs, and against the FP64 direct sum.  Trees and permutations must be
identical; values agree to round-off (same FP64 algorithm, other summation
order)."""
import glob
import os

import numpy as np
import pytest

import fgt2d_amd
from oracle import oracle as O
from tests import _pts

pytestmark = pytest.mark.gpu

NAMES = ["pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"]

# (recipe, ns, nt, delta, eps, nd, ifpgh, ifpghtarg)
CASES = [
    ("uniform", 3000, 1000, 1e-2, 1e-6, 1, 1, 1),      # small tree, npwlevel 2
    ("uniform", 40000, 5000, 1e-2, 1e-6, 1, 0, 1),     # targets only
    ("uniform", 30000, 5000, 1e-2, 1e-3, 2, 3, 2),     # nd = 2, gradients and Hessians
    ("shell", 20000, 4000, 2e-3, 1e-6, 1, 2, 1),       # clustered sources, level restriction
    ("uniform", 20000, 2000, 2e-2, 1e-3, 1, 1, 3),     # npw 12: direct plane-wave routes
    ("uniform", 2000, 700, 1e-3, 1e-6, 1, 1, 1),       # npwlevel == nlevels: near field only
    ("uniform", 1000, 300, 30.0, 1e-9, 1, 3, 3),       # large delta: one Taylor expansion
    ("cluster", 20000, 4000, 1e-5, 1e-9, 1, 1, 2),     # npw 38, Taylor orders 59/30/19 below the cutoff
]


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_fgt3d_matches_oracle(case):
    recipe, ns, nt, delta, eps, nd, ifs, ift = case
    print("case", case, flush=True)
    src, chg, tg = _pts.dlaran_points3d(recipe, ns, nt, nd)
    got = fgt2d_amd.fgt3d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift)
    want = O.oracle_fgt3d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift)
    for k in range(6):
        if (k < 3 and k < ifs) or (k >= 3 and k - 3 < ift):
            assert _rel(got[k], want[k]) < 1e-11, (NAMES[k], _rel(got[k], want[k]))


@pytest.mark.parametrize("case", CASES[:5], ids=[str(i) for i in range(5)])
def test_octree_matches_oracle(case):
    recipe, ns, nt, delta, eps, nd, ifs, ift = case
    src, chg, tg = _pts.dlaran_points3d(recipe, ns, max(nt, 1), nd)
    fgt2d_amd.fgt3d(delta, eps, src, chg, tg, ifpgh=1, ifpghtarg=1)
    g = fgt2d_amd.last_tree3d(ns, len(tg))
    t = O.oracle_tree3d(delta, eps, src, tg)
    assert g["nboxes"] == t["nboxes"] and g["nlevels"] == t["nlevels"]
    for k in ("laddr", "lev", "par", "nch", "ch", "ncol", "isrc", "isrcse", "itarg", "itargse"):
        assert np.array_equal(g[k], t[k]), k
    assert np.array_equal(g["cen"], t["cen"])
    for b in range(t["nboxes"]):
        n = t["ncol"][b]
        assert np.array_equal(g["col"][b, :n], t["col"][b, :n])


def test_deeper_octree_after_a_shallow_one():
    """As the 2D test: a shallow octree, one at least two levels deeper
    (clustered sources), the shallow one again, through one workspace."""
    cases = [("uniform", 2000, 500, 1e-2, 1e-6), ("cluster", 20000, 4000, 1e-5, 1e-9),
             ("uniform", 2000, 500, 1e-2, 1e-6)]
    depth = []
    for recipe, ns, nt, delta, eps in cases:
        print("case", recipe, ns, nt, delta, eps, flush=True)
        src, chg, tg = _pts.dlaran_points3d(recipe, ns, nt, 1)
        got = fgt2d_amd.fgt3d(delta, eps, src, chg, tg, ifpgh=1, ifpghtarg=1)
        want = O.oracle_fgt3d(delta, eps, src, chg, tg, ifpgh=1, ifpghtarg=1)
        assert _rel(got[0], want[0]) < 1e-11 and _rel(got[3], want[3]) < 1e-11
        g = fgt2d_amd.last_tree3d(ns, nt)
        t = O.oracle_tree3d(delta, eps, src, tg)
        assert g["nboxes"] == t["nboxes"] and g["nlevels"] == t["nlevels"]
        for k in ("laddr", "lev", "par", "nch", "ch", "isrc", "isrcse", "itarg", "itargse"):
            assert np.array_equal(g[k], t[k]), k
        depth.append(t["nlevels"])
    assert depth[1] >= depth[0] + 2, depth


GOLD3 = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
               if "3d" in os.path.basename(p))


@pytest.mark.parametrize("path", GOLD3, ids=[os.path.basename(p)[:-4] for p in GOLD3])
def test_fgt3d_matches_reference_golden(path):
    with np.load(path, allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    ns, nt, nd = int(g["ns"]), int(g["nt"]), int(g["nd"])
    print("golden", os.path.basename(path), ns, nt, nd, flush=True)
    src, chg, tg, rn, ds = _pts.golden_inputs3(g)
    delta, eps = float(g["delta"]), float(g["eps"])
    out = fgt2d_amd.fgt3d(delta, eps, src, chg, tg, ifpgh=int(g["ifpgh"]),
                          ifpghtarg=int(g["ifpghtarg"]), ifdipole=int(rn is not None),
                          rnormal=rn, dipstr=ds)
    for k, name in enumerate(NAMES):
        if name in g:
            sel = g["isample_src"] if k < 3 else g["isample_trg"]
            assert _rel(out[k][sel], g[name]) < 1e-11, name
    st = fgt2d_amd.last_stats()
    assert [st["npwlevel"], st["nlmax"], st["nlevels"], st["nboxes"]] == list(g["ints"])
    if "pot" in g:
        # dipoles: the reference's own error reaches ~100 eps
        assert O.derr(g["pot_direct20"], out[0][:20]) < (300 if rn is not None else 30) * eps


def test_fgt3d_accuracy_against_direct_sum():
    src, chg, tg = _pts.dlaran_points3d("uniform", 30000, 3000, 1)
    for delta, eps, tol in ((1e-2, 1e-6, 2e-5), (3e-3, 1e-9, 5e-8)):
        pt = fgt2d_amd.fgt3d(delta, eps, src, chg, tg, ifpgh=0, ifpghtarg=1)[3]
        ex = fgt2d_amd.fgt3dpart_direct(delta, 1e30, src, chg, tg[:300])[0]
        assert O.derr(ex, pt[:300]) < tol, (delta, eps)


def test_fgt3dpart_direct_matches_oracle():
    rng = np.random.default_rng(4)
    src, tg = rng.random((2000, 3)), rng.random((300, 3))
    q = rng.normal(size=(2000, 2))
    for ifpgh in (1, 2, 3):
        g = fgt2d_amd.fgt3dpart_direct(1e-2, 0.05, src, q, tg, ifpgh=ifpgh)
        w = O.direct3d(1e-2, 0.05, src, q, tg, ifpgh=ifpgh)
        for k in range(ifpgh):
            assert _rel(g[k], w[k]) < 1e-12


def test_fgt3d_device_entry_point_matches_host_entry_point():
    import torch
    src, chg, tg = _pts.dlaran_points3d("shell", 20000, 4000, 1)
    h = fgt2d_amd.fgt3d(2e-3, 1e-6, src, chg, tg, ifpgh=1, ifpghtarg=1)
    dev = torch.device("cuda:0")
    d = fgt2d_amd.fgt3d(2e-3, 1e-6, torch.from_numpy(src).to(dev), torch.from_numpy(chg).to(dev),
                        torch.from_numpy(tg).to(dev), ifpgh=1, ifpghtarg=1)
    torch.cuda.synchronize()
    assert np.array_equal(h[0], d[0].cpu().numpy())
    assert np.array_equal(h[3], d[3].cpu().numpy())


@pytest.mark.parametrize("cap", ["8", "37"])
def test_chunked_expansions_match_one_chunk(cap, monkeypatch):
    """Leaves formed, merged, split and evaluated in chunks of at most `cap`
    boxes (FGTPW_MAX_CHUNK) give the same bits as one chunk per level: the
    chunk region only changes where a leaf's expansion lives, not the order
    of any sum."""
    src, chg, tg = _pts.dlaran_points3d("shell", 50000, 4000, 2)
    for ifs, ift in ((2, 1), (1, 3)):
        monkeypatch.delenv("FGTPW_MAX_CHUNK", raising=False)
        one = fgt2d_amd.fgt3d(5e-3, 1e-6, src, chg, tg, ifpgh=ifs, ifpghtarg=ift)
        st = fgt2d_amd.last_stats()
        assert st["nlevels"] >= st["npwlevel"] + 2, st  # leaves below the top level
        monkeypatch.setenv("FGTPW_MAX_CHUNK", cap)
        print("cap", cap, "ifpgh", ifs, "ifpghtarg", ift, flush=True)
        got = fgt2d_amd.fgt3d(5e-3, 1e-6, src, chg, tg, ifpgh=ifs, ifpghtarg=ift)
        for k in range(6):
            if (k < 3 and k < ifs) or (k >= 3 and k - 3 < ift):
                assert np.array_equal(got[k], one[k]), NAMES[k]
