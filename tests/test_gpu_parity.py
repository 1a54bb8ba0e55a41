"""GPU parity: the HIP transform against the oracle (the C restatement,
itself pinned to the reference bit-for-bit) on seeded inputs, through the
C-ABI.  Tree and permutations must match exactly; potentials, gradients and
Hessians to round-off (the same FP64 algorithm, summed in another order)."""
import numpy as np
import pytest

import fgt2d_amd
from oracle import oracle as O
from tests import _pts

pytestmark = pytest.mark.gpu

# (kind, ns, nt, delta, eps, nd, ifpgh, ifpghtarg)
CASES = [
    ("uniform", 2000, 700, 1e-4, 1e-6, 1, 1, 1),      # npwlevel == nlevels: near field only
    ("mixture", 20000, 8000, 1e-5, 1e-6, 1, 0, 1),    # 2M/2M configuration, scaled down
    ("uniform", 100000, 0, 1e-4, 1e-9, 1, 1, 0),      # 100k configuration: PW evaluation at level 5
    ("mixture", 30000, 10000, 1e-4, 1e-9, 2, 3, 2),   # nd = 2, gradients and Hessians
    ("uniform", 5000, 2000, 1e-2, 1e-3, 1, 2, 3),
    ("uniform", 1000, 300, 5.0, 1e-6, 1, 1, 1),       # large delta: one Taylor expansion
    ("mixture", 50000, 20000, 3e-6, 1e-3, 3, 1, 1),   # deep tree, nd = 3
    # eps 1e-14 with Hessians: plane waves to eps/100, npw = 66 (above 64),
    # PW evaluation at the cutoff level (Taylor order 91)
    ("uniform", 20000, 5000, 1e-4, 1e-14, 1, 3, 3),
]


def _points(kind, rng, n):
    return _pts.uniform(rng, n) if kind == "uniform" else _pts.mixture(rng, n)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_transform_matches_oracle(case):
    kind, ns, nt, delta, eps, nd, ifs, ift = case
    print("case", case, flush=True)
    rng = np.random.default_rng(ns + nt)
    src = _points(kind, rng, ns)
    tg = _points(kind, rng, nt) if nt else None
    chg = rng.normal(size=(ns, nd))
    got = fgt2d_amd.fgt2d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift)
    want = O.oracle_fgt2d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift)
    for k, (g, w) in enumerate(zip(got, want)):
        if g is None or (k < 3 and k >= ifs) or (k >= 3 and k - 3 >= ift):
            continue
        assert _rel(g, w) < 1e-11, (k, _rel(g, w))


@pytest.mark.parametrize("case", CASES[:4], ids=[str(i) for i in range(4)])
def test_tree_matches_oracle(case):
    kind, ns, nt, delta, eps, nd, ifs, ift = case
    rng = np.random.default_rng(ns + nt)
    src = _points(kind, rng, ns)
    tg = _points(kind, rng, max(nt, 1))
    chg = rng.normal(size=(ns, nd))
    fgt2d_amd.fgt2d(delta, eps, src, chg, tg, ifpgh=1, ifpghtarg=1)
    g = fgt2d_amd.last_tree2d(ns, len(tg))
    t = O.oracle_tree2d(delta, eps, src, tg)
    assert g["nboxes"] == t["nboxes"] and g["nlevels"] == t["nlevels"]
    for k in ("laddr", "lev", "par", "nch", "ch", "ncol", "isrc", "isrcse", "itarg", "itargse"):
        assert np.array_equal(g[k], t[k]), k
    assert np.array_equal(g["cen"], t["cen"])


def test_deeper_tree_after_a_shallow_one():
    """Calls of different depth through one workspace (buffers, plan storage
    and leaf-record sizing kept from call to call): a shallow uniform
    transform, one several levels deeper (clustered, small delta), then the
    shallow one again, each against the oracle, trees and permutations
    exactly."""
    cases = [("uniform", 3000, 1000, 1e-4, 1e-6), ("mixture", 40000, 10000, 3e-6, 1e-3),
             ("uniform", 3000, 1000, 1e-4, 1e-6)]
    depth = []
    for kind, ns, nt, delta, eps in cases:
        print("case", kind, ns, nt, delta, eps, flush=True)
        rng = np.random.default_rng(ns + 7)
        src = _points(kind, rng, ns)
        tg = _points(kind, rng, nt)
        chg = rng.normal(size=(ns, 1))
        got = fgt2d_amd.fgt2d(delta, eps, src, chg, tg, ifpgh=1, ifpghtarg=1)
        want = O.oracle_fgt2d(delta, eps, src, chg, tg, ifpgh=1, ifpghtarg=1)
        assert _rel(got[0], want[0]) < 1e-11 and _rel(got[3], want[3]) < 1e-11
        g = fgt2d_amd.last_tree2d(ns, nt)
        t = O.oracle_tree2d(delta, eps, src, tg)
        assert g["nboxes"] == t["nboxes"] and g["nlevels"] == t["nlevels"]
        for k in ("laddr", "lev", "par", "nch", "ch", "isrc", "isrcse", "itarg", "itargse"):
            assert np.array_equal(g[k], t[k]), k
        depth.append(t["nlevels"])
    assert depth[1] >= depth[0] + 2, depth


def test_accuracy_against_direct_sum():
    """The requested precision against the FP64 direct sum with no cutoff
    (the reference test's oracle, test_fgt2d.f:161-188)."""
    rng = np.random.default_rng(11)
    for delta, eps, tol in ((1e-4, 1e-6, 2e-5), (1e-4, 1e-9, 5e-8), (1e-3, 1e-3, 5e-2)):
        src = _pts.uniform(rng, 20000)
        tg = _pts.uniform(rng, 5000)
        chg = rng.uniform(-1, 1, (20000, 1))
        pt = fgt2d_amd.fgt2d(delta, eps, src, chg, tg, ifpgh=0, ifpghtarg=1)[3]
        ex = O.direct2d(delta, 1e-100, src, chg, tg[:400])[0]
        err = O.derr(ex, pt[:400])
        assert err < tol, (delta, eps, err)


def test_device_entry_point_matches_host_entry_point():
    import torch
    rng = np.random.default_rng(5)
    src, tg = _pts.mixture(rng, 40000), _pts.mixture(rng, 10000)
    chg = rng.normal(size=(40000, 1))
    h = fgt2d_amd.fgt2d(1e-5, 1e-6, src, chg, tg, ifpgh=1, ifpghtarg=1)
    dev = torch.device("cuda:0")
    d = fgt2d_amd.fgt2d(1e-5, 1e-6, torch.from_numpy(src).to(dev), torch.from_numpy(chg).to(dev),
                        torch.from_numpy(tg).to(dev), ifpgh=1, ifpghtarg=1)
    torch.cuda.synchronize()
    assert np.array_equal(h[0], d[0].cpu().numpy())
    assert np.array_equal(h[3], d[3].cpu().numpy())


def test_part_direct_matches_oracle():
    rng = np.random.default_rng(9)
    src, tg = _pts.uniform(rng, 3000), _pts.uniform(rng, 500)
    chg = rng.normal(size=(3000, 2))
    for ifpgh in (1, 2, 3):
        g = fgt2d_amd.fgt2dpart_direct(1e-3, 1e-6, src, chg, tg, ifpgh=ifpgh)
        w = O.direct2d(1e-3, 1e-6, src, chg, tg, ifpgh=ifpgh)
        for k in range(ifpgh):
            assert _rel(g[k], w[k]) < 1e-12


def test_invalid_arguments_are_rejected():
    rng = np.random.default_rng(1)
    src = _pts.uniform(rng, 100)
    chg = np.ones((100, 1))
    with pytest.raises(fgt2d_amd.FGTError) as e:
        fgt2d_amd.fgt2d(1e-3, 1e-6, src, chg, None, ifpgh=4)
    assert e.value.code == fgt2d_amd.FGTPW_EINVAL
    bad = src.copy()
    bad[7, 0] = np.nan
    with pytest.raises(fgt2d_amd.FGTError) as e:
        fgt2d_amd.fgt2d(1e-3, 1e-6, bad, chg, None, ifpgh=1)
    assert e.value.code == fgt2d_amd.FGTPW_EINVAL
    with pytest.raises(fgt2d_amd.FGTError) as e:
        fgt2d_amd.fgt2d(1e-3, 1e-6, src, chg, None, ifpgh=1, ifcharge=2)
    assert e.value.code == fgt2d_amd.FGTPW_EINVAL
    # a dipole flag without dipole arrays, through the C-ABI itself
    lib = fgt2d_amd.load()
    z = np.zeros(8)
    p = fgt2d_amd._ptr(z)
    rc = lib.fgtpw_fgt2d(1, 1e-3, 1e-6, 100, fgt2d_amd._ptr(src), 1, fgt2d_amd._ptr(chg), 1, None,
                         None, 0, 1, fgt2d_amd._ptr(np.zeros(100)), p, p, 0, p, 0, p, p, p)
    assert rc == fgt2d_amd.FGTPW_EINVAL
    # the library still works after errors
    fgt2d_amd.fgt2d(1e-3, 1e-6, src, chg, None, ifpgh=1)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_finest_level_formed_beside_the_tree_is_bit_identical(precision):
    """The finest level's leaves are formed while the host builds the tree
    (boxes, ranges and centres from the leaf records); with timing on they
    are formed after the tree, with the other levels.  Both must give the
    same bits, in FP64 and on the mixed route, also through a call whose
    tree is deeper than the last one's (the expansion store sized from the
    last call is then too small and the level is formed again)."""
    prev = fgt2d_amd.set_precision(precision)
    try:
        rng = np.random.default_rng(21)
        cases = [(_pts.mixture(rng, 60000), _pts.mixture(rng, 20000), 1e-5, 1e-6),
                 (_pts.uniform(rng, 30000), _pts.uniform(rng, 10000), 1e-4, 1e-6),
                 (_pts.mixture(rng, 80000), _pts.mixture(rng, 20000), 3e-6, 1e-6)]
        for src, tg, delta, eps in cases:
            print("case", len(src), len(tg), delta, eps, precision, flush=True)
            chg = rng.normal(size=(len(src), 1))
            a = fgt2d_amd.fgt2d(delta, eps, src, chg, tg, ifpgh=1, ifpghtarg=1)
            fgt2d_amd.set_timing(True)
            try:
                b = fgt2d_amd.fgt2d(delta, eps, src, chg, tg, ifpgh=1, ifpghtarg=1)
            finally:
                fgt2d_amd.set_timing(False)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[3], b[3])
    finally:
        fgt2d_amd.set_precision(prev)
