"""The mixed-precision route (fgtpw_set_precision(1)): FP32 matrix-core
moments and Taylor evaluation, FP32 plane-wave expansions through merge,
m2l and split, for 2D transforms with eps >= 1e-6.

Tolerance: the route rounds every moment, expansion coefficient and Taylor
coefficient to FP32 (unit roundoff 6e-8) and accumulates a segment of at
most 256 points in FP32, so its results differ from the FP64 route's by a
few times 1e-7 (measured: 2.5e-7 to 2.8e-7 relative l2 on potentials at
eps 1e-6), inside the requested eps.  The tests require
  * relative l2 error against the FP64 direct sum (no cutoff) <= eps, the
    requested precision (eps = 1e-6 here), for potentials;
  * relative l2 difference from the FP64 oracle <= eps / 2 for potentials
    and <= eps for gradients and Hessians (the reference's own outputs
    within its requested eps);
  * the tree and point permutations identical to the FP64 route's (the
    route changes only arithmetic).
Transforms the route does not cover (eps < 1e-6, 3D) run in FP64 and must be
bit-identical to the FP64 mode's."""
import numpy as np
import pytest

import fgt2d_amd
from oracle import oracle as O
from tests import _pts

pytestmark = pytest.mark.gpu


def _rl2(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture
def mixed():
    prev = fgt2d_amd.set_precision("mixed")
    yield
    fgt2d_amd.set_precision(prev)


def test_mixed_c2_shape_within_eps_of_direct_sum(mixed):
    import bench
    # the benchmark's mixture (sources and targets around the same centres)
    src, chg, tg = bench.workload(200000, 200000, seed=21)
    got = fgt2d_amd.fgt2d(1e-5, 1e-6, src, chg, tg, ifpgh=0, ifpghtarg=1)[3]
    assert fgt2d_amd.last_precision() == "mixed"
    fgt2d_amd.set_precision("fp64")
    ref = fgt2d_amd.fgt2d(1e-5, 1e-6, src, chg, tg, ifpgh=0, ifpghtarg=1)[3]
    assert fgt2d_amd.last_precision() == "fp64"
    sel = np.random.default_rng(1).permutation(len(tg))[:3000]
    ex = fgt2d_amd.fgt2dpart_direct(1e-5, 1e-100, src, chg, tg[sel])[0]
    e_mixed, e_fp64 = O.derr(ex, got[sel]), O.derr(ex, ref[sel])
    print("rel l2 vs direct: mixed %.3g fp64 %.3g; mixed vs fp64 %.3g"
          % (e_mixed, e_fp64, _rl2(got, ref)), flush=True)
    assert e_mixed <= 1e-6, e_mixed
    assert _rl2(got, ref) <= 5e-7


CASES = [
    # (kind, ns, nt, delta, eps, nd, ifpgh, ifpghtarg, dipoles)
    ("mixture", 30000, 10000, 1e-5, 1e-6, 1, 1, 1, False),
    ("mixture", 30000, 10000, 1e-4, 1e-6, 2, 3, 2, False),   # nd = 2, gradients, Hessians
    ("uniform", 20000, 5000, 1e-3, 1e-3, 1, 2, 3, False),    # eps 1e-3
    ("mixture", 20000, 8000, 1e-5, 1e-6, 1, 2, 2, True),     # charges and dipoles
]


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_mixed_matches_oracle(case, mixed):
    kind, ns, nt, delta, eps, nd, ifs, ift, dip = case
    print("mixed case", case, flush=True)
    rng = np.random.default_rng(ns + 3 * nt)
    if kind == "mixture":
        import bench
        src, _, tg = bench.workload(ns, nt, seed=ns + nt)  # one set of centres
    else:
        src, tg = _pts.uniform(rng, ns), _pts.uniform(rng, nt)
    chg = rng.normal(size=(ns, nd))
    kw = {}
    if dip:
        kw = dict(rnormal=rng.normal(size=(ns, 2)), dipstr=rng.normal(size=(ns, nd)))
    got = fgt2d_amd.fgt2d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, ifdipole=int(dip),
                          **kw)
    assert fgt2d_amd.last_precision() == "mixed"
    want = O.oracle_fgt2d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, **kw)
    for k in range(6):
        if (k < 3 and k < ifs) or (k >= 3 and k - 3 < ift):
            tol = eps / 2 if k % 3 == 0 else eps
            err = _rl2(got[k], want[k])
            print("  output %d: rel l2 vs FP64 oracle %.3g" % (k, err), flush=True)
            assert err <= tol, (k, err)
    # the tree is the FP64 route's
    g = fgt2d_amd.last_tree2d(ns, nt)
    t = O.oracle_tree2d(delta, eps, src, tg)
    for key in ("laddr", "par", "nch", "isrc", "itarg", "isrcse", "itargse"):
        assert np.array_equal(g[key], t[key]), key


def test_routes_outside_mixed_stay_fp64(mixed):
    rng = np.random.default_rng(8)
    src, tg = _pts.uniform(rng, 20000), _pts.uniform(rng, 5000)
    chg = rng.normal(size=(20000, 1))
    a = fgt2d_amd.fgt2d(1e-4, 1e-9, src, chg, tg, ifpgh=1, ifpghtarg=1)
    assert fgt2d_amd.last_precision() == "fp64"
    s3, t3 = rng.random((20000, 3)), rng.random((4000, 3))
    a3 = fgt2d_amd.fgt3d(1e-2, 1e-6, s3, chg, t3, ifpgh=1, ifpghtarg=1)
    assert fgt2d_amd.last_precision() == "fp64"
    fgt2d_amd.set_precision("fp64")
    b = fgt2d_amd.fgt2d(1e-4, 1e-9, src, chg, tg, ifpgh=1, ifpghtarg=1)
    b3 = fgt2d_amd.fgt3d(1e-2, 1e-6, s3, chg, t3, ifpgh=1, ifpghtarg=1)
    for x, y in ((a[0], b[0]), (a[3], b[3]), (a3[0], b3[0]), (a3[3], b3[3])):
        assert np.array_equal(x, y)


def test_precision_mode_rejects_unknown_values():
    lib = fgt2d_amd.load()
    assert lib.fgtpw_set_precision(7) == fgt2d_amd.FGTPW_EINVAL
    with pytest.raises(ValueError):
        fgt2d_amd.set_precision("fp16")
