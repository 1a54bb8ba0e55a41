"""Dipole sources (ifdipole = 1) in the 2D and 3D oracles.

* The near-field kernels (g2d_directd*/directcd*_vec, g2ddirect.f:217-655)
  against what the mathematics fixes: a dipole of strength d and direction n
  at s is d (n . grad_s) of the charge kernel e^{-|t-s|^2/delta}, checked by
  central differences of the charge sums (pot, gradient and Hessian), and a
  charge+dipole sum is the sum of the two.
* The whole transform with dipoles against the direct sum without cutoff,
  within the requested eps (this reaches the dipole formation terms of
  g2dformhd_vec / g2dformhcd_vec and the large-delta g2dformld/lcd_vec).
* The oracle against the reference library itself where it is built."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import _pts

NAMES = ["pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"]


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("ifpgh", [1, 2, 3])
def test_dipole_kernel_is_the_source_derivative(ifpgh):
    rng = np.random.default_rng(ifpgh)
    delta = 2e-3
    src = rng.random((40, 2)) * 0.2 + 0.4
    tg = rng.random((30, 2)) * 0.2 + 0.4
    rn = rng.normal(size=(40, 2))
    ds = rng.normal(size=(40, 2))
    chg = rng.normal(size=(40, 2))
    got = O.direct2d_cd(delta, 1e-100, src, None, rn, ds, tg, ifpgh)
    h = 1e-5 * np.sqrt(delta)
    want = [np.zeros_like(v) for v in got]
    for i in range(40):
        for sgn in (1.0, -1.0):
            s1 = src[i:i + 1] + sgn * h * rn[i]
            part = O.direct2d(delta, 1e-100, s1, ds[i:i + 1], tg, ifpgh)
            for k in range(3):
                want[k] += sgn * part[k] / (2 * h)
    for k in range(ifpgh):
        assert _rel(got[k], want[k]) < 1e-7, NAMES[k]
    # charges and dipoles together are the sum of the two
    both = O.direct2d_cd(delta, 1e-100, src, chg, rn, ds, tg, ifpgh)
    only_c = O.direct2d(delta, 1e-100, src, chg, tg, ifpgh)
    for k in range(ifpgh):
        assert _rel(both[k], only_c[k] + got[k]) < 1e-13, NAMES[k]


@pytest.mark.parametrize("ifc", [1, 0])
@pytest.mark.parametrize("delta,eps,tol", [(1e-4, 1e-6, 3e-5), (1e-3, 1e-9, 3e-8), (5.0, 1e-6, 3e-5)])
def test_transform_with_dipoles_against_direct_sum(ifc, delta, eps, tol):
    rng = np.random.default_rng(7)
    src = _pts.uniform(rng, 4000)
    tg = _pts.uniform(rng, 1500)
    chg = rng.uniform(-1, 1, (4000, 1)) if ifc else None
    rn = rng.normal(size=(4000, 2))
    ds = rng.uniform(-1, 1, (4000, 1))
    out = O.oracle_fgt2d(delta, eps, src, chg, tg, ifpgh=2, ifpghtarg=3, rnormal=rn, dipstr=ds)
    ex = O.direct2d_cd(delta, 1e-100, src, chg, rn, ds, tg[:300], ifpgh=3)
    exs = O.direct2d_cd(delta, 1e-100, src, chg, rn, ds, src[:300], ifpgh=2)
    assert O.derr(ex[0], out[3][:300]) < tol
    assert O.derr(ex[1], out[4][:300]) < 10 * tol
    assert O.derr(exs[0], out[0][:300]) < tol


@pytest.mark.skipif(not O.have_ref(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("recipe,ns,nt,nd,delta,eps,ifs,ift,ifc", [
    ("cluster", 20000, 7000, 1, 1e-5, 1e-6, 1, 1, 1),
    ("uniform", 30000, 10000, 2, 1e-3, 1e-9, 3, 3, 0),
    ("cluster", 40000, 5000, 1, 3e-6, 1e-3, 2, 1, 1),
    ("uniform", 2000, 500, 2, 5.0, 1e-9, 3, 2, 1),
    ("uniform", 2000, 500, 1, 5.0, 1e-6, 1, 1, 0),
])
def test_oracle_with_dipoles_matches_reference_library(recipe, ns, nt, nd, delta, eps, ifs, ift, ifc):
    src, chg, tg, rn, ds = _pts.dlaran_points(recipe, ns, nt, nd, dipoles=True)
    chg = chg if ifc else None
    if not O.ref_defined(delta, eps, src, tg):
        pytest.skip("reference undefined (npwlevel == nlevels)")
    r = O.ref_fgt2d(delta, eps, src, chg, tg, ifs, ift, rnormal=rn, dipstr=ds)
    o = O.oracle_fgt2d(delta, eps, src, chg, tg, ifs, ift, rnormal=rn, dipstr=ds)
    for k in range(6):
        if (k < 3 and k < ifs) or (k >= 3 and k - 3 < ift):
            # same algorithm, same operation order
            assert _rel(o[k], r[k]) < 1e-13, NAMES[k]


@pytest.mark.parametrize("ifpgh", [1, 2, 3])
def test_dipole_kernel_3d_is_the_source_derivative(ifpgh):
    rng = np.random.default_rng(10 + ifpgh)
    delta = 5e-3
    src = rng.random((30, 3)) * 0.2 + 0.4
    tg = rng.random((20, 3)) * 0.2 + 0.4
    rn = rng.normal(size=(30, 3))
    ds = rng.normal(size=(30, 1))
    chg = rng.normal(size=(30, 1))
    got = O.direct3d_cd(delta, 1e30, src, None, rn, ds, tg, ifpgh)
    h = 1e-5 * np.sqrt(delta)
    want = [np.zeros_like(v) for v in got]
    for i in range(30):
        for sgn in (1.0, -1.0):
            part = O.direct3d(delta, 1e30, src[i:i + 1] + sgn * h * rn[i], ds[i:i + 1], tg, ifpgh)
            for k in range(3):
                want[k] += sgn * part[k] / (2 * h)
    for k in range(ifpgh):
        assert _rel(got[k], want[k]) < 1e-7, NAMES[k]
    both = O.direct3d_cd(delta, 1e30, src, chg, rn, ds, tg, ifpgh)
    only_c = O.direct3d(delta, 1e30, src, chg, tg, ifpgh)
    for k in range(ifpgh):
        assert _rel(both[k], only_c[k] + got[k]) < 1e-13, NAMES[k]


@pytest.mark.parametrize("ifc", [1, 0])
@pytest.mark.parametrize("n,delta,eps,tol", [(3000, 2e-2, 1e-6, 5e-5), (600, 5.0, 1e-6, 3e-5),
                                             (600, 5.0, 1e-9, 3e-8)])
def test_transform_3d_with_dipoles_against_direct_sum(ifc, n, delta, eps, tol):
    """Includes the large-delta branch, whose dipole terms this build
    computes as the mathematics fixes (the reference's g3dformld_vec indexes
    the x factors by the y order)."""
    rng = np.random.default_rng(8)
    src = rng.random((n, 3))
    tg = rng.random((400, 3))
    chg = rng.uniform(-1, 1, (n, 1)) if ifc else None
    rn = rng.normal(size=(n, 3))
    ds = rng.uniform(-1, 1, (n, 1))
    out = O.oracle_fgt3d(delta, eps, src, chg, tg, ifpgh=1, ifpghtarg=3, rnormal=rn, dipstr=ds)
    ex = O.direct3d_cd(delta, 1e30, src, chg, rn, ds, tg[:200], ifpgh=3)
    assert O.derr(ex[0], out[3][:200]) < tol
    assert O.derr(ex[1], out[4][:200]) < 10 * tol
    assert O.derr(ex[2], out[5][:200]) < 100 * tol


@pytest.mark.skipif(not O.have_ref3(), reason="3D reference build (oracle/_ref) not present")
@pytest.mark.parametrize("n,delta,eps,ifs,ift,ifc", [
    (40000, 1e-2, 1e-3, 3, 2, 1),
    (30000, 8e-3, 1e-6, 1, 1, 0),
    (30000, 3e-3, 1e-9, 2, 3, 1),
])
def test_oracle3d_with_dipoles_matches_reference_library(n, delta, eps, ifs, ift, ifc):
    src, chg, tg, rn, ds = _pts.dlaran_points3d("uniform", n, 3000, 2, dipoles=True)
    chg = chg if ifc else None
    if not O.ref3_defined(delta, eps, src, tg):
        pytest.skip("3D reference undefined for this input")
    r = O.ref_fgt3d(delta, eps, src, chg, tg, ifs, ift, rnormal=rn, dipstr=ds)
    o = O.oracle_fgt3d(delta, eps, src, chg, tg, ifs, ift, rnormal=rn, dipstr=ds)
    for k in range(6):
        if (k < 3 and k < ifs) or (k >= 3 and k - 3 < ift):
            assert _rel(o[k], r[k]) < 1e-13, NAMES[k]
