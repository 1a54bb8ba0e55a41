"""Dipole sources (ifdipole = 1) on the GPU, through the C-ABI, against the
oracle (itself bit-for-bit with the reference libraries on dipoles, apart
from the 3D large-delta branch, pinned by the direct sum): the whole
transform with charges+dipoles and dipoles only, gradients and Hessians,
nd = 2, the near-field-only, high-order (VALU formation) and large-delta
branches, in 2D and 3D (Hermite and direct plane-wave formation routes);
the all-pairs part_direct entries; the device entry against the host
entry; a dipole request in the request server."""
import numpy as np
import pytest

import fgt2d_amd
from oracle import oracle as O
from tests import _pts

pytestmark = pytest.mark.gpu

# (kind, ns, nt, delta, eps, nd, ifpgh, ifpghtarg, ifcharge)
CASES = [
    ("mixture", 20000, 8000, 1e-5, 1e-6, 1, 0, 1, 1),
    ("uniform", 30000, 10000, 1e-4, 1e-9, 2, 3, 2, 0),    # dipoles only, grad + Hessian
    ("uniform", 100000, 0, 1e-4, 1e-9, 1, 1, 0, 1),       # PW evaluation route at level 5
    ("uniform", 5000, 2000, 1e-2, 1e-3, 1, 2, 3, 1),
    ("uniform", 2000, 700, 1e-4, 1e-6, 1, 1, 1, 0),       # near field only
    ("uniform", 1000, 300, 5.0, 1e-6, 2, 3, 3, 1),        # large delta: one Taylor expansion
    ("uniform", 1000, 300, 5.0, 1e-9, 1, 1, 1, 0),        # large delta, dipoles only
    ("uniform", 20000, 2000, 1e-3, 1e-12, 1, 1, 1, 1),    # orders above 63: VALU formation
]


def _points(kind, rng, n):
    return _pts.uniform(rng, n) if kind == "uniform" else _pts.mixture(rng, n)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _inputs(case):
    kind, ns, nt, delta, eps, nd, ifs, ift, ifc = case
    rng = np.random.default_rng(ns + nt + 17)
    src = _points(kind, rng, ns)
    tg = _points(kind, rng, nt) if nt else None
    chg = rng.normal(size=(ns, nd)) if ifc else None
    rn = rng.normal(size=(ns, 2))
    ds = rng.normal(size=(ns, nd))
    return src, chg, tg, rn, ds


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_dipole_transform_matches_oracle(case):
    kind, ns, nt, delta, eps, nd, ifs, ift, ifc = case
    print("case", case, flush=True)
    src, chg, tg, rn, ds = _inputs(case)
    got = fgt2d_amd.fgt2d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, ifdipole=1,
                          rnormal=rn, dipstr=ds)
    want = O.oracle_fgt2d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, rnormal=rn,
                          dipstr=ds)
    for k, (g, w) in enumerate(zip(got, want)):
        if g is None or (k < 3 and k >= ifs) or (k >= 3 and k - 3 >= ift):
            continue
        # the same FP64 algorithm summed in another order
        assert _rel(g, w) < 1e-10, (k, _rel(g, w))


def test_dipole_part_direct_matches_oracle():
    rng = np.random.default_rng(3)
    src = rng.random((3000, 2))
    tg = rng.random((700, 2))
    chg = rng.normal(size=(3000, 2))
    rn = rng.normal(size=(3000, 2))
    ds = rng.normal(size=(3000, 2))
    for c in (chg, None):
        for ifp in (1, 2, 3):
            got = fgt2d_amd.fgt2dpart_direct(1e-3, 1e-6, src, c, tg, ifpgh=ifp, rnormal=rn,
                                             dipstr=ds)
            want = O.direct2d_cd(1e-3, 1e-6, src, c, rn, ds, tg, ifp)
            for k in range(ifp):
                assert _rel(got[k], want[k]) < 1e-12, (ifp, k)


def test_dipole_device_entry_matches_host_entry():
    import torch
    case = CASES[0]
    kind, ns, nt, delta, eps, nd, ifs, ift, ifc = case
    src, chg, tg, rn, ds = _inputs(case)
    host = fgt2d_amd.fgt2d(delta, eps, src, chg, tg, ifpgh=1, ifpghtarg=2, ifdipole=1,
                           rnormal=rn, dipstr=ds)
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(a).to(dev) for a in (src, chg, tg, rn, ds)]
    got = fgt2d_amd.fgt2d(delta, eps, t[0], t[1], t[2], ifpgh=1, ifpghtarg=2, ifdipole=1,
                          rnormal=t[3], dipstr=t[4])
    for k in (0, 3, 4):
        assert np.array_equal(got[k].cpu().numpy(), host[k])


def test_dipole_request_in_the_server():
    case = CASES[1]
    kind, ns, nt, delta, eps, nd, ifs, ift, ifc = case
    src, chg, tg, rn, ds = _inputs(case)
    alone = fgt2d_amd.fgt2d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, ifdipole=1,
                            rnormal=rn, dipstr=ds)
    with fgt2d_amd.Server(slots=2) as srv:
        t1 = srv.submit(2, delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, rnormal=rn,
                        dipstr=ds)
        t2 = srv.submit(2, 1e-3, 1e-6, src, rn[:, :1], None, ifpgh=1)
        served, _ = srv.wait(t1)
        srv.wait(t2)
    for k in range(6):
        if alone[k] is not None:
            assert np.array_equal(served[k], alone[k]), k


# 3D: (recipe, ns, nt, delta, eps, nd, ifpgh, ifpghtarg, ifcharge)
CASES3 = [
    ("uniform", 40000, 5000, 1e-2, 1e-3, 2, 3, 2, 0),     # Hermite route, dipoles only
    ("uniform", 30000, 3000, 3e-3, 1e-9, 1, 2, 3, 1),     # direct plane-wave route
    ("uniform", 20000, 4000, 1e-3, 1e-6, 1, 1, 1, 1),
    ("uniform", 4000, 1000, 1e-4, 1e-6, 1, 3, 3, 1),      # near field only
    ("uniform", 600, 300, 5.0, 1e-6, 2, 3, 3, 1),         # large delta
    ("cluster", 20000, 4000, 1e-5, 1e-9, 1, 1, 1, 1),     # npw 38, Taylor orders 59/30/19
]


@pytest.mark.parametrize("case", CASES3, ids=[str(i) for i in range(len(CASES3))])
def test_dipole_transform_3d_matches_oracle(case):
    recipe, ns, nt, delta, eps, nd, ifs, ift, ifc = case
    print("case3d", case, flush=True)
    src, chg, tg, rn, ds = _pts.dlaran_points3d(recipe, ns, nt, nd, dipoles=True)
    chg = chg if ifc else None
    got = fgt2d_amd.fgt3d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, ifdipole=1,
                          rnormal=rn, dipstr=ds)
    want = O.oracle_fgt3d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, rnormal=rn,
                          dipstr=ds)
    for k, (g, w) in enumerate(zip(got, want)):
        if g is None or (k < 3 and k >= ifs) or (k >= 3 and k - 3 >= ift):
            continue
        assert _rel(g, w) < 1e-10, (k, _rel(g, w))


def test_dipole_part_direct_3d_matches_oracle():
    rng = np.random.default_rng(4)
    src = rng.random((2000, 3))
    tg = rng.random((500, 3))
    chg = rng.normal(size=(2000, 1))
    rn = rng.normal(size=(2000, 3))
    ds = rng.normal(size=(2000, 1))
    lib = fgt2d_amd.load()
    for c in (chg, None):
        for ifp in (1, 2, 3):
            pot, grad, hess = np.zeros((500, 1)), np.zeros((500, 3, 1)), np.zeros((500, 6, 1))
            z = np.zeros(8)
            rc = lib.fgtpw_fgt3dpart_direct(
                1, 1e-2, 0.05, 1, 2000, 1, 500, fgt2d_amd._ptr(src), int(c is not None),
                fgt2d_amd._ptr(c if c is not None else z), 1, fgt2d_amd._ptr(rn),
                fgt2d_amd._ptr(ds), fgt2d_amd._ptr(tg), ifp, fgt2d_amd._ptr(pot),
                fgt2d_amd._ptr(grad), fgt2d_amd._ptr(hess))
            assert rc == 0
            want = O.direct3d_cd(1e-2, 0.05, src, c, rn, ds, tg, ifp)
            for k, g in enumerate((pot, grad, hess)[:ifp]):
                assert _rel(g, want[k]) < 1e-12, (ifp, k)
