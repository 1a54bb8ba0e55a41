"""Writes the golden fixtures of tests/golden/ from the reference itself.

Run in the container that holds the reference sources, after
``make -C oracle ref`` built oracle/_ref/libref2d.so and libref3d.so from
them:

    python tests/golden/make_golden.py        # 2D fixtures
    python tests/golden/make_golden.py dip    # 2D fixtures with dipole sources
    python tests/golden/make_golden.py 3d     # 3D fixtures (dim = 3)
    python tests/golden/make_golden.py 3ddip  # 3D fixtures with dipole sources

Every fixture is data only: the name of its input recipe
(tests/_pts.dlaran_points, DLARAN uniforms and exact arithmetic, so the
inputs regenerate bit-for-bit anywhere), its parameters, a sample of point
indices and the reference's outputs at those indices, plus the integers the
reference prints (npwlevel, nlmax, nlevels, nboxes) taken from its own tree
routines.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests import _pts  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

# name: (recipe, ns, nt, nd, delta, eps, ifpgh, ifpghtarg)
CASES = {
    # the reference test driver's own configuration (test_fgt2d.f:53-145),
    # with targets evaluated too
    "reftest_1m": ("reftest", 1_000_000, 1_000_000, 1, (1 / 64.) ** 2 / (1.5 * 1.5) * 2.0 ** -4,
                   0.5e-10, 1, 1),
    "cluster_pot": ("cluster", 200_000, 100_000, 1, 1e-5, 1e-6, 1, 1),
    "uniform_pw_eval": ("uniform", 100_000, 20_000, 1, 1e-4, 1e-9, 1, 1),
    "cluster_grad_hess_nd2": ("cluster", 60_000, 30_000, 2, 1e-4, 1e-9, 3, 2),
    "uniform_eps3": ("uniform", 50_000, 20_000, 1, 1e-3, 1e-3, 2, 3),
    "large_delta": ("uniform", 5_000, 2_000, 2, 5.0, 1e-6, 3, 3),
}
# 3D, from the 3D reference library (oracle/_ref/libref3d.so)
CASES3D = {
    # the reference 3D test driver's own configuration (test_fgt3d.f:30-120),
    # with targets evaluated too
    "reftest3d_300k": ("reftest", 300_000, 300_000, 1, 1.2e-3, 1e-6, 1, 1),
    "uniform3d_gh_nd2": ("uniform", 40_000, 5_000, 2, 1e-2, 1e-3, 3, 2),
    "uniform3d_eps9": ("uniform", 50_000, 3_000, 1, 3e-3, 1e-9, 2, 3),
    "shell3d": ("shell", 60_000, 20_000, 1, 2e-3, 1e-6, 1, 1),
}
# dipole sources (fgt2d.f:35, ifdipole = 1):
# name: (recipe, ns, nt, nd, delta, eps, ifpgh, ifpghtarg, ifcharge, ifdipole)
CASES_DIP = {
    "reftest_dip_200k": ("reftest", 200_000, 100_000, 1, 2e-5, 1e-9, 1, 1, 1, 1),
    "cluster_dip_only_gh_nd2": ("cluster", 60_000, 30_000, 2, 1e-4, 1e-6, 3, 2, 0, 1),
    "uniform_dip_eps3": ("uniform", 50_000, 20_000, 1, 1e-3, 1e-3, 2, 3, 1, 1),
    "large_delta_dip": ("uniform", 5_000, 2_000, 2, 5.0, 1e-6, 3, 3, 1, 1),
}
CASES3D_DIP = {
    "uniform3d_dip_only_gh_nd2": ("uniform", 40_000, 5_000, 2, 1e-2, 1e-3, 3, 2, 0, 1),
    "uniform3d_dip_eps9": ("uniform", 50_000, 3_000, 1, 3e-3, 1e-9, 1, 3, 1, 1),
    "reftest3d_dip_300k": ("reftest", 300_000, 100_000, 1, 1.2e-3, 1e-6, 1, 1, 1, 1),
}
NSAMPLE = 1000


def sample(n, seed):
    if n <= NSAMPLE:
        return np.arange(n)
    rng = np.random.default_rng(seed)
    return np.sort(np.concatenate([np.arange(20), 20 + rng.choice(n - 20, NSAMPLE - 20, replace=False)]))


def main3d(only=None):
    if not O.have_ref3():
        sys.exit("oracle/_ref/libref3d.so missing: run make -C oracle ref first")
    for name, (recipe, ns, nt, nd, delta, eps, ifs, ift) in CASES3D.items():
        if only and name not in only:
            continue
        src, chg, tg = _pts.dlaran_points3d(recipe, ns, nt, nd)
        if not O.ref3_defined(delta, eps, src, tg):
            raise SystemExit("%s: the 3D reference is undefined for this input" % name)
        out = O.ref_fgt3d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift)
        t = O.oracle_tree3d(delta, eps, src, tg)
        isrc, itrg = sample(ns, 1), sample(nt, 2)
        rec = dict(dim=3, recipe=recipe, ns=ns, nt=nt, nd=nd, delta=delta, eps=eps, ifpgh=ifs,
                   ifpghtarg=ift, isample_src=isrc, isample_trg=itrg,
                   ints=np.array([t["npwlevel"], t["nlmax"], t["nlevels"], t["nboxes"]]))
        names = ["pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"]
        for k, v in enumerate(out):
            if (k < 3 and k < ifs) or (k >= 3 and k - 3 < ift):
                rec[names[k]] = v[isrc if k < 3 else itrg]
        # the 3D reference test's oracle: direct sum with dmax = 1e30 at the
        # first 20 sources (test_fgt3d.f:142-152)
        rec["pot_direct20"] = O.direct3d(delta, 1e30, src, chg, src[:20])[0]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
        print("wrote", name, rec["ints"])


def main3d_dip(only=None):
    if not O.have_ref3():
        sys.exit("oracle/_ref/libref3d.so missing: run make -C oracle ref first")
    for name, (recipe, ns, nt, nd, delta, eps, ifs, ift, ifc, ifd) in CASES3D_DIP.items():
        if only and name not in only:
            continue
        src, chg, tg, rn, ds = _pts.dlaran_points3d(recipe, ns, nt, nd, dipoles=True)
        chg = chg if ifc else None
        if not O.ref3_defined(delta, eps, src, tg):
            raise SystemExit("%s: the 3D reference is undefined for this input" % name)
        out = O.ref_fgt3d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, rnormal=rn,
                          dipstr=ds)
        t = O.oracle_tree3d(delta, eps, src, tg)
        isrc, itrg = sample(ns, 1), sample(nt, 2)
        rec = dict(dim=3, recipe=recipe, ns=ns, nt=nt, nd=nd, delta=delta, eps=eps, ifpgh=ifs,
                   ifpghtarg=ift, ifcharge=ifc, ifdipole=ifd, isample_src=isrc,
                   isample_trg=itrg,
                   ints=np.array([t["npwlevel"], t["nlmax"], t["nlevels"], t["nboxes"]]))
        names = ["pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"]
        for k, v in enumerate(out):
            if (k < 3 and k < ifs) or (k >= 3 and k - 3 < ift):
                rec[names[k]] = v[isrc if k < 3 else itrg]
        rec["pot_direct20"] = O.direct3d_cd(delta, 1e30, src, chg, rn, ds, src[:20])[0]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
        print("wrote", name, rec["ints"])


def main_dip(only=None):
    if not O.have_ref():
        sys.exit("oracle/_ref/libref2d.so missing: run make -C oracle ref first")
    for name, (recipe, ns, nt, nd, delta, eps, ifs, ift, ifc, ifd) in CASES_DIP.items():
        if only and name not in only:
            continue
        src, chg, tg, rn, ds = _pts.dlaran_points(recipe, ns, nt, nd, dipoles=True)
        chg = chg if ifc else None
        if not O.ref_defined(delta, eps, src, tg):
            raise SystemExit("%s: the reference is undefined for this input" % name)
        out = O.ref_fgt2d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, rnormal=rn,
                          dipstr=ds)
        t = O.oracle_tree2d(delta, eps, src, tg)
        isrc, itrg = sample(ns, 1), sample(nt, 2)
        rec = dict(recipe=recipe, ns=ns, nt=nt, nd=nd, delta=delta, eps=eps, ifpgh=ifs,
                   ifpghtarg=ift, ifcharge=ifc, ifdipole=ifd, isample_src=isrc,
                   isample_trg=itrg,
                   ints=np.array([t["npwlevel"], t["nlmax"], t["nlevels"], t["nboxes"]]))
        names = ["pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"]
        for k, v in enumerate(out):
            if (k < 3 and k < ifs) or (k >= 3 and k - 3 < ift):
                rec[names[k]] = v[isrc if k < 3 else itrg]
        # direct sum without cutoff at the first 20 sources
        rec["pot_direct20"] = O.direct2d_cd(delta, 1e-100, src, chg, rn, ds, src[:20])[0]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
        print("wrote", name, rec["ints"])


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "3d":
        return main3d(sys.argv[2:])
    if len(sys.argv) > 1 and sys.argv[1] == "dip":
        return main_dip(sys.argv[2:])
    if len(sys.argv) > 1 and sys.argv[1] == "3ddip":
        return main3d_dip(sys.argv[2:])
    if not O.have_ref():
        sys.exit("oracle/_ref/libref2d.so missing: run make -C oracle ref first")
    for name, (recipe, ns, nt, nd, delta, eps, ifs, ift) in CASES.items():
        src, chg, tg = _pts.dlaran_points(recipe, ns, nt, nd)
        if not O.ref_defined(delta, eps, src, tg):
            raise SystemExit("%s: the reference is undefined for this input" % name)
        out = O.ref_fgt2d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift)
        t = O.oracle_tree2d(delta, eps, src, tg)
        rt = O.ref_tree2d(delta, eps, src, tg) if ns <= 200_000 else None
        if rt is not None:
            assert rt["nboxes"] == t["nboxes"] and np.array_equal(rt["par"], t["par"])
        isrc, itrg = sample(ns, 1), sample(nt, 2)
        rec = dict(recipe=recipe, ns=ns, nt=nt, nd=nd, delta=delta, eps=eps, ifpgh=ifs,
                   ifpghtarg=ift, isample_src=isrc, isample_trg=itrg,
                   ints=np.array([t["npwlevel"], t["nlmax"], t["nlevels"], t["nboxes"]]))
        names = ["pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"]
        for k, v in enumerate(out):
            if (k < 3 and k < ifs) or (k >= 3 and k - 3 < ift):
                rec[names[k]] = v[isrc if k < 3 else itrg]
        # the reference test's oracle: direct sum without cutoff at the
        # first 20 sources (test_fgt2d.f:161-169)
        rec["pot_direct20"] = O.direct2d(delta, 1e-100, src, chg, src[:20])[0]
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
        print("wrote", name, rec["ints"])


if __name__ == "__main__":
    main()
