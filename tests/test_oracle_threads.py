"""The OpenMP build of the oracle (oracle/liborc_omp.so, the CPU baseline on
all host cores) against the sequential build, bit for bit, and against the
reference's golden fixtures; its parallel loops keep every output's order of
additions (the near field runs per target box over the source boxes in the
sequential loop's order)."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import _pts
from tests.test_oracle_golden import NAMES, _rel, load

GOLD = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
              if "3d" not in os.path.basename(p))

CASES = [
    # (kind, ns, nt, delta, eps, nd, ifpgh, ifpghtarg, dipoles)
    ("mixture", 30000, 10000, 1e-4, 1e-9, 2, 3, 2, False),  # grad/Hess, nd = 2
    ("uniform", 60000, 0, 1e-4, 1e-9, 1, 1, 0, False),      # PW evaluation at the cutoff level
    ("mixture", 20000, 8000, 1e-5, 1e-6, 1, 2, 2, True),    # charges and dipoles
    ("uniform", 3000, 700, 1e-4, 1e-6, 1, 1, 1, False),     # near field only
]


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_threaded_oracle_equals_sequential(case):
    kind, ns, nt, delta, eps, nd, ifs, ift, dip = case
    rng = np.random.default_rng(ns + nt)
    gen = _pts.mixture if kind == "mixture" else _pts.uniform
    src = gen(rng, ns)
    tg = gen(rng, nt) if nt else None
    chg = rng.normal(size=(ns, nd))
    kw = dict(rnormal=rng.normal(size=(ns, 2)), dipstr=rng.normal(size=(ns, nd))) if dip else {}
    a = O.oracle_fgt2d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, **kw)
    b = O.oracle_fgt2d(delta, eps, src, chg, tg, ifpgh=ifs, ifpghtarg=ift, threads=4, **kw)
    for k in range(6):
        assert np.array_equal(a[k], b[k]), NAMES[k]


@pytest.mark.parametrize("path", GOLD[:4], ids=[os.path.basename(p)[:-4] for p in GOLD[:4]])
def test_threaded_oracle_matches_golden(path):
    g = load(path)
    src, chg, tg, rn, ds = _pts.golden_inputs(g)
    out = O.oracle_fgt2d(float(g["delta"]), float(g["eps"]), src, chg, tg, int(g["ifpgh"]),
                         int(g["ifpghtarg"]), rnormal=rn, dipstr=ds, threads=4)
    for k, name in enumerate(NAMES):
        if name in g:
            sel = g["isample_src"] if k < 3 else g["isample_trg"]
            assert _rel(out[k][sel], g[name]) < 1e-13, name


def test_thread_count_is_settable():
    assert O.threads_in_use(3) == 3
    assert O.threads_in_use(2) == 2
