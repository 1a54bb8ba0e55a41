"""The request server of the C-ABI (fgtpw_serve_start / fgtpw_submit /
fgtpw_wait).  A mix of 2D and 3D transforms of uneven size, delta and eps,
submitted together from several threads, gives every request the same
result as the same call made alone through fgtpw_fgt2d / fgtpw_fgt3d; the
kernels have no atomics and every reduction order is fixed by the tree, so
"the same" is bit for bit (tolerance 0).  Requests overlap (more than one
runs at a time), and a budget too small for two runs them one at a time
with the same results."""
import threading

import numpy as np
import pytest

import fgt2d_amd

pytestmark = pytest.mark.gpu


def request_mix(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        dim = 2 if k % 2 == 0 else 3
        n = int(10 ** rng.uniform(3, 4.8))
        delta = 10 ** rng.uniform(-6 if dim == 2 else -4.5, -2)
        eps = (1e-3, 1e-6, 1e-9)[k % 3]
        nd = 2 if k % 5 == 4 else 1
        src = rng.random((n, dim))
        chg = rng.normal(size=(n, nd))
        tg = rng.random((n // 3 + 1, dim)) if k % 4 != 1 else None
        ifpgh = (1, 0, 2, 3)[k % 4] if tg is not None else 1 + k % 3
        ifpghtarg = (1 + k % 3) if tg is not None else 0
        out.append(dict(dim=dim, delta=delta, eps=eps, sources=src, charge=chg, targ=tg,
                        ifpgh=ifpgh, ifpghtarg=ifpghtarg))
    return out


def alone(r):
    f = fgt2d_amd.fgt2d if r["dim"] == 2 else fgt2d_amd.fgt3d
    return f(r["delta"], r["eps"], r["sources"], r["charge"], r["targ"], ifpgh=r["ifpgh"],
             ifpghtarg=r["ifpghtarg"])


def same(a, b):
    for x, y in zip(a, b):
        if x is None or y is None:
            assert x is None and y is None
            continue
        assert x.shape == y.shape
        assert np.array_equal(x, y), np.abs(x - y).max()


def served(reqs, slots, budget_gib, threads=3):
    results = [None] * len(reqs)
    with fgt2d_amd.Server(slots=slots, hbm_budget_gib=budget_gib) as srv:
        tickets = [None] * len(reqs)

        def submit(part):
            for i in part:
                r = reqs[i]
                tickets[i] = srv.submit(r["dim"], r["delta"], r["eps"], r["sources"],
                                        r["charge"], r["targ"], ifpgh=r["ifpgh"],
                                        ifpghtarg=r["ifpghtarg"])

        ts = [threading.Thread(target=submit, args=(range(j, len(reqs), threads),))
              for j in range(threads)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        for i in range(len(reqs)):
            results[i], ms = srv.wait(tickets[i])
            assert ms[2] >= ms[1] >= 0 and ms[0] >= 0
        stats = srv.stats()
    return results, stats


def test_served_mix_matches_each_call_alone():
    reqs = request_mix(11, 18)
    ref = [alone(r) for r in reqs]
    got, stats = served(reqs, slots=4, budget_gib=0.0)
    for k, (a, b) in enumerate(zip(ref, got)):
        print("request", k, reqs[k]["dim"], len(reqs[k]["sources"]), reqs[k]["delta"],
              reqs[k]["eps"], flush=True)
        same(a, b)
    assert stats["peak_running"] >= 2


def test_small_budget_runs_one_at_a_time():
    reqs = request_mix(12, 6)
    ref = [alone(r) for r in reqs]
    got, stats = served(reqs, slots=3, budget_gib=1e-6, threads=2)
    for a, b in zip(ref, got):
        same(a, b)
    assert stats["peak_running"] == 1


def test_submit_errors_are_reported_and_the_server_goes_on():
    rng = np.random.default_rng(5)
    src = rng.random((500, 2))
    chg = rng.normal(size=(500, 1))
    with fgt2d_amd.Server(slots=2) as srv:
        with pytest.raises(fgt2d_amd.FGTError):
            srv.submit(2, -1.0, 1e-6, src, chg)
        t = srv.submit(2, 1e-3, 1e-6, src, chg)
        (pot, *_), _ = srv.wait(t)
    same((pot,), (alone(dict(dim=2, delta=1e-3, eps=1e-6, sources=src, charge=chg, targ=None,
                                  ifpgh=1, ifpghtarg=0))[0],))


def test_served_large_requests_match_each_call_alone():
    """Requests up to 10^6 points (the serving stream's range) among small
    ones, 2D and 3D, each served result bit-identical to the same call made
    alone; several run at once."""
    rng = np.random.default_rng(31)
    spec = [(2, 1_000_000, 1e-5, 1e-6), (3, 400_000, 1e-4, 1e-6), (2, 3_000, 1e-3, 1e-9),
            (2, 250_000, 3e-6, 1e-3), (3, 5_000, 1e-3, 1e-3), (2, 600_000, 1e-4, 1e-9)]
    reqs = []
    for dim, n, delta, eps in spec:
        reqs.append(dict(dim=dim, delta=delta, eps=eps, sources=rng.random((n, dim)),
                         charge=rng.normal(size=(n, 1)), targ=rng.random((n // 4 + 1, dim)),
                         ifpgh=1, ifpghtarg=1))
    ref = [alone(r) for r in reqs]
    got, stats = served(reqs, slots=4, budget_gib=0.0)
    for k, (a, b) in enumerate(zip(ref, got)):
        print("request", k, spec[k], flush=True)
        same(a, b)
    assert stats["peak_running"] >= 2
