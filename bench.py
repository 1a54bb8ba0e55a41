"""Benchmark of the 2D plane-wave FGT on one MI355X.

Workload: 2M sources and 2M separate targets, each from a 10-component
Gaussian mixture (centres uniform in [0.1,0.9]^2, sigma 0.01, clipped to the
unit square), charges N(0,1) in FP64, delta = 1e-5, eps = 1e-6, potentials at
the targets (ifpgh = 0, ifpghtarg = 1).  A step is one full transform --
tree, sorts, all six steps -- with inputs already resident in HBM, through
the device entry point of the C-ABI.  Throughput = (Ns + N_eval) / t.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NS = NT = 2_000_000
DELTA, EPS = 1e-5, 1e-6
SEED = 20240601


def workload(ns=NS, nt=NT, seed=SEED):
    rng = np.random.default_rng(seed)
    mu = rng.uniform(0.1, 0.9, (10, 2))

    def mix(n):
        k = rng.integers(0, 10, n)
        return np.clip(mu[k] + rng.normal(0.0, 0.01, (n, 2)), 0.0, 1.0)

    src = mix(ns)
    tg = mix(nt)
    chg = rng.normal(size=(ns, 1))
    return src, chg, tg


def cpu_baseline(src, chg, tg, budget_s=30.0):
    """The reference itself (oracle/_ref, single core: its OpenMP build does
    not compile) on the same workload, or the C restatement when the
    reference build is absent."""
    from oracle import oracle as O
    kind = "reference" if O.have_ref() else "port"
    f = O.ref_fgt2d if kind == "reference" else O.oracle_fgt2d
    # bounded sample: whole transforms of the first n points of the workload,
    # n halved until one call takes under budget_s / 3
    n = len(src)
    reps, elapsed = 0, 0.0
    while True:
        t0 = time.perf_counter()
        f(DELTA, EPS, src[:n], chg[:n], tg[:n], ifpgh=0, ifpghtarg=1)
        dt = time.perf_counter() - t0
        if dt < budget_s / 3 or n <= 100_000:
            break
        n //= 2
    reps, elapsed = 1, dt
    while elapsed < budget_s / 3 and reps < 3:
        t0 = time.perf_counter()
        f(DELTA, EPS, src[:n], chg[:n], tg[:n], ifpgh=0, ifpghtarg=1)
        elapsed += time.perf_counter() - t0
        reps += 1
    rate = 2 * n * reps / elapsed / 1e6
    return {"value": round(rate, 4), "unit": "Mpts/s", "cores": 1, "kind": kind,
            "sample": "%d full transforms of the first %d sources and %d targets of the workload, "
                      "%.1f s" % (reps, n, n, elapsed)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--stages", action="store_true", help="print per-stage times to stderr")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch

    import fgt2d_amd

    dev = torch.device("cuda:0")
    src, chg, tg = workload()
    d_src = torch.from_numpy(src).to(dev)
    d_chg = torch.from_numpy(chg).to(dev)
    d_tg = torch.from_numpy(tg).to(dev)
    torch.cuda.synchronize()

    def step():
        return fgt2d_amd.fgt2d(DELTA, EPS, d_src, d_chg, d_tg, ifpgh=0, ifpghtarg=1)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    stats = fgt2d_amd.last_stats()

    # per-stage times of one more call, HIP events on the call's stream
    fgt2d_amd.set_timing(True)
    step()
    torch.cuda.synchronize()
    st = fgt2d_amd.stage_times()
    fgt2d_amd.set_timing(False)
    if args.stages:
        print(json.dumps({"stages_ms": st, "stats": stats}), file=sys.stderr)

    # accuracy of this very output against the FP64 direct sum without
    # cutoff on 10k targets (the GPU all-pairs kernel, itself parity-tested
    # against the oracle), cross-checked by the oracle on 50 of them
    from oracle import oracle as O
    sel = np.random.default_rng(1).permutation(NT)[:10000]
    pt = out[3].cpu().numpy()[sel]
    ex = fgt2d_amd.fgt2dpart_direct(DELTA, 1e-100, src, chg, tg[sel])[0]
    err = O.derr(ex, pt)
    ex50 = O.direct2d(DELTA, 1e-100, src, chg, tg[sel[:50]])[0]
    direct_check = float(np.abs(ex50 - ex[:50]).max() / np.abs(ex50).max())

    npts = NS + NT
    rec = {
        "metric": "Mpts/s, 2D N=2M eps=1e-6 plane-wave FGT (one MI355X, FP64)",
        "value": round(npts / t / 1e6, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "fgt2d mixture 2M sources + 2M targets, delta=1e-5, eps=1e-6, pot at targets",
                   "ns": NS, "nt": NT, "delta": DELTA, "eps": EPS, "nlevels": stats.get("nlevels"),
                   "nboxes": stats.get("nboxes"), "npw": stats.get("npw")},
        "rel_l2_err_vs_direct": err,
        "direct_oracle_check": direct_check,
        "stages_ms": {k: round(v, 4) for k, v in st.items()},
    }
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(src, chg, tg)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
