"""Tile-shape autotuner of the 2D transform.

Measures the transform of a workload (default: 2M + 2M Gaussian-mixture
sources and targets, delta 1e-5, eps 1e-6, potentials at the targets, inputs
resident on the GPU) under candidate tile shapes and writes the best to
``fgt2d_amd/tuning.txt``, the table the library reads at load time (beside
``libfgtpw.so``; ``FGTPW_TUNE`` overrides it).  The shapes, keyed by the order
tile count TM of a tree level (orders up to 16 TM):

* ``segfTM``  points per formation segment (one wavefront's moments);
* ``segeTM``  points per evaluation segment;
* ``spwTM``   consecutive segments per wavefront of the fused PW -> Taylor
  evaluation;
* ``sort``    32-bit key radix sort: 4 and 5 the library's stable LSD sort
  (tiles of 2048 or 4096 keys), 0..3 the block shape of rocPRIM's onesweep
  sort.

Search: coordinate descent from the current table, one key at a time; a
candidate replaces the incumbent when its median step time is more than
``--margin`` lower.  Every measurement runs in a child process (the library
reads the table once per process).

    python -m fgt2d_amd.autotune            # tune, write tuning.txt
    python -m fgt2d_amd.autotune --dry-run  # print the table only
"""
import argparse
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TABLE = os.path.join(HERE, "tuning.txt")

# candidate values per key, the first one of each the built-in default
SPACE = [
    ("sort", [4, 5, 0, 1, 2, 3]),
    ("segf1", [256, 128, 512]),
    ("sege1", [256, 128]),
    ("spw1", [2, 1, 4]),
    ("segf2", [256, 128]),
    ("sege2", [256, 128]),
    ("spw2", [2, 1, 4]),
    ("segf3", [128, 256]),
]


def read_table(path=TABLE):
    """key -> value of a tuning table (missing file: empty)."""
    out = {}
    if os.path.exists(path):
        for line in open(path):
            line = line.split("#", 1)[0].strip()
            if "=" in line:
                k, v = line.split("=", 1)
                out[k.strip()] = int(v)
    return out


def measure_here(ns, nt, delta, eps, steps, warmup, reps):
    """Median milliseconds per transform in this process (FGTPW_TUNE set by
    the parent)."""
    import numpy as np
    import torch

    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import fgt2d_amd

    rng = np.random.default_rng(20240601)
    mu = rng.uniform(0.1, 0.9, (10, 2))

    def mix(n):
        k = rng.integers(0, 10, n)
        return np.clip(mu[k] + rng.normal(0.0, 0.01, (n, 2)), 0.0, 1.0)

    dev = torch.device("cuda:0")
    src = torch.from_numpy(mix(ns)).to(dev)
    tg = torch.from_numpy(mix(nt)).to(dev)
    chg = torch.from_numpy(rng.normal(size=(ns, 1))).to(dev)

    def step():
        return fgt2d_amd.fgt2d(delta, eps, src, chg, tg, ifpgh=0, ifpghtarg=1)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        times.append((time.perf_counter() - t0) / steps * 1e3)
    times.sort()
    return times[len(times) // 2]


def measure(table, args):
    """Runs one measurement of `table` in a child process."""
    env = dict(os.environ)
    env["FGTPW_TUNE"] = ",".join("%s=%d" % kv for kv in sorted(table.items()))
    cmd = [sys.executable, "-m", "fgt2d_amd.autotune", "--child", "--ns", str(args.ns), "--nt",
           str(args.nt), "--delta", str(args.delta), "--eps", str(args.eps), "--steps",
           str(args.steps), "--warmup", str(args.warmup), "--reps", str(args.reps)]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError("measurement failed for %s:\n%s" % (env["FGTPW_TUNE"], r.stderr[-2000:]))
    return float(r.stdout.strip().splitlines()[-1])


def tune(args):
    best = {k: vals[0] for k, vals in SPACE}
    best.update(read_table())
    log = []
    best_ms = measure(best, args)
    log.append((dict(best), best_ms))
    print("start %s: %.4f ms" % (best, best_ms), flush=True)
    for key, vals in SPACE:
        for v in vals:
            if v == best[key]:
                continue
            cand = dict(best)
            cand[key] = v
            ms = measure(cand, args)
            log.append((dict(cand), ms))
            print("  %s=%d: %.4f ms" % (key, v, ms), flush=True)
            if ms < best_ms * (1 - args.margin):
                best, best_ms = cand, ms
                print("  -> keep %s=%d" % (key, v), flush=True)
    return best, best_ms, log


def write_table(best, best_ms, log, args, path=TABLE):
    try:
        import torch
        device = torch.cuda.get_device_name(0)
    except Exception:
        device = "unknown device"
    lines = ["# Tile shapes of the 2D transform, written by python -m fgt2d_amd.autotune",
             "# device: %s; %s" % (device, time.strftime("%Y-%m-%d")),
             "# workload: %d + %d Gaussian-mixture points, delta %g, eps %g, pot at targets, FP64"
             % (args.ns, args.nt, args.delta, args.eps),
             "# median ms per transform of %d x %d steps per candidate; a candidate is kept when "
             "%.1f%% faster" % (args.reps, args.steps, 100 * args.margin),
             "# measured:"]
    for t, ms in log:
        lines.append("#   %s  %.4f ms" % (json.dumps(t, sort_keys=True), ms))
    lines.append("# best: %.4f ms" % best_ms)
    for k, _ in SPACE:
        lines.append("%s=%d" % (k, best[k]))
    open(path, "w").write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ns", type=int, default=2_000_000)
    ap.add_argument("--nt", type=int, default=2_000_000)
    ap.add_argument("--delta", type=float, default=1e-5)
    ap.add_argument("--eps", type=float, default=1e-6)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--margin", type=float, default=0.015)
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--dry-run", action="store_true")
    ap.add_argument("--child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.child:
        print("%.6f" % measure_here(args.ns, args.nt, args.delta, args.eps, args.steps,
                                    args.warmup, args.reps))
        return
    if args.dry_run:
        print(read_table(args.out))
        return
    best, best_ms, log = tune(args)
    write_table(best, best_ms, log, args, args.out)
    print("wrote %s" % args.out)


if __name__ == "__main__":
    main()
