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


# MI355X published peaks (dense; FP32 matrix and vector 157.3 TF, FP64 78.6 TF).
# The near field is VALU-bound: 256 CUs x 4 SIMDs x 16 lanes at 2.4 GHz issue
# 39.3 T lane-instructions/s, and the compiled pair loop of k_direct<1, false>
# (the FP64 exp path of one source-target pair) is 42 VALU instructions, so
# at most 39.3e12 / 42 = 936 G pairs/s.
VALU_PER_PAIR = 42
PEAK = {"mfma": (78.6, "TFLOP/s"), "mfma32": (157.3, "TFLOP/s"), "hbm": (8000.0, "GB/s"),
        "valu": (round(256 * 4 * 16 * 2.4e9 / VALU_PER_PAIR / 1e9, 1), "Gpair/s")}


def sort_passes(keybits):
    """Radix passes of the 32-bit key sort: digits of 9, 11 or 10 bits for
    keys of up to 18, 22 or 30 bits (None: 64-bit keys)."""
    if keybits <= 22:
        return 2
    return 3 if keybits <= 30 else None


def per_unit(kernel, npw, nd, p, level=-1, mixed=False, keybits=18):
    """Algorithmic work of one unit of an instrumented kernel on the route it
    runs, (bound, amount, what), for expansion order p at its level.
    Hermite formation accumulates one FMA per coefficient (j, k <= p) and
    density per source, 2(p+1)^2 nd flop; Taylor evaluation the same per
    evaluation point (potential); plane-wave evaluation contracts the
    npw/2 x npw half-spectrum per point, 8 (npw/2) npw + 8 npw flop per
    density; translations move 48 bytes per complex coefficient and box
    (16 B read, 32 B read-modify-write), 24 with FP32 expansions; the
    source gather reads and writes a point (16 B each way), its densities
    (8 nd each way) and reads its index (4 B); the target gather (level -2)
    the point and index only; the scatter (the points that no evaluation
    covers) reads, writes and indexes their outputs."""
    half = npw // 2
    mf = "mfma32" if mixed else "mfma"
    if kernel in ("form_moments", "form_valu"):
        return mf if kernel == "form_moments" else "mfma", 2.0 * (p + 1) ** 2 * nd, \
            "flop per source"
    if kernel == "taylor_eval":
        return (mf if p + 1 <= 32 else "mfma"), 2.0 * (p + 1) ** 2 * nd, "flop per evaluation point"
    if kernel == "pw_eval":
        return "mfma", (8.0 * half * npw + 8.0 * npw) * nd, "flop per evaluation point"
    if kernel in ("merge", "m2l", "split"):
        return "hbm", (24.0 if mixed else 48.0) * half * npw * nd, "bytes per box"
    if kernel == "gather":
        if level == -2:
            return "hbm", 2 * 16.0 + 4, "bytes per point"
        return "hbm", 2 * 8.0 * (2 + nd) + 4, "bytes per point"
    if kernel == "scatter":
        # points no evaluation covers: the near field's sorted output read,
        # written in input order, the permutation index
        return "hbm", 2 * 8.0 * nd + 4, "bytes per point"
    if kernel == "sort":
        # per pass: key and index read and written (16 B); the digit
        # histograms of all passes: one read of the keys (4 B)
        np_ = sort_passes(keybits)
        if np_ is None:
            return None
        return "hbm", 16.0 * np_ + 4, "bytes per key (%d radix passes)" % np_
    if kernel == "direct":
        return "valu", 1.0, "source-target pair (%d VALU instructions)" % VALU_PER_PAIR
    return None


def pw_equivalent(kernel, npw, nd):
    """The plane-wave-direct flop count of one source (formation) or one
    evaluation point: the work the Hermite/Taylor route replaces."""
    half = npw // 2
    if kernel in ("form_moments", "form_valu"):
        return 8.0 * half * npw * nd
    if kernel in ("taylor_eval", "pw_eval"):
        return (8.0 * half * npw + 8.0 * npw) * nd
    return None


def rocprof_name(kernel, nl, ifpgh, mixed=False, npw=0, sort=4):
    """Symbol of the HIP kernel an instrumented launch runs (as rocprofv3
    lists it, template arguments included), for the Taylor order nl of its
    level; `sort` is the 32-bit key sort in use (tuning key sort)."""
    tm = min(4, (nl + 1 + 15) // 16)
    f32 = "_f32" if mixed else ""
    if kernel == "form_moments":
        return "k_form_wave%s<%d, false>" % (f32, tm)
    if kernel == "taylor_eval":
        if nl + 1 <= 32:
            return "k_taylor_wave%s<%d, %d>" % (f32 if mixed else "", tm, ifpgh)
        return "k_taylor_mfma<%d, %d>" % (tm, ifpgh)
    if kernel == "form_h2pw":
        return "k_form_h2pw<%s>" % ("float" if mixed else "double")
    if kernel in ("merge", "m2l", "split"):
        return "k_%s<%s>" % (kernel, "float" if mixed else "double")
    if kernel == "pw2local":
        tt = "float, float" if mixed and nl + 1 <= 32 else ("float, double" if mixed else
                                                           "double, double")
        if nl + 1 <= 48 and npw <= 64:
            return "k_pw2local_wave<%d, %d, %s>" % (tm, (npw // 2 + 15) // 16, tt)
        return "k_pw2local<%s>" % tt
    if kernel == "pw_eval":
        return "k_pw_eval_mfma<%d, %s>" % (ifpgh, "float" if mixed else "double")
    if kernel == "gather":
        return "k_gather<true>"
    if kernel == "direct":
        return "k_direct<%d, false>" % ifpgh
    if kernel == "scatter":
        # the uncovered leaves' points (k_scatter when a transform has no
        # expansions at all)
        return "k_scatter_boxes"
    if kernel == "sort":
        if sort in (4, 5):
            return "k_rs_count+k_rs_offsets+k_rs_scatter"
        return "rocprim radix sort (onesweep passes, histogram, fills)"
    return "k_" + kernel


def read_traffic(name, workload="c2"):
    """HBM bytes per launch of kernel `name` from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel factor),
    or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_%s.txt" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = dict(kv.split("=", 1) for kv in line.split() if "=" in kv)
        if f.get("kernel") == name and "hbm_bytes" in f:
            return float(f["hbm_bytes"])
    return None


def kernel_profile(step, torch, reps=5):
    """Runs `reps` more transforms with serial timing: the side streams' work
    queued on the call's stream (no two kernels overlap), every instrumented
    kernel dispatched with a start and an end event that the dispatch itself
    stamps, so each record is the kernel's own duration (as rocprofv3's
    kernel trace of a run with FGTPW_SERIAL=1 measures it); per (kernel,
    level): total ms, launches and units, plus the stage times of the last
    transform."""
    import fgt2d_amd
    fgt2d_amd.set_timing("serial")
    agg = {}
    try:
        for _ in range(reps):
            step()
            torch.cuda.synchronize()
            for r in fgt2d_amd.kernel_launches():
                a = agg.setdefault((r["kernel"], r["level"]), [0.0, 0, 0])
                a[0] += r["ms"]
                a[1] += 1
                a[2] += r["units"]
        st = fgt2d_amd.stage_times()
    finally:
        fgt2d_amd.set_timing(False)
    return agg, st


def rooflines(agg, stats, nd, count=5, mixed=False):
    """Roofline records of the instrumented kernels with an algorithmic
    per-unit figure, largest total time first."""
    import fgt2d_amd
    npw, nlocal = stats["npw"], stats["nlocal"]
    sort = fgt2d_amd.tuning_table().get("sort", 4)
    ranked = sorted(agg.items(), key=lambda kv: -kv[1][0])
    total = sum(v[0] for v in agg.values())
    out = []
    for (kernel, level), (ms, n, units) in ranked:
        nl = nlocal[level] if 0 <= level < len(nlocal) else 0
        pu = per_unit(kernel, npw, nd, nl, level, mixed, keybits=max(2, 2 * stats["nlmax"]))
        if pu is None:
            continue
        bound, amount, what = pu
        launch_s = ms / n / 1e3
        work = amount * units / n
        peak, unit = PEAK[bound]
        scale = 1e12 if bound.startswith("mfma") else 1e9
        achieved = work / launch_s / scale
        f32 = mixed and (bound == "mfma32" or kernel in ("merge", "m2l", "split"))
        name = rocprof_name(kernel, nl, 1, f32, npw, sort)
        rec = {"bound": "mfma" if bound.startswith("mfma") else bound,
               "achieved": round(achieved, 3), "peak": peak, "unit": unit,
               "frac": round(achieved / peak, 4),
               "traffic": read_traffic(name + ({-1: "@sources", -2: "@targets"}[level]
                                               if kernel == "gather" else "")),
               "kernel": name, "level": level, "launch_us": round(launch_s * 1e6, 2),
               "units_per_launch": units // n, "per_unit": "%g %s" % (amount, what),
               "share_of_kernel_time": round(ms / total, 3)}
        if bound == "mfma32":
            rec["precision"] = "fp32"
        eq = pw_equivalent(kernel, npw, nd)
        if eq is not None:
            rec["pw_equivalent_rate"] = round(eq * units / n / launch_s / scale, 3)
        out.append(rec)
        if len(out) == count:
            break
    return out


def cpu_baseline():
    """Runs the CPU leg in a child process (the reference prints through the
    Fortran runtime, which flushes at process exit) and returns its record."""
    import subprocess
    import tempfile
    fd, path = tempfile.mkstemp(suffix=".json")
    os.close(fd)
    try:
        subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", path],
                       stdout=sys.stderr, stderr=sys.stderr, check=True, timeout=600)
        with open(path) as f:
            return json.load(f)
    finally:
        os.unlink(path)


def _timed_transforms(f, src, chg, tg, budget_s, **kw):
    """Whole transforms of the first n points of the workload, n halved
    until one call takes under budget_s / 3; then up to 3 calls within
    budget_s / 3 seconds.  Returns (n, reps, elapsed seconds)."""
    n = len(src)
    while True:
        t0 = time.perf_counter()
        f(DELTA, EPS, src[:n], chg[:n], tg[:n], ifpgh=0, ifpghtarg=1, **kw)
        dt = time.perf_counter() - t0
        if dt < budget_s / 3 or n <= 100_000:
            break
        n //= 2
    reps, elapsed = 1, dt
    while elapsed < budget_s / 3 and reps < 3:
        t0 = time.perf_counter()
        f(DELTA, EPS, src[:n], chg[:n], tg[:n], ifpgh=0, ifpghtarg=1, **kw)
        elapsed += time.perf_counter() - t0
        reps += 1
    return n, reps, elapsed


def cpu_leg(src, chg, tg, budget_s=30.0):
    """The CPU baseline on the same workload.  Headline: the oracle's C
    restatement built with OpenMP (oracle/liborc_omp.so: its per-box loops
    threaded, results bit-identical to the sequential build) on all the
    host's cores given to this process (OMP_NUM_THREADS, else the CPU
    count).  Beside it: the reference itself (oracle/_ref, one core: its own
    OpenMP build does not compile), when built."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 1
    used = O.threads_in_use(threads)
    n, reps, el = _timed_transforms(O.oracle_fgt2d, src, chg, tg, budget_s, threads=used)
    rec = {"value": round(2 * n * reps / el / 1e6, 4), "unit": "Mpts/s", "cores": used,
           "kind": "port", "nproc": os.cpu_count(),
           "sample": "%d full transforms of the first %d sources and %d targets of the workload "
                     "on %d OpenMP threads, %.1f s" % (reps, n, n, used, el)}
    if O.have_ref():
        n, reps, el = _timed_transforms(O.ref_fgt2d, src, chg, tg, budget_s)
        rec["reference_1core"] = {
            "value": round(2 * n * reps / el / 1e6, 4), "unit": "Mpts/s", "cores": 1,
            "kind": "reference",
            "sample": "%d full transforms of the first %d sources and %d targets of the workload, "
                      "%.1f s" % (reps, n, n, el)}
    return rec


N3 = 1_000_000
DELTA3, EPS3 = 1e-3, 1e-6
N5 = 8_000_000
DELTA5, EPS5 = 1e-5, 1e-9


def shell_workload(n=N5, seed=SEED):
    """8M sources on the sphere shell of radius 0.4 about (0.5,0.5,0.5) with
    radial noise of standard deviation 1e-3 (directions uniform on the
    sphere), 8M targets uniform in [0,1)^3, charges U[-1,1]."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    src = 0.5 + d * (0.4 + rng.normal(0.0, 1e-3, n))[:, None]
    tg = rng.random((n, 3))
    chg = rng.uniform(-1, 1, (n, 1))
    return src, chg, tg


def bench3d(args, torch, fgt2d_amd):
    """3D.  c3: 1M sources = targets uniform in [0,1)^3, charges U[-1,1],
    delta 1e-3, eps 1e-6, potentials at the sources (ifpgh = 1).  c5: the
    8M + 8M shell workload, delta 1e-5, eps 1e-9, potentials at the targets
    (ifpghtarg = 1)."""
    c5 = args.workload == "c5"
    if c5:
        src, chg, tg = shell_workload()
        delta, eps, nsel = DELTA5, EPS5, 10000
        npts = 2 * N5
    else:
        rng = np.random.default_rng(SEED)
        src = rng.random((N3, 3))
        chg = rng.uniform(-1, 1, (N3, 1))
        tg = None
        delta, eps, nsel = DELTA3, EPS3, 5000
        npts = 2 * N3  # Ns + N_eval, potentials evaluated at the 1M sources
    dev = torch.device("cuda:0")
    d_src = torch.from_numpy(src).to(dev)
    d_chg = torch.from_numpy(chg).to(dev)
    d_tg = torch.from_numpy(tg).to(dev) if c5 else None
    torch.cuda.synchronize()

    def step():
        if c5:
            return fgt2d_amd.fgt3d(delta, eps, d_src, d_chg, d_tg, ifpgh=0, ifpghtarg=1)
        return fgt2d_amd.fgt3d(delta, eps, d_src, d_chg, None, ifpgh=1, ifpghtarg=0)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    stats = fgt2d_amd.last_stats()
    agg, st = kernel_profile(step, torch, reps=2)
    if args.stages:
        ks = {"%s@%d" % k: [round(v[0] / v[1] * 1e3, 2), v[1], v[2] // v[1]]
              for k, v in sorted(agg.items(), key=lambda kv: -kv[1][0])}
        print(json.dumps({"stages_ms": st, "stats": stats, "kernels_us_n_units": ks}),
              file=sys.stderr)
    pts = tg if c5 else src
    sel = np.random.default_rng(1).permutation(len(pts))[:nsel]
    ex = fgt2d_amd.fgt3dpart_direct(delta, 1e30, src, chg, pts[sel])[0]
    from oracle import oracle as O
    err = O.derr(ex, out[3 if c5 else 0].cpu().numpy()[sel])
    if c5:
        metric = "Mpts/s, 3D 8M shell sources + 8M targets eps=1e-9 plane-wave FGT (one MI355X, FP64)"
        wl = ("fgt3d 8M sources on a sphere shell (r=0.4, radial sd 1e-3) + 8M uniform targets, "
              "delta=1e-5, eps=1e-9, pot at targets")
    else:
        metric = "Mpts/s, 3D N=1M eps=1e-6 plane-wave FGT (one MI355X, FP64)"
        wl = "fgt3d uniform 1M sources = targets, delta=1e-3, eps=1e-6, pot at sources"
    rec = {"metric": metric,
           "value": round(npts / t / 1e6, 3), "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(t * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": wl, "ns": len(src), "nt": 0 if tg is None else len(tg),
                      "delta": delta, "eps": eps, "nlevels": stats.get("nlevels"),
                      "nboxes": stats.get("nboxes"), "npw": stats.get("npw")},
           "rel_l2_err_vs_direct": err,
           "stages_ms": {k: round(v, 4) for k, v in st.items()}}
    print(json.dumps(rec))


N1 = 100_000
DELTA1, EPS1 = 1e-4, 1e-9


def bench_c1(args, torch, fgt2d_amd):
    """2D, 100k sources = targets uniform in [0,1)^2, charges U[-1,1],
    delta 1e-4, eps 1e-9, potentials at the sources: the size of the
    reference's own test.  The cutoff level's Taylor order is above 50, so
    its leaves are evaluated from their plane waves (k_pw_eval_mfma)."""
    rng = np.random.default_rng(SEED)
    src = rng.random((N1, 2))
    chg = rng.uniform(-1, 1, (N1, 1))
    dev = torch.device("cuda:0")
    d_src = torch.from_numpy(src).to(dev)
    d_chg = torch.from_numpy(chg).to(dev)
    torch.cuda.synchronize()

    def step():
        return fgt2d_amd.fgt2d(DELTA1, EPS1, d_src, d_chg, None, ifpgh=1, ifpghtarg=0)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    stats = fgt2d_amd.last_stats()
    agg, st = kernel_profile(step, torch, reps=3)
    roofs = rooflines(agg, stats, 1, count=5)
    sel = np.random.default_rng(1).permutation(N1)[:5000]
    ex = fgt2d_amd.fgt2dpart_direct(DELTA1, 1e-100, src, chg, src[sel])[0]
    from oracle import oracle as O
    err = O.derr(ex, out[0].cpu().numpy()[sel])
    rec = {"metric": "Mpts/s, 2D N=100k eps=1e-9 plane-wave FGT (one MI355X, FP64)",
           "value": round(2 * N1 / t / 1e6, 3), "unit": "Mpts/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t * 1e3, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "fgt2d uniform 100k sources = targets, delta=1e-4, eps=1e-9, "
                                  "pot at sources", "ns": N1, "nt": 0, "delta": DELTA1,
                      "eps": EPS1, "nlevels": stats.get("nlevels"), "nboxes": stats.get("nboxes"),
                      "npw": stats.get("npw")},
           "rel_l2_err_vs_direct": err,
           "roofline": roofs[0] if roofs else None, "roofline_next": roofs[1:],
           "stages_ms": {k: round(v, 4) for k, v in st.items()}}
    print(json.dumps(rec))


POOL = 1_000_000


def serve_stream(fgt2d_amd, nreq, rate=200.0, slots=4, seed=SEED, check=24):
    """Serving stream: `nreq` requests with seed-fixed Poisson arrivals at
    `rate` per second; dimension 2 or 3 (p = 1/2), N = round(10^U[3,6])
    points uniform in the unit box with sources = targets and potentials at
    the sources, charges N(0,1), delta = 10^U[-6,-2], eps from {1e-3, 1e-6,
    1e-9}.  The points of a request are a random slice of a pool of 10^6
    uniform points per dimension (no generation inside the timed stream).
    Latency runs from the scheduled arrival to completion.  The rel. l2 error
    of `check` requests of at most 2*10^5 points is taken against the GPU
    direct sum on 1000 of their points."""
    import queue
    import threading
    rng = np.random.default_rng(seed)
    pools = {d: rng.random((POOL, d)) for d in (2, 3)}
    qpool = rng.normal(size=(POOL, 1))
    arrive = np.cumsum(rng.exponential(1.0 / rate, nreq))
    dims = rng.choice([2, 3], nreq)
    npts = np.rint(10 ** rng.uniform(3, 6, nreq)).astype(np.int64)
    deltas = 10 ** rng.uniform(-6, -2, nreq)
    epss = rng.choice([1e-3, 1e-6, 1e-9], nreq)
    offs = rng.integers(0, POOL - npts + 1)
    small = np.flatnonzero(npts <= 200_000)
    checked = set(rng.choice(small, min(check, len(small)), replace=False).tolist())

    def req(i):
        o, n = int(offs[i]), int(npts[i])
        return pools[int(dims[i])][o:o + n], qpool[o:o + n]

    srv = fgt2d_amd.Server(slots=slots)
    tickets = queue.Queue()
    submitted = np.zeros(nreq)
    t0 = time.perf_counter()

    def submitter():
        for i in range(nreq):
            dt = t0 + arrive[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            src, chg = req(i)
            submitted[i] = time.perf_counter() - t0
            tickets.put((i, srv.submit(int(dims[i]), deltas[i], epss[i], src, chg, None,
                                       ifpgh=1, ifpghtarg=0)))

    th = threading.Thread(target=submitter)
    th.start()
    lat = np.zeros(nreq)
    done = np.zeros(nreq)
    qr = np.zeros((nreq, 2))
    errs = {}
    try:
        for _ in range(nreq):
            i, t = tickets.get()
            outs, ms = srv.wait(t)
            lat[i] = submitted[i] - arrive[i] + ms[2] / 1e3
            done[i] = submitted[i] + ms[2] / 1e3
            qr[i] = ms[0], ms[1]
            if i in checked:
                src, chg = req(i)
                sel = np.random.default_rng(i).permutation(len(src))[:1000]
                f = fgt2d_amd.fgt2dpart_direct if dims[i] == 2 else fgt2d_amd.fgt3dpart_direct
                # no cutoff: eps 1e-300 (2D) and dmax 1e300 (3D)
                ex = f(deltas[i], 1e-300 if dims[i] == 2 else 1e300, src, chg, src[sel])[0]
                e = float(np.linalg.norm(ex - outs[0][sel]) / np.linalg.norm(ex))
                k = "%g" % epss[i]
                errs[k] = max(errs.get(k, 0.0), e)
            del outs
    finally:
        th.join()
        stats = srv.stats()
        srv.close()
    span = done.max()
    pts = float(2 * npts.sum())
    lat_ms = lat * 1e3
    return {"requests": nreq, "rate_per_s": rate, "slots": slots,
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "mean_ms": round(float(lat_ms.mean()), 3), "max_ms": round(float(lat_ms.max()), 3),
            "Mpts_s": round(pts / span / 1e6, 3),
            "offered_Mpts_s": round(pts / arrive[-1] / 1e6, 3), "span_s": round(float(span), 3),
            "peak_running": stats.get("peak_running"),
            "max_rel_l2_err_by_eps": {k: float("%.3g" % v) for k, v in sorted(errs.items())},
            "checked_requests": len(checked),
            # the slowest requests: where their latency went
            "worst": [{"arrive_s": round(float(arrive[i]), 3), "dim": int(dims[i]), "n": int(npts[i]),
                       "eps": float(epss[i]),
                       "submit_delay_ms": round(float(submitted[i] - arrive[i]) * 1e3, 2),
                       "queue_ms": round(float(qr[i, 0]), 2), "run_ms": round(float(qr[i, 1]), 2)}
                      for i in np.argsort(-lat_ms)[:5]]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--stages", action="store_true", help="print per-stage times to stderr")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"],
                    help="c2: the 2D 2M+2M mixture (default, the headline metric); "
                         "c1: the 2D 100k uniform configuration at eps 1e-9 (the reference's "
                         "own test size); "
                         "c3: the 3D 1M uniform configuration; "
                         "c4: the serving stream alone (10k requests at 200/s); "
                         "c5: the 3D 8M shell + 8M targets at eps 1e-9")
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests of the serving stream reported beside c2 (0: none)")
    ap.add_argument("--slots", type=int, default=4, help="server slots of the serving stream")
    ap.add_argument("--serve-capacity", type=int, default=500,
                    help="requests of the same mix submitted at once, for the server's capacity "
                         "(0: none)")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"],
                    help="c2 legs: both (default; the mixed-precision route is the headline), "
                         "or one of them")
    ap.add_argument("--host-steps", type=int, default=5,
                    help="c2: timed transforms through the host-array entry point (0: none)")
    ap.add_argument("--cpu-leg", metavar="OUT", help=argparse.SUPPRESS)
    ap.add_argument("--copy-calib", action="store_true",
                    help="end with one 256 MiB device copy (the reference kernel for PMC byte counts)")
    args = ap.parse_args()
    if args.cpu_leg:  # child process of cpu_baseline(): CPU only, no GPU
        with open(args.cpu_leg, "w") as f:
            json.dump(cpu_leg(*workload()), f)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch

    import fgt2d_amd

    if args.workload in ("c3", "c5"):
        return bench3d(args, torch, fgt2d_amd)
    if args.workload == "c1":
        return bench_c1(args, torch, fgt2d_amd)
    if args.workload == "c4":
        rec = serve_stream(fgt2d_amd, 10_000, slots=args.slots)
        print(json.dumps({"metric": "p50/p99 request latency, serving stream (one MI355X, FP64)",
                          "value": rec["p50_ms"], "unit": "ms", "n_gpus": 1,
                          "higher_is_better": False, "dtype": "f64", "data": "synthetic",
                          "config": {"workload": "serving: 10k requests, Poisson 200/s, 2D/3D, "
                                                 "N 10^U[3,6], delta 10^U[-6,-2], eps 1e-3/6/9"},
                          "serving": rec}))
        return
    dev = torch.device("cuda:0")
    src, chg, tg = workload()
    d_src = torch.from_numpy(src).to(dev)
    d_chg = torch.from_numpy(chg).to(dev)
    d_tg = torch.from_numpy(tg).to(dev)
    torch.cuda.synchronize()
    from oracle import oracle as O

    def step():
        return fgt2d_amd.fgt2d(DELTA, EPS, d_src, d_chg, d_tg, ifpgh=0, ifpghtarg=1)

    # the FP64 direct sum without cutoff on 10k targets (the GPU all-pairs
    # kernel, itself parity-tested against the oracle), cross-checked by the
    # oracle on 50 of them
    sel = np.random.default_rng(1).permutation(NT)[:10000]
    ex = fgt2d_amd.fgt2dpart_direct(DELTA, 1e-100, src, chg, tg[sel])[0]
    ex50 = O.direct2d(DELTA, 1e-100, src, chg, tg[sel[:50]])[0]
    direct_check = float(np.abs(ex50 - ex[:50]).max() / np.abs(ex50).max())

    def leg(precision):
        """W untimed and K timed transforms in one precision mode; the rel.
        l2 error of the last output against the direct sum; per-kernel times
        of 5 more transforms (serial timing)."""
        fgt2d_amd.set_precision(precision)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = step()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        ran = fgt2d_amd.last_precision()
        stats = fgt2d_amd.last_stats()
        err = O.derr(ex, out[3].cpu().numpy()[sel])
        del out
        agg, st = kernel_profile(step, torch)
        if args.stages:
            ks = {"%s@%d" % k: [round(v[0] / v[1] * 1e3, 2), v[1], v[2] // v[1]]
                  for k, v in sorted(agg.items(), key=lambda kv: -kv[1][0])}
            print(json.dumps({"leg": precision, "stages_ms": st, "stats": stats,
                              "kernels_us_n_units": ks}), file=sys.stderr)
        fgt2d_amd.set_precision("fp64")
        return dict(t=t, ran=ran, stats=stats, err=err, agg=agg, st=st)

    npts = NS + NT
    legs = {}
    for p in (["fp64", "mixed"] if args.precision == "both" else [args.precision]):
        legs[p] = leg(p)
    head = legs["mixed"] if "mixed" in legs else legs["fp64"]
    mixed = head["ran"] == "mixed"
    stats, t = head["stats"], head["t"]
    rec = {
        "metric": "Mpts/s, 2D N=2M eps=1e-6 plane-wave FGT, mixed-precision MFMA route vs FP64 "
                  "(one MI355X)",
        "value": round(npts / t / 1e6, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mfma+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": "fgt2d mixture 2M sources + 2M targets, delta=1e-5, eps=1e-6, "
                               "pot at targets",
                   "ns": NS, "nt": NT, "delta": DELTA, "eps": EPS, "nlevels": stats.get("nlevels"),
                   "nboxes": stats.get("nboxes"), "npw": stats.get("npw"),
                   "route": head["ran"]},
        "rel_l2_err_vs_direct": head["err"],
        "direct_oracle_check": direct_check,
        "stages_ms": {k: round(v, 4) for k, v in head["st"].items()},
    }
    for p, L in legs.items():
        rec["leg_" + p] = {"Mpts_s": round(npts / L["t"] / 1e6, 3),
                           "ms_per_step": round(L["t"] * 1e3, 4), "route": L["ran"],
                           "rel_l2_err_vs_direct": L["err"]}
    if "fp64" in legs and "mixed" in legs:
        rec["mixed_vs_fp64"] = round(legs["fp64"]["t"] / legs["mixed"]["t"], 4)
    # the headline record: the hand-written kernel with the largest total
    # time (rocPRIM's sort, when chosen, is not one); the rest follow in
    # roofline_next
    rl = rooflines(head["agg"], stats, 1, count=9, mixed=mixed)
    own = [r for r in rl if not r["kernel"].startswith("rocprim")]
    rec["roofline"] = own[0] if own else None
    rec["roofline_next"] = [r for r in rl if r is not rec["roofline"]]
    if args.host_steps > 0:
        # the host-array entry (what the Fortran drop-in calls): transfers
        # in and out included, FP64
        for _ in range(2):
            fgt2d_amd.fgt2d(DELTA, EPS, src, chg, tg, ifpgh=0, ifpghtarg=1)
        t0 = time.perf_counter()
        for _ in range(args.host_steps):
            fgt2d_amd.fgt2d(DELTA, EPS, src, chg, tg, ifpgh=0, ifpghtarg=1)
        th = (time.perf_counter() - t0) / args.host_steps
        rec["host_arrays"] = {"Mpts_s": round(npts / th / 1e6, 3), "ms_per_step": round(th * 1e3, 4),
                              "steps": args.host_steps, "route": "fp64",
                              "what": "fgtpw_fgt2d on host arrays: H2D of sources, charges and "
                                      "targets, the transform, D2H of the target potentials"}
    if args.copy_calib:
        a = torch.empty(32 << 20, dtype=torch.float64, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
    if args.serve_requests > 0:
        rec["serving"] = serve_stream(fgt2d_amd, args.serve_requests, slots=args.slots)
    if args.serve_capacity > 0:
        # the same request mix submitted all at once: delivered Mpts/s over
        # the makespan is the server's capacity (latencies are queueing)
        cap = serve_stream(fgt2d_amd, args.serve_capacity, rate=1e9, slots=args.slots, check=0)
        rec["serving_capacity"] = {k: cap[k] for k in ("requests", "slots", "Mpts_s", "span_s",
                                                       "peak_running")}
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline()
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
