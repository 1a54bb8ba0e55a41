"""Test helpers: seeded inputs of the benchmark configurations and a numpy
statement of the GPU's key / leaf-list stage (for host-only tree tests)."""
import numpy as np


def uniform(rng, n, d=2):
    return rng.random((n, d))


def mixture(rng, n, ncent=10, sigma=0.01, lo=0.1, hi=0.9):
    """Gaussian mixture clipped to the unit square (the 2M/2M configuration)."""
    mu = rng.uniform(lo, hi, (ncent, 2))
    k = rng.integers(0, ncent, n)
    return np.clip(mu[k] + rng.normal(0.0, sigma, (n, 2)), 0.0, 1.0)


def keys2d(xy, cen0, bs0, nlmax):
    """Full-depth Morton keys from the reference's comparisons p - c < 0 with
    centres c +- bs/2 built level by level (pts_tree2d4.f:594-689)."""
    x, y = xy[:, 0], xy[:, 1]
    cx = np.full(len(x), float(cen0[0]))
    cy = np.full(len(x), float(cen0[1]))
    bs = float(bs0)
    k = np.zeros(len(x), dtype=np.uint64)
    for _ in range(nlmax):
        bs = bs / 2
        h = bs / 2
        xb = ~(x - cx < 0)
        yb = ~(y - cy < 0)
        k = (k << np.uint64(2)) | (xb.astype(np.uint64) + np.uint64(2) * yb.astype(np.uint64))
        cx = cx + np.where(xb, h, -h)
        cy = cy + np.where(yb, h, -h)
    return k


def leaf_runs(keys_sorted, nlmax, levcut, ndiv0=10, ndiv=80):
    """Non-empty leaves of the adaptive tree (before level restriction):
    a level-l box splits when it holds more than ndiv0 (l < levcut) or ndiv
    sources (pts_tree2d4.f:507-523), up to nlmax."""
    n = len(keys_sorted)
    L = np.full(n, -1)
    for l in range(nlmax + 1):
        pre = keys_sorted >> np.uint64(2 * (nlmax - l))
        _, start, counts = np.unique(pre, return_index=True, return_counts=True)
        cnt = np.repeat(counts, counts)
        lim = ndiv0 if l < levcut else ndiv
        here = (L < 0) & ((cnt <= lim) | (l == nlmax))
        L[here] = l
    keys, levels = [], []
    prev = None
    for i in range(n):
        pre = int(keys_sorted[i]) >> (2 * (nlmax - L[i]))
        tag = (L[i], pre)
        if tag != prev:
            keys.append(pre)
            levels.append(L[i])
            prev = tag
    return np.array(keys, dtype=np.uint64), np.array(levels, dtype=np.int32)


# ---------------------------------------------------------------------------
# inputs regenerable bit-for-bit on any host: DLARAN uniforms (the reference
# test's generator) combined with exact IEEE arithmetic only
# ---------------------------------------------------------------------------
def dlaran_points(kind, ns, nt, nd=1, dipoles=False):
    """Sources, charges (ns, nd) and targets drawn from one DLARAN stream.

    uniform  : points uniform in [0,1)^2, charges 2u-1.
    cluster  : 8 centres in [0.2,0.8]^2; offsets 0.01*(u1+u2+u3+u4-2) per axis
               (a bell of width ~0.01); charges 2u-1.
    reftest  : the reference test driver's own inputs (test_fgt2d.f:81-120).

    dipoles=True also returns normals (ns, 2) and dipole strengths (ns, nd):
    for reftest the driver's own draws, otherwise 2u-1 drawn after the
    targets from the same stream.
    """
    from oracle import oracle as O
    if kind == "reftest":
        if dipoles:
            src, chg, tg, rn, ds = O.reftest_inputs(ns, nt, dipoles=True)
            return src, chg, tg, rn, np.repeat(ds, nd, axis=1) if nd > 1 else ds
        src, chg, tg = O.reftest_inputs(ns, nt)
        return src, chg, tg
    r = O.Hkrand()

    def pts(n):
        if kind == "uniform":
            return r.draw(2 * n).reshape(n, 2)
        cen = 0.2 + 0.6 * r.draw(16).reshape(8, 2)
        k = np.minimum((r.draw(n) * 8).astype(np.int64), 7)
        off = r.draw(8 * n).reshape(n, 2, 4).sum(axis=2) - 2.0
        return cen[k] + 0.01 * off

    src = pts(ns)
    chg = 2.0 * r.draw(ns * nd).reshape(ns, nd) - 1.0
    tg = pts(nt)
    if dipoles:
        rn = 2.0 * r.draw(2 * ns).reshape(ns, 2) - 1.0
        ds = 2.0 * r.draw(ns * nd).reshape(ns, nd) - 1.0
        return src, chg, tg, rn, ds
    return src, chg, tg


def golden_inputs(g):
    """(sources, charge or None, targets, rnormal or None, dipstr or None)
    of a 2D golden fixture."""
    ns, nt, nd = int(g["ns"]), int(g["nt"]), int(g["nd"])
    ifc = int(g["ifcharge"]) if "ifcharge" in g else 1
    ifd = int(g["ifdipole"]) if "ifdipole" in g else 0
    if not ifd:
        src, chg, tg = dlaran_points(str(g["recipe"]), ns, nt, nd)
        return src, chg, tg, None, None
    src, chg, tg, rn, ds = dlaran_points(str(g["recipe"]), ns, nt, nd, dipoles=True)
    return src, chg if ifc else None, tg, rn, ds


def dlaran_points3d(kind, ns, nt, nd=1, dipoles=False):
    """3D sources, charges (ns, nd) and targets from one DLARAN stream.

    uniform : points uniform in [0,1)^3, charges 2u-1.
    shell   : sources on a shell of radius 0.4 about (0.5,0.5,0.5) with
              radial offsets 0.002*(u1+u2+u3+u4-2) (directions from three
              centred sums of four uniforms, normalised); targets uniform;
              charges 2u-1.
    cluster : sources uniform in the cube of side 0.03 centred at
              (0.5,0.5,0.5), targets uniform in the cube of side 0.06 about
              the same centre; charges 2u-1 (at delta 1e-5 the tree refines
              two levels below the cutoff level with few points).
    reftest : the reference 3D test driver's own inputs (test_fgt3d.f:48-98).

    dipoles=True also returns normals (ns, 3) and dipole strengths (ns, nd):
    for reftest the driver's own draws, otherwise 2u-1 drawn after the
    targets from the same stream.
    """
    from oracle import oracle as O
    if kind == "reftest":
        if dipoles:
            src, chg, tg, rn, ds = O.reftest3_inputs(ns, nt, dipoles=True)
            return src, chg, tg, rn, np.repeat(ds, nd, axis=1) if nd > 1 else ds
        return O.reftest3_inputs(ns, nt)
    r = O.Hkrand()
    if kind == "uniform":
        src = r.draw(3 * ns).reshape(ns, 3)
    elif kind == "cluster":
        src = 0.485 + 0.03 * r.draw(3 * ns).reshape(ns, 3)
    else:
        v = r.draw(12 * ns).reshape(ns, 3, 4).sum(axis=2) - 2.0
        v = v / np.sqrt((v * v).sum(axis=1))[:, None]
        rad = 0.4 + 0.002 * (r.draw(4 * ns).reshape(ns, 4).sum(axis=1) - 2.0)
        src = 0.5 + rad[:, None] * v
    chg = 2.0 * r.draw(ns * nd).reshape(ns, nd) - 1.0
    tg = r.draw(3 * nt).reshape(nt, 3)
    if kind == "cluster":
        tg = 0.47 + 0.06 * tg
    if dipoles:
        rn = 2.0 * r.draw(3 * ns).reshape(ns, 3) - 1.0
        ds = 2.0 * r.draw(ns * nd).reshape(ns, nd) - 1.0
        return src, chg, tg, rn, ds
    return src, chg, tg


def golden_inputs3(g):
    """(sources, charge or None, targets, rnormal or None, dipstr or None)
    of a 3D golden fixture."""
    ns, nt, nd = int(g["ns"]), int(g["nt"]), int(g["nd"])
    ifc = int(g["ifcharge"]) if "ifcharge" in g else 1
    ifd = int(g["ifdipole"]) if "ifdipole" in g else 0
    if not ifd:
        src, chg, tg = dlaran_points3d(str(g["recipe"]), ns, nt, nd)
        return src, chg, tg, None, None
    src, chg, tg, rn, ds = dlaran_points3d(str(g["recipe"]), ns, nt, nd, dipoles=True)
    return src, chg if ifc else None, tg, rn, ds
