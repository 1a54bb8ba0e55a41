"""ctypes front end of the test oracle.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(fgt2d_amd) never imports this module.

Two libraries are wrapped:

* ``liborc.so`` -- the C restatement in ``fgt2d_oracle.c`` (built by
  ``make -C oracle``), and ``liborc_omp.so``, the same source built with
  OpenMP (its per-box loops threaded, results bit-identical): the CPU
  baseline on all host cores (``oracle_fgt2d(..., threads=n)``);
* ``_ref/libref2d.so`` -- the reference Fortran itself, compiled from
  ``2d/planewave/src`` by ``make -C oracle ref`` where the reference sources
  exist.  On hosts without it, ``have_ref()`` is False and the callers skip.

Arrays follow the reference layout: points are (2, n) column-major, i.e. a
C-contiguous numpy array of shape (n, 2); densities are (nd, n) column-major,
i.e. C-contiguous (n, nd).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORC = None
_REF = None


_ORC_OMP = None


def _orc(threaded=False):
    global _ORC, _ORC_OMP
    if (_ORC_OMP if threaded else _ORC) is None:
        name = "liborc_omp.so" if threaded else "liborc.so"
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError("oracle/%s is not built (run make -C oracle)" % name)
        lib = C.CDLL(path)
        lib.orc_set_threads.restype = C.c_int
        lib.orc_set_threads.argtypes = [C.c_int]
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        lib.orc_fgt2d.restype = C.c_int
        lib.orc_fgt2d.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, dp, C.c_int, dp, C.c_int,
                                  dp, dp, C.c_int, dp, dp, dp, C.c_int, dp, C.c_int, dp, dp, dp]
        lib.orc_direct2d.restype = None
        lib.orc_direct2d.argtypes = [C.c_int, C.c_double, C.c_double, dp, C.c_int, dp, dp, C.c_int,
                                     C.c_int, dp, dp, dp]
        lib.orc_direct2d_cd.restype = None
        lib.orc_direct2d_cd.argtypes = [C.c_int, C.c_double, C.c_double, dp, C.c_int, dp, dp, dp,
                                        dp, C.c_int, C.c_int, dp, dp, dp]
        lib.orc_hlterms.restype = C.c_int
        lib.orc_hlterms.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        lib.orc_pwterms.restype = None
        lib.orc_pwterms.argtypes = [C.c_double, C.c_double, C.c_double, dp, ip]
        lib.orc_pwnodes.restype = None
        lib.orc_pwnodes.argtypes = [C.c_double, C.c_int, dp, dp]
        lib.orc_boxsize0.restype = None
        lib.orc_boxsize0.argtypes = [C.c_double, C.c_double, dp, C.c_int, dp, C.c_int, ip, dp, dp]
        lib.orc_tree2d_build.restype = C.c_void_p
        lib.orc_tree2d_build.argtypes = [C.c_double, C.c_double, C.c_int, dp, C.c_int, dp, ip]
        lib.orc_tree2d_get.restype = None
        lib.orc_tree2d_get.argtypes = [C.c_void_p] + [ip] * 7 + [dp, dp] + [ip] * 4 + [dp]
        lib.orc_tree2d_free.restype = None
        lib.orc_tree2d_free.argtypes = [C.c_void_p]
        lib.orc_last_stats.restype = None
        lib.orc_last_stats.argtypes = [ip]
        lib.orc_last_time.restype = C.c_double
        lib.orc_derr.restype = C.c_double
        lib.orc_derr.argtypes = [dp, dp, C.c_int]
        lib.orc_dlaran.restype = None
        lib.orc_dlaran.argtypes = [C.POINTER(C.c_uint64), C.c_int, dp]
        if threaded:
            _ORC_OMP = lib
        else:
            _ORC = lib
    return _ORC_OMP if threaded else _ORC


def threads_in_use(n=0):
    """Sets (n > 0) and returns the thread count of the OpenMP oracle."""
    return _orc(True).orc_set_threads(n)


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _pts(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert a.ndim == 2 and a.shape[1] == 2
    return a


def _outs(nd, n, ifp):
    return (np.zeros((n, nd)), np.zeros((n, 2, nd)), np.zeros((n, 3, nd))) if ifp > 0 else \
           (np.zeros(nd), np.zeros(2 * nd), np.zeros(3 * nd))


def _strengths(ns, charge, rnormal, dipstr):
    """(nd, charge (ns, nd) or None, rnormal (ns, dim) or None, dipstr
    (ns, nd) or None) from the arguments of a transform."""
    chg = None if charge is None else np.ascontiguousarray(charge, dtype=np.float64).reshape(ns, -1)
    rn = ds = None
    if rnormal is not None:
        rn = np.ascontiguousarray(rnormal, dtype=np.float64).reshape(ns, -1)
        ds = np.ascontiguousarray(dipstr, dtype=np.float64).reshape(ns, -1)
    nd = chg.shape[1] if chg is not None else ds.shape[1]
    return nd, chg, rn, ds


def _ptr_or(a, n):
    return _d(a) if a is not None else _d(np.zeros(max(n, 1)))


def oracle_fgt2d(delta, eps, sources, charge, targ=None, ifpgh=1, ifpghtarg=0, rnormal=None,
                 dipstr=None, threads=None):
    """The C restatement of fgt2d (fgt2d.f:19).  charge=None: no charges;
    rnormal/dipstr (ns, 2) and (ns, nd): dipoles.  Returns (pot, grad, hess,
    pottarg, gradtarg, hesstarg) shaped (n, nd), (n, 2, nd), (n, 3, nd) when
    requested.  threads: run the OpenMP build on that many threads (0: its
    default) instead of the sequential one; the results are the same."""
    lib = _orc(threads is not None)
    if threads is not None:
        lib.orc_set_threads(threads)
    src = _pts(sources)
    ns = src.shape[0]
    nd, chg, rn, ds = _strengths(ns, charge, rnormal, dipstr)
    tg = _pts(targ) if targ is not None else np.zeros((0, 2))
    nt = tg.shape[0]
    pot, grad, hess = _outs(nd, ns, ifpgh)
    pt, gt, ht = _outs(nd, max(nt, 1), ifpghtarg)
    rc = lib.orc_fgt2d(nd, delta, eps, ns, _d(src), int(chg is not None), _ptr_or(chg, nd * ns),
                       int(rn is not None), _ptr_or(rn, 2 * ns), _ptr_or(ds, nd * ns), ifpgh,
                       _d(pot), _d(grad), _d(hess), nt, _d(tg), ifpghtarg, _d(pt), _d(gt), _d(ht))
    if rc != 0:
        raise RuntimeError("orc_fgt2d returned %d" % rc)
    return pot, grad, hess, pt[:nt], gt[:nt], ht[:nt]


def oracle_stats():
    lib = _orc()
    out = np.zeros(80, dtype=np.int32)
    lib.orc_last_stats(_i(out))
    nlev = int(out[2])
    return dict(npwlevel=int(out[0]), nlmax=int(out[1]), nlevels=nlev, nboxes=int(out[3]),
                npw=int(out[4]), ntermmax=int(out[5]), nlocal=[int(v) for v in out[6:6 + nlev + 1]],
                time=float(lib.orc_last_time()))


def direct2d(delta, eps, sources, charge, targ, ifpgh=1):
    """FP64 direct sum with the reference's cutoff r^2 < ln(1/eps) delta
    (g2ddirect.f:32-76); eps=1e-100 is the reference test's no-cutoff oracle."""
    lib = _orc()
    src = _pts(sources)
    chg = np.ascontiguousarray(charge, dtype=np.float64).reshape(src.shape[0], -1)
    nd = chg.shape[1]
    tg = _pts(targ)
    nt = tg.shape[0]
    pot, grad, hess = np.zeros((nt, nd)), np.zeros((nt, 2, nd)), np.zeros((nt, 3, nd))
    lib.orc_direct2d(nd, delta, eps, _d(src), src.shape[0], _d(chg), _d(tg), nt, ifpgh,
                     _d(pot), _d(grad), _d(hess))
    return pot, grad, hess


def direct2d_cd(delta, eps, sources, charge, rnormal, dipstr, targ, ifpgh=1):
    """The reference's near-field sums with dipoles (g2d_directdp/dg/dh and
    g2d_directcdp/cdg/cdh, g2ddirect.f:217-655); charge=None: dipoles only."""
    lib = _orc()
    src = _pts(sources)
    ns = src.shape[0]
    nd, chg, rn, ds = _strengths(ns, charge, rnormal, dipstr)
    tg = _pts(targ)
    nt = tg.shape[0]
    pot, grad, hess = np.zeros((nt, nd)), np.zeros((nt, 2, nd)), np.zeros((nt, 3, nd))
    lib.orc_direct2d_cd(nd, delta, eps, _d(src), ns, _d(chg) if chg is not None else None,
                        _ptr_or(rn, 2 * ns), _ptr_or(ds, nd * ns), _d(tg), nt, ifpgh,
                        _d(pot), _d(grad), _d(hess))
    return pot, grad, hess


def derr(vref, vcomp):
    """Relative l2 error of the reference test (test_fgt2d.f:204-225)."""
    a = np.ascontiguousarray(vref, dtype=np.float64).ravel()
    b = np.ascontiguousarray(vcomp, dtype=np.float64).ravel()
    return float(_orc().orc_derr(_d(a), _d(b), a.size))


class Hkrand:
    """The reference test's random stream (hkrand -> DLARAN, seed 1)."""

    def __init__(self):
        self.state = C.c_uint64(1)

    def draw(self, n):
        out = np.zeros(n)
        _orc().orc_dlaran(C.byref(self.state), n, _d(out))
        return out


def reftest_inputs(nsrc, ntarg, dipoles=False):
    """Inputs of the reference's own test driver (2d/planewave/test/
    test_fgt2d.f:81-120) for nsrc sources and ntarg targets: per source seven
    draws (angle, x, y, two normal components, charge, dipole strength), the
    corner points (0,0) and (1,1) with zero charge and a unit charge at
    (0.05, 0.05), then two draws per target.  dipoles=True also returns the
    normals and dipole strengths of those draws."""
    r = Hkrand()
    d = r.draw(7 * nsrc).reshape(nsrc, 7)
    src = d[:, 1:3].copy()
    chg = d[:, 5:6].copy()
    chg[0:2] = 0.0
    chg[2] = 1.0
    src[0] = (0.0, 0.0)
    src[1] = (1.0, 1.0)
    src[2] = (0.05, 0.05)
    tg = r.draw(2 * ntarg).reshape(ntarg, 2)
    if dipoles:
        return src, chg, tg, d[:, 3:5].copy(), d[:, 6:7].copy()
    return src, chg, tg


def ref_hkrand(n):
    """n draws of the reference's own hkrand(0) from oracle/_ref (fresh
    process state only: hkrand keeps its seed between calls)."""
    lib = _ref()
    lib.hkrand_.restype = C.c_double
    return np.array([lib.hkrand_(_ri(0)) for _ in range(n)])


def hlterms(ndim, bsize, delta, eps):
    return int(_orc().orc_hlterms(ndim, bsize, delta, eps))


def pwterms(bsize, delta, eps):
    pm, n = C.c_double(), C.c_int()
    _orc().orc_pwterms(bsize, delta, eps, C.byref(pm), C.byref(n))
    return pm.value, n.value


def pwnodes(pmax, npw):
    ws, ts = np.zeros(npw), np.zeros(npw)
    _orc().orc_pwnodes(pmax, npw, _d(ws), _d(ts))
    return ws, ts


def boxsize0(delta, eps, sources, targ):
    src, tg = _pts(sources), _pts(targ)
    lc, bs0, cen = C.c_int(), C.c_double(), np.zeros(2)
    _orc().orc_boxsize0(delta, eps, _d(src), src.shape[0], _d(tg), tg.shape[0], C.byref(lc),
                        C.byref(bs0), _d(cen))
    return lc.value, bs0.value, cen


def oracle_tree2d(delta, eps, sources, targ):
    """The tree fgt2d builds (fgt2d.f:124-221), 0-based."""
    lib = _orc()
    src, tg = _pts(sources), _pts(targ)
    info = np.zeros(4, dtype=np.int32)
    h = lib.orc_tree2d_build(delta, eps, src.shape[0], _d(src), tg.shape[0], _d(tg), _i(info))
    npwlevel, nlmax, nlevels, nboxes = (int(v) for v in info)
    out = dict(npwlevel=npwlevel, nlmax=nlmax, nlevels=nlevels, nboxes=nboxes,
               laddr=np.zeros((nlevels + 1, 2), np.int32), lev=np.zeros(nboxes, np.int32),
               par=np.zeros(nboxes, np.int32), nch=np.zeros(nboxes, np.int32),
               ch=np.zeros((nboxes, 4), np.int32), ncol=np.zeros(nboxes, np.int32),
               col=np.zeros((nboxes, 9), np.int32), cen=np.zeros((nboxes, 2)),
               bsz=np.zeros(nlevels + 1), isrc=np.zeros(src.shape[0], np.int32),
               isrcse=np.zeros((nboxes, 2), np.int32), itarg=np.zeros(tg.shape[0], np.int32),
               itargse=np.zeros((nboxes, 2), np.int32), bs0cen0=np.zeros(3))
    o = out
    lib.orc_tree2d_get(h, _i(o["laddr"]), _i(o["lev"]), _i(o["par"]), _i(o["nch"]), _i(o["ch"]),
                       _i(o["ncol"]), _i(o["col"]), _d(o["cen"]), _d(o["bsz"]), _i(o["isrc"]),
                       _i(o["isrcse"]), _i(o["itarg"]), _i(o["itargse"]), _d(o["bs0cen0"]))
    lib.orc_tree2d_free(h)
    return out


# --------------------------------------------------------------------------
# the reference itself (oracle/_ref/libref2d.so)
# --------------------------------------------------------------------------

def ref_path():
    return os.path.join(_HERE, "_ref", "libref2d.so")


def have_ref():
    return os.path.exists(ref_path())


def _ref():
    global _REF
    if _REF is None:
        if not have_ref():
            raise RuntimeError("oracle/_ref/libref2d.so is not built")
        _REF = C.CDLL(ref_path())
    return _REF


def _ri(v):
    return C.byref(C.c_int(int(v)))


def _rd(v):
    return C.byref(C.c_double(float(v)))


class RefUndefined(RuntimeError):
    """The reference's behaviour is undefined for this input."""


def ref_defined(delta, eps, sources, targ):
    """False when the reference reads uninitialised memory: with
    npwlevel == nlevels g2dmpalloc returns before setting lmptot
    (fgt2d.f:1567) and fgt2d allocates and zeroes rmlexp(lmptot) with that
    garbage length (fgt2d.f:397-400)."""
    t = oracle_tree2d(delta, eps, sources, targ if targ is not None else np.zeros((0, 2)))
    return t["npwlevel"] != t["nlevels"]


def ref_fgt2d(delta, eps, sources, charge, targ=None, ifpgh=1, ifpghtarg=0, rnormal=None,
              dipstr=None):
    """Calls the reference's own fgt2d_ (fgt2d.f:19) through its Fortran ABI.
    Raises RefUndefined where the reference reads uninitialised memory."""
    lib = _ref()
    if not ref_defined(delta, eps, sources, targ):
        raise RefUndefined("npwlevel == nlevels: the reference's lmptot is uninitialised")
    src = _pts(sources)
    ns = src.shape[0]
    nd, chg, rn, ds = _strengths(ns, charge, rnormal, dipstr)
    ifc, ifd = int(chg is not None), int(rn is not None)
    if chg is None:
        chg = np.zeros((ns, nd))
    if rn is None:
        rn, ds = np.zeros((ns, 2)), np.zeros((ns, nd))
    tg = _pts(targ) if targ is not None else np.zeros((1, 2))
    nt = tg.shape[0] if targ is not None else 0
    pot, grad, hess = _outs(nd, ns, ifpgh)
    pt, gt, ht = _outs(nd, max(nt, 1), ifpghtarg)
    lc, bs0, cen0 = C.c_int(), C.c_double(), np.zeros(2)
    lib.pts_tree_boxsize0_(_rd(delta), _rd(eps), _d(src), _ri(ns), _d(tg), _ri(nt), C.byref(lc),
                           C.byref(bs0), _d(cen0))
    if lc.value < -2:
        # fgt2d_large_delta (fgt2d.f:1616-1735) allocates local(0:nlocal,
        # 0:nlocal,nd) and lets g2dformlc_vec add into it without zeroing
        # it first (fgt2d.f:1678-1690, g2drouts.f:621-682), so its result
        # depends on leftover heap contents.  Run its steps -- the same
        # reference routines, in the same order -- with local zeroed.
        nl = C.c_int()
        lib.g2dhlterms_(_ri(2), _rd(bs0.value), _rd(delta), _rd(eps), C.byref(nl))
        nloc = nl.value + max(ifpgh, ifpghtarg) - 1          # fgt2d.f:134-135
        local = np.zeros(nd * (nloc + 1) ** 2)
        chg_f = np.ascontiguousarray(chg)                    # (ns, nd) = charge(nd, ns)
        if ifc and not ifd:                                  # fgt2d.f:1688-1696
            lib.g2dformlc_vec_(_ri(nd), _rd(delta), _d(src), _ri(ns), _d(chg_f), _d(cen0),
                               _ri(nloc), _d(local))
        elif ifd and not ifc:
            lib.g2dformld_vec_(_ri(nd), _rd(delta), _d(src), _ri(ns), _d(rn), _d(ds), _d(cen0),
                               _ri(nloc), _d(local))
        elif ifd and ifc:
            lib.g2dformlcd_vec_(_ri(nd), _rd(delta), _d(src), _ri(ns), _d(chg_f), _d(rn), _d(ds),
                                _d(cen0), _ri(nloc), _d(local))
        legs = [(ifpghtarg, tg, nt, pt, gt, ht), (ifpgh, src, ns, pot, grad, hess)]
        for ifp, x, n, p, g, h in legs:                      # targets first, fgt2d.f:1697-1732
            if ifp == 1:
                lib.g2dlevalp_vec_(_ri(nd), _rd(delta), _d(cen0), _ri(nloc), _d(local), _d(x),
                                   _ri(n), _d(p))
            elif ifp == 2:
                lib.g2dlevalg_vec_(_ri(nd), _rd(delta), _d(cen0), _ri(nloc), _d(local), _d(x),
                                   _ri(n), _d(p), _d(g))
            elif ifp == 3:
                lib.g2dlevalh_vec_(_ri(nd), _rd(delta), _d(cen0), _ri(nloc), _d(local), _d(x),
                                   _ri(n), _d(p), _d(g), _d(h))
        return pot, grad, hess, pt[:nt], gt[:nt], ht[:nt]
    lib.fgt2d_(_ri(nd), _rd(delta), _rd(eps), _ri(ns), _d(src), _ri(ifc), _d(chg), _ri(ifd), _d(rn),
               _d(ds), _ri(0), _ri(ifpgh), _d(pot), _d(grad), _d(hess), _ri(nt), _d(tg),
               _ri(ifpghtarg), _d(pt), _d(gt), _d(ht))
    return pot, grad, hess, pt[:nt], gt[:nt], ht[:nt]


def ref_tree2d(delta, eps, sources, targ):
    """The reference's own tree routines, called as fgt2d calls them
    (fgt2d.f:124-221): pts_tree_boxsize0, nlmax loop, pts_tree_mem,
    pts_tree_build, pts_tree_sort.  Returned 0-based."""
    lib = _ref()
    src, tg = _pts(sources), _pts(targ)
    ns, nt = src.shape[0], tg.shape[0]
    lc, bs0, cen0 = C.c_int(), C.c_double(), np.zeros(2)
    lib.pts_tree_boxsize0_(_rd(delta), _rd(eps), _d(src), _ri(ns), _d(tg), _ri(nt), C.byref(lc),
                           C.byref(bs0), _d(cen0))
    npwlevel = lc.value
    nlevstart = max(npwlevel, 0)
    nlmax = npwlevel + 4
    bsize = bs0.value / (2 ** nlevstart)
    for i in range(nlevstart, 101):
        nl = C.c_int()
        lib.g2dhlterms_(_ri(2), _rd(bsize), _rd(delta), _rd(eps), C.byref(nl))
        if nl.value <= 18:
            nlmax = i
            break
        bsize = bsize / 2
    nlevels, nboxes, ltree = C.c_int(), C.c_int(), C.c_int()
    lib.pts_tree_mem_(_d(src), _ri(ns), _d(tg), _ri(nt), _ri(0), _ri(80), _ri(0), _ri(nlmax), _ri(0),
                      _ri(10), _ri(npwlevel), C.byref(bs0), _d(cen0), C.byref(nlevels), C.byref(nboxes),
                      C.byref(ltree))
    nl_, nb_, lt_ = nlevels.value, nboxes.value, ltree.value
    itree = np.zeros(lt_, np.int32)
    iptr = np.zeros(8, np.int32)
    cen = np.zeros((nb_, 2))
    bsz = np.zeros(nl_ + 1)
    lib.pts_tree_build_(_d(src), _ri(ns), _d(tg), _ri(nt), _ri(0), _ri(80), _ri(0), _ri(0), C.byref(nlevels),
                        C.byref(nboxes), _ri(10), _ri(npwlevel), C.byref(bs0), _d(cen0), C.byref(ltree),
                        _i(itree), _i(iptr), _d(cen), _d(bsz))
    isrc, isrcse = np.zeros(ns, np.int32), np.zeros((nb_, 2), np.int32)
    itarg, itargse = np.zeros(nt, np.int32), np.zeros((nb_, 2), np.int32)
    lib.pts_tree_sort_(_ri(ns), _d(src), _i(itree), C.byref(ltree), C.byref(nboxes), C.byref(nlevels),
                       _i(iptr), _d(cen), _i(isrc), _i(isrcse))
    lib.pts_tree_sort_(_ri(nt), _d(tg), _i(itree), C.byref(ltree), C.byref(nboxes), C.byref(nlevels),
                       _i(iptr), _d(cen), _i(itarg), _i(itargse))
    p = [int(v) - 1 for v in iptr]

    def z(v):  # 1-based box numbers to 0-based, "none" (<= 0) to -1
        v = np.asarray(v, dtype=np.int32)
        return np.where(v > 0, v - 1, -1).astype(np.int32)

    nb = nb_
    laddr = itree[p[0]:p[0] + 2 * (nl_ + 1)].reshape(nl_ + 1, 2) - 1
    se_s = isrcse.copy()
    se_s[:, 0] -= 1
    se_t = itargse.copy()
    se_t[:, 0] -= 1
    return dict(npwlevel=npwlevel, nlmax=nlmax, nlevels=nl_, nboxes=nb, laddr=laddr,
                lev=itree[p[1]:p[1] + nb].copy(), par=z(itree[p[2]:p[2] + nb]),
                nch=itree[p[3]:p[3] + nb].copy(), ch=z(itree[p[4]:p[4] + 4 * nb]).reshape(nb, 4),
                ncol=itree[p[5]:p[5] + nb].copy(), col=z(itree[p[6]:p[6] + 9 * nb]).reshape(nb, 9),
                cen=cen, bsz=bsz, isrc=isrc - 1, isrcse=se_s, itarg=itarg - 1, itargse=se_t,
                bs0cen0=np.array([bs0.value, cen0[0], cen0[1]]))


# --------------------------------------------------------------------------
# 3D: the C restatement of fgt3d (fgt3d_oracle.c) and the reference's own
# 3D library (oracle/_ref/libref3d.so)
# --------------------------------------------------------------------------

def reftest3_inputs(nsrc, ntarg, dipoles=False):
    """Inputs of the reference's own 3D test driver (3d/planewave/test/
    test_fgt3d.f:48-98): per source ten draws (theta, phi, x, y, z, three
    normal components, charge, dipole strength) with (x, y, z) kept, the
    corners (0,0,0) and (1,1,1) with zero charge and a unit charge at
    (0.05,0.05,0.05), then three draws per target.  dipoles=True also
    returns the normals and dipole strengths of those draws."""
    r = Hkrand()
    d = r.draw(10 * nsrc).reshape(nsrc, 10)
    src = d[:, 2:5].copy()
    chg = d[:, 8:9].copy()
    chg[0:2] = 0.0
    chg[2] = 1.0
    src[0] = (0.0, 0.0, 0.0)
    src[1] = (1.0, 1.0, 1.0)
    src[2] = (0.05, 0.05, 0.05)
    tg = r.draw(3 * ntarg).reshape(ntarg, 3)
    if dipoles:
        return src, chg, tg, d[:, 5:8].copy(), d[:, 9:10].copy()
    return src, chg, tg


def _pts3(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert a.ndim == 2 and a.shape[1] == 3
    return a


def _outs3(nd, n, ifp):
    return (np.zeros((n, nd)), np.zeros((n, 3, nd)), np.zeros((n, 6, nd))) if ifp > 0 else \
           (np.zeros(nd), np.zeros(3 * nd), np.zeros(6 * nd))


_ORC3 = False


def _orc3():
    global _ORC3
    lib = _orc()
    if not _ORC3:
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        lib.orc_fgt3d.restype = C.c_int
        lib.orc_fgt3d.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, dp, C.c_int, dp, C.c_int,
                                  dp, dp, C.c_int, dp, dp, dp, C.c_int, dp, C.c_int, dp, dp, dp]
        lib.orc_direct3d.restype = None
        lib.orc_direct3d.argtypes = [C.c_int, C.c_double, C.c_double, dp, C.c_int, dp, dp, C.c_int,
                                     C.c_int, dp, dp, dp]
        lib.orc_direct3d_cd.restype = None
        lib.orc_direct3d_cd.argtypes = [C.c_int, C.c_double, C.c_double, dp, C.c_int, dp, dp, dp,
                                        dp, C.c_int, C.c_int, dp, dp, dp]
        lib.orc3_boxsize0.restype = None
        lib.orc3_boxsize0.argtypes = [C.c_double, C.c_double, dp, C.c_int, dp, C.c_int, ip, dp, dp]
        lib.orc3_pwnodes.restype = None
        lib.orc3_pwnodes.argtypes = [C.c_double, C.c_int, dp, dp]
        lib.orc_tree3d_build.restype = C.c_void_p
        lib.orc_tree3d_build.argtypes = [C.c_double, C.c_double, C.c_int, dp, C.c_int, dp, ip]
        lib.orc_tree3d_get.restype = None
        lib.orc_tree3d_get.argtypes = [C.c_void_p] + [ip] * 7 + [dp, dp] + [ip] * 4 + [dp]
        lib.orc_tree3d_free.restype = None
        lib.orc_tree3d_free.argtypes = [C.c_void_p]
        lib.orc3_last_stats.restype = None
        lib.orc3_last_stats.argtypes = [ip]
        _ORC3 = True
    return lib


def oracle_fgt3d(delta, eps, sources, charge, targ=None, ifpgh=1, ifpghtarg=0, rnormal=None,
                 dipstr=None):
    """The C restatement of fgt3d (fgt3d.f:19).  charge=None: no charges;
    rnormal/dipstr (ns, 3) and (ns, nd): dipoles.  Returns (pot, grad, hess,
    pottarg, gradtarg, hesstarg) shaped (n, nd), (n, 3, nd), (n, 6, nd);
    Hessians ordered xx, yy, zz, xy, xz, yz."""
    lib = _orc3()
    src = _pts3(sources)
    ns = src.shape[0]
    nd, chg, rn, ds = _strengths(ns, charge, rnormal, dipstr)
    tg = _pts3(targ) if targ is not None else np.zeros((0, 3))
    nt = tg.shape[0]
    pot, grad, hess = _outs3(nd, ns, ifpgh)
    pt, gt, ht = _outs3(nd, max(nt, 1), ifpghtarg)
    rc = lib.orc_fgt3d(nd, delta, eps, ns, _d(src), int(chg is not None), _ptr_or(chg, nd * ns),
                       int(rn is not None), _ptr_or(rn, 3 * ns), _ptr_or(ds, nd * ns), ifpgh,
                       _d(pot), _d(grad), _d(hess), nt, _d(tg), ifpghtarg, _d(pt), _d(gt), _d(ht))
    if rc != 0:
        raise RuntimeError("orc_fgt3d returned %d" % rc)
    return pot, grad, hess, pt[:nt], gt[:nt], ht[:nt]


def oracle_stats3d():
    out = np.zeros(8, dtype=np.int32)
    _orc3().orc3_last_stats(_i(out))
    return dict(npwlevel=int(out[0]), nlmax=int(out[1]), nlevels=int(out[2]), nboxes=int(out[3]),
                npw=int(out[4]), nleaves=int(out[5]))


def direct3d(delta, dmax, sources, charge, targ, ifpgh=1):
    """FP64 3D direct sum over pairs with r^2 < dmax (g3ddirect.f:3-206,
    fgt3dpart_direct_vec takes dmax itself); dmax=1e30 is the reference
    3D test's no-cutoff oracle."""
    lib = _orc3()
    src = _pts3(sources)
    chg = np.ascontiguousarray(charge, dtype=np.float64).reshape(src.shape[0], -1)
    nd = chg.shape[1]
    tg = _pts3(targ)
    nt = tg.shape[0]
    pot, grad, hess = np.zeros((nt, nd)), np.zeros((nt, 3, nd)), np.zeros((nt, 6, nd))
    lib.orc_direct3d(nd, delta, dmax, _d(src), src.shape[0], _d(chg), _d(tg), nt, ifpgh,
                     _d(pot), _d(grad), _d(hess))
    return pot, grad, hess


def direct3d_cd(delta, dmax, sources, charge, rnormal, dipstr, targ, ifpgh=1):
    """The 3D reference's near-field sums with dipoles (g3d_directdp/dg/dh
    and g3d_directcdp/cdg/cdh, g3ddirect.f:207-700); charge=None: dipoles
    only."""
    lib = _orc3()
    src = _pts3(sources)
    ns = src.shape[0]
    nd, chg, rn, ds = _strengths(ns, charge, rnormal, dipstr)
    tg = _pts3(targ)
    nt = tg.shape[0]
    pot, grad, hess = np.zeros((nt, nd)), np.zeros((nt, 3, nd)), np.zeros((nt, 6, nd))
    lib.orc_direct3d_cd(nd, delta, dmax, _d(src), ns, _d(chg) if chg is not None else None,
                        _ptr_or(rn, 3 * ns), _ptr_or(ds, nd * ns), _d(tg), nt, ifpgh,
                        _d(pot), _d(grad), _d(hess))
    return pot, grad, hess


def boxsize0_3d(delta, eps, sources, targ):
    src, tg = _pts3(sources), _pts3(targ)
    lc, bs0, cen = C.c_int(), C.c_double(), np.zeros(3)
    _orc3().orc3_boxsize0(delta, eps, _d(src), src.shape[0], _d(tg), tg.shape[0], C.byref(lc),
                          C.byref(bs0), _d(cen))
    return lc.value, bs0.value, cen


def pwnodes3d(pmax, npw):
    ws, ts = np.zeros(npw), np.zeros(npw)
    _orc3().orc3_pwnodes(pmax, npw, _d(ws), _d(ts))
    return ws, ts


def oracle_tree3d(delta, eps, sources, targ):
    """The octree fgt3d builds (fgt3d.f:124-221), 0-based."""
    lib = _orc3()
    src, tg = _pts3(sources), _pts3(targ)
    info = np.zeros(4, dtype=np.int32)
    h = lib.orc_tree3d_build(delta, eps, src.shape[0], _d(src), tg.shape[0], _d(tg), _i(info))
    npwlevel, nlmax, nlevels, nboxes = (int(v) for v in info)
    o = dict(npwlevel=npwlevel, nlmax=nlmax, nlevels=nlevels, nboxes=nboxes,
             laddr=np.zeros((nlevels + 1, 2), np.int32), lev=np.zeros(nboxes, np.int32),
             par=np.zeros(nboxes, np.int32), nch=np.zeros(nboxes, np.int32),
             ch=np.zeros((nboxes, 8), np.int32), ncol=np.zeros(nboxes, np.int32),
             col=np.zeros((nboxes, 27), np.int32), cen=np.zeros((nboxes, 3)),
             bsz=np.zeros(nlevels + 1), isrc=np.zeros(src.shape[0], np.int32),
             isrcse=np.zeros((nboxes, 2), np.int32), itarg=np.zeros(tg.shape[0], np.int32),
             itargse=np.zeros((nboxes, 2), np.int32), bs0cen0=np.zeros(4))
    lib.orc_tree3d_get(h, _i(o["laddr"]), _i(o["lev"]), _i(o["par"]), _i(o["nch"]), _i(o["ch"]),
                       _i(o["ncol"]), _i(o["col"]), _d(o["cen"]), _d(o["bsz"]), _i(o["isrc"]),
                       _i(o["isrcse"]), _i(o["itarg"]), _i(o["itargse"]), _d(o["bs0cen0"]))
    lib.orc_tree3d_free(h)
    return o


_REF3 = None


def ref3_path():
    return os.path.join(_HERE, "_ref", "libref3d.so")


def have_ref3():
    return os.path.exists(ref3_path())


def _ref3():
    global _REF3
    if _REF3 is None:
        if not have_ref3():
            raise RuntimeError("oracle/_ref/libref3d.so is not built")
        _REF3 = C.CDLL(ref3_path())
    return _REF3


def ref3_defined(delta, eps, sources, targ):
    """False where the 3D reference reads uninitialised memory: npwlevel ==
    nlevels (g3dmpalloc returns before setting lmptot, fgt3d.f:1752, and
    fgt3d zeroes rmlexp(lmptot), :402-405), and the large-delta branch
    (g3dformlc_vec leaves z undefined, g3drouts.f:803-804)."""
    tg = targ if targ is not None else np.zeros((0, 3))
    lc = boxsize0_3d(delta, eps, sources, tg)[0]
    if lc < -2:
        return False
    t = oracle_tree3d(delta, eps, sources, tg)
    return t["npwlevel"] != t["nlevels"]


def ref_fgt3d(delta, eps, sources, charge, targ=None, ifpgh=1, ifpghtarg=0, rnormal=None,
              dipstr=None):
    """Calls the reference's own fgt3d_ (fgt3d.f:19) through its Fortran ABI."""
    lib = _ref3()
    if not ref3_defined(delta, eps, sources, targ):
        raise RefUndefined("the 3D reference reads uninitialised memory for this input")
    src = _pts3(sources)
    ns = src.shape[0]
    nd, chg, rn, ds = _strengths(ns, charge, rnormal, dipstr)
    ifc, ifd = int(chg is not None), int(rn is not None)
    if chg is None:
        chg = np.zeros((ns, nd))
    if rn is None:
        rn, ds = np.zeros((ns, 3)), np.zeros((ns, nd))
    tg = _pts3(targ) if targ is not None else np.zeros((1, 3))
    nt = tg.shape[0] if targ is not None else 0
    pot, grad, hess = _outs3(nd, ns, ifpgh)
    pt, gt, ht = _outs3(nd, max(nt, 1), ifpghtarg)
    lib.fgt3d_(_ri(nd), _rd(delta), _rd(eps), _ri(ns), _d(src), _ri(ifc), _d(chg), _ri(ifd), _d(rn),
               _d(ds), _ri(0), _ri(ifpgh), _d(pot), _d(grad), _d(hess), _ri(nt), _d(tg),
               _ri(ifpghtarg), _d(pt), _d(gt), _d(ht))
    return pot, grad, hess, pt[:nt], gt[:nt], ht[:nt]


def ref_tree3d(delta, eps, sources, targ):
    """The 3D reference's own tree routines, called as fgt3d calls them
    (fgt3d.f:124-221; ltree and iptr are integer*8 there).  0-based."""
    lib = _ref3()
    src, tg = _pts3(sources), _pts3(targ)
    ns, nt = src.shape[0], tg.shape[0]
    lc, bs0, cen0 = C.c_int(), C.c_double(), np.zeros(3)
    lib.pts_tree_boxsize0_(_rd(delta), _rd(eps), _d(src), _ri(ns), _d(tg), _ri(nt), C.byref(lc),
                           C.byref(bs0), _d(cen0))
    npwlevel = lc.value
    nlevstart = max(npwlevel, 0)
    nlmax = npwlevel + 4
    bsize = bs0.value / (2 ** nlevstart)
    for i in range(nlevstart, 101):
        nl = C.c_int()
        lib.g3dhlterms_(_ri(3), _rd(bsize), _rd(delta), _rd(eps), C.byref(nl))
        if nl.value <= 15:
            nlmax = i
            break
        bsize = bsize / 2
    nlevels, nboxes, ltree = C.c_int(), C.c_int(), C.c_int64()
    lib.pts_tree_mem_(_d(src), _ri(ns), _d(tg), _ri(nt), _ri(0), _ri(100), _ri(0), _ri(nlmax), _ri(0),
                      _ri(0), _ri(40), _ri(npwlevel), C.byref(bs0), _d(cen0), C.byref(nlevels),
                      C.byref(nboxes), C.byref(ltree))
    nl_, nb_, lt_ = nlevels.value, nboxes.value, ltree.value
    itree = np.zeros(lt_, np.int32)
    iptr = np.zeros(8, np.int64)
    cen = np.zeros((nb_, 3))
    bsz = np.zeros(nl_ + 1)
    lib.pts_tree_build_(_d(src), _ri(ns), _d(tg), _ri(nt), _ri(0), _ri(100), _ri(0), _ri(nlmax), _ri(0),
                        _ri(0), C.byref(nlevels), C.byref(nboxes), _ri(40), _ri(npwlevel),
                        C.byref(bs0), _d(cen0), C.byref(ltree), _i(itree),
                        iptr.ctypes.data_as(C.POINTER(C.c_int64)), _d(cen), _d(bsz))
    isrc, isrcse = np.zeros(ns, np.int32), np.zeros((nb_, 2), np.int32)
    itarg, itargse = np.zeros(nt, np.int32), np.zeros((nb_, 2), np.int32)
    ip8 = iptr.ctypes.data_as(C.POINTER(C.c_int64))
    lib.pts_tree_sort_(_ri(ns), _d(src), _i(itree), C.byref(ltree), C.byref(nboxes), C.byref(nlevels),
                       ip8, _d(cen), _i(isrc), _i(isrcse))
    lib.pts_tree_sort_(_ri(nt), _d(tg), _i(itree), C.byref(ltree), C.byref(nboxes), C.byref(nlevels),
                       ip8, _d(cen), _i(itarg), _i(itargse))
    p = [int(v) - 1 for v in iptr]

    def z(v):
        v = np.asarray(v, dtype=np.int32)
        return np.where(v > 0, v - 1, -1).astype(np.int32)

    nb = nb_
    laddr = itree[p[0]:p[0] + 2 * (nl_ + 1)].reshape(nl_ + 1, 2) - 1
    se_s = isrcse.copy()
    se_s[:, 0] -= 1
    se_t = itargse.copy()
    se_t[:, 0] -= 1
    return dict(npwlevel=npwlevel, nlmax=nlmax, nlevels=nl_, nboxes=nb, laddr=laddr,
                lev=itree[p[1]:p[1] + nb].copy(), par=z(itree[p[2]:p[2] + nb]),
                nch=itree[p[3]:p[3] + nb].copy(), ch=z(itree[p[4]:p[4] + 8 * nb]).reshape(nb, 8),
                ncol=itree[p[5]:p[5] + nb].copy(), col=z(itree[p[6]:p[6] + 27 * nb]).reshape(nb, 27),
                cen=cen, bsz=bsz, isrc=isrc - 1, isrcse=se_s, itarg=itarg - 1, itargse=se_t,
                bs0cen0=np.array([bs0.value, cen0[0], cen0[1], cen0[2]]))
