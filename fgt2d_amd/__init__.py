"""fgt2d_amd -- plane-wave fast Gauss transform on AMD MI355X (gfx950).

Python side of the drop-in for the reference's planewave driver
(2d/planewave/src/fgt2d.f:19-22).  The transform runs in the HIP library
``libfgtpw.so`` (C-ABI in include/fgtpw.h); this module only marshals arrays.
There is no CPU fallback: when the library is missing the calls raise.

Arrays use the reference's layouts: ``sources`` (ns, 2) C-contiguous float64
(= Fortran (2, ns)), ``charge`` (ns, nd) or (ns,), outputs (n, nd),
(n, 2, nd), (n, 3, nd).  numpy arrays go through the host entry point;
torch tensors on a GPU go through the device entry point on torch's current
stream, with no host copies.
"""
import ctypes as _C
import os as _os

import numpy as _np

__all__ = ["fgt2d", "fgt3d", "fgt2dpart_direct", "fgt3dpart_direct", "FGTError", "Server",
           "last_stats", "last_tree2d", "last_tree3d", "tree2d_from_leaves", "set_timing",
           "stage_times", "kernel_launches", "library_path", "load", "set_precision",
           "last_precision", "tuning_table"]

_HERE = _os.path.dirname(_os.path.abspath(__file__))
_LIB = None

FGTPW_OK, FGTPW_EINVAL, FGTPW_ENOMEM, FGTPW_EUNSUPPORTED, FGTPW_EDEVICE = 0, -1, -2, -3, -4


class FGTError(RuntimeError):
    """A non-zero status of the C-ABI; ``code`` is the FGTPW_* value."""

    def __init__(self, code, msg):
        super().__init__("fgtpw error %d: %s" % (code, msg))
        self.code = code


def library_path():
    return _os.path.join(_HERE, "libfgtpw.so")


def load():
    """Loads libfgtpw.so (built by ``python -m fgt2d_amd.build``)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not _os.path.exists(path):
        raise ImportError("fgt2d_amd: %s is not built; run python -m fgt2d_amd.build" % path)
    lib = _C.CDLL(path)
    dp, ip, vp, i, d = _C.c_void_p, _C.POINTER(_C.c_int), _C.c_void_p, _C.c_int, _C.c_double
    lib.fgtpw_last_error.restype = _C.c_char_p
    lib.fgtpw_init.argtypes = [i]
    lib.fgtpw_fgt2d.argtypes = [i, d, d, i, dp, i, dp, i, dp, dp, i, i, dp, dp, dp, i, dp, i, dp, dp, dp]
    lib.fgtpw_fgt2d_device.argtypes = lib.fgtpw_fgt2d.argtypes + [vp]
    lib.fgtpw_fgt2d_device_noalloc.argtypes = lib.fgtpw_fgt2d_device.argtypes
    lib.fgtpw_fgt2dpart_direct.argtypes = [i, d, d, i, i, i, i, dp, i, dp, i, dp, dp, dp, i, dp, dp, dp]
    lib.fgtpw_last_stats.argtypes = [ip, i]
    lib.fgtpw_set_timing.argtypes = [i]
    lib.fgtpw_set_precision.argtypes = [i]
    lib.fgtpw_last_precision.argtypes = []
    lib.fgtpw_tuning_table.restype = _C.c_char_p
    lib.fgtpw_tuning_table.argtypes = []
    lib.fgtpw_stage_times.argtypes = [_C.POINTER(_C.c_double), i]
    lib.fgtpw_stage_name.restype = _C.c_char_p
    lib.fgtpw_stage_name.argtypes = [i]
    lib.fgtpw_kernel_launches.argtypes = [ip, ip, _C.POINTER(_C.c_double),
                                          _C.POINTER(_C.c_longlong), i]
    lib.fgtpw_kernel_name.restype = _C.c_char_p
    lib.fgtpw_kernel_name.argtypes = [i]
    lib.fgtpw_last_tree2d.argtypes = [ip, i] + [dp] * 12
    lib.fgtpw_fgt3d.argtypes = lib.fgtpw_fgt2d.argtypes
    lib.fgtpw_fgt3d_device.argtypes = lib.fgtpw_fgt2d_device.argtypes
    lib.fgtpw_fgt3dpart_direct.argtypes = lib.fgtpw_fgt2dpart_direct.argtypes
    lib.fgtpw_last_tree3d.argtypes = [ip, i] + [dp] * 12
    lib.fgtpw_tree2d_from_leaves.argtypes = [i, i, d, dp, i, dp, dp, ip, i] + [dp] * 8
    lib.fgtpw_serve_start.argtypes = [i, d]
    lib.fgtpw_submit.argtypes = [_C.POINTER(_Call)]
    lib.fgtpw_submit.restype = _C.c_longlong
    lib.fgtpw_wait.argtypes = [_C.c_longlong, _C.POINTER(_C.c_double)]
    lib.fgtpw_serve_stats.argtypes = [_C.POINTER(_C.c_longlong), i]
    for f in ("fgtpw_serve_start", "fgtpw_wait", "fgtpw_serve_stats", "fgtpw_serve_stop",
              "fgtpw_init", "fgtpw_fgt2d", "fgtpw_fgt2d_device", "fgtpw_fgt2d_device_noalloc",
              "fgtpw_fgt2dpart_direct",
              "fgtpw_last_stats", "fgtpw_set_timing", "fgtpw_stage_times", "fgtpw_last_tree2d",
              "fgtpw_set_precision", "fgtpw_last_precision",
              "fgtpw_tree2d_from_leaves", "fgtpw_fgt3d", "fgtpw_fgt3d_device",
              "fgtpw_fgt3dpart_direct", "fgtpw_last_tree3d"):
        getattr(lib, f).restype = _C.c_int
    _LIB = lib
    return lib


class _Call(_C.Structure):
    """struct fgtpw_call of include/fgtpw.h."""
    _fields_ = [("dim", _C.c_int), ("nd", _C.c_int), ("delta", _C.c_double),
                ("eps", _C.c_double), ("ns", _C.c_int), ("sources", _C.c_void_p),
                ("ifcharge", _C.c_int), ("charge", _C.c_void_p), ("ifdipole", _C.c_int),
                ("rnormal", _C.c_void_p), ("dipstr", _C.c_void_p), ("ifpgh", _C.c_int),
                ("pot", _C.c_void_p), ("grad", _C.c_void_p), ("hess", _C.c_void_p),
                ("nt", _C.c_int), ("targ", _C.c_void_p), ("ifpghtarg", _C.c_int),
                ("pottarg", _C.c_void_p), ("gradtarg", _C.c_void_p), ("hesstarg", _C.c_void_p)]


def _check(rc):
    if rc != FGTPW_OK:
        raise FGTError(rc, load().fgtpw_last_error().decode())


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return _C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(_C.c_void_p)


def _is_torch_gpu(a):
    return hasattr(a, "is_cuda") and a.is_cuda


def _outputs(mod, n, nd, ifp, like, dim=2):
    # not initialised: the transform overwrites every element of a requested
    # output (no fill on the caller's stream ahead of it)
    shapes = [(n, nd), (n, dim, nd), (n, 3 if dim == 2 else 6, nd)]
    if mod is None:
        return [(_np.empty(s) if k < ifp else None) for k, s in enumerate(shapes)]
    return [(mod.empty(s, dtype=mod.float64, device=like.device) if k < ifp else None)
            for k, s in enumerate(shapes)]


_DUMMY = {}


def _dummy(mod, like):
    """A small array for the pointer arguments of quantities not requested
    (never read or written), one per device, kept."""
    key = "host" if mod is None else str(like.device)
    if key not in _DUMMY:
        _DUMMY[key] = (_np.zeros(8) if mod is None
                       else mod.zeros(8, dtype=mod.float64, device=like.device))
    return _DUMMY[key]


def fgt2d(delta, eps, sources, charge, targ=None, ifpgh=1, ifpghtarg=0, ifcharge=1,
          ifdipole=0, rnormal=None, dipstr=None, iper=0):
    """The transform of fgt2d (fgt2d.f:19): potentials (and gradients,
    Hessians for ifpgh/ifpghtarg = 2, 3) of sum_j q_j exp(-|x - y_j|^2/delta)
    at the sources and/or targets, to precision eps.

    Returns (pot, grad, hess, pottarg, gradtarg, hesstarg), None for what was
    not requested.  numpy in -> numpy out; torch GPU tensors in -> torch out.
    """
    return _transform(2, delta, eps, sources, charge, targ, ifpgh, ifpghtarg, ifcharge,
                      ifdipole, rnormal, dipstr, iper)


def fgt3d(delta, eps, sources, charge, targ=None, ifpgh=1, ifpghtarg=0, ifcharge=1,
          ifdipole=0, rnormal=None, dipstr=None, iper=0):
    """The transform of fgt3d (fgt3d.f:19) for points of shape (n, 3):
    gradients (n, 3, nd), Hessians (n, 6, nd) ordered xx, yy, zz, xy, xz, yz.
    Same conventions as fgt2d."""
    return _transform(3, delta, eps, sources, charge, targ, ifpgh, ifpghtarg, ifcharge,
                      ifdipole, rnormal, dipstr, iper)


def _transform(dim, delta, eps, sources, charge, targ, ifpgh, ifpghtarg, ifcharge, ifdipole,
               rnormal, dipstr, iper):
    lib = load()
    gpu = _is_torch_gpu(sources)
    if charge is None:
        ifcharge = 0
    if rnormal is None or dipstr is None:
        ifdipole = 0
    if gpu:
        import torch as mod

        def arr(a, n):
            return a.contiguous().to(mod.float64).reshape(n, -1)
        src = sources.contiguous().to(mod.float64)
        tg = targ.contiguous().to(mod.float64) if targ is not None else None
    else:
        mod = None

        def arr(a, n):
            return _np.ascontiguousarray(a, dtype=_np.float64).reshape(n, -1)
        src = _np.ascontiguousarray(sources, dtype=_np.float64)
        tg = _np.ascontiguousarray(targ, dtype=_np.float64) if targ is not None else None
    ns = src.shape[0]
    chg = arr(charge, ns) if ifcharge else None
    rn = arr(rnormal, ns) if ifdipole else None
    ds = arr(dipstr, ns) if ifdipole else None
    if src.ndim != 2 or src.shape[1] != dim:
        raise ValueError("sources must have shape (ns, %d)" % dim)
    if rn is not None and rn.shape[1] != dim:
        raise ValueError("rnormal must have shape (ns, %d)" % dim)
    if chg is None and ds is None:
        raise ValueError("no charges and no dipoles")
    nd = chg.shape[1] if chg is not None else ds.shape[1]
    if ds is not None and ds.shape[1] != nd:
        raise ValueError("dipstr must have shape (ns, nd)")
    nt = 0 if tg is None else tg.shape[0]
    if tg is not None and (tg.ndim != 2 or tg.shape[1] != dim):
        raise ValueError("targ must have shape (nt, %d)" % dim)
    outs = (_outputs(mod, ns, nd, ifpgh, src, dim) +
            _outputs(mod, max(nt, 1), nd, ifpghtarg, src, dim))
    dummy = _dummy(mod, src)
    P = [_ptr(o if o is not None else dummy) for o in outs]
    args = [nd, float(delta), float(eps), ns, _ptr(src), ifcharge,
            _ptr(chg if chg is not None else dummy), ifdipole,
            _ptr(rn if rn is not None else dummy), _ptr(ds if ds is not None else dummy),
            iper, ifpgh, P[0], P[1], P[2], nt, _ptr(tg if tg is not None else dummy), ifpghtarg,
            P[3], P[4], P[5]]
    if gpu:
        stream = mod.cuda.current_stream(src.device).cuda_stream
        f = lib.fgtpw_fgt2d_device if dim == 2 else lib.fgtpw_fgt3d_device
        _check(f(*args, _C.c_void_p(stream)))
    else:
        _check((lib.fgtpw_fgt2d if dim == 2 else lib.fgtpw_fgt3d)(*args))
    if nt == 0:
        outs[3:] = [None, None, None]
    return tuple(outs)


def fgt3dpart_direct(delta, dmax, sources, charge, targ, ifpgh=1, pot=None, grad=None, hess=None):
    """fgt3dpart_direct_vec (fgt3d.f:1533): adds the direct sum over pairs with
    r^2 < dmax over all sources to pot (grad, hess) at all targets, on the
    GPU; returns the arrays."""
    lib = load()
    src = _np.ascontiguousarray(sources, dtype=_np.float64)
    ns = src.shape[0]
    chg = _np.ascontiguousarray(charge, dtype=_np.float64).reshape(ns, -1)
    nd = chg.shape[1]
    tg = _np.ascontiguousarray(targ, dtype=_np.float64)
    nt = tg.shape[0]
    pot = _np.zeros((nt, nd)) if pot is None else pot
    grad = _np.zeros((nt, 3, nd)) if grad is None else grad
    hess = _np.zeros((nt, 6, nd)) if hess is None else hess
    dummy = _np.zeros(8)
    _check(lib.fgtpw_fgt3dpart_direct(nd, float(delta), float(dmax), 1, ns, 1, nt, _ptr(src), 1,
                                      _ptr(chg), 0, _ptr(dummy), _ptr(dummy), _ptr(tg), ifpgh,
                                      _ptr(pot), _ptr(grad), _ptr(hess)))
    return pot, grad, hess


def fgt2dpart_direct(delta, eps, sources, charge, targ, ifpgh=1, pot=None, grad=None, hess=None,
                     rnormal=None, dipstr=None):
    """fgt2dpart_direct_vec (fgt2d.f:1354): adds the direct sum with cutoff
    r^2 < ln(1/eps) delta over all sources to pot (grad, hess) at all
    targets, on the GPU; returns the arrays.  charge=None: no charges;
    rnormal (ns, 2) and dipstr (ns, nd): dipoles."""
    lib = load()
    src = _np.ascontiguousarray(sources, dtype=_np.float64)
    ns = src.shape[0]
    chg = None if charge is None else _np.ascontiguousarray(charge, dtype=_np.float64).reshape(ns, -1)
    rn = ds = None
    if rnormal is not None:
        rn = _np.ascontiguousarray(rnormal, dtype=_np.float64).reshape(ns, 2)
        ds = _np.ascontiguousarray(dipstr, dtype=_np.float64).reshape(ns, -1)
    tg = _np.ascontiguousarray(targ, dtype=_np.float64)
    nd = chg.shape[1] if chg is not None else ds.shape[1]
    nt = tg.shape[0]
    pot = _np.zeros((nt, nd)) if pot is None else pot
    grad = _np.zeros((nt, 2, nd)) if grad is None else grad
    hess = _np.zeros((nt, 3, nd)) if hess is None else hess
    dummy = _np.zeros(8)
    _check(lib.fgtpw_fgt2dpart_direct(
        nd, float(delta), float(eps), 1, ns, 1, nt, _ptr(src), int(chg is not None),
        _ptr(chg if chg is not None else dummy), int(rn is not None),
        _ptr(rn if rn is not None else dummy), _ptr(ds if ds is not None else dummy), _ptr(tg),
        ifpgh, _ptr(pot), _ptr(grad), _ptr(hess)))
    return pot, grad, hess


def last_stats():
    """npwlevel, nlmax, nlevels, nboxes, npw, leaves and nlocal per level of
    the last transform (the integers fgt2d.f prints)."""
    out = _np.zeros(80, dtype=_np.int32)
    k = load().fgtpw_last_stats(out.ctypes.data_as(_C.POINTER(_C.c_int)), 80)
    o = [int(v) for v in out[:k]]
    keys = ["npwlevel", "nlmax", "nlevels", "nboxes", "npw", "nleaves"]
    r = dict(zip(keys, o[:6]))
    r["nlocal"] = o[6:]
    return r


def set_timing(on=True):
    """Per-stage and per-kernel timing of the next transforms: True (events
    on the streams as they run), "serial" (the side streams' work queued on
    the call's stream, so that each kernel's time is its own) or False."""
    load().fgtpw_set_timing(2 if on == "serial" else (1 if on else 0))


def set_precision(mode):
    """Arithmetic of the next transforms: "fp64" (default) or "mixed" (FP32
    matrix-core moments and Taylor evaluation and FP32 plane-wave expansions
    for 2D transforms with eps >= 1e-6; see include/fgtpw.h).  Returns the
    previous mode."""
    modes = {"fp64": 0, "mixed": 1}
    if mode not in modes:
        raise ValueError("precision mode must be 'fp64' or 'mixed'")
    prev = load().fgtpw_set_precision(modes[mode])
    if prev < 0:
        _check(prev)
    return "mixed" if prev == 1 else "fp64"


def tuning_table():
    """The tile shapes in use (see include/fgtpw.h, fgtpw_tuning_table), as a
    dict with the key "source" naming where they came from."""
    text = load().fgtpw_tuning_table().decode()
    rest, src = text.split(";", 1)
    out = {"source": src.split("=", 1)[1]}
    for kv in rest.split():
        k, v = kv.split("=")
        out[k] = int(v)
    return out


def last_precision():
    """'mixed' when the last transform of this thread took the mixed route."""
    return "mixed" if load().fgtpw_last_precision() == 1 else "fp64"


def stage_times():
    """Milliseconds per stage of the last transform (set_timing(True) first)."""
    lib = load()
    buf = (_C.c_double * 32)()
    k = lib.fgtpw_stage_times(buf, 32)
    return {lib.fgtpw_stage_name(i).decode(): buf[i] for i in range(k)}


def kernel_launches():
    """Instrumented kernel launches of the last transform made with timing
    on: list of dicts (kernel, level, ms, units), in launch order."""
    lib = load()
    n = lib.fgtpw_kernel_launches(None, None, None, None, 0)
    ids, lev = _np.zeros(max(n, 1), _np.int32), _np.zeros(max(n, 1), _np.int32)
    ms, un = _np.zeros(max(n, 1)), _np.zeros(max(n, 1), _np.int64)
    lib.fgtpw_kernel_launches(ids.ctypes.data_as(_C.POINTER(_C.c_int)),
                              lev.ctypes.data_as(_C.POINTER(_C.c_int)),
                              ms.ctypes.data_as(_C.POINTER(_C.c_double)),
                              un.ctypes.data_as(_C.POINTER(_C.c_longlong)), n)
    return [dict(kernel=lib.fgtpw_kernel_name(int(ids[k])).decode(), level=int(lev[k]),
                 ms=float(ms[k]), units=int(un[k])) for k in range(n)]


def _tree_arrays(nb, nlevels):
    return dict(laddr=_np.zeros((nlevels + 1, 2), _np.int32), lev=_np.zeros(nb, _np.int32),
                par=_np.zeros(nb, _np.int32), nch=_np.zeros(nb, _np.int32),
                ch=_np.zeros((nb, 4), _np.int32), ncol=_np.zeros(nb, _np.int32),
                col=_np.zeros((nb, 9), _np.int32), cen=_np.zeros((nb, 2)))


def last_tree2d(ns, nt):
    """The quadtree of the last transform (0-based, itree layout)."""
    lib = load()
    info = _np.zeros(4, _np.int32)
    ip = info.ctypes.data_as(_C.POINTER(_C.c_int))
    _check(lib.fgtpw_last_tree2d(ip, 0, *([None] * 12)))
    nlevels, nb, nleaves, nlmax = (int(v) for v in info)
    t = _tree_arrays(nb, nlevels)
    t.update(isrc=_np.zeros(ns, _np.int32), isrcse=_np.zeros((nb, 2), _np.int32),
             itarg=_np.zeros(max(nt, 1), _np.int32), itargse=_np.zeros((nb, 2), _np.int32))
    _check(lib.fgtpw_last_tree2d(ip, nb, _ptr(t["laddr"]), _ptr(t["lev"]), _ptr(t["par"]),
                                 _ptr(t["nch"]), _ptr(t["ch"]), _ptr(t["ncol"]), _ptr(t["col"]),
                                 _ptr(t["cen"]), _ptr(t["isrc"]), _ptr(t["isrcse"]),
                                 _ptr(t["itarg"]), _ptr(t["itargse"])))
    t["itarg"] = t["itarg"][:nt]
    t.update(nlevels=nlevels, nboxes=nb, nleaves=nleaves, nlmax=nlmax)
    return t


def tree2d_from_leaves(nlmax, npwlevel, bs0, cen0, keys, levels):
    """Host-only tree construction from the non-empty leaves of the adaptive
    tree (Morton prefixes and levels, Morton order)."""
    lib = load()
    keys = _np.ascontiguousarray(keys, dtype=_np.uint64)
    levels = _np.ascontiguousarray(levels, dtype=_np.int32)
    c0 = _np.ascontiguousarray(cen0, dtype=_np.float64)
    info = _np.zeros(4, _np.int32)
    ip = info.ctypes.data_as(_C.POINTER(_C.c_int))
    _check(lib.fgtpw_tree2d_from_leaves(nlmax, npwlevel, bs0, _ptr(c0), len(keys), _ptr(keys),
                                        _ptr(levels), ip, 0, *([None] * 8)))
    nlevels, nb = int(info[0]), int(info[1])
    t = _tree_arrays(nb, nlevels)
    _check(lib.fgtpw_tree2d_from_leaves(nlmax, npwlevel, bs0, _ptr(c0), len(keys), _ptr(keys),
                                        _ptr(levels), ip, nb, _ptr(t["laddr"]), _ptr(t["lev"]),
                                        _ptr(t["par"]), _ptr(t["nch"]), _ptr(t["ch"]),
                                        _ptr(t["ncol"]), _ptr(t["col"]), _ptr(t["cen"])))
    t.update(nlevels=nlevels, nboxes=nb, nleaves=int(info[2]), nlmax=int(info[3]))
    return t


def last_tree3d(ns, nt):
    """The octree of the last 3D transform (0-based, itree layout)."""
    lib = load()
    info = _np.zeros(4, _np.int32)
    ip = info.ctypes.data_as(_C.POINTER(_C.c_int))
    _check(lib.fgtpw_last_tree3d(ip, 0, *([None] * 12)))
    nlevels, nb, nleaves, nlmax = (int(v) for v in info)
    t = dict(laddr=_np.zeros((nlevels + 1, 2), _np.int32), lev=_np.zeros(nb, _np.int32),
             par=_np.zeros(nb, _np.int32), nch=_np.zeros(nb, _np.int32),
             ch=_np.zeros((nb, 8), _np.int32), ncol=_np.zeros(nb, _np.int32),
             col=_np.zeros((nb, 27), _np.int32), cen=_np.zeros((nb, 3)),
             isrc=_np.zeros(ns, _np.int32), isrcse=_np.zeros((nb, 2), _np.int32),
             itarg=_np.zeros(max(nt, 1), _np.int32), itargse=_np.zeros((nb, 2), _np.int32))
    _check(lib.fgtpw_last_tree3d(ip, nb, _ptr(t["laddr"]), _ptr(t["lev"]), _ptr(t["par"]),
                                 _ptr(t["nch"]), _ptr(t["ch"]), _ptr(t["ncol"]), _ptr(t["col"]),
                                 _ptr(t["cen"]), _ptr(t["isrc"]), _ptr(t["isrcse"]),
                                 _ptr(t["itarg"]), _ptr(t["itargse"])))
    t["itarg"] = t["itarg"][:nt]
    t.update(nlevels=nlevels, nboxes=nb, nleaves=nleaves, nlmax=nlmax)
    return t


class Server:
    """Request server of the C-ABI (fgtpw_serve_start / fgtpw_submit /
    fgtpw_wait): transforms of any dimension, size, delta and eps are queued,
    admitted by estimated HBM footprint into `slots` concurrent slots, and
    run with their kernels sharing the device.  Host (numpy) arrays; the
    inputs are copied at submit, the outputs belong to the ticket.

        with Server(slots=4) as srv:
            t = srv.submit(2, delta, eps, src, chg, tg, ifpgh=1, ifpghtarg=1)
            (pot, grad, hess, pt, gt, ht), ms = srv.wait(t)
    """

    def __init__(self, slots=4, hbm_budget_gib=0.0):
        self._lib = load()
        _check(self._lib.fgtpw_serve_start(int(slots), float(hbm_budget_gib)))
        self._out = {}

    def submit(self, dim, delta, eps, sources, charge, targ=None, ifpgh=1, ifpghtarg=0,
               rnormal=None, dipstr=None):
        """Queues one transform; returns its ticket.  charge=None: no
        charges; rnormal (ns, dim) and dipstr (ns, nd): dipoles."""
        src = _np.ascontiguousarray(sources, dtype=_np.float64)
        if src.ndim != 2 or src.shape[1] != dim:
            raise ValueError("sources must have shape (ns, %d)" % dim)
        ns = src.shape[0]
        chg = None if charge is None else _np.ascontiguousarray(charge, dtype=_np.float64).reshape(ns, -1)
        rn = ds = None
        if rnormal is not None:
            rn = _np.ascontiguousarray(rnormal, dtype=_np.float64).reshape(ns, dim)
            ds = _np.ascontiguousarray(dipstr, dtype=_np.float64).reshape(ns, -1)
        nd = chg.shape[1] if chg is not None else ds.shape[1]
        tg = None if targ is None else _np.ascontiguousarray(targ, dtype=_np.float64)
        nt = 0 if tg is None else tg.shape[0]
        if tg is not None and (tg.ndim != 2 or tg.shape[1] != dim):
            raise ValueError("targ must have shape (nt, %d)" % dim)
        outs = (_outputs(None, ns, nd, ifpgh, None, dim) +
                _outputs(None, max(nt, 1), nd, ifpghtarg if nt else 0, None, dim))
        c = _Call(dim=dim, nd=nd, delta=float(delta), eps=float(eps), ns=ns,
                  sources=src.ctypes.data, ifcharge=int(chg is not None),
                  charge=None if chg is None else chg.ctypes.data, ifdipole=int(rn is not None),
                  rnormal=None if rn is None else rn.ctypes.data,
                  dipstr=None if ds is None else ds.ctypes.data,
                  ifpgh=ifpgh, nt=nt, targ=None if tg is None else tg.ctypes.data,
                  ifpghtarg=ifpghtarg if nt else 0)
        for name, o in zip(("pot", "grad", "hess", "pottarg", "gradtarg", "hesstarg"), outs):
            setattr(c, name, None if o is None else o.ctypes.data)
        ticket = self._lib.fgtpw_submit(_C.byref(c))
        if ticket <= 0:
            _check(int(ticket))
        self._out[ticket] = outs
        return ticket

    def wait(self, ticket):
        """Blocks until the ticket's transform is done; returns (outputs, ms)
        with outputs as fgt2d returns them and ms = (queue, run, total)."""
        ms = (_C.c_double * 3)()
        rc = self._lib.fgtpw_wait(_C.c_longlong(ticket), ms)
        outs = self._out.pop(ticket, None)
        _check(rc)
        return tuple(outs), (ms[0], ms[1], ms[2])

    def stats(self):
        """slots, budget bytes, queued, running, bytes in flight, peak
        running, peak bytes in flight."""
        buf = (_C.c_longlong * 7)()
        k = self._lib.fgtpw_serve_stats(buf, 7)
        keys = ["slots", "budget", "queued", "running", "in_flight", "peak_running",
                "peak_in_flight"]
        return dict(zip(keys, [int(v) for v in buf[:k]]))

    def close(self):
        if self._lib is not None:
            lib, self._lib = self._lib, None
            _check(lib.fgtpw_serve_stop())

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
