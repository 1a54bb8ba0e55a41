"""The no-allocation entry point fgtpw_fgt2d_device_noalloc against
fgtpw_fgt2d_device, through the C-ABI, on two calls in a row (the second
smaller, with other eps and gradients), bit for bit; free device memory
unchanged by a call; inputs too large for its scratch rejected before any
device work."""
import ctypes

import numpy as np
import pytest

import fgt2d_amd
from tests import _pts

pytestmark = pytest.mark.gpu


def _call(fn, delta, eps, src, chg, tg, ifpgh, ifpghtarg):
    import torch
    dev = torch.device("cuda:0")
    s = torch.from_numpy(src).to(dev)
    q = torch.from_numpy(chg).to(dev)
    t = torch.from_numpy(tg).to(dev)
    ns, nt, nd = len(src), len(tg), chg.shape[1]
    out = [torch.full((ns, nd), np.nan, dtype=torch.float64, device=dev),
           torch.full((ns, 2, nd), np.nan, dtype=torch.float64, device=dev),
           torch.full((ns, 3, nd), np.nan, dtype=torch.float64, device=dev),
           torch.full((nt, nd), np.nan, dtype=torch.float64, device=dev),
           torch.full((nt, 2, nd), np.nan, dtype=torch.float64, device=dev),
           torch.full((nt, 3, nd), np.nan, dtype=torch.float64, device=dev)]
    dummy = torch.zeros(8, dtype=torch.float64, device=dev)
    p = [ctypes.c_void_p(o.data_ptr()) for o in out]
    stream = torch.cuda.current_stream(dev)
    torch.cuda.synchronize()
    rc = fn(nd, delta, eps, ns, ctypes.c_void_p(s.data_ptr()), 1, ctypes.c_void_p(q.data_ptr()),
            0, ctypes.c_void_p(dummy.data_ptr()), ctypes.c_void_p(dummy.data_ptr()), 0, ifpgh,
            p[0], p[1], p[2], nt, ctypes.c_void_p(t.data_ptr()), ifpghtarg, p[3], p[4], p[5],
            ctypes.c_void_p(stream.cuda_stream))
    torch.cuda.synchronize()
    assert rc == 0, (rc, fgt2d_amd.load().fgtpw_last_error().decode())
    res = []
    for k, o in enumerate(out):
        want = (k < 3 and k < ifpgh) or (k >= 3 and k - 3 < ifpghtarg)
        res.append(o.cpu().numpy() if want else None)
    return res


CASES = [
    # (kind, ns, nt, delta, eps, nd, ifpgh, ifpghtarg)
    ("mixture", 200000, 60000, 1e-5, 1e-6, 1, 1, 1),
    ("uniform", 20000, 5000, 1e-4, 1e-9, 2, 2, 1),
]


def test_noalloc_matches_device_entry_point_twice():
    lib = fgt2d_amd.load()
    for case in CASES:
        kind, ns, nt, delta, eps, nd, ifs, ift = case
        print("no-alloc case", case, flush=True)
        rng = np.random.default_rng(ns)
        gen = _pts.mixture if kind == "mixture" else _pts.uniform
        src, tg = gen(rng, ns), gen(rng, nt)
        chg = rng.normal(size=(ns, nd))
        a = _call(lib.fgtpw_fgt2d_device_noalloc, delta, eps, src, chg, tg, ifs, ift)
        b = _call(lib.fgtpw_fgt2d_device, delta, eps, src, chg, tg, ifs, ift)
        for k in range(6):
            if b[k] is not None:
                assert np.all(np.isfinite(a[k])), k
                assert np.array_equal(a[k], b[k]), (k, float(np.abs(a[k] - b[k]).max()))


def test_noalloc_call_allocates_no_device_memory():
    import torch
    lib = fgt2d_amd.load()
    rng = np.random.default_rng(2)
    src, tg = _pts.mixture(rng, 50000), _pts.mixture(rng, 20000)
    chg = rng.normal(size=(50000, 1))
    _call(lib.fgtpw_fgt2d_device_noalloc, 1e-5, 1e-6, src, chg, tg, 1, 1)
    free0 = torch.cuda.mem_get_info()[0]
    _call(lib.fgtpw_fgt2d_device_noalloc, 1e-5, 1e-6, src, chg, tg, 1, 1)
    assert torch.cuda.mem_get_info()[0] == free0


def test_noalloc_rejects_inputs_beyond_its_scratch():
    import torch
    lib = fgt2d_amd.load()
    dev = torch.device("cuda:0")
    x = torch.zeros(64, dtype=torch.float64, device=dev)
    p = ctypes.c_void_p(x.data_ptr())
    # 20M sources: rejected by the host-side size check before any device
    # access (the arrays behind the pointers are never read)
    rc = lib.fgtpw_fgt2d_device_noalloc(1, 1e-5, 1e-6, 20_000_000, p, 1, p, 0, p, p, 0, 1, p, p,
                                        p, 0, p, 0, p, p, p, None)
    assert rc == fgt2d_amd.FGTPW_EINVAL
    assert "scratch" in lib.fgtpw_last_error().decode()
