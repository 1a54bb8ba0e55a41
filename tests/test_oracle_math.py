"""Closed forms that pin the oracle's precomputation independently of the
reference: plane-wave quadrature, term counts, cut-off level."""
import math

import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("eps", [1e-3, 1e-6, 1e-9, 5e-11])
def test_planewave_quadrature_reproduces_gaussian(eps):
    """sum_k ws_k e^{i ts_k u} = e^{-u^2} for |u| <= rmax (Poisson summation
    with truncation at pmax, g2dterms.f:77-104).  The nodes are spaced
    pmax/(npw/2) >= the h of the bound (npw rounds down), so the aliased
    image creeps up near |u| = rmax: within 3 eps up to 0.9 rmax, within
    200 eps at the very edge (where the Gaussian itself is ~eps^4)."""
    delta = 1e-4
    dc = math.sqrt(delta * math.log(1 / eps))
    bsize = 2 * dc  # the plane-wave box of the cutoff level, fgt2d.f:351
    pmax, npw = O.pwterms(bsize, delta, eps)
    ws, ts = O.pwnodes(pmax, npw)
    rmax = bsize / math.sqrt(delta)
    u = np.linspace(-rmax, rmax, 2001)
    approx = (ws[None, :] * np.exp(1j * ts[None, :] * u[:, None])).sum(axis=1)
    assert np.abs(approx.imag).max() < 1e-14
    err = np.abs(approx.real - np.exp(-u ** 2))
    assert err[np.abs(u) <= 0.9 * rmax].max() < 3 * eps
    assert err.max() < 200 * eps


@pytest.mark.parametrize("eps,npw", [(1e-3, 12), (1e-6, 26), (1e-9, 38)])
def test_planewave_length_closed_form(eps, npw):
    """At the cutoff level rmax = 2 sqrt(ln 1/eps), so npw = 2 int(3 ln(1/eps)/pi)."""
    delta = 1e-5
    dc = math.sqrt(delta * math.log(1 / eps))
    _, n = O.pwterms(2 * dc, delta, eps)
    assert n == npw == 2 * int(3 * math.log(1 / eps) / math.pi)


def test_hermite_terms_bound_and_monotone():
    delta, eps = 1e-4, 1e-6
    prev = 0
    for bsize in np.geomspace(1e-3, 2e-2, 12):
        p = O.hlterms(2, bsize, delta, eps)
        r = bsize / math.sqrt(2 * delta)
        rp = r * math.sqrt(math.e / p)
        rk = float(np.float32(1.09)) * (2 * math.pi) ** -0.25
        assert 2 * rk * p ** -0.25 * rp ** p / (1 - rp) < eps           # the bound holds at p
        if p - 1 > int(math.e * r * r):
            rq = r * math.sqrt(math.e / (p - 1))
            assert 2 * rk * (p - 1) ** -0.25 * rq ** (p - 1) / (1 - rq) >= eps  # and not at p-1
        assert p >= prev
        prev = p


def test_cutoff_level_box_is_dc():
    """bs0 is enlarged to 2^(levcut+1) dc so that a cutoff-level box has side
    dc = sqrt(delta ln 1/eps) exactly (pts_tree2d4.f:1245-1254)."""
    rng = np.random.default_rng(0)
    src, tg = rng.random((100, 2)), rng.random((50, 2))
    for delta, eps in ((1e-5, 1e-6), (1e-3, 1e-9), (1e-4, 1e-3)):
        lev, bs0, cen = O.boxsize0(delta, eps, src, tg)
        dc = math.sqrt(delta * math.log(1 / eps))
        assert bs0 / 2 ** lev == dc
        ext = max(np.ptp(np.concatenate([src, tg])[:, 0]), np.ptp(np.concatenate([src, tg])[:, 1]))
        assert bs0 >= ext and bs0 / 2 < ext
