"""Explicit per-layer media (the input of the Jacobian path) on the CPU oracle.

A generator instance (tau_R, tau_a, ssa_a, scale height on US62 altitudes,
rt_atm.c:272-400) and the same medium written out layer by layer must give the
same solve; absorption renormalises the scattering fractions as
rt_atm.c:521-538 does."""
import ctypes as C

import numpy as np
import pytest

from conftest import oracle_solve
from ocrt_amd.abi import make_instance
from ocrt_amd import synthetic as S


def medium(oracle, d):
    L = d.n_layers + 1
    h, ch, x, y, z = (np.zeros(L) for _ in range(5))
    rc = oracle.oracle_medium(C.byref(d), *(a.ctypes.data_as(C.c_void_p) for a in (h, ch, x, y, z)))
    assert rc == 0
    return h, ch, x, y


def layers_of(h, x, y, ssa):
    dt = np.diff(h)
    aer = x[1:] * dt / ssa
    ray = y[1:] * dt
    absorb = np.clip(dt - ray - aer, 0.0, None)
    return np.stack([ray, aer, np.full_like(dt, ssa), absorb])


@pytest.mark.parametrize("surface,wind", [("flat", 0.0), ("black_fresnel_ocean", 6.0), ("black", 0.0)])
def test_layers_reproduce_generator(oracle, surface, wind):
    gk = S.hg_greek(0.7, 30)
    d = make_instance(surface=surface, wind=wind, tau_a=0.2, ssa_a=0.93, greek=gk, n_mu=10, n_layers=30,
                      m_max=4, sza=35.0, vza=50.0, raa=70.0)
    h, ch, x, y = medium(oracle, d)
    lay = layers_of(h, x, y, 0.93)
    dl = make_instance(surface=surface, wind=wind, greek=gk, n_mu=10, m_max=4, sza=35.0, vza=50.0, raa=70.0,
                       layers=lay)
    hl, chl, xl, yl = medium(oracle, dl)
    assert np.max(np.abs(hl - h)) <= 1e-15 * h[-1]
    assert np.max(np.abs(xl - x)) <= 1e-13 and np.max(np.abs(yl - y)) <= 1e-13
    rc, r, modes, _ = oracle_solve(oracle, d)
    rcl, rl, model, _ = oracle_solve(oracle, dl)
    assert rc == 0 and rcl == 0
    for f in ("rho_I", "rho_Q", "rho_U"):
        assert abs(getattr(rl, f) - getattr(r, f)) <= 1e-12 * abs(r.rho_I), f
    assert rl.n_orders_used == r.n_orders_used
    assert abs(rl.tau_R_total - r.tau_R_total) <= 1e-14


def test_absorption_renormalises_fractions(oracle):
    """A pure absorber added to every layer: h grows by the absorption, the
    scattering optical depths stay, xdel and ydel scale by dtau_old / dtau_new."""
    gk = S.hg_greek(0.7, 20)
    nt = 12
    ray = np.linspace(0.001, 0.01, nt)
    aer = np.linspace(0.02, 0.002, nt)
    ssa = np.full(nt, 0.9)
    gas = np.full(nt, 0.004)
    d0 = make_instance(greek=gk, layers=np.stack([ray, aer, ssa, np.zeros(nt)]), n_mu=8, m_max=2)
    d1 = make_instance(greek=gk, layers=np.stack([ray, aer, ssa, gas]), n_mu=8, m_max=2)
    h0, _, x0, y0 = medium(oracle, d0)
    h1, _, x1, y1 = medium(oracle, d1)
    assert np.allclose(h1, h0 + np.concatenate([[0.0], np.cumsum(gas)]), rtol=0, atol=1e-15)
    r = np.diff(h0) / np.diff(h1)
    assert np.allclose(x1[1:], x0[1:] * r, rtol=1e-14, atol=0)
    assert np.allclose(y1[1:], y0[1:] * r, rtol=1e-14, atol=0)
    assert x1[0] == x0[0] == 0.0 and y1[0] == y0[0] == 1.0
    # absorption darkens the scene
    rc0, r0, _, _ = oracle_solve(oracle, d0)
    rc1, r1, _, _ = oracle_solve(oracle, d1)
    assert rc0 == 0 and rc1 == 0 and r1.rho_I < r0.rho_I


def test_layers_reject_ipss_and_bad_layers(oracle):
    gk = S.hg_greek(0.7, 10)
    lay = np.stack([np.full(5, 0.01), np.full(5, 0.02), np.full(5, 0.9), np.zeros(5)])
    d = make_instance(greek=gk, layers=lay, ipss=1, n_mu=6)
    rc, *_ = oracle_solve(oracle, d)
    assert rc == -1
    bad = lay.copy()
    bad[2, 3] = 1.5
    rc, *_ = oracle_solve(oracle, make_instance(greek=gk, layers=bad, n_mu=6))
    assert rc == -1
    empty = lay.copy()
    empty[:, 2] = 0.0
    rc, *_ = oracle_solve(oracle, make_instance(greek=gk, layers=empty, n_mu=6))
    assert rc == -1
